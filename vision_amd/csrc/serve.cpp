// Batched serving: the host scheduler behind vs_server_* (include/vision_amd.h).
//
// A worker thread owns three HIP streams.  Each loop it
//   1. retires finished result copies (the job is done, its slot and HBM
//      range return to the pool),
//   2. admits queued jobs first-come first-served while a slot and a range
//      of the HBM arena are free (film cleared, job record uploaded),
//   3. plans a step when fewer than two are in flight: every admitted job
//      with frames left gets frames in admission order until the step's
//      sample budget (step_ms at the measured rate) is used, each at least
//      one; the plan is written into one of two pinned step tables and the
//      step's graph (table upload, counter reset, one persistent kernel per
//      kernel variant on forked streams) is replayed,
//   4. otherwise waits for the oldest step, credits its frames, and starts
//      the output pass and device-to-host copy of each job it finished on
//      the copy stream.
// Two steps in flight keep the device busy while the host plans; the copy
// stream overlaps result downloads with the next step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "context.h"
#include "serve.h"

using namespace vs_internal;
using Clock = std::chrono::steady_clock;

namespace {

thread_local std::string g_server_error;

constexpr uint32_t kMaxScenes = 64;
constexpr uint32_t kStreams = 3;
constexpr size_t kArenaAlign = 256;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// first-fit allocator over one device range; free ranges coalesce
class Arena {
public:
    void reset(size_t size) {
        free_.clear();
        if (size) free_[0] = size;
        used_ = peak_ = 0;
    }
    bool alloc(size_t n, size_t &off) {
        n = (n + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            off = it->first;
            size_t rest = it->second - n;
            free_.erase(it);
            if (rest) free_[off + n] = rest;
            used_ += n;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n) {
        n = (n + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t peak() const { return peak_; }
    void reset_peak() { peak_ = used_; }

private:
    std::map<size_t, size_t> free_;
    size_t used_{0}, peak_{0};
};

struct JobRec {
    uint64_t id{0};
    vs_job_desc d{};
    int32_t state{VS_JOB_QUEUED};
    Clock::time_point t_submit, t_admit, t_done;
    uint32_t slot{0};
    size_t off{0}, bytes{0};
    uint32_t variant{0};
    uint32_t frames_sched{0}, frames_done{0};
    vs::RenderView rv{};
    float *h_rt{nullptr}, *h_out{nullptr};// pinned host copies of the films
    hipEvent_t ev_copy{nullptr};
    uint32_t route{0};           // 0 persistent megakernel, 1 wavefront lane
    std::string error;
};

struct StepRec {
    hipEvent_t ev_start{nullptr}, ev_end{nullptr};
    uint32_t buf{0};
    uint64_t samples{0};
    std::vector<std::pair<JobRec *, uint32_t>> work;
};

struct SceneSlot {
    vs_scene *scene{nullptr};
    vs::CameraView cam{};   // the scene's camera when it joined (jobs re-aspect it)
    vs::RenderView rv{};    // and its integrator / film settings
    vs::SceneView view[2]{};// sRGB and hero-wavelength views
    uint32_t route[2]{0, 0};// per mode: 0 persistent megakernel, 1 wavefront lane
    bool diffuse_only{false};
};

}// namespace

struct vs_server {
    vs_context *ctx{nullptr};
    vs_server_config cfg{};
    uint32_t cu_count{0};
    hipStream_t st[kStreams]{};
    hipStream_t st_copy{nullptr};
    hipEvent_t ev_fork{nullptr}, ev_join[kStreams]{};
    // resident scenes: views[2 * i + (hero ? 1 : 0)]
    std::vector<SceneSlot> scenes;
    DevBuf d_views;
    uint32_t max_stack{1};
    // job films
    DevBuf arena_mem;
    Arena arena;
    // slots
    DevBuf d_jobs;
    vs::ServeJob *h_jobs{nullptr};// pinned
    std::vector<JobRec *> slot_job;
    // step tables: two pinned host copies (steps in flight) and their device copies
    DevBuf d_tables[2], d_next, d_counters;
    vs::StepTable *h_tables[2]{nullptr, nullptr};
    std::map<uint32_t, hipGraphExec_t> graphs;// key: variant mask | buffer << 8
    // scheduling state (worker thread only)
    std::vector<JobRec *> active;// admission order
    std::deque<StepRec> inflight;
    std::vector<hipEvent_t> free_events;
    std::vector<JobRec *> copying;
    // wavefront lanes: each its own thread, stream and path buffers; jobs
    // routed to them render one at a time per lane in the wavefront form
    // beside the persistent steps; several lanes fill each other's
    // per-bounce host waits and short last bounces (measured on the serving
    // mix, jobs of 1M samples and more on the lanes: one lane 583, two 635,
    // three 651 Msamples/s)
    static constexpr int kLanes = 3;
    std::thread lane[kLanes];
    hipStream_t st_lane[kLanes]{};
    WavefrontBuffers lane_wb[kLanes];
    hipEvent_t lane_ev[kLanes][4]{};// stage timing of each lane's bounces
    DevBuf lane_ctr[kLanes];         // counting runs: 8 counters per lane
    // lane totals (under mu)
    uint64_t lane_samples{0}, lane_stage_launches[3]{}, lane_counts[6]{};
    double lane_stage_ms[3]{};
    std::deque<JobRec *> lane_queue;   // under mu
    std::vector<JobRec *> lane_done;   // under mu
    std::condition_variable cv_lane;
    double rate{0.0};// samples per ms, measured
    uint64_t step_count{0};
    // shared with callers (under mu)
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<JobRec *> queue;
    std::unordered_map<uint64_t, std::unique_ptr<JobRec>> jobs;
    std::multimap<size_t, float *> pinned_pool;// released result buffers by size
    uint64_t next_id{1};
    uint64_t submit_seq{0};// bumped by each submit (under mu): wakes a waiting worker
    bool stop{false};
    std::string fatal;
    // statistics (under mu)
    uint64_t n_submitted{0}, n_done{0}, n_failed{0}, samples_done{0}, graph_launches{0}, steps{0};
    double busy_ms{0.0};
    bool have_first{false};
    Clock::time_point t_first, t_last;
    std::vector<double> latencies;
    uint32_t max_active{0};
    std::thread worker;
};

namespace {

vs_status server_fail(vs_status code, const std::string &msg) {
    g_server_error = msg;
    return code;
}

template <typename F>
vs_status server_guarded(F &&f) {
    try {
        f();
        g_server_error.clear();
        return VS_OK;
    } catch (const DeviceError &e) {
        return server_fail(VS_ERROR_DEVICE, e.what());
    } catch (const std::bad_alloc &) {
        return server_fail(VS_ERROR_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::invalid_argument &e) {
        return server_fail(VS_ERROR_INVALID_ARGUMENT, e.what());
    } catch (const std::logic_error &e) {
        return server_fail(VS_ERROR_STATE, e.what());
    } catch (const std::exception &e) {
        return server_fail(VS_ERROR_INVALID_ARGUMENT, e.what());
    }
}

size_t job_bytes(const vs_job_desc &d) { return size_t(d.width) * d.height * 32; }// rt + output buffer

uint32_t tile_h(const vs_server *S) { return 64u / S->cfg.tile_w; }

hipEvent_t take_event(vs_server *S) {
    if (!S->free_events.empty()) {
        hipEvent_t e = S->free_events.back();
        S->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    hip_check(hipEventCreate(&e), "event create");
    return e;
}

void drop_graphs(vs_server *S) {
    for (auto &g : S->graphs) (void)hipGraphExecDestroy(g.second);
    S->graphs.clear();
}

uint32_t variant_blocks(vs_server *S, uint32_t v) {
    int occ = vs::serve_max_blocks_per_cu(v, S->cfg.min_waves_per_eu, S->max_stack);
    if (occ <= 0) occ = 1;
    uint32_t wpc = S->cfg.waves_per_cu ? std::min<uint32_t>(S->cfg.waves_per_cu, uint32_t(occ)) : uint32_t(occ);
    return S->cu_count * wpc;
}

// the step's device work: upload the table, reset the claim counters, one
// persistent kernel per variant in `mask`, spread over the forked streams
void enqueue_step(vs_server *S, uint32_t mask, uint32_t buf) {
    hipStream_t s0 = S->st[0];
    hip_check(hipMemcpyAsync(S->d_tables[buf].ptr, S->h_tables[buf], sizeof(vs::StepTable) * vs::SV_COUNT,
                             hipMemcpyHostToDevice, s0),
              "step table upload");
    hip_check(vs::serve_reset(S->d_next.as<uint32_t>(), vs::SV_COUNT, s0), "serve reset");
    hip_check(hipEventRecord(S->ev_fork, s0), "fork");
    uint32_t used[kStreams] = {0, 0, 0};
    uint32_t k = 0;
    const vs::StepTable *tables = S->d_tables[buf].as<vs::StepTable>();
    unsigned long long *ctr = S->cfg.count_rays ? S->d_counters.as<unsigned long long>() : nullptr;
    for (uint32_t v = 0; v < vs::SV_COUNT; ++v) {
        if (!(mask & (1u << v))) continue;
        uint32_t si = k++ % kStreams;
        hipStream_t s = S->st[si];
        if (si && !used[si]) hip_check(hipStreamWaitEvent(s, S->ev_fork, 0), "fork wait");
        used[si] = 1;
        hip_check(vs::serve_launch(v, S->cfg.min_waves_per_eu, S->d_views.as<vs::SceneView>(), S->d_jobs.as<vs::ServeJob>(), tables + v,
                                   S->d_next.as<uint32_t>() + v, variant_blocks(S, v), S->cfg.tile_w,
                                   S->max_stack, ctr, s),
                  "serve launch");
    }
    for (uint32_t si = 1; si < kStreams; ++si) {
        if (!used[si]) continue;
        hip_check(hipEventRecord(S->ev_join[si], S->st[si]), "join");
        hip_check(hipStreamWaitEvent(s0, S->ev_join[si], 0), "join wait");
    }
}

void launch_step(vs_server *S, uint32_t mask, uint32_t buf) {
    if (!S->cfg.use_graphs) {
        enqueue_step(S, mask, buf);
        return;
    }
    uint32_t key = mask | (buf << 8);
    auto it = S->graphs.find(key);
    if (it == S->graphs.end()) {
        hipGraph_t g = nullptr;
        hip_check(hipStreamBeginCapture(S->st[0], hipStreamCaptureModeThreadLocal), "begin capture");
        try {
            enqueue_step(S, mask, buf);
        } catch (...) {
            (void)hipStreamEndCapture(S->st[0], &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(S->st[0], &g), "end capture");
        hipGraphExec_t ex = nullptr;
        hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        hip_check(e, "graph instantiate");
        it = S->graphs.emplace(key, ex).first;
    }
    hip_check(hipGraphLaunch(it->second, S->st[0]), "graph launch");
    std::lock_guard<std::mutex> lk(S->mu);
    ++S->graph_launches;
}

// ---- worker stages ---------------------------------------------------------

void finish_job(vs_server *S, JobRec *j, bool ok, const std::string &err) {
    Clock::time_point now = Clock::now();
    if (j->off != SIZE_MAX) S->arena.release(j->off, j->bytes);
    j->off = SIZE_MAX;
    S->slot_job[j->slot] = nullptr;
    std::lock_guard<std::mutex> lk(S->mu);
    j->t_done = now;
    j->state = ok ? VS_JOB_DONE : VS_JOB_FAILED;
    j->error = err;
    if (ok) {
        ++S->n_done;
        S->samples_done += uint64_t(j->d.width) * j->d.height * j->d.spp;
        S->latencies.push_back(ms_between(j->t_submit, now));
        S->t_last = now;
    } else {
        ++S->n_failed;
    }
    S->cv_done.notify_all();
}

void retire_copies(vs_server *S, bool block) {
    for (size_t i = 0; i < S->copying.size();) {
        JobRec *j = S->copying[i];
        hipError_t e = block ? hipEventSynchronize(j->ev_copy) : hipEventQuery(j->ev_copy);
        if (e == hipErrorNotReady) {
            ++i;
            continue;
        }
        hip_check(e, "result copy");
        S->free_events.push_back(j->ev_copy);
        j->ev_copy = nullptr;
        S->copying.erase(S->copying.begin() + long(i));
        finish_job(S, j, true, "");
        block = false;// one is enough to make progress
    }
}

void admit(vs_server *S) {
    while (true) {
        JobRec *j = nullptr;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (S->queue.empty()) return;
            j = S->queue.front();
        }
        auto free_slot = std::find(S->slot_job.begin(), S->slot_job.end(), nullptr);
        if (free_slot == S->slot_job.end()) return;
        size_t off = 0;
        if (!S->arena.alloc(j->bytes, off)) return;// first come, first served: no overtaking
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->queue.pop_front();
            j->state = VS_JOB_RUNNING;
            j->t_admit = Clock::now();
        }
        j->off = off;
        {
            // pinned host copy of the films, recycled by size
            float *buf = nullptr;
            {
                std::lock_guard<std::mutex> lk(S->mu);
                auto it = S->pinned_pool.find(j->bytes);
                if (it != S->pinned_pool.end()) {
                    buf = it->second;
                    S->pinned_pool.erase(it);
                }
            }
            if (!buf) {
                void *hp = nullptr;
                hip_check(hipHostMalloc(&hp, j->bytes, hipHostMallocDefault), "pinned result");
                buf = static_cast<float *>(hp);
            }
            j->h_rt = buf;
            j->h_out = buf + size_t(j->d.width) * j->d.height * 4;
        }
        j->slot = uint32_t(free_slot - S->slot_job.begin());
        *free_slot = j;
        vs::CameraView cam;
        uint32_t route;
        {
            // the scene table never reallocates (reserved at creation), but
            // a route may change under the lock
            std::lock_guard<std::mutex> lk(S->mu);
            const SceneSlot &sc = S->scenes[j->d.scene];
            cam = sc.cam;
            route = sc.route[j->d.spectral ? 1 : 0];
            if (S->cfg.lane_min_samples &&
                uint64_t(j->d.width) * j->d.height * j->d.spp >= S->cfg.lane_min_samples)
                route = 1;
        }
        vs::ServeJob &sj = S->h_jobs[j->slot];
        std::memset(&sj, 0, sizeof(sj));
        sj.cam = cam;
        camera_set_resolution(sj.cam, j->d.width, j->d.height);
        sj.rv = j->rv;
        sj.rt = reinterpret_cast<float4 *>(static_cast<char *>(S->arena_mem.ptr) + off);
        sj.scene = 2u * j->d.scene + (j->d.spectral ? 1u : 0u);
        sj.tiles_x = (j->d.width + S->cfg.tile_w - 1) / S->cfg.tile_w;
        uint32_t px = j->d.width * j->d.height;
        j->route = route;
        if (j->route == 1) {
            // the lane clears the film on its own stream: nothing of a lane
            // job waits on st[0], which may be capturing a step graph (a
            // wait on an event of a capturing stream is an error)
            std::lock_guard<std::mutex> lk(S->mu);
            S->lane_queue.push_back(j);
            S->cv_lane.notify_one();
            continue;
        }
        hip_check(vs::serve_clear(sj.rt, px, S->st[0]), "film clear");
        hip_check(hipMemcpyAsync(S->d_jobs.as<vs::ServeJob>() + j->slot, &sj, sizeof(sj), hipMemcpyHostToDevice,
                                 S->st[0]),
                  "job upload");
        S->active.push_back(j);
        std::lock_guard<std::mutex> lk(S->mu);
        S->max_active = std::max<uint32_t>(S->max_active, uint32_t(S->active.size()));
    }
}

bool has_unscheduled(const vs_server *S) {
    for (const JobRec *j : S->active)
        if (j->frames_sched < j->d.spp) return true;
    return false;
}

void plan_and_launch(vs_server *S) {
    const uint32_t buf = uint32_t(S->step_count & 1u);
    vs::StepTable *tabs = S->h_tables[buf];
    for (uint32_t v = 0; v < vs::SV_COUNT; ++v) {
        tabs[v].num_entries = 0;
        tabs[v].num_items = 0;
    }
    StepRec step;
    step.buf = buf;
    // sample budget of the step from the measured rate (first step: 1 frame each)
    double budget = S->rate > 0.0 ? S->rate * double(S->cfg.step_ms) : 0.0;
    struct Pick {
        JobRec *j;
        uint32_t frames;
    };
    std::vector<Pick> picks;
    for (JobRec *j : S->active) {
        uint32_t left = j->d.spp - j->frames_sched;
        if (!left) continue;
        double px = double(j->d.width) * j->d.height;
        uint32_t f = 1;
        if (budget > px) f = uint32_t(std::min<double>(double(left), std::floor(budget / px)));
        f = std::max<uint32_t>(1u, std::min(f, left));
        budget -= px * f;
        picks.push_back({j, f});
    }
    // longest items first within each variant, so the step's tail is short
    std::stable_sort(picks.begin(), picks.end(), [](const Pick &a, const Pick &b) { return a.frames > b.frames; });
    uint32_t mask = 0;
    for (const Pick &p : picks) {
        vs::StepTable &t = tabs[p.j->variant];
        if (t.num_entries == vs::kServeMaxEntries) continue;// next step
        uint32_t th = tile_h(S);
        uint32_t tiles = ((p.j->d.width + S->cfg.tile_w - 1) / S->cfg.tile_w) * ((p.j->d.height + th - 1) / th);
        vs::StepEntry &e = t.e[t.num_entries++];
        e.job = p.j->slot;
        e.frame_start = p.j->frames_sched;
        e.frames = p.frames;
        e.first_item = t.num_items;
        t.num_items += tiles;
        p.j->frames_sched += p.frames;
        step.work.emplace_back(p.j, p.frames);
        step.samples += uint64_t(p.j->d.width) * p.j->d.height * p.frames;
        mask |= 1u << p.j->variant;
    }
    step.ev_start = take_event(S);
    step.ev_end = take_event(S);
    hip_check(hipEventRecord(step.ev_start, S->st[0]), "step start");
    launch_step(S, mask, buf);
    hip_check(hipEventRecord(step.ev_end, S->st[0]), "step end");
    ++S->step_count;
    S->inflight.push_back(std::move(step));
}

void complete_oldest_step(vs_server *S) {
    StepRec step = std::move(S->inflight.front());
    S->inflight.pop_front();
    hip_check(hipEventSynchronize(step.ev_end), "step");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, step.ev_start, step.ev_end), "step time");
    S->free_events.push_back(step.ev_start);
    S->free_events.push_back(step.ev_end);
    // the step's rate steers the next budgets (device time only, so a step
    // that waited behind its predecessor does not count the wait: its start
    // event fires when the device reaches it)
    if (ms > 0.f && step.samples) {
        double r = double(step.samples) / double(ms);
        S->rate = S->rate > 0.0 ? 0.7 * S->rate + 0.3 * r : r;
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->busy_ms += ms;
        ++S->steps;
    }
    for (auto &w : step.work) {
        JobRec *j = w.first;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            j->frames_done += w.second;
        }
        if (j->frames_done < j->d.spp) continue;
        // finished: output pass and download on the copy stream
        S->active.erase(std::find(S->active.begin(), S->active.end(), j));
        uint32_t px = j->d.width * j->d.height;
        float4 *rt = reinterpret_cast<float4 *>(static_cast<char *>(S->arena_mem.ptr) + j->off);
        float4 *out = rt + px;
        hip_check(hipStreamWaitEvent(S->st_copy, step.ev_end, 0), "copy wait");
        hip_check(vs::serve_output(j->rv, rt, out, px, S->st_copy), "output pass");
        hip_check(hipMemcpyAsync(j->h_rt, rt, size_t(px) * 16, hipMemcpyDeviceToHost, S->st_copy), "result D2H");
        hip_check(hipMemcpyAsync(j->h_out, out, size_t(px) * 16, hipMemcpyDeviceToHost, S->st_copy), "result D2H");
        j->ev_copy = take_event(S);
        hip_check(hipEventRecord(j->ev_copy, S->st_copy), "copy event");
        S->copying.push_back(j);
    }
}

void fail_everything(vs_server *S, const std::string &msg) {
    for (JobRec *j : S->active) finish_job(S, j, false, msg);
    S->active.clear();
    for (JobRec *j : S->copying) finish_job(S, j, false, msg);
    S->copying.clear();
    S->inflight.clear();
    std::lock_guard<std::mutex> lk(S->mu);
    // queued jobs and those waiting for the wavefront lane or for their
    // retirement by this (ending) worker fail too; a job the lane is
    // rendering now fails through vs_server_wait's check of `fatal`
    auto fail = [&](JobRec *j) {
        j->state = VS_JOB_FAILED;
        j->error = msg;
        ++S->n_failed;
    };
    for (JobRec *j : S->queue) fail(j);
    S->queue.clear();
    for (JobRec *j : S->lane_queue) fail(j);
    S->lane_queue.clear();
    for (JobRec *j : S->lane_done) fail(j);
    S->lane_done.clear();
    S->fatal = msg;
    S->cv_done.notify_all();
}

// jobs the wavefront lane has finished: release their slot and HBM range
void retire_lane(vs_server *S) {
    std::vector<JobRec *> done;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        done.swap(S->lane_done);
    }
    for (JobRec *j : done) {
        std::string err = j->error;
        finish_job(S, j, err.empty(), err);
    }
}

// a wavefront lane: one job at a time, all its frames, output pass and
// download on the lane's stream; the worker retires it
void lane_main(vs_server *S, int lane) {
    (void)hipSetDevice(S->ctx->device);
    while (true) {
        JobRec *j = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_lane.wait(lk, [&] { return S->stop || !S->lane_queue.empty(); });
            if (S->lane_queue.empty()) return;// stopping
            j = S->lane_queue.front();
            S->lane_queue.pop_front();
        }
        try {
            vs::SceneView view;
            vs::CameraView cam;
            bool diffuse_only;
            {
                std::lock_guard<std::mutex> lk(S->mu);
                const SceneSlot &sc = S->scenes[j->d.scene];
                view = sc.view[j->d.spectral ? 1 : 0];
                cam = sc.cam;
                diffuse_only = sc.diffuse_only;
            }
            const uint32_t px = j->d.width * j->d.height;
            float4 *rt = reinterpret_cast<float4 *>(static_cast<char *>(S->arena_mem.ptr) + j->off);
            float4 *out = rt + px;
            camera_set_resolution(cam, j->d.width, j->d.height);
            // frames per chunk: up to 32 M paths of state (5.9 GB) at a time
            const uint32_t chunk = std::max<uint32_t>(1u, std::min<uint32_t>(64u, (32u << 20) / std::max(px, 1u)));
            hipStream_t ls = S->st_lane[lane];
            hip_check(vs::serve_clear(rt, px, ls), "film clear");
            unsigned long long *ctr = nullptr;
            if (S->cfg.count_rays) {
                ctr = S->lane_ctr[lane].as<unsigned long long>();
                hip_check(hipMemsetAsync(ctr, 0, 8 * sizeof(unsigned long long), ls), "lane counters");
            }
            vs_render_stats ts{};
            wavefront_render(view, cam, j->rv, !j->d.spectral, diffuse_only, rt, out, 0, j->d.spp, chunk,
                             64, S->lane_wb[lane], ls, ctr, &ts, S->lane_ev[lane]);
            unsigned long long c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (ctr) hip_check(hipMemcpyAsync(c, ctr, sizeof(c), hipMemcpyDeviceToHost, ls), "lane counters D2H");
            hip_check(vs::serve_output(j->rv, rt, out, px, ls), "output pass");
            hip_check(hipMemcpyAsync(j->h_rt, rt, size_t(px) * 16, hipMemcpyDeviceToHost, ls), "result D2H");
            hip_check(hipMemcpyAsync(j->h_out, out, size_t(px) * 16, hipMemcpyDeviceToHost, ls), "result D2H");
            hip_check(hipStreamSynchronize(ls), "lane sync");
            std::lock_guard<std::mutex> lk(S->mu);
            j->frames_done = j->d.spp;
            S->lane_samples += uint64_t(px) * j->d.spp;
            for (int k = 0; k < 3; ++k) {
                S->lane_stage_ms[k] += ts.stage_ms[k];
                S->lane_stage_launches[k] += ts.stage_launches[k];
            }
            // [0..2] closest-hit rays / nodes / triangles, [4..6] any-hit
            for (int k = 0; k < 3; ++k) {
                S->lane_counts[k] += c[k];
                S->lane_counts[3 + k] += c[4 + k];
            }
        } catch (const std::exception &e) {
            std::lock_guard<std::mutex> lk(S->mu);
            j->error = e.what();
        }
        std::lock_guard<std::mutex> lk(S->mu);
        S->lane_done.push_back(j);
        S->cv_work.notify_all();
    }
}

void worker_main(vs_server *S) {
    try {
        hip_check(hipSetDevice(S->ctx->device), "hipSetDevice");
        while (true) {
            uint64_t seen;
            {
                std::lock_guard<std::mutex> lk(S->mu);
                seen = S->submit_seq;
            }
            retire_lane(S);
            retire_copies(S, false);
            admit(S);
            if (S->inflight.size() < 2 && has_unscheduled(S)) {
                plan_and_launch(S);
                continue;
            }
            if (!S->inflight.empty()) {
                complete_oldest_step(S);
                continue;
            }
            if (!S->copying.empty()) {
                retire_copies(S, true);
                continue;
            }
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->stop) break;
            if (!S->queue.empty()) {
                // the head job was not admitted and no step or copy is in
                // flight: the wavefront lane holds the slots or the HBM it
                // needs (a job within the aligned budget always fits an
                // empty arena), so wait for the lane or for a new submit
                S->cv_work.wait(lk, [&] { return S->stop || !S->lane_done.empty() || S->submit_seq != seen; });
                continue;
            }
            if (!S->lane_done.empty()) continue;
            S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty() || !S->lane_done.empty(); });
        }
    } catch (const std::exception &e) {
        fail_everything(S, e.what());
    }
}

void validate_config(const vs_server_config &c) {
    if (c.max_jobs == 0 || c.max_jobs > vs::kServeMaxEntries) throw std::invalid_argument("max_jobs must be in 1..256");
    if (!(c.tile_w == 8 || c.tile_w == 16 || c.tile_w == 32 || c.tile_w == 64))
        throw std::invalid_argument("tile_w must be 8, 16, 32 or 64");
    if (!(c.step_ms > 0.f && c.step_ms <= 10000.f)) throw std::invalid_argument("step_ms must be in (0, 10000]");
    if (c.waves_per_cu > 64) throw std::invalid_argument("waves_per_cu must be at most 64");
    if (!vs::serve_occ_supported(c.min_waves_per_eu)) throw std::invalid_argument("min_waves_per_eu must be 0 or 6");
}

vs_server_config with_defaults(const vs_server_config *in) {
    vs_server_config c{};
    if (in) c = *in;
    if (!c.max_jobs) c.max_jobs = 64;
    if (!(c.step_ms > 0.f)) c.step_ms = 16.f;
    if (!c.tile_w) c.tile_w = 8;
    return c;
}

bool idle(vs_server *S) {
    std::lock_guard<std::mutex> lk(S->mu);
    for (auto &kv : S->jobs)
        if (kv.second->state == VS_JOB_QUEUED || kv.second->state == VS_JOB_RUNNING) return false;
    return true;
}

}// namespace

extern "C" {

const char *vs_server_last_error(void) { return g_server_error.c_str(); }

vs_status vs_server_create(vs_context *ctx, const vs_server_config *cfg, vs_server **out) {
    if (!ctx || !out) return server_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    *out = nullptr;
    if (ctx->device < 0) return server_fail(VS_ERROR_STATE, "host-only context: no device work possible");
    std::unique_ptr<vs_server> S(new vs_server());
    vs_status st = server_guarded([&] {
        S->ctx = ctx;
        S->cfg = with_defaults(cfg);
        validate_config(S->cfg);
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        int cus = 0;
        hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device), "device attribute");
        S->cu_count = uint32_t(std::max(cus, 1));
        if (!S->cfg.hbm_budget) {
            size_t fr = 0, tot = 0;
            hip_check(hipMemGetInfo(&fr, &tot), "memory info");
            S->cfg.hbm_budget = fr / 4;
        }
        // the arena hands out kArenaAlign-byte multiples
        S->cfg.hbm_budget = S->cfg.hbm_budget / kArenaAlign * kArenaAlign;
        if (!S->cfg.hbm_budget) throw std::invalid_argument("hbm_budget below the arena alignment (256 B)");
        // scene slots are referenced while jobs run: never reallocate them
        S->scenes.reserve(kMaxScenes);
        for (auto &s : S->st) hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream");
        hip_check(hipStreamCreateWithFlags(&S->st_copy, hipStreamNonBlocking), "stream");
        hip_check(hipEventCreateWithFlags(&S->ev_fork, hipEventDisableTiming), "event");
        for (auto &e : S->ev_join) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
        S->d_views.alloc(sizeof(vs::SceneView) * 2 * kMaxScenes);
        S->arena_mem.alloc(S->cfg.hbm_budget);
        S->arena.reset(S->cfg.hbm_budget);
        S->d_jobs.alloc(sizeof(vs::ServeJob) * S->cfg.max_jobs);
        void *hp = nullptr;
        hip_check(hipHostMalloc(&hp, sizeof(vs::ServeJob) * S->cfg.max_jobs, hipHostMallocDefault), "pinned jobs");
        S->h_jobs = static_cast<vs::ServeJob *>(hp);
        for (int b = 0; b < 2; ++b) {
            S->d_tables[b].alloc(sizeof(vs::StepTable) * vs::SV_COUNT);
            hip_check(hipHostMalloc(&hp, sizeof(vs::StepTable) * vs::SV_COUNT, hipHostMallocDefault), "pinned table");
            S->h_tables[b] = static_cast<vs::StepTable *>(hp);
            std::memset(hp, 0, sizeof(vs::StepTable) * vs::SV_COUNT);
        }
        S->d_next.alloc(64 * sizeof(uint32_t));
        S->d_counters.alloc(8 * sizeof(unsigned long long));
        hip_check(hipMemset(S->d_counters.ptr, 0, 8 * sizeof(unsigned long long)), "counter clear");
        S->slot_job.assign(S->cfg.max_jobs, nullptr);
        for (auto &ls : S->st_lane) hip_check(hipStreamCreateWithFlags(&ls, hipStreamNonBlocking), "stream");
        for (int l = 0; l < vs_server::kLanes; ++l) {
            for (auto &e : S->lane_ev[l]) hip_check(hipEventCreate(&e), "lane event");
            S->lane_ctr[l].alloc(8 * sizeof(unsigned long long));
        }
        S->worker = std::thread(worker_main, S.get());
        for (int l = 0; l < vs_server::kLanes; ++l) S->lane[l] = std::thread(lane_main, S.get(), l);
    });
    if (st != VS_OK) {
        vs_server *raw = S.release();
        vs_server_destroy(raw);
        return st;
    }
    *out = S.release();
    return VS_OK;
}

vs_status vs_server_add_scene(vs_server *S, vs_scene *scene, uint32_t *index) {
    if (!S || !scene || !index) return server_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    if (scene->ctx != S->ctx) return server_fail(VS_ERROR_INVALID_ARGUMENT, "scene belongs to another context");
    if (!scene->built) return server_fail(VS_ERROR_STATE, "accel not built: call vs_scene_build_accel first");
    if (scene->host.integrator.type == "rt")
        return server_fail(VS_ERROR_UNSUPPORTED, "serving renders path-tracing jobs; \"rt\" scenes are not supported");
    return server_guarded([&] {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->scenes.size() >= kMaxScenes) throw std::invalid_argument("too many resident scenes (64)");
        hip_check(hipSetDevice(S->ctx->device), "hipSetDevice");
        SceneSlot slot;
        slot.scene = scene;
        slot.cam = scene->cam;
        slot.rv = render_view(scene);
        slot.diffuse_only = scene->diffuse_only;
        // both spectrum modes share the geometry and BVH buffers
        vs::SceneView views[2] = {view_in_mode(scene, true), view_in_mode(scene, false)};
        slot.view[0] = views[0];
        slot.view[1] = views[1];
        if (scene->view.stack_size > S->max_stack) {
            for (auto &kv : S->jobs)
                if (kv.second->state == VS_JOB_QUEUED || kv.second->state == VS_JOB_RUNNING)
                    throw std::logic_error("a scene with a deeper BVH cannot join while jobs run");
        }
        uint32_t i = uint32_t(S->scenes.size());
        hip_check(hipMemcpy(S->d_views.as<vs::SceneView>() + 2 * i, views, sizeof(views), hipMemcpyHostToDevice),
                  "view upload");
        S->scenes.push_back(std::move(slot));
        if (scene->view.stack_size > S->max_stack) {
            S->max_stack = scene->view.stack_size;
            drop_graphs(S);
        }
        *index = i;
    });
}

vs_status vs_server_submit(vs_server *S, const vs_job_desc *d, uint64_t *job_id) {
    if (!S || !d || !job_id) return server_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    return server_guarded([&] {
        std::unique_ptr<JobRec> j(new JobRec());
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) throw std::logic_error("server is shutting down");
        if (!S->fatal.empty()) throw DeviceError("server failed: " + S->fatal);
        if (d->scene >= S->scenes.size()) throw std::invalid_argument("unknown scene index");
        if (d->width == 0 || d->height == 0 || d->width > 16384 || d->height > 16384)
            throw std::invalid_argument("bad job resolution");
        if (d->spp == 0 || d->spp > 65536) throw std::invalid_argument("spp must be in 1..65536");
        if (d->max_depth > 1024) throw std::invalid_argument("max_depth must be at most 1024");
        uint32_t dim = d->spectral ? (d->spectrum_dim ? d->spectrum_dim : 4u) : 3u;
        if (dim != 3 && dim != 4) throw std::invalid_argument("spectrum_dim must be 3 or 4");
        j->d = *d;
        j->d.spectrum_dim = dim;
        j->bytes = job_bytes(*d);
        if ((j->bytes + kArenaAlign - 1) / kArenaAlign * kArenaAlign > S->cfg.hbm_budget) throw std::bad_alloc();
        const SceneSlot &sc = S->scenes[d->scene];
        j->rv = sc.rv;
        if (d->max_depth) j->rv.max_depth = d->max_depth;
        j->rv.spectrum_dim = dim;
        j->rv.write_output = 0;
        j->rv.frame_start = 0;
        j->rv.spp = d->spp;
        uint32_t base = !d->spectral ? vs::SV_RGB : (dim == 4 ? vs::SV_HERO4 : vs::SV_HERO3);
        j->variant = base + (sc.diffuse_only ? 1u : 0u);
        j->off = SIZE_MAX;
        j->id = S->next_id++;
        j->t_submit = Clock::now();
        if (!S->have_first) {
            S->have_first = true;
            S->t_first = j->t_submit;
        }
        ++S->n_submitted;
        *job_id = j->id;
        S->queue.push_back(j.get());
        S->jobs.emplace(j->id, std::move(j));
        ++S->submit_seq;
        S->cv_work.notify_one();
    });
}

vs_status vs_server_wait(vs_server *S, uint64_t id, double timeout_ms, int32_t *done) {
    if (!S || !done) return server_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->jobs.find(id);
    if (it == S->jobs.end()) return server_fail(VS_ERROR_INVALID_ARGUMENT, "unknown job");
    JobRec *j = it->second.get();
    auto finished = [&] { return j->state == VS_JOB_DONE || j->state == VS_JOB_FAILED || !S->fatal.empty(); };
    if (timeout_ms < 0) {
        S->cv_done.wait(lk, finished);
    } else {
        S->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), finished);
    }
    *done = finished() ? 1 : 0;
    if (j->state == VS_JOB_FAILED) return server_fail(VS_ERROR_DEVICE, "job failed: " + j->error);
    if (j->state != VS_JOB_DONE && !S->fatal.empty()) return server_fail(VS_ERROR_DEVICE, "server failed: " + S->fatal);
    return VS_OK;
}

vs_status vs_server_job_info(vs_server *S, uint64_t id, vs_job_info *info) {
    if (!S || !info) return server_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->jobs.find(id);
    if (it == S->jobs.end()) return server_fail(VS_ERROR_INVALID_ARGUMENT, "unknown job");
    const JobRec *j = it->second.get();
    std::memset(info, 0, sizeof(*info));
    info->state = j->state;
    info->frames_done = j->frames_done;
    info->route = j->route;
    info->samples = uint64_t(j->d.width) * j->d.height * j->d.spp;
    if (j->state != VS_JOB_QUEUED) info->queue_ms = ms_between(j->t_submit, j->t_admit);
    if (j->state == VS_JOB_DONE) info->latency_ms = ms_between(j->t_submit, j->t_done);
    return VS_OK;
}

vs_status vs_server_result(vs_server *S, uint64_t id, int32_t which, float *rgba) {
    if (!S || !rgba || (which != 0 && which != 1)) return server_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->jobs.find(id);
    if (it == S->jobs.end()) return server_fail(VS_ERROR_INVALID_ARGUMENT, "unknown job");
    const JobRec *j = it->second.get();
    if (j->state != VS_JOB_DONE) return server_fail(VS_ERROR_STATE, "job not finished");
    size_t px = size_t(j->d.width) * j->d.height;
    std::memcpy(rgba, which == 0 ? j->h_rt : j->h_out, px * 16);
    return VS_OK;
}

vs_status vs_server_release(vs_server *S, uint64_t id) {
    if (!S) return server_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->jobs.find(id);
    if (it == S->jobs.end()) return server_fail(VS_ERROR_INVALID_ARGUMENT, "unknown job");
    JobRec *j = it->second.get();
    if (j->state != VS_JOB_DONE && j->state != VS_JOB_FAILED) return server_fail(VS_ERROR_STATE, "job not finished");
    if (j->h_rt) S->pinned_pool.emplace(j->bytes, j->h_rt);
    S->jobs.erase(it);
    return VS_OK;
}

vs_status vs_server_drain(vs_server *S, double timeout_ms) {
    if (!S) return server_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    std::unique_lock<std::mutex> lk(S->mu);
    auto drained = [&] {
        if (!S->fatal.empty()) return true;
        for (auto &kv : S->jobs)
            if (kv.second->state == VS_JOB_QUEUED || kv.second->state == VS_JOB_RUNNING) return false;
        return true;
    };
    bool ok = timeout_ms < 0 ? (S->cv_done.wait(lk, drained), true)
                             : S->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), drained);
    if (!S->fatal.empty()) return server_fail(VS_ERROR_DEVICE, "server failed: " + S->fatal);
    if (!ok) return server_fail(VS_ERROR_STATE, "drain timed out");
    return VS_OK;
}

vs_status vs_server_stats_get(vs_server *S, vs_server_stats *o) {
    if (!S || !o) return server_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    std::memset(o, 0, sizeof(*o));
    {
        std::lock_guard<std::mutex> lk(S->mu);
        o->jobs_submitted = S->n_submitted;
        o->jobs_done = S->n_done;
        o->jobs_failed = S->n_failed;
        o->samples_done = S->samples_done;
        o->steps = S->steps;
        o->graph_launches = S->graph_launches;
        o->busy_ms = S->busy_ms;
        o->wall_ms = (S->have_first && S->n_done) ? ms_between(S->t_first, S->t_last) : 0.0;
        o->msamples_per_s = o->wall_ms > 0 ? double(S->samples_done) / (o->wall_ms * 1e3) : 0.0;
        std::vector<double> l = S->latencies;
        if (!l.empty()) {
            std::sort(l.begin(), l.end());
            auto pct = [&](double q) {
                // nearest-rank percentile
                size_t r = size_t(std::ceil(q * double(l.size())));
                return l[std::min(l.size() - 1, r ? r - 1 : 0)];
            };
            o->latency_p50_ms = pct(0.50);
            o->latency_p99_ms = pct(0.99);
            double sum = 0;
            for (double x : l) sum += x;
            o->latency_mean_ms = sum / double(l.size());
            o->latency_max_ms = l.back();
        }
        o->max_active = S->max_active;
        o->queued = uint32_t(S->queue.size());
        uint32_t running = 0;
        for (auto &kv : S->jobs) running += kv.second->state == VS_JOB_RUNNING ? 1u : 0u;
        o->active = running;
        o->hbm_budget = S->cfg.hbm_budget;
        o->hbm_peak = S->arena.peak();
        o->lane_samples = S->lane_samples;
        for (int k = 0; k < 3; ++k) {
            o->lane_stage_ms[k] = S->lane_stage_ms[k];
            o->lane_stage_launches[k] = S->lane_stage_launches[k];
        }
        for (int k = 0; k < 2; ++k) {
            o->lane_rays[k] = S->lane_counts[3 * k];
            o->lane_node_fetches[k] = S->lane_counts[3 * k + 1];
            o->lane_tri_fetches[k] = S->lane_counts[3 * k + 2];
        }
    }
    if (S->cfg.count_rays) {
        unsigned long long c[8];
        vs_status st = server_guarded([&] {
            hip_check(hipSetDevice(S->ctx->device), "hipSetDevice");
            hip_check(hipMemcpy(c, S->d_counters.ptr, sizeof(c), hipMemcpyDeviceToHost), "counters D2H");
        });
        if (st != VS_OK) return st;
        o->rays = c[0];
        o->node_fetches = c[1];
        o->tri_fetches = c[2];
    }
    return VS_OK;
}

vs_status vs_server_reset_stats(vs_server *S) {
    if (!S) return server_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    if (!idle(S)) return server_fail(VS_ERROR_STATE, "server busy");
    return server_guarded([&] {
        std::lock_guard<std::mutex> lk(S->mu);
        S->n_submitted = S->n_done = S->n_failed = S->samples_done = S->graph_launches = S->steps = 0;
        S->busy_ms = 0.0;
        S->lane_samples = 0;
        for (int k = 0; k < 3; ++k) {
            S->lane_stage_ms[k] = 0.0;
            S->lane_stage_launches[k] = 0;
        }
        for (auto &c : S->lane_counts) c = 0;
        S->have_first = false;
        S->latencies.clear();
        S->max_active = 0;
        S->arena.reset_peak();
        hip_check(hipSetDevice(S->ctx->device), "hipSetDevice");
        hip_check(hipMemset(S->d_counters.ptr, 0, 8 * sizeof(unsigned long long)), "counter clear");
    });
}

vs_status vs_server_set_config(vs_server *S, const vs_server_config *cfg) {
    if (!S || !cfg) return server_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    if (!idle(S)) return server_fail(VS_ERROR_STATE, "server busy");
    return server_guarded([&] {
        std::lock_guard<std::mutex> lk(S->mu);
        vs_server_config c = with_defaults(cfg);
        c.hbm_budget = S->cfg.hbm_budget;
        c.max_jobs = S->cfg.max_jobs;
        validate_config(c);
        // the worker reads the config only while it has jobs, and it has none
        bool relaunch = c.tile_w != S->cfg.tile_w || c.waves_per_cu != S->cfg.waves_per_cu ||
                        c.count_rays != S->cfg.count_rays || c.min_waves_per_eu != S->cfg.min_waves_per_eu;
        S->cfg = c;
        if (relaunch) drop_graphs(S);
    });
}

vs_status vs_server_get_config(vs_server *S, vs_server_config *cfg) {
    if (!S || !cfg) return server_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    std::lock_guard<std::mutex> lk(S->mu);
    *cfg = S->cfg;
    return VS_OK;
}

vs_status vs_server_set_route(vs_server *S, uint32_t scene, int32_t spectral, uint32_t route) {
    if (!S || route > 1) return server_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    std::lock_guard<std::mutex> lk(S->mu);
    if (scene >= S->scenes.size()) return server_fail(VS_ERROR_INVALID_ARGUMENT, "unknown scene index");
    S->scenes[scene].route[spectral ? 1 : 0] = route;
    return VS_OK;
}

vs_status vs_server_get_route(vs_server *S, uint32_t scene, int32_t spectral, uint32_t *route) {
    if (!S || !route) return server_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    std::lock_guard<std::mutex> lk(S->mu);
    if (scene >= S->scenes.size()) return server_fail(VS_ERROR_INVALID_ARGUMENT, "unknown scene index");
    *route = S->scenes[scene].route[spectral ? 1 : 0];
    return VS_OK;
}

vs_status vs_server_autotune(vs_server *S, const vs_job_desc *probe, uint32_t n, vs_autotune_result *out) {
    if (!S || !probe || !n || !out) return server_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    if (!idle(S)) return server_fail(VS_ERROR_STATE, "server busy");
    std::memset(out, 0, sizeof(*out));
    vs_server_config base{};
    vs_server_get_config(S, &base);
    for (uint32_t i = 0; i < n; ++i)
        if (probe[i].scene >= S->scenes.size()) return server_fail(VS_ERROR_INVALID_ARGUMENT, "unknown scene index");
    // candidate shapes: register budget x tile width x persistent waves per CU
    // (the occupancy limit of the probe's variants, half and a quarter of it)
    struct Shape {
        uint32_t regs, tw, wpc;
    };
    std::vector<Shape> shapes;
    for (uint32_t regs : {0u, 6u}) {
        int occ = 64;
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t dim = probe[i].spectral ? (probe[i].spectrum_dim ? probe[i].spectrum_dim : 4u) : 3u;
            uint32_t v = !probe[i].spectral ? vs::SV_RGB : (dim == 4 ? vs::SV_HERO4 : vs::SV_HERO3);
            v += S->scenes[probe[i].scene].diffuse_only ? 1u : 0u;
            occ = std::min(occ, std::max(1, vs::serve_max_blocks_per_cu(v, regs, S->max_stack)));
        }
        for (uint32_t tw : {8u, 16u})
            for (uint32_t wpc : {uint32_t(occ), uint32_t(std::max(1, occ / 2))}) shapes.push_back({regs, tw, wpc});
    }
    double best = -1.0;
    vs_server_config best_cfg = base;
    for (const Shape &sh : shapes) {
        {
            const uint32_t tw = sh.tw, wpc = sh.wpc;
            vs_server_config c = base;
            c.tile_w = tw;
            c.waves_per_cu = wpc;
            c.min_waves_per_eu = sh.regs;
            vs_status st = vs_server_set_config(S, &c);
            if (st == VS_OK) st = vs_server_reset_stats(S);
            std::vector<uint64_t> ids(n);
            for (uint32_t i = 0; i < n && st == VS_OK; ++i) st = vs_server_submit(S, &probe[i], &ids[i]);
            if (st == VS_OK) st = vs_server_drain(S, -1.0);
            vs_server_stats ss{};
            if (st == VS_OK) st = vs_server_stats_get(S, &ss);
            for (uint64_t id : ids)
                if (id) vs_server_release(S, id);
            if (st != VS_OK) {
                vs_server_set_config(S, &base);
                return st;
            }
            if (out->trials < VS_AUTOTUNE_MAX_TRIALS) {
                out->trial_tile_w[out->trials] = tw;
                out->trial_waves_per_cu[out->trials] = wpc;
                out->trial_min_waves_per_eu[out->trials] = sh.regs;
                out->trial_msamples_per_s[out->trials] = ss.msamples_per_s;
                ++out->trials;
            }
            if (ss.msamples_per_s > best) {
                best = ss.msamples_per_s;
                best_cfg = c;
            }
        }
    }
    vs_status st = vs_server_set_config(S, &best_cfg);
    // routes: the largest probe job of each (scene, mode), alone on each route
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    for (uint32_t i = 0; i < n && st == VS_OK; ++i) {
        std::pair<uint32_t, uint32_t> key{probe[i].scene, probe[i].spectral ? 1u : 0u};
        if (std::find(pairs.begin(), pairs.end(), key) != pairs.end()) continue;
        pairs.push_back(key);
        const vs_job_desc *big = &probe[i];
        for (uint32_t k = i; k < n; ++k)
            if (probe[k].scene == key.first && (probe[k].spectral ? 1u : 0u) == key.second &&
                uint64_t(probe[k].width) * probe[k].height * probe[k].spp > uint64_t(big->width) * big->height * big->spp)
                big = &probe[k];
        double ms[2] = {0.0, 0.0};
        for (uint32_t route = 0; route < 2 && st == VS_OK; ++route) {
            st = vs_server_set_route(S, key.first, int32_t(key.second), route);
            uint64_t id = 0;
            if (st == VS_OK) st = vs_server_submit(S, big, &id);
            int32_t done = 0;
            if (st == VS_OK) st = vs_server_wait(S, id, -1.0, &done);
            vs_job_info info{};
            if (st == VS_OK) st = vs_server_job_info(S, id, &info);
            if (id) vs_server_release(S, id);
            ms[route] = info.latency_ms;
        }
        if (st != VS_OK) break;
        // the lane renders one job at a time beside the persistent steps:
        // under load it only pays where the wavefront form is more than
        // twice as fast alone
        uint32_t choice = ms[1] * 2.0 < ms[0] ? 1u : 0u;
        st = vs_server_set_route(S, key.first, int32_t(key.second), choice);
        if (out->routes < VS_AUTOTUNE_MAX_TRIALS) {
            uint32_t r = out->routes++;
            out->route_scene[r] = key.first;
            out->route_spectral[r] = key.second;
            out->route_choice[r] = choice;
            out->route_megakernel_ms[r] = ms[0];
            out->route_wavefront_ms[r] = ms[1];
        }
    }
    if (st == VS_OK) st = vs_server_reset_stats(S);
    out->tile_w = best_cfg.tile_w;
    out->waves_per_cu = best_cfg.waves_per_cu;
    out->min_waves_per_eu = best_cfg.min_waves_per_eu;
    out->msamples_per_s = best;
    return st;
}

void vs_server_destroy(vs_server *S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        // queued jobs never start
        while (!S->queue.empty()) {
            JobRec *j = S->queue.front();
            S->queue.pop_front();
            j->state = VS_JOB_FAILED;
            j->error = "server destroyed";
        }
        S->cv_work.notify_all();
        S->cv_done.notify_all();
    }
    if (S->worker.joinable()) S->worker.join();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->cv_lane.notify_all();
    }
    for (auto &t : S->lane)
        if (t.joinable()) t.join();
    if (S->ctx && S->ctx->device >= 0) {
        (void)hipSetDevice(S->ctx->device);
        for (auto &s : S->st)
            if (s) (void)hipStreamSynchronize(s);
        if (S->st_copy) (void)hipStreamSynchronize(S->st_copy);
        drop_graphs(S);
        for (auto &kv : S->pinned_pool) (void)hipHostFree(kv.second);
        for (auto &kv : S->jobs) {
            if (kv.second->h_rt) (void)hipHostFree(kv.second->h_rt);
            if (kv.second->ev_copy) (void)hipEventDestroy(kv.second->ev_copy);
        }
        for (hipEvent_t e : S->free_events) (void)hipEventDestroy(e);
        for (auto &st : S->inflight) {
            (void)hipEventDestroy(st.ev_start);
            (void)hipEventDestroy(st.ev_end);
        }
        if (S->h_jobs) (void)hipHostFree(S->h_jobs);
        for (auto *t : S->h_tables)
            if (t) (void)hipHostFree(t);
        if (S->ev_fork) (void)hipEventDestroy(S->ev_fork);
        for (auto &e : S->ev_join)
            if (e) (void)hipEventDestroy(e);
        for (auto &s : S->st)
            if (s) (void)hipStreamDestroy(s);
        if (S->st_copy) (void)hipStreamDestroy(S->st_copy);
        for (auto &ls : S->st_lane) {
            if (!ls) continue;
            (void)hipStreamSynchronize(ls);
            (void)hipStreamDestroy(ls);
        }
        for (auto &evs : S->lane_ev)
            for (auto &e : evs)
                if (e) (void)hipEventDestroy(e);
    }
    delete S;
}

}// extern "C"
