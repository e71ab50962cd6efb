// Radiance cache: host driver behind vs_nrc_* (include/vision_amd.h).
//
// Each frame: paths of every pixel run to the termination bounce (radiance
// so far into a per-sample buffer, a query where they stop); training paths
// of every stride-th pixel run to full length and emit one sample per
// surface vertex; the current network answers the queries; the training
// samples take `train_steps` Adam steps; the frame's samples fold into the
// film's running mean.  Everything is enqueued on the scene's stream with
// device-side counts, so a frame needs no host synchronisation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "context.h"
#include "nrc.h"

using namespace vs_internal;

struct vs_nrc {
    vs_scene *scene{nullptr};
    vs_nrc_config cfg{};
    DevBuf mlp, mlp_bf16, grid, grid_f16, mlp_m, mlp_v, mlp_g, grid_m, grid_v, grid_g;
    DevBuf L, queries, counts, staging, samples, loss, io_a, io_b;
    size_t L_cap{0}, q_cap{0};
    uint32_t sample_cap{0}, staging_paths{0};
    vs::NrcView view{};
    uint32_t adam_t{0};
    // per-frame stage events of vs_nrc_render (6 a frame), made as needed:
    // the frames run back to back and their times are read once at the end
    std::vector<hipEvent_t> fev;
    vs_nrc_stats stats{};
};

namespace {

thread_local std::string g_nrc_error;

vs_status nrc_fail(vs_status code, const std::string &m) {
    g_nrc_error = m;
    return code;
}

template <typename F>
vs_status nrc_guarded(vs_nrc *n, F &&f) {
    vs_status st = guarded(n ? n->scene->ctx : nullptr, std::forward<F>(f));
    if (st != VS_OK && n) g_nrc_error = n->scene->ctx->last_error;
    return st;
}

void upload_params(vs_nrc *n, const float *mlp, const float *grid, hipStream_t st) {
    hip_check(hipMemcpyAsync(n->mlp.ptr, mlp, sizeof(float) * vs::kNrcMlpParams, hipMemcpyHostToDevice, st), "mlp");
    if (grid)
        hip_check(hipMemcpyAsync(n->grid.ptr, grid, sizeof(float) * vs::kNrcGridParams, hipMemcpyHostToDevice, st),
                  "grid");
    for (DevBuf *b : {&n->mlp_m, &n->mlp_v, &n->mlp_g})
        hip_check(hipMemsetAsync(b->ptr, 0, sizeof(float) * vs::kNrcMlpParams, st), "clear");
    for (DevBuf *b : {&n->grid_m, &n->grid_v, &n->grid_g})
        hip_check(hipMemsetAsync(b->ptr, 0, sizeof(float) * vs::kNrcGridParams, st), "clear");
    hip_check(vs::nrc_sync_copies(n->view, st), "sync copies");
    n->adam_t = 0;
}

void adam_step(vs_nrc *n, hipStream_t st) {
    ++n->adam_t;
    float c1 = float(1.0 / (1.0 - std::pow(0.9, double(n->adam_t))));
    float c2 = float(1.0 / (1.0 - std::pow(0.99, double(n->adam_t))));
    hip_check(vs::nrc_adam(n->view, n->cfg.learning_rate, c1, c2, st), "adam");
}

uint32_t train_stride(const vs_nrc *n, uint32_t pixels) {
    if (!n->cfg.train_paths) return 0;
    return std::max<uint32_t>(1u, pixels / n->cfg.train_paths);
}

}// namespace

extern "C" {

void vs_nrc_default_config(vs_nrc_config *c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->term_bounce = 2;
    c->train_paths = 65536;
    c->train_steps = 4;
    c->train_batch = 65536;
    c->learning_rate = 1e-2f;
    c->min_res = 16;
    c->max_res = 2048;
    c->seed = 1;
}

const char *vs_nrc_last_error(void) { return g_nrc_error.c_str(); }

vs_status vs_nrc_create(vs_scene *s, const vs_nrc_config *cfg, vs_nrc **out) {
    if (!s || !out) return nrc_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    *out = nullptr;
    std::unique_ptr<vs_nrc> n(new vs_nrc());
    n->scene = s;
    if (cfg) {
        n->cfg = *cfg;
    } else {
        vs_nrc_default_config(&n->cfg);
    }
    const vs_nrc_config &c = n->cfg;
    if (c.min_res < 2 || c.max_res < c.min_res || c.max_res > 1u << 20)
        return nrc_fail(VS_ERROR_INVALID_ARGUMENT, "need 2 <= min_res <= max_res <= 2^20");
    if (!(c.learning_rate > 0.f) || c.train_steps > 64 || c.term_bounce > 64)
        return nrc_fail(VS_ERROR_INVALID_ARGUMENT, "bad learning rate, train_steps or term_bounce");
    vs_status st = nrc_guarded(n.get(), [&] {
        require_built(s);
        if (!s->packed.rgb) throw std::invalid_argument("the radiance cache renders sRGB-mode scenes");
        hip_check(hipSetDevice(s->ctx->device), "hipSetDevice");
        hipStream_t stream = s->ctx->stream;
        vs::NrcView &v = n->view;
        // levels (the oracle restates this in or_nrc_levels)
        double b = std::exp((std::log(double(c.max_res)) - std::log(double(c.min_res))) / double(vs::kNrcLevels - 1));
        v.dense_mask = 0;
        for (int l = 0; l < vs::kNrcLevels; ++l) {
            v.res[l] = uint32_t(std::floor(double(c.min_res) * std::pow(b, double(l)) + 1e-9));
            uint64_t m = uint64_t(v.res[l]) + 1u;
            if (m * m * m <= vs::kNrcTable) v.dense_mask |= 1u << l;
        }
        // the scene's bounds, padded by 0.1% of the extent
        const vs::SceneHost &h = s->host;
        float lo[3] = {h.aabb_lo.x, h.aabb_lo.y, h.aabb_lo.z}, hi[3] = {h.aabb_hi.x, h.aabb_hi.y, h.aabb_hi.z};
        for (int a = 0; a < 3; ++a) {
            float ext = std::max(hi[a] - lo[a], 1e-4f);
            v.lo[a] = lo[a] - 1e-3f * ext;
            v.inv_ext[a] = 1.f / (ext * 1.002f);
        }
        n->mlp.alloc(sizeof(float) * vs::kNrcMlpParams);
        n->mlp_bf16.alloc(sizeof(uint16_t) * vs::kNrcMlpParams);
        n->grid.alloc(sizeof(float) * vs::kNrcGridParams);
        n->grid_f16.alloc(sizeof(uint16_t) * vs::kNrcGridParams);
        for (DevBuf *bb : {&n->mlp_m, &n->mlp_v, &n->mlp_g}) bb->alloc(sizeof(float) * vs::kNrcMlpParams);
        for (DevBuf *bb : {&n->grid_m, &n->grid_v, &n->grid_g}) bb->alloc(sizeof(float) * vs::kNrcGridParams);
        n->counts.alloc(64);
        n->loss.alloc(64);
        v.mlp = n->mlp.as<float>();
        v.mlp_bf16 = n->mlp_bf16.as<uint16_t>();
        v.grid = n->grid.as<float>();
        v.grid_f16 = n->grid_f16.as<uint16_t>();
        v.mlp_m = n->mlp_m.as<float>();
        v.mlp_v = n->mlp_v.as<float>();
        v.mlp_g = n->mlp_g.as<float>();
        v.grid_m = n->grid_m.as<float>();
        v.grid_v = n->grid_v.as<float>();
        v.grid_g = n->grid_g.as<float>();
        // initial parameters: Xavier-uniform layers, grid uniform in +-1e-4
        std::mt19937 rng(c.seed);
        std::vector<float> mlp(vs::kNrcMlpParams), grid(vs::kNrcGridParams);
        auto fill = [&](int off, int fan_in, int fan_out) {
            float a = std::sqrt(6.f / float(fan_in + fan_out));
            std::uniform_real_distribution<float> d(-a, a);
            for (int i = 0; i < fan_in * fan_out; ++i) mlp[size_t(off + i)] = d(rng);
        };
        fill(vs::kNrcW1, vs::kNrcIn, vs::kNrcHidden);
        fill(vs::kNrcW2, vs::kNrcHidden, vs::kNrcHidden);
        fill(vs::kNrcW3, vs::kNrcHidden, vs::kNrcHidden);
        fill(vs::kNrcW4, vs::kNrcHidden, vs::kNrcOut);
        std::uniform_real_distribution<float> dg(-1e-4f, 1e-4f);
        for (float &g : grid) g = dg(rng);
        upload_params(n.get(), mlp.data(), grid.data(), stream);
        hip_check(hipStreamSynchronize(stream), "nrc init");
    });
    if (st != VS_OK) return st;
    *out = n.release();
    return VS_OK;
}

void vs_nrc_destroy(vs_nrc *n) {
    if (!n) return;
    for (auto &e : n->fev)
        if (e) (void)hipEventDestroy(e);
    delete n;
}

vs_status vs_nrc_levels(vs_nrc *n, uint32_t *res, uint32_t *dense_mask, float *lo, float *inv_ext) {
    if (!n || !res || !dense_mask || !lo || !inv_ext) return nrc_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    std::memcpy(res, n->view.res, sizeof(n->view.res));
    *dense_mask = n->view.dense_mask;
    std::memcpy(lo, n->view.lo, sizeof(n->view.lo));
    std::memcpy(inv_ext, n->view.inv_ext, sizeof(n->view.inv_ext));
    return VS_OK;
}

vs_status vs_nrc_get_params(vs_nrc *n, float *mlp, float *grid) {
    if (!n || !mlp) return nrc_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    return nrc_guarded(n, [&] {
        hip_check(hipSetDevice(n->scene->ctx->device), "hipSetDevice");
        hipStream_t st = n->scene->ctx->stream;
        hip_check(hipMemcpyAsync(mlp, n->mlp.ptr, sizeof(float) * vs::kNrcMlpParams, hipMemcpyDeviceToHost, st),
                  "mlp D2H");
        if (grid)
            hip_check(hipMemcpyAsync(grid, n->grid.ptr, sizeof(float) * vs::kNrcGridParams, hipMemcpyDeviceToHost, st),
                      "grid D2H");
        hip_check(hipStreamSynchronize(st), "sync");
    });
}

vs_status vs_nrc_set_params(vs_nrc *n, const float *mlp, const float *grid) {
    if (!n || !mlp) return nrc_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    return nrc_guarded(n, [&] {
        hip_check(hipSetDevice(n->scene->ctx->device), "hipSetDevice");
        hipStream_t st = n->scene->ctx->stream;
        upload_params(n, mlp, grid, st);
        hip_check(hipStreamSynchronize(st), "sync");
    });
}

vs_status vs_nrc_forward(vs_nrc *n, const float *pos, uint32_t count, float *out) {
    if (!n || (count && (!pos || !out))) return nrc_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    return nrc_guarded(n, [&] {
        if (!count) return;
        hip_check(hipSetDevice(n->scene->ctx->device), "hipSetDevice");
        hipStream_t st = n->scene->ctx->stream;
        if (n->io_a.bytes < size_t(count) * 12) n->io_a.alloc(size_t(count) * 12);
        if (n->io_b.bytes < size_t(count) * 12) n->io_b.alloc(size_t(count) * 12);
        hip_check(hipMemcpyAsync(n->io_a.ptr, pos, size_t(count) * 12, hipMemcpyHostToDevice, st), "pos H2D");
        hip_check(vs::nrc_forward(n->view, n->io_a.as<float>(), count, n->io_b.as<float>(), st), "forward");
        hip_check(hipMemcpyAsync(out, n->io_b.ptr, size_t(count) * 12, hipMemcpyDeviceToHost, st), "out D2H");
        hip_check(hipStreamSynchronize(st), "sync");
    });
}

vs_status vs_nrc_gradients(vs_nrc *n, const float *samples, uint32_t count, float *loss, float *mlp_grad,
                           float *grid_grad) {
    if (!n || !samples || !count || !loss || !mlp_grad) return nrc_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    return nrc_guarded(n, [&] {
        hip_check(hipSetDevice(n->scene->ctx->device), "hipSetDevice");
        hipStream_t st = n->scene->ctx->stream;
        if (n->io_a.bytes < size_t(count) * 24) n->io_a.alloc(size_t(count) * 24);
        hip_check(hipMemcpyAsync(n->io_a.ptr, samples, size_t(count) * 24, hipMemcpyHostToDevice, st), "H2D");
        hip_check(hipMemcpyAsync(n->counts.ptr, &count, 4, hipMemcpyHostToDevice, st), "count");
        hip_check(hipMemsetAsync(n->loss.ptr, 0, 4, st), "loss clear");
        hip_check(hipMemsetAsync(n->mlp_g.ptr, 0, sizeof(float) * vs::kNrcMlpParams, st), "clear");
        hip_check(hipMemsetAsync(n->grid_g.ptr, 0, sizeof(float) * vs::kNrcGridParams, st), "clear");
        hip_check(vs::nrc_backward(n->view, n->io_a.as<vs::NrcSample>(), n->counts.as<uint32_t>(), count, 0, count,
                                   n->loss.as<float>(), 0, st),
                  "backward");
        hip_check(hipMemcpyAsync(loss, n->loss.ptr, 4, hipMemcpyDeviceToHost, st), "loss D2H");
        hip_check(hipMemcpyAsync(mlp_grad, n->mlp_g.ptr, sizeof(float) * vs::kNrcMlpParams, hipMemcpyDeviceToHost, st),
                  "grad D2H");
        if (grid_grad)
            hip_check(hipMemcpyAsync(grid_grad, n->grid_g.ptr, sizeof(float) * vs::kNrcGridParams,
                                     hipMemcpyDeviceToHost, st),
                      "grad D2H");
        hip_check(hipMemsetAsync(n->mlp_g.ptr, 0, sizeof(float) * vs::kNrcMlpParams, st), "clear");
        hip_check(hipMemsetAsync(n->grid_g.ptr, 0, sizeof(float) * vs::kNrcGridParams, st), "clear");
        hip_check(hipStreamSynchronize(st), "sync");
    });
}

vs_status vs_nrc_train(vs_nrc *n, const float *samples, uint32_t count, uint32_t steps, float *last_loss) {
    if (!n || !samples || !count || !steps) return nrc_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    return nrc_guarded(n, [&] {
        hip_check(hipSetDevice(n->scene->ctx->device), "hipSetDevice");
        hipStream_t st = n->scene->ctx->stream;
        if (n->io_a.bytes < size_t(count) * 24) n->io_a.alloc(size_t(count) * 24);
        hip_check(hipMemcpyAsync(n->io_a.ptr, samples, size_t(count) * 24, hipMemcpyHostToDevice, st), "H2D");
        hip_check(hipMemcpyAsync(n->counts.ptr, &count, 4, hipMemcpyHostToDevice, st), "count");
        for (uint32_t i = 0; i < steps; ++i) {
            hip_check(hipMemsetAsync(n->loss.ptr, 0, 4, st), "loss clear");
            hip_check(vs::nrc_backward(n->view, n->io_a.as<vs::NrcSample>(), n->counts.as<uint32_t>(), count, 0,
                                       count, n->loss.as<float>(), 0, st),
                      "backward");
            adam_step(n, st);
        }
        float l = 0.f;
        hip_check(hipMemcpyAsync(&l, n->loss.ptr, 4, hipMemcpyDeviceToHost, st), "loss D2H");
        hip_check(hipStreamSynchronize(st), "sync");
        n->stats.last_loss = l;
        if (last_loss) *last_loss = l;
    });
}

vs_status vs_nrc_render(vs_nrc *n, uint32_t frames, uint32_t spp, uint32_t flags) {
    if (!n || !frames || !spp || spp > 1024) return nrc_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    vs_scene *s = n->scene;
    return nrc_guarded(n, [&] {
        require_built(s);
        if (!s->packed.rgb) throw std::invalid_argument("the radiance cache renders sRGB-mode scenes");
        if (s->host.integrator.type == "rt") throw std::invalid_argument("radiance cache renders path-tracing scenes");
        hip_check(hipSetDevice(s->ctx->device), "hipSetDevice");
        hipStream_t st = s->ctx->stream;
        const uint32_t P = s->cam.width * s->cam.height;
        const size_t slots = size_t(P) * spp;
        if (slots > 0xFFFFFFFFull) throw std::invalid_argument("too many samples per frame");
        if (n->L_cap < slots) {
            n->L.alloc(slots * sizeof(float4));
            n->queries.alloc(slots * sizeof(vs::NrcQuery));
            n->L_cap = n->q_cap = slots;
        }
        vs::RenderView rv = render_view(s);
        const bool terminate = (flags & VS_NRC_TERMINATE) != 0 && n->cfg.term_bounce > 0;
        const bool train = (flags & VS_NRC_TRAIN) != 0 && n->cfg.train_steps > 0;
        const uint32_t stride = train_stride(n, P);
        const uint32_t paths = stride ? (P + stride - 1) / stride : 0u;
        const uint32_t maxv = std::max<uint32_t>(rv.max_depth, 1u);
        if (train && (n->staging_paths < paths || n->sample_cap < paths * maxv)) {
            n->staging.alloc(size_t(paths) * maxv * 9 * sizeof(float));
            n->samples.alloc(size_t(paths) * maxv * sizeof(vs::NrcSample));
            n->staging_paths = paths;
            n->sample_cap = paths * maxv;
        }
        uint32_t *qcount = n->counts.as<uint32_t>(), *tcount = qcount + 1;
        float ms[6] = {0, 0, 0, 0, 0, 0};
        while (n->fev.size() < size_t(frames) * 6) {
            hipEvent_t e = nullptr;
            hip_check(hipEventCreate(&e), "event");
            n->fev.push_back(e);
        }
        for (uint32_t f = 0; f < frames; ++f) {
            hipEvent_t *ev = n->fev.data() + size_t(f) * 6;
            const uint32_t frame0 = s->frame_index;
            hip_check(vs::nrc_zero(qcount, 2, st), "zero");
            hip_check(hipEventRecord(ev[0], st), "event");
            hip_check(vs::nrc_render(s->view, s->cam, rv, frame0, spp, terminate ? n->cfg.term_bounce : 0u,
                                     n->L.as<float4>(), n->queries.as<vs::NrcQuery>(), qcount, uint32_t(slots),
                                     s->diffuse_only, st),
                      "nrc render");
            hip_check(hipEventRecord(ev[1], st), "event");
            if (train)
                hip_check(vs::nrc_train_paths(s->view, s->cam, rv, frame0 / spp, stride, n->staging.as<float>(),
                                              n->samples.as<vs::NrcSample>(), tcount, n->sample_cap, s->diffuse_only,
                                              st),
                          "train paths");
            hip_check(hipEventRecord(ev[2], st), "event");
            if (terminate)
                hip_check(vs::nrc_infer(n->view, n->queries.as<vs::NrcQuery>(), qcount, uint32_t(slots),
                                        n->L.as<float4>(), st),
                          "infer");
            hip_check(hipEventRecord(ev[3], st), "event");
            if (train) {
                // steps over consecutive batches of the frame's samples (all of
                // them in each step when train_batch is 0)
                const uint32_t batch = n->cfg.train_batch ? n->cfg.train_batch : n->sample_cap;
                for (uint32_t k = 0; k < n->cfg.train_steps; ++k) {
                    const uint32_t off = n->cfg.train_batch ? k * batch : 0u;// wrapped by the kernel
                    hip_check(hipMemsetAsync(n->loss.ptr, 0, 4, st), "loss clear");
                    hip_check(vs::nrc_backward(n->view, n->samples.as<vs::NrcSample>(), tcount, n->sample_cap, off,
                                               batch, n->loss.as<float>(), 0, st),
                              "backward");
                    adam_step(n, st);
                }
            }
            hip_check(hipEventRecord(ev[4], st), "event");
            hip_check(vs::nrc_accumulate(rv, n->L.as<float4>(), P, spp, frame0, s->film_rt.as<float4>(),
                                         s->film_out.as<float4>(), st),
                      "accumulate");
            hip_check(hipEventRecord(ev[5], st), "event");
            s->frame_index += spp;
        }
        if (frames) hip_check(hipEventSynchronize(n->fev[size_t(frames) * 6 - 1]), "frames");
        for (uint32_t f = 0; f < frames; ++f) {
            const hipEvent_t *ev = n->fev.data() + size_t(f) * 6;
            for (int k = 0; k < 5; ++k) {
                float t = 0.f;
                hip_check(hipEventElapsedTime(&t, ev[k], ev[k + 1]), "event time");
                ms[k] += t;
            }
        }
        uint32_t c[2];
        float l = 0.f;
        hip_check(hipMemcpy(c, qcount, 8, hipMemcpyDeviceToHost), "counts");
        hip_check(hipMemcpy(&l, n->loss.ptr, 4, hipMemcpyDeviceToHost), "loss");
        vs_nrc_stats &o = n->stats;
        o.frames += frames;
        o.adam_steps = n->adam_t;
        o.queries_last_frame = c[0];
        o.train_samples_last_frame = std::min(c[1], n->sample_cap);
        o.train_paths_per_frame = train ? paths : 0u;
        if (train) o.last_loss = l;
        o.render_ms = ms[0];
        o.train_paths_ms = ms[1];
        o.infer_ms = ms[2];
        o.train_ms = ms[3];
        o.accumulate_ms = ms[4];
    });
}

vs_status vs_nrc_stats_get(vs_nrc *n, vs_nrc_stats *out) {
    if (!n || !out) return nrc_fail(VS_ERROR_INVALID_ARGUMENT, "bad argument");
    *out = n->stats;
    return VS_OK;
}

}// extern "C"
