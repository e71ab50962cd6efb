// vs_trace_closest_noalloc: closest hits of a batch of host rays, staged
// through a fixed device array instead of the per-call device buffers that
// vs_trace_closest allocates.  For callers that must not allocate device
// memory (hipGraph capture of the surrounding work, serving loops, harnesses
// that forbid hipMalloc).
//
// The array `scratch` is a file-scope __device__ array: every GPU that loads
// this library holds its own copy until the process ends.  Host code never
// passes the array's name to a kernel or a copy (on the host it is only a
// stand-in); it uses the device address that hipGetSymbolAddress returns,
// resolved once per device with that device current (by vs_create, or by the
// first call on a device whose context was made before) and kept in
// g_scratch.
//
// Calls that overlap on one GPU (two host threads, or two contexts of the
// same device on different streams) race on the array: the library does not
// serialise them.  Each call writes every hit it returns, so nothing in the
// array needs clearing between calls.
#include <mutex>

#include "context.h"
#include "launch.h"

namespace {

using namespace vs_internal;

constexpr uint32_t kMaxRays = 1u << 18;
constexpr size_t kRayBytes = size_t(kMaxRays) * sizeof(vs::RayRec);  // 8 MiB
constexpr size_t kHitBytes = size_t(kMaxRays) * sizeof(vs::HitRec);  // 6 MiB
constexpr size_t kScratchBytes = kRayBytes + kHitBytes;
static_assert(kRayBytes % 256 == 0, "hits must start 256-byte aligned");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

void *scratch_for(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_scratch[device];
}

}// namespace

namespace vs_internal {

// Resolves and keeps this device's address of `scratch`; `device` must be
// current.  Idempotent.
void noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) throw DeviceError("device index beyond the staging table (64)");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return;
    void *p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "staging array address");
    if (!p) throw DeviceError("staging array address: null");
    g_scratch[device] = p;
}

}// namespace vs_internal

extern "C" {

vs_status vs_trace_closest_noalloc(vs_scene *s, const vs_ray *rays, uint32_t n, vs_hit *hits) {
    static_assert(sizeof(vs_ray) == sizeof(vs::RayRec), "ray layout");
    static_assert(sizeof(vs_hit) == sizeof(vs::HitRec), "hit layout");
    if (!s || (n && (!rays || !hits))) return fail(s ? s->ctx : nullptr, VS_ERROR_INVALID_ARGUMENT, "null argument");
    if (n > kMaxRays)
        return fail(s->ctx, VS_ERROR_INVALID_ARGUMENT,
                    "vs_trace_closest_noalloc takes at most 262144 rays per call (" + std::to_string(n) + " given)");
    return guarded(s->ctx, [&] {
        require_built(s);
        if (n == 0) return;
        const int dev = s->ctx->device;
        hip_check(hipSetDevice(dev), "hipSetDevice");
        void *base = scratch_for(dev);
        if (!base) {
            noalloc_setup(dev);
            base = scratch_for(dev);
        }
        auto *d_rays = static_cast<vs::RayRec *>(base);
        auto *d_hits = reinterpret_cast<vs::HitRec *>(static_cast<unsigned char *>(base) + kRayBytes);
        hipStream_t st = s->ctx->stream;
        hip_check(hipMemcpyAsync(d_rays, rays, size_t(n) * sizeof(vs::RayRec), hipMemcpyHostToDevice, st),
                  "rays H2D");
        hip_check(vs::launch_trace(s->view, d_rays, d_hits, nullptr, n, false, nullptr, st), "trace launch");
        hip_check(hipMemcpyAsync(hits, d_hits, size_t(n) * sizeof(vs::HitRec), hipMemcpyDeviceToHost, st),
                  "hits D2H");
        hip_check(hipStreamSynchronize(st), "trace sync");
    });
}

}// extern "C"
