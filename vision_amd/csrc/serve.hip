// Persistent path-tracing kernels for batched serving (gfx950).
//
// One workgroup is one 64-lane wave.  Lane 0 claims the next work item with
// an atomic on a device counter and broadcasts it; the item's job, scene and
// camera are read through wave-uniform addresses, so they stay in scalar
// registers.  Each lane owns one pixel of the item's tile and runs the
// step's frames of that pixel in order, with the film's running mean in
// registers (the arithmetic of the single-job megakernel, kernels.hip), so
// a job's film does not depend on what else shares the launch.
#include <hip/hip_runtime.h>

#include "dev_pt.h"
#include "serve.h"

namespace vs {

using namespace dev;

// OCC > 0 asks the compiler for a number of resident waves per SIMD (a
// register budget of 512 / waves, spilling the rest): OCC for the diffuse
// family, four for every material (measured on one MI355X, mixed-material
// serving: 4 waves 615 against 6 waves 602 Msamples/s); 0 leaves it free
// COUNT: rays and BVH fetches tallied into counters (the counting runs)
template <int NS, bool RGB, int MS, int OCC, bool COUNT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(OCC > 0 ? (MS == MS_ALL ? 4 : OCC) : 1))) serve_kernel(const SceneView *__restrict__ scenes,
                                                     const ServeJob *__restrict__ jobs,
                                                     const StepTable *__restrict__ table, uint32_t *__restrict__ next,
                                                     uint32_t tile_w, unsigned long long *__restrict__ counters) {
    extern __shared__ uint2 lds_stack[];
    const uint32_t n_items = table->num_items;
    const uint32_t n_entries = table->num_entries;
    const uint32_t tile_h = 64u / tile_w;
    const uint32_t lx = threadIdx.x % tile_w, ly = threadIdx.x / tile_w;
    PathCounters pc{0, 0, 0};
    while (true) {
        uint32_t k = 0;
        if (threadIdx.x == 0) k = atomicAdd(next, 1u);
        k = __builtin_amdgcn_readfirstlane(__shfl(k, 0));
        if (k >= n_items) break;
        // the entry holding item k: the last one whose first_item <= k
        uint32_t lo = 0, hi = n_entries;
        while (hi - lo > 1u) {
            uint32_t mid = (lo + hi) >> 1;
            if (table->e[mid].first_item <= k) {
                lo = mid;
            } else {
                hi = mid;
            }
        }
        const StepEntry e = table->e[lo];
        const ServeJob &job = jobs[e.job];
        const SceneView &sc = scenes[job.scene];
        const uint32_t t = k - e.first_item;
        // one pixel of the item's tile per lane, its frames back to back
        const uint32_t px = (t % job.tiles_x) * tile_w + lx;
        const uint32_t py = (t / job.tiles_x) * tile_h + ly;
        render_column<NS, RGB, MS, false, COUNT>(sc, job.cam, job.rv, job.rt, nullptr, px, py, 1u, 1u, e.frame_start,
                                                 e.frames, lds_stack + threadIdx.x, 64u, pc);
    }
    if (COUNT && counters) {
        atomicAdd(counters + 0, (unsigned long long)pc.rays);
        atomicAdd(counters + 1, (unsigned long long)pc.nodes);
        atomicAdd(counters + 2, (unsigned long long)pc.tris);
    }
}

__global__ void __launch_bounds__(64) serve_reset_kernel(uint32_t *next, uint32_t n) {
    if (threadIdx.x < n) next[threadIdx.x] = 0u;
}

__global__ void __launch_bounds__(256) serve_clear_kernel(float4 *film, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) film[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void __launch_bounds__(256) serve_output_kernel(RenderView rv, const float4 *__restrict__ rt,
                                                            float4 *__restrict__ out, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = film_output(rv, rt[i]);
}

namespace {

using ServeFn = void (*)(const SceneView *, const ServeJob *, const StepTable *, uint32_t *, uint32_t,
                         unsigned long long *);

template <int OCC, bool COUNT>
ServeFn serve_fn_occ(uint32_t variant) {
    switch (variant) {
    case SV_RGB: return serve_kernel<3, true, MS_ALL, OCC, COUNT>;
    case SV_RGB_DIFFUSE: return serve_kernel<3, true, MS_DIFFUSE, OCC, COUNT>;
    case SV_HERO4: return serve_kernel<4, false, MS_ALL, OCC, COUNT>;
    case SV_HERO4_DIFFUSE: return serve_kernel<4, false, MS_DIFFUSE, OCC, COUNT>;
    case SV_HERO3: return serve_kernel<3, false, MS_ALL, OCC, COUNT>;
    case SV_HERO3_DIFFUSE: return serve_kernel<3, false, MS_DIFFUSE, OCC, COUNT>;
    default: return nullptr;
    }
}

// count: the counting form (occupancy queries use the other, the one steps
// run when not counting)
ServeFn serve_fn(uint32_t variant, uint32_t occ, bool count = false) {
    if (occ == 0) return count ? serve_fn_occ<0, true>(variant) : serve_fn_occ<0, false>(variant);
    if (occ == 6) return count ? serve_fn_occ<6, true>(variant) : serve_fn_occ<6, false>(variant);
    return nullptr;
}

}// namespace

hipError_t serve_launch(uint32_t variant, uint32_t occ, const SceneView *scenes, const ServeJob *jobs,
                        const StepTable *table, uint32_t *next, uint32_t blocks, uint32_t tile_w,
                        uint32_t stack_entries, unsigned long long *counters, hipStream_t stream) {
    ServeFn fn = serve_fn(variant, occ, counters != nullptr);
    if (!fn || blocks == 0 || !(tile_w == 8 || tile_w == 16 || tile_w == 32 || tile_w == 64) || stack_entries == 0 ||
        stack_entries > 64)
        return hipErrorInvalidValue;
    size_t lds = size_t(stack_entries) * 64 * sizeof(uint2);
    hipLaunchKernelGGL(fn, dim3(blocks), dim3(64), lds, stream, scenes, jobs, table, next, tile_w, counters);
    return hipGetLastError();
}

hipError_t serve_reset(uint32_t *next, uint32_t n, hipStream_t stream) {
    if (n > 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(serve_reset_kernel, dim3(1), dim3(64), 0, stream, next, n);
    return hipGetLastError();
}

hipError_t serve_clear(float4 *film, uint32_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(serve_clear_kernel, dim3((n + 255u) / 256u), dim3(256), 0, stream, film, n);
    return hipGetLastError();
}

hipError_t serve_output(const RenderView &rv, const float4 *rt, float4 *out, uint32_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(serve_output_kernel, dim3((n + 255u) / 256u), dim3(256), 0, stream, rv, rt, out, n);
    return hipGetLastError();
}

int serve_max_blocks_per_cu(uint32_t variant, uint32_t occ, uint32_t stack_entries) {
    ServeFn fn = serve_fn(variant, occ);
    if (!fn) return -1;
    int n = 0;
    size_t lds = size_t(stack_entries) * 64 * sizeof(uint2);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void *>(fn), 64, lds) != hipSuccess)
        return -1;
    return n;
}

int serve_kernel_vgprs(uint32_t variant, uint32_t occ) {
    ServeFn fn = serve_fn(variant, occ);
    if (!fn) return -1;
    hipFuncAttributes a{};
    if (hipFuncGetAttributes(&a, reinterpret_cast<const void *>(fn)) != hipSuccess) return -1;
    return a.numRegs;
}

}// namespace vs
