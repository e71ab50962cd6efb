// Wavefront path tracer for gfx950: the bounce loop of the megakernel split at
// its traversals.  Per bounce: closest-hit over the live-path queue (lean
// kernel, high occupancy), shading (path_shade: emission, NEE set-up, BSDF
// sample, RR; appends live paths and shadow rays, one atomic per block and
// buffer), any-hit over the shadow rays.  Paths run the same path_shade
// arithmetic as the megakernel, so both forms give the same film.
#include <hip/hip_runtime.h>

#include "dev_pt.h"
#include "wavefront.h"

namespace vs {

using namespace dev;

namespace {

template <int NS, bool RGB>
__device__ __forceinline__ void load_state(const WavefrontView &w, uint32_t i, PathState<NS> &ps) {
    float4 a = w.o[i], b = w.d[i], t = w.thr[i], l = w.L[i];
    uint4 m = w.meta[i];
    ps.o = mk3(a.x, a.y, a.z);
    ps.scatter_pdf = a.w;
    ps.d = mk3(b.x, b.y, b.z);
    ps.eta_scale = b.w;
    const float tv[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
    for (int k = 0; k < NS; ++k) ps.thr.v[k] = tv[k];
    ps.L = mk3(l.x, l.y, l.z);
    ps.fw = l.w;
    ps.sampler.state = m.x;
    ps.bounce = m.y & 0xFFFFu;
    ps.skips = m.y >> 16;
    ps.alive = m.z;
    ps.tail = m.w;
    if (RGB) {
        ps.swl = sample_wavelengths<NS, true>(0u, 0u, 0u);
    } else {
        float4 lm = w.lam[i], lp = w.lpdf[i];
        const float lv[4] = {lm.x, lm.y, lm.z, lm.w}, pv[4] = {lp.x, lp.y, lp.z, lp.w};
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            ps.swl.lambda[k] = lv[k];
            ps.swl.pdf[k] = pv[k];
        }
    }
}

template <int NS, bool RGB>
__device__ __forceinline__ void store_state(const WavefrontView &w, uint32_t i, const PathState<NS> &ps) {
    w.o[i] = make_float4(ps.o.x, ps.o.y, ps.o.z, ps.scatter_pdf);
    w.d[i] = make_float4(ps.d.x, ps.d.y, ps.d.z, ps.eta_scale);
    float tv[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < NS; ++k) tv[k] = ps.thr.v[k];
    w.thr[i] = make_float4(tv[0], tv[1], tv[2], tv[3]);
    w.L[i] = make_float4(ps.L.x, ps.L.y, ps.L.z, ps.fw);
    w.meta[i] = make_uint4(ps.sampler.state, ps.bounce | (ps.skips << 16), ps.alive, ps.tail);
    if (!RGB) {
        float lv[4] = {0.f, 0.f, 0.f, 0.f}, pv[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            lv[k] = ps.swl.lambda[k];
            pv[k] = ps.swl.pdf[k];
        }
        w.lam[i] = make_float4(lv[0], lv[1], lv[2], lv[3]);
        w.lpdf[i] = make_float4(pv[0], pv[1], pv[2], pv[3]);
    }
}

template <int NS, bool RGB>
__global__ void __launch_bounds__(256) begin_kernel(CameraView cam, RenderView rv, WavefrontView w, uint32_t frame0,
                                                    uint32_t n, uint32_t *__restrict__ queue) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t p = i % w.pixels, s = i / w.pixels;
    PathState<NS> ps;
    path_begin<NS, RGB>(cam, rv, p % w.width, p / w.width, frame0 + s, ps);
    store_state<NS, RGB>(w, i, ps);
    queue[i] = i;
}

__global__ void __launch_bounds__(256) trace_queue_kernel(SceneView sc, WavefrontView w,
                                                          const uint32_t *__restrict__ queue, uint32_t count,
                                                          unsigned long long *__restrict__ counters) {
    extern __shared__ uint2 lds_stack[];
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    uint32_t i = queue[k];
    float4 a = w.o[i], b = w.d[i];
    uint32_t nf = 0, tf = 0;
    TraceHit h = traverse<false>(sc, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), 0.f, 3.402823466e38f,
                                 lds_stack + threadIdx.x, blockDim.x, nf, tf);
    w.hit[i] = make_float4(h.t, __uint_as_float(h.tri), h.b1, h.b2);
    if (counters) {
        atomicAdd(counters + 0, 1ull);
        atomicAdd(counters + 1, (unsigned long long)nf);
        atomicAdd(counters + 2, (unsigned long long)tf);
    }
}

// Three resident waves per SIMD (168 VGPRs, 250-420 B of scratch per lane):
// left free the compiler takes 256 VGPRs and one wave.  Measured on the
// procedural 1080p 64-spp hero-4 job: 2.64 ms per launch free, 2.06 ms at
// three waves, 2.21 ms at four, 2.61 ms at five.
// The survivors and shadow rays of a block are appended together: each wave
// reserves its lanes' share in LDS, the block one range per buffer with one
// global atomic.  Per-wave atomics on the two shared counters serialise in
// L2: with them (64-thread blocks) the stage took 130.8 ms per procedural
// job, with block ranges 106.1 ms.
// Before shading, a block sorts its paths by what they hit (the miss branch,
// a material-less surface, or the material's kind) in LDS, so that a wave
// mostly runs one branch of path_shade; the block's paths stay the same 256
// queue entries, so their state loads stay within the same cache lines.
constexpr uint32_t kShadeBlock = 256;
constexpr uint32_t kShadeKeys = 16;// the last key: no path (past the queue's end)

__device__ __forceinline__ uint32_t shade_key(const SceneView &sc, uint32_t tri) {
    if (tri == kInvalid) return 0u;
    const uint32_t m = sc.inst[sc.tri_inst[tri]].mat;
    if (m == kInvalid) return 1u;
    return 2u + min(sc.mat[m].type, kShadeKeys - 3u);
}

// register budget of the shading kernel, in waves per SIMD: three for every
// material (measured on the procedural job: 2 -> 97.1, 3 -> 93.5, 4 -> 105.0
// ms of shading per 64-frame job), more for the diffuse family, whose
// smaller code needs fewer registers
constexpr int shade_waves(int ms) { return ms == MS_ALL ? 3 : 4; }

template <int NS, bool RGB, int MS>
__global__ void __launch_bounds__(kShadeBlock) __attribute__((amdgpu_waves_per_eu(shade_waves(MS)))) shade_kernel(SceneView sc, RenderView rv, WavefrontView w,
                                                   const uint32_t *__restrict__ qin, uint32_t count,
                                                   uint32_t *__restrict__ qout, uint32_t *__restrict__ counts) {
    __shared__ uint32_t s_n[2], s_base[2], s_key[kShadeKeys];
    __shared__ uint32_t s_path[kShadeBlock];
    __shared__ float4 s_hit[kShadeBlock];
    if (threadIdx.x < 2) s_n[threadIdx.x] = 0u;
    if (threadIdx.x < kShadeKeys) s_key[threadIdx.x] = 0u;
    __syncthreads();
    {
        const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
        uint32_t i0 = 0u, key = kShadeKeys - 1u;
        float4 hv0 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (k < count) {
            i0 = qin[k];
            hv0 = w.hit[i0];
            key = shade_key(sc, __float_as_uint(hv0.y));
        }
        const uint32_t rank = atomicAdd(&s_key[key], 1u);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t sum = 0u;
            for (uint32_t j = 0; j < kShadeKeys; ++j) {
                const uint32_t c = s_key[j];
                s_key[j] = sum;
                sum += c;
            }
        }
        __syncthreads();
        const uint32_t pos = s_key[key] + rank;
        s_path[pos] = key == kShadeKeys - 1u ? kInvalid : i0;
        s_hit[pos] = hv0;
        __syncthreads();
    }
    const uint32_t i = s_path[threadIdx.x];
    bool live = false;
    ShadowRay sr;
    sr.valid = 0u;
    if (i != kInvalid) {
        PathState<NS> ps;
        load_state<NS, RGB>(w, i, ps);
        float4 hv = s_hit[threadIdx.x];
        TraceHit h;
        h.t = hv.x;
        h.tri = __float_as_uint(hv.y);
        h.b1 = hv.z;
        h.b2 = hv.w;
        path_shade<NS, RGB, MS>(sc, rv, ps, h, sr);
        store_state<NS, RGB>(w, i, ps);
        live = ps.alive != 0u;
    }
    const bool shadow = sr.valid != 0u;
    const unsigned long long below = (1ull << __lane_id()) - 1ull;
    const unsigned long long ml = __ballot(live), ms = __ballot(shadow);
    uint32_t wl = 0u, wsh = 0u;
    if (__lane_id() == 0u) {
        if (ml) wl = atomicAdd(&s_n[0], uint32_t(__popcll(ml)));
        if (ms) wsh = atomicAdd(&s_n[1], uint32_t(__popcll(ms)));
    }
    wl = __shfl(wl, 0);
    wsh = __shfl(wsh, 0);
    __syncthreads();
    if (threadIdx.x == 0) {
        s_base[0] = s_n[0] ? atomicAdd(counts + WF_LIVE, s_n[0]) : 0u;
        s_base[1] = s_n[1] ? atomicAdd(counts + WF_SHADOW, s_n[1]) : 0u;
    }
    __syncthreads();
    if (shadow) {
        const uint32_t slot = s_base[1] + wsh + uint32_t(__popcll(ms & below));
        w.sh_o[slot] = make_float4(sr.o.x, sr.o.y, sr.o.z, __uint_as_float(i));
        w.sh_d[slot] = make_float4(sr.d.x, sr.d.y, sr.d.z, 0.f);
        w.sh_c[slot] = make_float4(sr.c.x, sr.c.y, sr.c.z, 0.f);
    }
    if (live) qout[s_base[0] + wl + uint32_t(__popcll(ml & below))] = i;
}

__global__ void __launch_bounds__(256) shadow_kernel(SceneView sc, WavefrontView w,
                                                     const uint32_t *__restrict__ counts, uint32_t max_count,
                                                     unsigned long long *__restrict__ counters) {
    extern __shared__ uint2 lds_stack[];
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= max_count || k >= counts[WF_SHADOW]) return;
    float4 a = w.sh_o[k], b = w.sh_d[k];
    uint32_t nf = 0, tf = 0;
    TraceHit h = traverse<true>(sc, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), 0.f, 1.f - kShadowEpsilon,
                                lds_stack + threadIdx.x, blockDim.x, nf, tf);
    if (h.tri == kInvalid) {
        uint32_t i = __float_as_uint(a.w);
        float4 c = w.sh_c[k], l = w.L[i];
        f3 L = mk3(l.x, l.y, l.z);
        L += mk3(c.x, c.y, c.z);
        w.L[i] = make_float4(L.x, L.y, L.z, l.w);
    }
    if (counters) {
        atomicAdd(counters + 0, 1ull);
        atomicAdd(counters + 1, (unsigned long long)nf);
        atomicAdd(counters + 2, (unsigned long long)tf);
    }
}

__global__ void zero_kernel(uint32_t *counts) {
    if (threadIdx.x < 2) counts[threadIdx.x] = 0u;
}

// fold the chunk's samples into the film in frame order, as the megakernel's
// per-pixel loop does
__global__ void __launch_bounds__(256) accumulate_kernel(RenderView rv, WavefrontView w, uint32_t frame0,
                                                         uint32_t frames, float4 *__restrict__ rt,
                                                         float4 *__restrict__ out) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= w.pixels) return;
    float4 acc = rt[p];
    for (uint32_t s = 0; s < frames; ++s) {
        float4 l = w.L[size_t(s) * w.pixels + p];
        acc = film_add(rv, acc, mk3(l.x, l.y, l.z) * l.w, frame0 + s);
    }
    rt[p] = acc;
    if (rv.write_output) out[p] = film_output(rv, acc);
}

uint32_t blocks(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

}// namespace

hipError_t wf_begin(const CameraView &cam, const RenderView &rv, const WavefrontView &w, uint32_t spectrum_dim,
                    bool rgb, uint32_t frame0, uint32_t n, uint32_t *queue, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    dim3 g(blocks(n, 256)), b(256);
    if (rgb) hipLaunchKernelGGL((begin_kernel<3, true>), g, b, 0, stream, cam, rv, w, frame0, n, queue);
    else if (spectrum_dim == 4) hipLaunchKernelGGL((begin_kernel<4, false>), g, b, 0, stream, cam, rv, w, frame0, n, queue);
    else if (spectrum_dim == 3) hipLaunchKernelGGL((begin_kernel<3, false>), g, b, 0, stream, cam, rv, w, frame0, n, queue);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t wf_trace(const SceneView &sc, const WavefrontView &w, const uint32_t *queue, uint32_t count,
                    unsigned long long *counters, uint32_t block, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    if (block != 64 && block != 128 && block != 256) return hipErrorInvalidValue;
    size_t lds = size_t(sc.stack_size) * block * sizeof(uint2);
    hipLaunchKernelGGL(trace_queue_kernel, dim3(blocks(count, block)), dim3(block), lds, stream, sc, w, queue, count,
                       counters);
    return hipGetLastError();
}

namespace {
template <int MS>
hipError_t shade_ms(const SceneView &sc, const RenderView &rv, const WavefrontView &w, uint32_t spectrum_dim,
                    bool rgb, const uint32_t *qin, uint32_t count, uint32_t *qout, uint32_t *counts,
                    hipStream_t stream) {
    dim3 g(blocks(count, kShadeBlock)), b(kShadeBlock);
    if (rgb) hipLaunchKernelGGL((shade_kernel<3, true, MS>), g, b, 0, stream, sc, rv, w, qin, count, qout, counts);
    else if (spectrum_dim == 4)
        hipLaunchKernelGGL((shade_kernel<4, false, MS>), g, b, 0, stream, sc, rv, w, qin, count, qout, counts);
    else if (spectrum_dim == 3)
        hipLaunchKernelGGL((shade_kernel<3, false, MS>), g, b, 0, stream, sc, rv, w, qin, count, qout, counts);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}
}// namespace

hipError_t wf_shade(const SceneView &sc, const RenderView &rv, const WavefrontView &w, uint32_t spectrum_dim,
                    bool rgb, int material_set, const uint32_t *qin, uint32_t count, uint32_t *qout,
                    uint32_t *counts, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    if (material_set == MS_DIFFUSE)
        return shade_ms<MS_DIFFUSE>(sc, rv, w, spectrum_dim, rgb, qin, count, qout, counts, stream);
    if (material_set != MS_ALL) return hipErrorInvalidValue;
    return shade_ms<MS_ALL>(sc, rv, w, spectrum_dim, rgb, qin, count, qout, counts, stream);
}

hipError_t wf_shadow(const SceneView &sc, const WavefrontView &w, const uint32_t *counts, uint32_t max_count,
                     unsigned long long *counters, uint32_t block, hipStream_t stream) {
    if (max_count == 0) return hipSuccess;
    if (block != 64 && block != 128 && block != 256) return hipErrorInvalidValue;
    size_t lds = size_t(sc.stack_size) * block * sizeof(uint2);
    hipLaunchKernelGGL(shadow_kernel, dim3(blocks(max_count, block)), dim3(block), lds, stream, sc, w, counts,
                       max_count, counters);
    return hipGetLastError();
}

hipError_t wf_zero(uint32_t *counts, hipStream_t stream) {
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, stream, counts);
    return hipGetLastError();
}

hipError_t wf_accumulate(const RenderView &rv, const WavefrontView &w, uint32_t frame0, uint32_t frames, float4 *rt,
                         float4 *out, hipStream_t stream) {
    hipLaunchKernelGGL(accumulate_kernel, dim3(blocks(w.pixels, 256)), dim3(256), 0, stream, rv, w, frame0, frames,
                       rt, out);
    return hipGetLastError();
}

}// namespace vs
