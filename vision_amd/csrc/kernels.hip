// HIP kernels for gfx950: BVH8 trace entry points and the PT megakernel.
#include <hip/hip_runtime.h>

#include "dev_pt.h"
#include "launch.h"

namespace vs {

using namespace dev;

template <bool ANY>
__global__ void __launch_bounds__(256) trace_kernel(SceneView sc, const RayRec *__restrict__ rays,
                                                     HitRec *__restrict__ hits, uint8_t *__restrict__ occ,
                                                     uint32_t n, unsigned long long *__restrict__ counters) {
    extern __shared__ uint2 lds_stack[];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    RayRec r = rays[i];
    uint32_t nf = 0, tf = 0;
    TraceHit h = traverse<ANY>(sc, mk3(r.o[0], r.o[1], r.o[2]), mk3(r.d[0], r.d[1], r.d[2]), r.tmin, r.tmax,
                               lds_stack + threadIdx.x, blockDim.x, nf, tf);
    if (ANY) {
        occ[i] = h.tri != kInvalid ? 1 : 0;
    } else {
        HitRec out;
        if (h.tri == kInvalid) {
            out.inst = kInvalid;
            out.prim = kInvalid;
            out.bary[0] = out.bary[1] = 0.f;
            out.t = r.tmax;
            out.tri = kInvalid;
        } else {
            uint32_t inst = sc.tri_inst[h.tri];
            out.inst = inst;
            out.prim = h.tri - sc.inst[inst].tri_base;
            out.bary[0] = h.b1;
            out.bary[1] = h.b2;
            out.t = h.t;
            out.tri = h.tri;
        }
        hits[i] = out;
    }
    if (counters) {
        atomicAdd(counters + 0, 1ull);
        atomicAdd(counters + 1, (unsigned long long)nf);
        atomicAdd(counters + 2, (unsigned long long)tf);
    }
}

// one thread per pixel; all `spp` frames of the launch accumulate in registers
// with the film's running-mean arithmetic (normal.cpp:87-101), each frame's
// path starting as soon as the previous one ends (render_column).  NS: wavelengths
// carried per path; RGB: sRGB mode (three fixed channels) vs hero wavelengths.
// The kernel asks for a number of resident waves per SIMD that spills some
// state to scratch: traversal latency is better hidden by more waves than
// lost to the spills.  Measured on one MI355X: the diffuse family at six
// waves (80 VGPRs; 1.25-1.55x over the compiler's own choice on diffuse
// scenes, four waves 5 % slower); every material at four (128 VGPRs; 6 waves:
// 466 ms, 5: 441 ms, 4: 428 ms, 3: 473 ms for the procedural 1080p
// 64-spp hero-4 job).
// ADAPT: adaptive sampling on (rv.vstats holds the per-pixel statistics).
// COUNT: rays and BVH fetches tallied into counters.
template <int NS, bool RGB, int MS, bool ADAPT, bool COUNT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MS == MS_ALL ? 4 : 6))) pt_kernel(SceneView sc, CameraView cam, RenderView rv,
                                                  float4 *__restrict__ rt, float4 *__restrict__ out,
                                                  unsigned long long *__restrict__ counters) {
    extern __shared__ uint2 lds_stack[];
    uint32_t px = blockIdx.x * 8u + (threadIdx.x & 7u);
    uint32_t py = blockIdx.y * 8u + (threadIdx.x >> 3);
    if (px >= cam.width || py >= cam.height) return;
    PathCounters pc{0, 0, 0};
    render_column<NS, RGB, MS, ADAPT, COUNT>(sc, cam, rv, rt, rv.write_output ? out : nullptr, px, py, 1u, 1u,
                                             rv.frame_start, rv.spp, lds_stack + threadIdx.x, blockDim.x, pc);
    if (COUNT && counters) {
        atomicAdd(counters + 0, (unsigned long long)pc.rays);
        atomicAdd(counters + 1, (unsigned long long)pc.nodes);
        atomicAdd(counters + 2, (unsigned long long)pc.tris);
    }
}

// material-level parity hook: evaluate or sample material `mat` at a surface
// whose shading frame is the world axes and whose geometric normal is +z
template <int NS, bool RGB>
__global__ void __launch_bounds__(256) bsdf_kernel(SceneView sc, uint32_t mat, const float *__restrict__ wo,
                                                    const float *__restrict__ wi, const float *__restrict__ lambda,
                                                    uint32_t n, int sample, uint32_t frame, float *__restrict__ wi_out,
                                                    float *__restrict__ f, float *__restrict__ pdf) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Swl<NS> swl;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        swl.lambda[k] = RGB ? (k == 0 ? 602.785f : (k == 1 ? 539.285f : 445.772f)) : lambda[4 * i + k];
        swl.pdf[k] = RGB ? 1.f : visible_wavelength_pdf(swl.lambda[k]);
    }
    Frame fr{mk3(1.f, 0.f, 0.f), mk3(0.f, 1.f, 0.f), mk3(0.f, 0.f, 1.f)};
    bool dispersive;
    MatCtx<NS> c = make_material<NS, RGB>(sc, sc.mat[mat], fr, mk3(0.f, 0.f, 1.f),
                                          mk3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), swl, dispersive,
                                          mk2(0.f, 0.f));
    Spec<NS> fv;
    float p;
    if (sample) {
        Sampler sp;
        sp.start(0u, i, frame, 0u);
        SampleN<NS> s = mat_sample_any<NS, RGB>(sc, c, sp, swl);
        wi_out[3 * i] = s.wi.x;
        wi_out[3 * i + 1] = s.wi.y;
        wi_out[3 * i + 2] = s.wi.z;
        fv = s.e.f;
        p = s.e.pdf;
    } else {
        EvalN<NS> e = mat_eval_any<NS, RGB>(sc, c, mk3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]), nullptr, swl);
        fv = e.f;
        p = e.pdf;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) f[4 * i + k] = k < NS ? fv.v[k] : 0.f;
    pdf[i] = p;
}

hipError_t launch_bsdf(const SceneView &sc, uint32_t spectrum_dim, uint32_t mat, const float *wo, const float *wi,
                       const float *lambda, uint32_t n, int sample, uint32_t frame, float *wi_out, float *f,
                       float *pdf, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    dim3 block(256), grid((n + 255) / 256);
    if (sc.spectrum_rgb) {
        hipLaunchKernelGGL((bsdf_kernel<3, true>), grid, block, 0, stream, sc, mat, wo, wi, lambda, n, sample, frame,
                           wi_out, f, pdf);
    } else if (spectrum_dim == 4) {
        hipLaunchKernelGGL((bsdf_kernel<4, false>), grid, block, 0, stream, sc, mat, wo, wi, lambda, n, sample,
                           frame, wi_out, f, pdf);
    } else if (spectrum_dim == 3) {
        hipLaunchKernelGGL((bsdf_kernel<3, false>), grid, block, 0, stream, sc, mat, wo, wi, lambda, n, sample,
                           frame, wi_out, f, pdf);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_trace(const SceneView &sc, const RayRec *rays, HitRec *hits, uint8_t *occ, uint32_t n, bool any,
                        unsigned long long *counters, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t block = 128;
    size_t lds = size_t(sc.stack_size) * block * sizeof(uint2);
    dim3 grid((n + block - 1) / block);
    if (any) {
        hipLaunchKernelGGL(trace_kernel<true>, grid, dim3(block), lds, stream, sc, rays, hits, occ, n, counters);
    } else {
        hipLaunchKernelGGL(trace_kernel<false>, grid, dim3(block), lds, stream, sc, rays, hits, occ, n, counters);
    }
    return hipGetLastError();
}

hipError_t launch_pt(const SceneView &sc, const CameraView &cam, const RenderView &rv, float4 *rt, float4 *out,
                     unsigned long long *counters, bool diffuse_only, hipStream_t stream) {
    dim3 block(64);
    dim3 grid((cam.width + 7) / 8, (cam.height + 7) / 8);
    size_t lds = size_t(sc.stack_size) * 64 * sizeof(uint2);
#define VS_PT2(NS_, RGB_, AD_, CT_)                                                                             \
    do {                                                                                                        \
        if (diffuse_only)                                                                                       \
            hipLaunchKernelGGL((pt_kernel<NS_, RGB_, MS_DIFFUSE, AD_, CT_>), grid, block, lds, stream, sc, cam, rv, \
                               rt, out, counters);                                                              \
        else                                                                                                    \
            hipLaunchKernelGGL((pt_kernel<NS_, RGB_, MS_ALL, AD_, CT_>), grid, block, lds, stream, sc, cam, rv, rt, \
                               out, counters);                                                                  \
    } while (0)
#define VS_PT(NS_, RGB_)                                                                                        \
    do {                                                                                                        \
        if (rv.vstats) VS_PT2(NS_, RGB_, true, true);                                                           \
        else if (counters) VS_PT2(NS_, RGB_, false, true);                                                      \
        else VS_PT2(NS_, RGB_, false, false);                                                                   \
    } while (0)
    if (sc.spectrum_rgb) {
        VS_PT(3, true);
    } else if (rv.spectrum_dim == 4) {
        VS_PT(4, false);
    } else if (rv.spectrum_dim == 3) {
        VS_PT(3, false);
    } else {
        return hipErrorInvalidValue;
    }
#undef VS_PT
#undef VS_PT2
    return hipGetLastError();
}

}// namespace vs
