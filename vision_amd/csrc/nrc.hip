// Radiance-cache kernels for gfx950: rendering with early termination into
// the cache, training paths, the MLP forward (inference) and backward on
// BF16 MFMA tiles, Adam.
#include <hip/hip_runtime.h>

#include "dev_nrc.h"
#include "dev_pt.h"
#include "nrc.h"

#include <algorithm>

namespace vs {

using namespace dev;

namespace {

constexpr int kBlock = 256;// 4 waves
constexpr uint32_t kTrainStream = 1u << 30;

// per-wave LDS of the forward-only kernels: one 16-row tile at a time
// (features, two hidden buffers used in turn, outputs), so that twelve
// waves per CU fit beside the block's weights
struct NrcFwdLds {
    uint16_t x[16 * kXs];
    uint16_t h[2][16 * kHs];
    float y[16 * 4];
};

// forward of one row tile (the arithmetic of nrc_forward_tile with the
// hidden buffers reused: H3 overwrites H1 once layer 2 has read it)
__device__ __forceinline__ void fwd_tile(const NrcWeightsLds &w, NrcFwdLds &s, int lane) {
    nrc_fwd_tile<kNrcIn, kNrcHidden>(s.x, kXs, w.w1t, kXs, s.h[0], lane);
    __builtin_amdgcn_wave_barrier();
    nrc_fwd_tile<kNrcHidden, kNrcHidden>(s.h[0], kHs, w.w2t, kHs, s.h[1], lane);
    __builtin_amdgcn_wave_barrier();
    nrc_fwd_tile<kNrcHidden, kNrcHidden>(s.h[1], kHs, w.w3t, kHs, s.h[0], lane);
    __builtin_amdgcn_wave_barrier();
    if (lane < 48) {
        int row = lane / 3, c = lane % 3;
        float acc = 0.f;
        for (int k = 0; k < kNrcHidden; ++k) acc += bf2f(s.h[0][row * kHs + k]) * w.w4[k * kNrcOut + c];
        s.y[row * 4 + c] = acc;
    }
    __builtin_amdgcn_wave_barrier();
}

// INFER: rows are queries (count on the device), their radiance goes into
// L; otherwise rows are n positions and outputs go to out[n][3]
template <bool INFER>
__global__ void __launch_bounds__(kBlock) nrc_forward_kernel(NrcView v, const float *__restrict__ pos, uint32_t n_pos,
                                                             const NrcQuery *__restrict__ q,
                                                             const uint32_t *__restrict__ count, uint32_t cap,
                                                             float *__restrict__ out, float4 *__restrict__ L) {
    __shared__ NrcWeightsLds w;
    __shared__ NrcFwdLds ws[kBlock / 64];
    nrc_load_weights(v, w, threadIdx.x, kBlock);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    NrcFwdLds &s = ws[wave];
    const uint32_t n = INFER ? min(*count, cap) : n_pos;
    for (uint32_t base = (blockIdx.x * (kBlock / 64) + wave) * 64u; base < n; base += gridDim.x * kBlock) {
        const uint32_t row = base + lane;
        const bool valid = row < n;
        float p[3] = {0.f, 0.f, 0.f};
        if (valid) {
            const float *src = INFER ? q[row].p : pos + 3 * size_t(row);
            p[0] = src[0];
            p[1] = src[1];
            p[2] = src[2];
        }
        uint16_t x[kNrcIn];
        nrc_encode(v, p, x);
        for (int rt = 0; rt < 4; ++rt) {
            if ((lane >> 4) == rt) {
#pragma unroll
                for (int k = 0; k < kNrcIn; k += 4)
                    *reinterpret_cast<bf16x4 *>(s.x + (lane & 15) * kXs + k) =
                        bf16x4{short(x[k]), short(x[k + 1]), short(x[k + 2]), short(x[k + 3])};
            }
            __builtin_amdgcn_wave_barrier();
            fwd_tile(w, s, lane);
            if ((lane >> 4) == rt && valid) {
                const int i = lane & 15;
                float y0 = s.y[i * 4], y1 = s.y[i * 4 + 1], y2 = s.y[i * 4 + 2];
                if (INFER) {
                    const NrcQuery &qq = q[row];
                    float4 l = L[qq.slot];
                    l.x += qq.beta[0] * fmaxf(y0, 0.f);
                    l.y += qq.beta[1] * fmaxf(y1, 0.f);
                    l.z += qq.beta[2] * fmaxf(y2, 0.f);
                    L[qq.slot] = l;
                } else {
                    out[3 * size_t(row)] = y0;
                    out[3 * size_t(row) + 1] = y1;
                    out[3 * size_t(row) + 2] = y2;
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// forward + backward over the training samples; one block per CU loops
// over chunks of 64 samples per wave, weight gradients gather in LDS and
// reach global memory once per block
__global__ void __launch_bounds__(kBlock) nrc_backward_kernel(NrcView v, const NrcSample *__restrict__ samples,
                                                              const uint32_t *__restrict__ count, uint32_t cap,
                                                              uint32_t offset, uint32_t batch,
                                                              float *__restrict__ loss) {
    __shared__ NrcWeightsLds w;
    __shared__ NrcGradLds gl;
    __shared__ NrcWaveLds ws[kBlock / 64];
    nrc_load_weights(v, w, threadIdx.x, kBlock);
    for (int i = threadIdx.x; i < kNrcIn * kNrcHidden; i += kBlock) gl.w1[i] = 0.f;
    for (int i = threadIdx.x; i < kNrcHidden * kNrcHidden; i += kBlock) {
        gl.w2[i] = 0.f;
        gl.w3[i] = 0.f;
    }
    for (int i = threadIdx.x; i < kNrcHidden * kNrcOut; i += kBlock) gl.w4[i] = 0.f;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    NrcWaveLds &s = ws[wave];
    // this step's samples: [offset, offset + batch) of the min(*count, cap)
    // present, the offset taken modulo that count
    const uint32_t total = min(*count, cap);
    if (total) offset %= total;
    const uint32_t n = min(total - offset, batch);
    samples += offset;
    const float inv = n ? 1.f / (3.f * float(n)) : 0.f;
    float loss_acc = 0.f;
    for (uint32_t base = (blockIdx.x * (kBlock / 64) + wave) * 64u; base < n; base += gridDim.x * kBlock) {
        const uint32_t row = base + lane;
        const bool valid = row < n;
        float p[3] = {0.f, 0.f, 0.f}, t[3] = {0.f, 0.f, 0.f};
        if (valid) {
            const NrcSample &sm = samples[row];
            p[0] = sm.p[0];
            p[1] = sm.p[1];
            p[2] = sm.p[2];
            t[0] = sm.t[0];
            t[1] = sm.t[1];
            t[2] = sm.t[2];
        }
        uint16_t x[kNrcIn];
        nrc_encode(v, p, x);
#pragma unroll
        for (int k = 0; k < kNrcIn; k += 4)
            *reinterpret_cast<bf16x4 *>(s.x + lane * kXs + k) = bf16x4{short(x[k]), short(x[k + 1]),
                                                                         short(x[k + 2]), short(x[k + 3])};
        __builtin_amdgcn_wave_barrier();
        for (int rt = 0; rt < 4; ++rt) {
            nrc_forward_tile(w, s, rt, lane);
            if ((lane >> 4) == rt) {
                const int i = lane & 15;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    float y = s.y[i * 4 + c];
                    float d = y - t[c];
                    float den = y * y + 0.01f;
                    loss_acc += valid ? d * d / den * inv : 0.f;
                    s.y[i * 4 + c] = valid ? 2.f * d / den * inv : 0.f;
                }
            }
            __builtin_amdgcn_wave_barrier();
            nrc_backward_tile(w, s, gl, rt, lane);
            if ((lane >> 4) == rt && valid) nrc_encode_backward(v, p, s.dx + (lane & 15) * (kNrcIn + 1));
            __builtin_amdgcn_wave_barrier();
        }
    }
    // loss: wave sum, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) loss_acc += __shfl_xor(loss_acc, off);
    if (lane == 0 && loss_acc != 0.f) atomicAdd(loss, loss_acc);
    __syncthreads();
    for (int i = threadIdx.x; i < kNrcIn * kNrcHidden; i += kBlock)
        if (gl.w1[i] != 0.f) atomicAdd(v.mlp_g + kNrcW1 + i, gl.w1[i]);
    for (int i = threadIdx.x; i < kNrcHidden * kNrcHidden; i += kBlock) {
        if (gl.w2[i] != 0.f) atomicAdd(v.mlp_g + kNrcW2 + i, gl.w2[i]);
        if (gl.w3[i] != 0.f) atomicAdd(v.mlp_g + kNrcW3 + i, gl.w3[i]);
    }
    for (int i = threadIdx.x; i < kNrcHidden * kNrcOut; i += kBlock)
        if (gl.w4[i] != 0.f) atomicAdd(v.mlp_g + kNrcW4 + i, gl.w4[i]);
}

__device__ __forceinline__ float adam_update(float &p, float &m, float &vv, float &g, float lr, float c1, float c2) {
    const float b1 = 0.9f, b2 = 0.99f, eps = 1e-15f;
    float gr = g;
    m = b1 * m + (1.f - b1) * gr;
    vv = b2 * vv + (1.f - b2) * gr * gr;
    p -= lr * (m * c1) / (sqrtf(vv * c2) + eps);
    g = 0.f;
    return p;
}

__global__ void __launch_bounds__(256) nrc_adam_kernel(NrcView v, float lr, float c1, float c2) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < kNrcGridParams + kNrcMlpParams; i += stride) {
        if (i < kNrcGridParams) {
            float p = adam_update(v.grid[i], v.grid_m[i], v.grid_v[i], v.grid_g[i], lr, c1, c2);
            v.grid_f16[i] = f2h(p);
        } else {
            size_t j = i - kNrcGridParams;
            float p = adam_update(v.mlp[j], v.mlp_m[j], v.mlp_v[j], v.mlp_g[j], lr, c1, c2);
            v.mlp_bf16[j] = f2bf(p);
        }
    }
}

__global__ void __launch_bounds__(256) nrc_sync_kernel(NrcView v) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < kNrcGridParams + kNrcMlpParams; i += stride) {
        if (i < kNrcGridParams) {
            v.grid_f16[i] = f2h(v.grid[i]);
        } else {
            size_t j = i - kNrcGridParams;
            v.mlp_bf16[j] = f2bf(v.mlp[j]);
        }
    }
}

// lanes with `active` get consecutive slots of *counter (one atomic per wave)
__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool active) {
    const uint64_t mask = __ballot(active);
    const uint32_t lane = __lane_id();
    uint32_t base = 0;
    if (mask) {
        const uint32_t leader = uint32_t(__ffsll((unsigned long long)mask) - 1);
        if (lane == leader) base = atomicAdd(counter, uint32_t(__popcll(mask)));
        base = __shfl(base, int(leader));
    }
    return base + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
}

// the cache learns radiance divided by the surface's diffuse albedo (sharp
// colour edges then need no resolution in the grid): the constant albedo of
// diffuse and Oren-Nayar materials, 1 for the others, for textured colours
// and for black surfaces
__device__ __forceinline__ f3 nrc_albedo(const SceneView &sc, uint32_t mat) {
    const MaterialRecord &m = sc.mat[mat];
    if ((m.type == MAT_DIFFUSE || m.type == MAT_OREN_NAYAR) && !(m.flags & MATF_TEXTURE_SLOT0)) {
        f3 a = mk3(m.p[0], m.p[1], m.p[2]);
        if (a.x > 1e-4f && a.y > 1e-4f && a.z > 1e-4f) return a;
    }
    return mk3(1.f, 1.f, 1.f);
}

// the emitter term of a hit, as path_shade adds it (integrator.cpp:188-196)
template <int MS>
__device__ __forceinline__ void emission_at(const SceneView &sc, const RenderView &rv, PathState<3> &ps,
                                            const TraceHit &h, const Surface &it, const InstanceRecord &inst) {
    if (inst.light == kInvalid) return;
    Spec<3> Le;
    float lpdf;
    eval_hit_light<3, true>(sc, it, inst.light, h.tri, ps.o, ps.swl, Le, lpdf);
    float w = mis_weight(ps.scatter_pdf, lpdf);
    if (rv.mis_mode == 2) w = 1.f;
    else if (rv.mis_mode == 1) w = ps.bounce == 0 ? w : 0.f;
    ps.L += to_linear_srgb<3, true>(sc, Le * ps.thr * w, ps.swl);
}

template <int MS>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) nrc_render_kernel(SceneView sc, CameraView cam, RenderView rv, uint32_t frame0,
                                                        uint32_t spp, uint32_t term, float4 *__restrict__ L,
                                                        NrcQuery *__restrict__ q, uint32_t *__restrict__ qcount,
                                                        uint32_t qcap) {
    extern __shared__ uint2 lds_stack[];
    const uint32_t px = blockIdx.x * 8u + (threadIdx.x & 7u);
    const uint32_t py = blockIdx.y * 8u + (threadIdx.x >> 3);
    const bool inside = px < cam.width && py < cam.height;
    const uint32_t pix = inside ? py * cam.width + px : 0u;
    uint32_t nf = 0, tf = 0;
    for (uint32_t s = 0; s < spp; ++s) {
        PathState<3> ps;
        path_begin<3, true>(cam, rv, inside ? px : 0u, inside ? py : 0u, frame0 + s, ps);
        if (!inside) ps.alive = 0u;
        bool has_q = false;
        f3 qp = mk3(0.f, 0.f, 0.f), qa = mk3(1.f, 1.f, 1.f);
        while (ps.alive) {
            TraceHit h = traverse<false>(sc, ps.o, ps.d, 0.f, 3.402823466e38f, lds_stack + threadIdx.x, 64u, nf, tf);
            if (term && ps.bounce == term && !ps.tail && h.tri != kInvalid) {
                Surface it = surface_at(sc, h.tri, h.b1, h.b2, false);
                const InstanceRecord inst = sc.inst[it.inst];
                if (inst.mat != kInvalid) {
                    emission_at<MS>(sc, rv, ps, h, it, inst);
                    has_q = true;
                    qp = it.pos;
                    qa = nrc_albedo(sc, inst.mat);
                    ps.alive = 0u;
                    break;
                }
            }
            ShadowRay sr;
            path_shade<3, true, MS>(sc, rv, ps, h, sr);
            if (sr.valid) {
                TraceHit sh = traverse<true>(sc, sr.o, sr.d, 0.f, 1.f - kShadowEpsilon, lds_stack + threadIdx.x, 64u,
                                             nf, tf);
                if (sh.tri == kInvalid) ps.L += sr.c;
            }
        }
        const uint32_t slot = pix * spp + s;
        if (inside) {
            f3 l = ps.L * ps.fw;
            L[slot] = make_float4(l.x, l.y, l.z, 0.f);
        }
        const uint32_t k = wave_append(qcount, has_q);
        if (has_q && k < qcap) {
            NrcQuery &o = q[k];
            o.p[0] = qp.x;
            o.p[1] = qp.y;
            o.p[2] = qp.z;
            o.slot = slot;
            o.beta[0] = ps.thr.v[0] * qa.x * ps.fw;
            o.beta[1] = ps.thr.v[1] * qa.y * ps.fw;
            o.beta[2] = ps.thr.v[2] * qa.z * ps.fw;
            o.pad = 0u;
        }
    }
}

template <int MS>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) nrc_train_paths_kernel(SceneView sc, CameraView cam, RenderView rv,
                                                             uint32_t frame, uint32_t stride, uint32_t offset,
                                                             uint32_t n_paths, float *__restrict__ staging,
                                                             NrcSample *__restrict__ out, uint32_t *__restrict__ count,
                                                             uint32_t cap) {
    extern __shared__ uint2 lds_stack[];
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    if (i >= n_paths) return;
    const uint32_t pixel = i * stride + offset;
    const uint32_t px = pixel % cam.width, py = pixel / cam.width;
    if (py >= cam.height) return;
    const uint32_t maxv = rv.max_depth;
    float *st = staging + size_t(i) * maxv * 9u;
    PathState<3> ps;
    path_begin<3, true>(cam, rv, px, py, kTrainStream + frame, ps);
    uint32_t nv = 0, nf = 0, tf = 0;
    while (ps.alive) {
        TraceHit h = traverse<false>(sc, ps.o, ps.d, 0.f, 3.402823466e38f, lds_stack + threadIdx.x, 64u, nf, tf);
        bool record = false;
        f3 pos = mk3(0.f, 0.f, 0.f);
        f3 beta = mk3(ps.thr.v[0], ps.thr.v[1], ps.thr.v[2]);
        if (h.tri != kInvalid && !ps.tail) {
            Surface it = surface_at(sc, h.tri, h.b1, h.b2, false);
            const uint32_t mat = sc.inst[it.inst].mat;
            if (mat != kInvalid) {
                record = true;
                pos = it.pos;
                f3 a = nrc_albedo(sc, mat);
                beta = mk3(beta.x * a.x, beta.y * a.y, beta.z * a.z);
            }
        }
        ShadowRay sr;
        path_shade<3, true, MS>(sc, rv, ps, h, sr);
        if (record && nv < maxv) {
            // the radiance gathered so far includes this vertex's emission
            float *r = st + nv * 9u;
            r[0] = pos.x;
            r[1] = pos.y;
            r[2] = pos.z;
            r[3] = beta.x;
            r[4] = beta.y;
            r[5] = beta.z;
            r[6] = ps.L.x;
            r[7] = ps.L.y;
            r[8] = ps.L.z;
            ++nv;
        }
        if (sr.valid) {
            TraceHit sh = traverse<true>(sc, sr.o, sr.d, 0.f, 1.f - kShadowEpsilon, lds_stack + threadIdx.x, 64u,
                                         nf, tf);
            if (sh.tri == kInvalid) ps.L += sr.c;
        }
    }
    uint32_t good = 0;
    for (uint32_t j = 0; j < nv; ++j) {
        const float *r = st + j * 9u;
        good += (r[3] > 0.f && r[4] > 0.f && r[5] > 0.f) ? 1u : 0u;
    }
    if (!good) return;
    uint32_t k = atomicAdd(count, good);
    for (uint32_t j = 0; j < nv; ++j) {
        const float *r = st + j * 9u;
        if (!(r[3] > 0.f && r[4] > 0.f && r[5] > 0.f)) continue;
        if (k < cap) {
            NrcSample &o = out[k];
            o.p[0] = r[0];
            o.p[1] = r[1];
            o.p[2] = r[2];
            o.t[0] = (ps.L.x - r[6]) / r[3];
            o.t[1] = (ps.L.y - r[7]) / r[4];
            o.t[2] = (ps.L.z - r[8]) / r[5];
        }
        ++k;
    }
}

__global__ void __launch_bounds__(256) nrc_accumulate_kernel(RenderView rv, const float4 *__restrict__ L,
                                                             uint32_t pixels, uint32_t spp, uint32_t frame0,
                                                             float4 *__restrict__ rt, float4 *__restrict__ out) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    float4 acc = rt[p];
    for (uint32_t s = 0; s < spp; ++s) {
        float4 l = L[size_t(p) * spp + s];
        acc = film_add(rv, acc, mk3(l.x, l.y, l.z), frame0 + s);
    }
    rt[p] = acc;
    out[p] = film_output(rv, acc);
}

__global__ void nrc_zero_kernel(uint32_t *p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

uint32_t grid_blocks() {
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    return uint32_t(cus > 0 ? cus : 64);
}

}// namespace

hipError_t nrc_forward(const NrcView &v, const float *pos, uint32_t n, float *out, hipStream_t s) {
    if (!n) return hipSuccess;
    uint32_t blocks = std::min<uint32_t>((n + kBlock - 1) / kBlock, 8u * grid_blocks());
    hipLaunchKernelGGL(nrc_forward_kernel<false>, dim3(blocks), dim3(kBlock), 0, s, v, pos, n, nullptr, nullptr, 0u,
                       out, nullptr);
    return hipGetLastError();
}

hipError_t nrc_infer(const NrcView &v, const NrcQuery *q, const uint32_t *count, uint32_t capacity, float4 *L,
                     hipStream_t s) {
    if (!capacity) return hipSuccess;
    uint32_t blocks = std::min<uint32_t>((capacity + kBlock - 1) / kBlock, 8u * grid_blocks());
    hipLaunchKernelGGL(nrc_forward_kernel<true>, dim3(blocks), dim3(kBlock), 0, s, v, nullptr, 0u, q, count,
                       capacity, nullptr, L);
    return hipGetLastError();
}

hipError_t nrc_backward(const NrcView &v, const NrcSample *samples, const uint32_t *count, uint32_t capacity,
                        uint32_t offset, uint32_t batch, float *loss, uint32_t blocks, hipStream_t s) {
    if (!capacity || !batch) return hipSuccess;
    if (!blocks) blocks = grid_blocks();
    hipLaunchKernelGGL(nrc_backward_kernel, dim3(blocks), dim3(kBlock), 0, s, v, samples, count, capacity, offset,
                       batch, loss);
    return hipGetLastError();
}

hipError_t nrc_adam(const NrcView &v, float lr, float c1, float c2, hipStream_t s) {
    hipLaunchKernelGGL(nrc_adam_kernel, dim3(8u * grid_blocks()), dim3(256), 0, s, v, lr, c1, c2);
    return hipGetLastError();
}

hipError_t nrc_sync_copies(const NrcView &v, hipStream_t s) {
    hipLaunchKernelGGL(nrc_sync_kernel, dim3(8u * grid_blocks()), dim3(256), 0, s, v);
    return hipGetLastError();
}

hipError_t nrc_render(const SceneView &sc, const CameraView &cam, const RenderView &rv, uint32_t frame0,
                      uint32_t spp, uint32_t term, float4 *L, NrcQuery *q, uint32_t *qcount, uint32_t qcap,
                      bool diffuse_only, hipStream_t s) {
    if (!sc.spectrum_rgb) return hipErrorInvalidValue;
    dim3 grid((cam.width + 7) / 8, (cam.height + 7) / 8);
    size_t lds = size_t(sc.stack_size) * 64 * sizeof(uint2);
    if (diffuse_only)
        hipLaunchKernelGGL(nrc_render_kernel<MS_DIFFUSE>, grid, dim3(64), lds, s, sc, cam, rv, frame0, spp, term, L,
                           q, qcount, qcap);
    else
        hipLaunchKernelGGL(nrc_render_kernel<MS_ALL>, grid, dim3(64), lds, s, sc, cam, rv, frame0, spp, term, L, q,
                           qcount, qcap);
    return hipGetLastError();
}

hipError_t nrc_train_paths(const SceneView &sc, const CameraView &cam, const RenderView &rv, uint32_t frame,
                           uint32_t stride, float *staging, NrcSample *out, uint32_t *count, uint32_t cap,
                           bool diffuse_only, hipStream_t s) {
    if (!sc.spectrum_rgb || !stride || !rv.max_depth) return hipErrorInvalidValue;
    const uint32_t pixels = cam.width * cam.height;
    const uint32_t offset = (frame * 2654435761u >> 8) % stride;
    const uint32_t n_paths = pixels > offset ? (pixels - offset + stride - 1) / stride : 0u;
    if (!n_paths) return hipSuccess;
    size_t lds = size_t(sc.stack_size) * 64 * sizeof(uint2);
    dim3 grid((n_paths + 63) / 64);
    if (diffuse_only)
        hipLaunchKernelGGL(nrc_train_paths_kernel<MS_DIFFUSE>, grid, dim3(64), lds, s, sc, cam, rv, frame, stride,
                           offset, n_paths, staging, out, count, cap);
    else
        hipLaunchKernelGGL(nrc_train_paths_kernel<MS_ALL>, grid, dim3(64), lds, s, sc, cam, rv, frame, stride, offset,
                           n_paths, staging, out, count, cap);
    return hipGetLastError();
}

hipError_t nrc_accumulate(const RenderView &rv, const float4 *L, uint32_t pixels, uint32_t spp, uint32_t frame0,
                          float4 *rt, float4 *out, hipStream_t s) {
    if (!pixels) return hipSuccess;
    hipLaunchKernelGGL(nrc_accumulate_kernel, dim3((pixels + 255) / 256), dim3(256), 0, s, rv, L, pixels, spp, frame0,
                       rt, out);
    return hipGetLastError();
}

hipError_t nrc_zero(uint32_t *p, uint32_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(nrc_zero_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

}// namespace vs
