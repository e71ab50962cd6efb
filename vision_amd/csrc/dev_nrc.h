// Device code of the radiance cache: hash-grid encoding, the MLP on BF16
// MFMA tiles (16x16x16, FP32 accumulation) and its backward pass.
//
// A wave owns 64 rows (queries or training samples).  Each lane encodes its
// own row; the rows then go through the network in tiles of 16.  Tile
// operands live in LDS: activations row-major (bf16), weights transposed
// (wt[n][k] = W[k][n]) so that a lane's four consecutive k of an MFMA
// operand are one 8-byte load.  MFMA lane layout (16x16x16 bf16): the A
// operand of lane l holds row l % 16, k = 4 (l / 16) + e; the B operand
// holds column l % 16, k = 4 (l / 16) + e; the accumulator holds rows
// 4 (l / 16) + e of column l % 16, e = 0..3.
#pragma once

#include "dev_common.h"
#include "nrc.h"

namespace vs {
namespace dev {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x) {// round to nearest even
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }
__device__ __forceinline__ float h2f(uint16_t h) { return float(__builtin_bit_cast(_Float16, h)); }
__device__ __forceinline__ uint16_t f2h(float x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }

__device__ __forceinline__ f32x4 mma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x4 ld4(const uint16_t *p) { return *reinterpret_cast<const bf16x4 *>(p); }
__device__ __forceinline__ bf16x4 gather4(const uint16_t *p, int stride) {
    bf16x4 r;
    r.x = short(p[0]);
    r.y = short(p[stride]);
    r.z = short(p[2 * stride]);
    r.w = short(p[3 * stride]);
    return r;
}

// ---- encoding ----------------------------------------------------------------

__device__ __forceinline__ uint32_t nrc_index(const NrcView &v, int l, uint32_t x, uint32_t y, uint32_t z) {
    if (v.dense_mask & (1u << l)) {
        uint32_t n = v.res[l] + 1u;
        return x + n * (y + n * z);
    }
    return (x ^ (y * 2654435761u) ^ (z * 805459861u)) & (kNrcTable - 1u);
}

// the cell of a position at level l: integer corner and fractions
__device__ __forceinline__ void nrc_cell(const NrcView &v, int l, const float u[3], uint32_t c[3], float f[3]) {
    float s = float(v.res[l]);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float q = u[a] * s;
        float fl = floorf(q);
        uint32_t ci = uint32_t(fl);
        if (ci >= v.res[l]) ci = v.res[l] - 1u;
        c[a] = ci;
        f[a] = q - float(ci);
    }
}

__device__ __forceinline__ void nrc_unit(const NrcView &v, const float p[3], float u[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) u[a] = fminf(fmaxf((p[a] - v.lo[a]) * v.inv_ext[a], 0.f), 1.f);
}

// trilinear weight of corner k (bit 0: x, bit 1: y, bit 2: z)
__device__ __forceinline__ float nrc_corner_weight(int k, const float f[3]) {
    float wx = (k & 1) ? f[0] : 1.f - f[0];
    float wy = (k & 2) ? f[1] : 1.f - f[1];
    float wz = (k & 4) ? f[2] : 1.f - f[2];
    return (wx * wy) * wz;
}

// 32 features of a position, rounded to bf16 into x[0..31]
__device__ __forceinline__ void nrc_encode(const NrcView &v, const float p[3], uint16_t *x) {
    float u[3];
    nrc_unit(v, p, u);
    // four levels at a time: their 32 gathers are in flight together
#pragma unroll 4
    for (int l = 0; l < kNrcLevels; ++l) {
        uint32_t c[3];
        float f[3];
        nrc_cell(v, l, u, c, f);
        float a0 = 0.f, a1 = 0.f;
        const uint32_t *g = reinterpret_cast<const uint32_t *>(v.grid_f16) + size_t(l) * kNrcTable;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            uint32_t idx = nrc_index(v, l, c[0] + (k & 1), c[1] + ((k >> 1) & 1), c[2] + ((k >> 2) & 1));
            uint32_t pair = g[idx];
            float w = nrc_corner_weight(k, f);
            a0 += w * h2f(uint16_t(pair & 0xFFFFu));
            a1 += w * h2f(uint16_t(pair >> 16));
        }
        x[2 * l] = f2bf(a0);
        x[2 * l + 1] = f2bf(a1);
    }
}

// scatter dL/dx of one row into the grid gradients
__device__ __forceinline__ void nrc_encode_backward(const NrcView &v, const float p[3], const float *dx) {
    float u[3];
    nrc_unit(v, p, u);
    for (int l = 0; l < kNrcLevels; ++l) {
        uint32_t c[3];
        float f[3];
        nrc_cell(v, l, u, c, f);
        float g0 = dx[2 * l], g1 = dx[2 * l + 1];
        if (g0 == 0.f && g1 == 0.f) continue;
        float *gg = v.grid_g + size_t(l) * kNrcTable * kNrcFeatures;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            uint32_t idx = nrc_index(v, l, c[0] + (k & 1), c[1] + ((k >> 1) & 1), c[2] + ((k >> 2) & 1));
            float w = nrc_corner_weight(k, f);
            atomicAdd(gg + 2 * idx, w * g0);
            atomicAdd(gg + 2 * idx + 1, w * g1);
        }
    }
}

// ---- network tiles -------------------------------------------------------------

constexpr int kXs = 40;// row stride of the feature tile (32 + pad)
constexpr int kHs = 72;// row stride of hidden tiles (64 + pad)

// weights of one block in LDS
struct NrcWeightsLds {
    uint16_t w1t[kNrcHidden * kXs];// [n][k], k < 32
    uint16_t w2t[kNrcHidden * kHs];
    uint16_t w3t[kNrcHidden * kHs];
    float w4[kNrcHidden * kNrcOut];// [k][c]
};

// per-wave scratch
struct NrcWaveLds {
    uint16_t x[64 * kXs];     // features of the wave's 64 rows
    uint16_t h[3][16 * kHs];  // H1..H3 of the current row tile (post-ReLU)
    uint16_t g[2][16 * kHs];  // gradient tiles
    float y[16 * 4];          // outputs, then dL/dy, of the tile
    float dx[16 * (kNrcIn + 1)];
};

__device__ inline void nrc_load_weights(const NrcView &v, NrcWeightsLds &w, int tid, int nthreads) {
    for (int i = tid; i < kNrcIn * kNrcHidden; i += nthreads) {
        int k = i / kNrcHidden, n = i % kNrcHidden;
        w.w1t[n * kXs + k] = v.mlp_bf16[kNrcW1 + i];
    }
    for (int i = tid; i < kNrcHidden * kNrcHidden; i += nthreads) {
        int k = i / kNrcHidden, n = i % kNrcHidden;
        w.w2t[n * kHs + k] = v.mlp_bf16[kNrcW2 + i];
        w.w3t[n * kHs + k] = v.mlp_bf16[kNrcW3 + i];
    }
    for (int i = tid; i < kNrcHidden * kNrcOut; i += nthreads) w.w4[i] = v.mlp[kNrcW4 + i];
}

// out[16][N] = relu(in[16][K] W), in rows at `in` (stride is), wt = W^T (stride ws)
template <int K, int N>
__device__ __forceinline__ void nrc_fwd_tile(const uint16_t *in, int is, const uint16_t *wt, int ws, uint16_t *out,
                                             int lane) {
    const int r = lane & 15, q = lane >> 4;
#pragma unroll
    for (int nt = 0; nt < N / 16; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < K / 16; ++kt)
            acc = mma16(ld4(in + r * is + kt * 16 + 4 * q), ld4(wt + (nt * 16 + r) * ws + kt * 16 + 4 * q), acc);
#pragma unroll
        for (int e = 0; e < 4; ++e) out[(4 * q + e) * kHs + nt * 16 + r] = f2bf(fmaxf(acc[e], 0.f));
    }
}

// forward of row tile rt (rows 16 rt .. 16 rt + 15 of the wave): hidden
// tiles into s.h, outputs y[16][3] into s.y
__device__ __forceinline__ void nrc_forward_tile(const NrcWeightsLds &w, NrcWaveLds &s, int rt, int lane) {
    nrc_fwd_tile<kNrcIn, kNrcHidden>(s.x + rt * 16 * kXs, kXs, w.w1t, kXs, s.h[0], lane);
    __builtin_amdgcn_wave_barrier();
    nrc_fwd_tile<kNrcHidden, kNrcHidden>(s.h[0], kHs, w.w2t, kHs, s.h[1], lane);
    __builtin_amdgcn_wave_barrier();
    nrc_fwd_tile<kNrcHidden, kNrcHidden>(s.h[1], kHs, w.w3t, kHs, s.h[2], lane);
    __builtin_amdgcn_wave_barrier();
    if (lane < 48) {
        int row = lane / 3, c = lane % 3;
        float acc = 0.f;
        for (int k = 0; k < kNrcHidden; ++k) acc += bf2f(s.h[2][row * kHs + k]) * w.w4[k * kNrcOut + c];
        s.y[row * 4 + c] = acc;
    }
    __builtin_amdgcn_wave_barrier();
}

// dW[i][j] += sum over the tile's 16 rows of A[row][i] G[row][j]; A rows at
// `a` (stride as, I columns), G at `g` (stride kHs, 64 columns); dW in LDS (stride 64)
template <int I>
__device__ __forceinline__ void nrc_weight_grad(const uint16_t *a, int as, const uint16_t *g, float *dw, int lane) {
    const int r = lane & 15, q = lane >> 4;
#pragma unroll
    for (int it = 0; it < I / 16; ++it) {
        bf16x4 av = gather4(a + (4 * q) * as + it * 16 + r, as);
#pragma unroll
        for (int jt = 0; jt < 4; ++jt) {
            bf16x4 gv = gather4(g + (4 * q) * kHs + jt * 16 + r, kHs);
            f32x4 acc = mma16(av, gv, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
            for (int e = 0; e < 4; ++e) atomicAdd(dw + (it * 16 + 4 * q + e) * kNrcHidden + jt * 16 + r, acc[e]);
        }
    }
}

// out[16][J] = (G[16][64] W^T) (x relu'(h) when h != nullptr); W^T's B
// operand is read from wt[n][k] = W[k][n] as B[n][j] = wt[n][j]
template <int J>
__device__ __forceinline__ void nrc_input_grad(const uint16_t *g, const uint16_t *wt, int ws, const uint16_t *h,
                                               uint16_t *out_bf, float *out_f, int out_fs, int lane) {
    const int r = lane & 15, q = lane >> 4;
#pragma unroll
    for (int jt = 0; jt < J / 16; ++jt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt)
            acc = mma16(ld4(g + r * kHs + kt * 16 + 4 * q), gather4(wt + (kt * 16 + 4 * q) * ws + jt * 16 + r, ws), acc);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int row = 4 * q + e, col = jt * 16 + r;
            if (h) {
                float v = bf2f(h[row * kHs + col]) > 0.f ? acc[e] : 0.f;
                out_bf[row * kHs + col] = f2bf(v);
            } else {
                out_f[row * out_fs + col] = acc[e];
            }
        }
    }
}

// block-level accumulators of the weight gradients (fp32, LDS)
struct NrcGradLds {
    float w1[kNrcIn * kNrcHidden];
    float w2[kNrcHidden * kNrcHidden];
    float w3[kNrcHidden * kNrcHidden];
    float w4[kNrcHidden * kNrcOut];
};

// backward of row tile rt after nrc_forward_tile: s.y holds dL/dy of the
// tile's rows (0 for masked rows); weight gradients into gl, dL/dx into s.dx
__device__ __forceinline__ void nrc_backward_tile(const NrcWeightsLds &w, NrcWaveLds &s, NrcGradLds &gl, int rt,
                                                  int lane) {
    // output layer: dW4[k][c] = sum_rows h3[row][k] dy[row][c]; G3 = dy W4^T x relu'(h3)
    {
        int k = lane;
        float a0 = 0.f, a1 = 0.f, a2 = 0.f;
        for (int row = 0; row < 16; ++row) {
            float hv = bf2f(s.h[2][row * kHs + k]);
            a0 += hv * s.y[row * 4 + 0];
            a1 += hv * s.y[row * 4 + 1];
            a2 += hv * s.y[row * 4 + 2];
            float gsum = s.y[row * 4 + 0] * w.w4[k * 3 + 0] + s.y[row * 4 + 1] * w.w4[k * 3 + 1] +
                         s.y[row * 4 + 2] * w.w4[k * 3 + 2];
            s.g[0][row * kHs + k] = f2bf(hv > 0.f ? gsum : 0.f);
        }
        atomicAdd(gl.w4 + k * 3 + 0, a0);
        atomicAdd(gl.w4 + k * 3 + 1, a1);
        atomicAdd(gl.w4 + k * 3 + 2, a2);
    }
    __builtin_amdgcn_wave_barrier();
    nrc_weight_grad<kNrcHidden>(s.h[1], kHs, s.g[0], gl.w3, lane);                    // dW3 = H2^T G3
    nrc_input_grad<kNrcHidden>(s.g[0], w.w3t, kHs, s.h[1], s.g[1], nullptr, 0, lane); // G2
    __builtin_amdgcn_wave_barrier();
    nrc_weight_grad<kNrcHidden>(s.h[0], kHs, s.g[1], gl.w2, lane);                    // dW2 = H1^T G2
    nrc_input_grad<kNrcHidden>(s.g[1], w.w2t, kHs, s.h[0], s.g[0], nullptr, 0, lane); // G1
    __builtin_amdgcn_wave_barrier();
    nrc_weight_grad<kNrcIn>(s.x + rt * 16 * kXs, kXs, s.g[0], gl.w1, lane);           // dW1 = X^T G1
    nrc_input_grad<kNrcIn>(s.g[0], w.w1t, kXs, nullptr, nullptr, s.dx, kNrcIn + 1, lane);// dX
    __builtin_amdgcn_wave_barrier();
}

}// namespace dev
}// namespace vs
