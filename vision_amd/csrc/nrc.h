// Radiance cache: a multiresolution hash-grid encoding of position followed
// by a small MLP, trained online from the path tracer's own paths and
// queried where paths terminate early.  Shared by the host driver (nrc.cpp)
// and the kernels (nrc.hip); the oracle restates the same arithmetic.
//
// Encoding: 16 levels of 2 features; level l has res[l] cells per axis over
// the scene's bounding box; its (res+1)^3 grid vertices index a table of
// 2^19 entries directly while they fit, else through the spatial hash
// (x * 1) ^ (y * 2654435761) ^ (z * 805459861) mod 2^19.  A position's
// features are the trilinear blend of its cell's 8 vertices.
// Network: 32 -> 64 -> 64 -> 64 (ReLU) -> 3 (linear), no biases; BF16
// inputs and weights on the MFMA layers with FP32 accumulation, FP32
// output layer.  Loss: relative L2, (y - t)^2 / (y^2 + 0.01) with y held
// constant in the denominator, averaged over samples and channels.
// Optimiser: Adam (beta1 0.9, beta2 0.99, eps 1e-15) on FP32 master
// copies; the grid is read as FP16, the hidden layers as BF16.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "dev_scene.h"

namespace vs {

constexpr int kNrcLevels = 16;
constexpr int kNrcFeatures = 2;
constexpr int kNrcLog2Table = 19;
constexpr uint32_t kNrcTable = 1u << kNrcLog2Table;
constexpr int kNrcIn = kNrcLevels * kNrcFeatures;// 32
constexpr int kNrcHidden = 64;
constexpr int kNrcOut = 3;
// parameter offsets in the flat MLP parameter array (row-major W[k][n])
constexpr int kNrcW1 = 0;                                   // 32 x 64
constexpr int kNrcW2 = kNrcW1 + kNrcIn * kNrcHidden;         // 64 x 64
constexpr int kNrcW3 = kNrcW2 + kNrcHidden * kNrcHidden;     // 64 x 64
constexpr int kNrcW4 = kNrcW3 + kNrcHidden * kNrcHidden;     // 64 x 3
constexpr int kNrcMlpParams = kNrcW4 + kNrcHidden * kNrcOut; // 10432
constexpr size_t kNrcGridParams = size_t(kNrcLevels) * kNrcTable * kNrcFeatures;

struct NrcView {
    // parameters: fp32 masters, the copies the forward pass reads, Adam
    // moments and gradient accumulators
    float *mlp;              // kNrcMlpParams
    uint16_t *mlp_bf16;      // bf16 copies of W1..W3 (same offsets)
    float *grid;             // kNrcGridParams, [level][entry][feature]
    uint16_t *grid_f16;      // fp16 bit patterns
    float *mlp_m, *mlp_v, *mlp_g;
    float *grid_m, *grid_v, *grid_g;
    uint32_t res[kNrcLevels];// cells per axis of each level
    uint32_t dense_mask;     // bit l: level l indexes its vertices directly
    float lo[3], inv_ext[3]; // position -> [0, 1]^3
};

// one training sample: position and the radiance scattered from it,
// divided by the surface's diffuse albedo (the query multiplies it back)
struct NrcSample {
    float p[3];
    float t[3];
};

// a path terminated into the cache: position, throughput and the film slot
// its radiance adds to
struct NrcQuery {
    float p[3];
    uint32_t slot;
    float beta[3];
    uint32_t pad;
};

// ---- kernels (nrc.hip) ----
// parity and utility entry points over n positions / samples
hipError_t nrc_forward(const NrcView &v, const float *pos, uint32_t n, float *out, hipStream_t s);
// forward + backward over samples [offset, offset + batch) of the
// min(*count, capacity) present: gradients of their mean loss accumulate
// into v.mlp_g / v.grid_g, the mean loss into *loss (all must start at zero)
hipError_t nrc_backward(const NrcView &v, const NrcSample *samples, const uint32_t *count, uint32_t capacity,
                        uint32_t offset, uint32_t batch, float *loss, uint32_t blocks, hipStream_t s);
// Adam step over all parameters with bias corrections c1 = 1 / (1 - 0.9^t),
// c2 = 1 / (1 - 0.99^t); clears the gradients and refreshes the fp16 /
// bf16 copies
hipError_t nrc_adam(const NrcView &v, float lr, float c1, float c2, hipStream_t s);
// refresh the fp16 / bf16 copies from the masters (after an upload)
hipError_t nrc_sync_copies(const NrcView &v, hipStream_t s);
// queries: out radiance added into L[slot] += beta * max(y, 0)
hipError_t nrc_infer(const NrcView &v, const NrcQuery *q, const uint32_t *count, uint32_t capacity, float4 *L,
                     hipStream_t s);

// rendering with the cache (sRGB mode): one thread per (pixel, sample) of
// a frame; paths stop at bounce `term` and append a query, radiance so far
// goes to L[slot]; slot = pixel * spp + sample
hipError_t nrc_render(const SceneView &sc, const CameraView &cam, const RenderView &rv, uint32_t frame0,
                      uint32_t spp, uint32_t term, float4 *L, NrcQuery *q, uint32_t *qcount, uint32_t qcap,
                      bool diffuse_only, hipStream_t s);
// training paths: full-length paths of every `stride`-th pixel (offset by
// frame) on their own sampler stream (frame index 2^30 + frame), each
// surface vertex giving one sample (position, radiance scattered from it:
// the path's radiance gathered after the vertex's own emission, divided by
// the throughput reaching it); `staging` holds max_depth records per path
hipError_t nrc_train_paths(const SceneView &sc, const CameraView &cam, const RenderView &rv, uint32_t frame,
                           uint32_t stride, float *staging, NrcSample *out, uint32_t *count, uint32_t cap,
                           bool diffuse_only, hipStream_t s);
// fold the frame's samples into the film's running mean in order
hipError_t nrc_accumulate(const RenderView &rv, const float4 *L, uint32_t pixels, uint32_t spp, uint32_t frame0,
                          float4 *rt, float4 *out, hipStream_t s);
hipError_t nrc_zero(uint32_t *p, uint32_t n, hipStream_t s);

}// namespace vs
