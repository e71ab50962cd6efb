// Wavefront form of the path tracer: per-path state in HBM (one array per
// field, so a wave's loads coalesce), queues of live paths between the
// kernels of a bounce.  Shared by the host driver (capi.cpp) and the kernels
// (wavefront.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "dev_scene.h"

namespace vs {

struct WavefrontView {
    float4 *o;    // ray origin xyz, scatter pdf
    float4 *d;    // ray direction xyz, eta scale
    float4 *thr;  // throughput (up to 4 wavelengths)
    float4 *L;    // radiance gathered xyz, film weight
    uint4 *meta;  // sampler state, bounce | skips << 16, alive, tail bounce
    float4 *lam;  // hero wavelengths
    float4 *lpdf; // their pdfs
    float4 *hit;  // t, triangle id bits, b1, b2
    float4 *sh_o; // shadow ray origin xyz, path index bits
    float4 *sh_d; // shadow ray direction xyz
    float4 *sh_c; // shadow ray contribution xyz
    uint32_t pixels;
    uint32_t width;
};

// counters the kernels append to: [0] live paths written by shade, [1]
// shadow rays written by shade
constexpr int WF_LIVE = 0;
constexpr int WF_SHADOW = 1;

hipError_t wf_begin(const CameraView &cam, const RenderView &rv, const WavefrontView &w, uint32_t spectrum_dim,
                    bool rgb, uint32_t frame0, uint32_t n, uint32_t *queue, hipStream_t stream);
hipError_t wf_trace(const SceneView &sc, const WavefrontView &w, const uint32_t *queue, uint32_t count,
                    unsigned long long *counters, uint32_t block, hipStream_t stream);
// material_set: 0 every material, 1 the diffuse family only (a smaller
// kernel for scenes of diffuse / Oren-Nayar / black materials)
hipError_t wf_shade(const SceneView &sc, const RenderView &rv, const WavefrontView &w, uint32_t spectrum_dim,
                    bool rgb, int material_set, const uint32_t *qin, uint32_t count, uint32_t *qout,
                    uint32_t *counts, hipStream_t stream);
hipError_t wf_shadow(const SceneView &sc, const WavefrontView &w, const uint32_t *counts, uint32_t max_count,
                     unsigned long long *counters, uint32_t block, hipStream_t stream);
hipError_t wf_zero(uint32_t *counts, hipStream_t stream);
hipError_t wf_accumulate(const RenderView &rv, const WavefrontView &w, uint32_t frame0, uint32_t frames,
                         float4 *rt, float4 *out, hipStream_t stream);

}// namespace vs
