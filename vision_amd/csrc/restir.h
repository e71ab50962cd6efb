// ReSTIR DI state shared by the host driver (capi.cpp) and the kernels
// (restir.hip): device buffers, the DIParam settings
// (src/render_core/integrators/ReSTIR/direct.h:10-34) and the camera of the
// current and the previous frame.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "dev_scene.h"

namespace vs {

struct RestirView {
    float4 *rsv; // reservoirs: 4 buffers x 3 fields x pixels
    float4 *surf;// surfaces:   2 buffers x 3 fields x pixels
    float4 *ext; // surface extensions: 2 buffers x 2 fields x pixels
    uint32_t width, height, pixels;
    uint32_t frame;
    uint32_t M_light, M_bsdf, max_age;
    float diff_factor;
    uint32_t spatial, spatial_num;
    float s_dot, s_depth, s_radius;
    uint32_t temporal, history_limit;
    float t_dot, t_depth, t_radius;
    uint32_t temporal_N, temporal_mis, pairwise, max_recursion;
    // camera now and one frame earlier (Sensor::store_prev_data,
    // sensor.cpp:99-103); w2c maps world points to camera space (3x4 rows)
    float cam_pos[3], prev_cam_pos[3];
    float w2c[12], prev_w2c[12];
    // with ReSTIR GI open: shader0 leaves its per-pixel record for GI in
    // gi_aux (3 fields x pixels: HitBSDF wi + pdf, f, motion vector) and
    // shader1 writes the direct radiance to gi_direct instead of the film;
    // both null otherwise
    float4 *gi_aux;
    float4 *gi_direct;
};

// ReSTIR GI state and settings (GIParam, ReSTIR/indirect.h:10-28, with
// construct_param, indirect.cpp:335-352).  HBM layout, float4 field-major:
//   rsv:    3 buffers (previous / current by frame parity, pass-through)
//           x 4 fields x pixels: (weight_sum, C, W, age), sample position,
//           sample normal, Lo
//   aux:    shader0's record (see RestirView::gi_aux)
//   direct: shader1's radiance, rad: the indirect radiance (pixels each)
struct RestirGiView {
    float4 *rsv;
    const float4 *aux;
    const float4 *direct;
    float4 *rad;
    uint32_t max_age;
    float diff_factor;
    uint32_t spatial, s_num;
    float s_dot, s_depth, s_radius;
    uint32_t temporal, t_limit, t_N;
    float t_dot, t_depth, t_radius;
    // path depth of the initial samples: the rt integrator's max_depth - 1
    // (rt.cpp:30)
    uint32_t depth;
};
constexpr uint32_t kGiReservoirBuffers = 3;

// reservoir buffers: previous / current by frame parity, the pass-through of
// shader0 -> shader1, and a second pass-through for extra spatial passes
constexpr uint32_t kRestirReservoirBuffers = 4;

// material_set: 0 every material, 1 the diffuse family only (a smaller
// kernel for scenes without metal, glass or principled materials)
hipError_t restir_clear(const RestirView &v, hipStream_t stream);
hipError_t restir_initial(const SceneView &sc, const CameraView &cam, const RestirView &v, uint32_t spectrum_dim,
                          bool rgb, int material_set, unsigned long long *counters, hipStream_t stream);
hipError_t restir_spatial(const SceneView &sc, const RestirView &v, uint32_t spectrum_dim, bool rgb,
                          int material_set, uint32_t in, uint32_t out, uint32_t dim, hipStream_t stream);
hipError_t restir_final(const SceneView &sc, const CameraView &cam, const RestirView &v, const RenderView &rv,
                        uint32_t spectrum_dim, bool rgb, int material_set, uint32_t in, float4 *rt, float4 *out,
                        unsigned long long *counters, hipStream_t stream);
// ReSTIR GI (ReSTIRGI::dispatch, indirect.cpp:354-362) in two launches:
// initial samples fused with temporal reuse, then spatial reuse, shading
// and the combine of the direct and indirect halves into the film
// (rt.cpp:119-126).  sRGB mode only.
hipError_t restir_gi_clear(const RestirGiView &g, uint32_t pixels, hipStream_t stream);
hipError_t restir_gi_initial_temporal(const SceneView &sc, const CameraView &cam, const RestirView &v,
                                      const RestirGiView &g, const RenderView &rv, int material_set,
                                      unsigned long long *counters, hipStream_t stream);
hipError_t restir_gi_spatial_shade(const SceneView &sc, const RestirView &v, const RestirGiView &g,
                                   const RenderView &rv, int material_set, float4 *rt, float4 *out,
                                   unsigned long long *counters, hipStream_t stream);

}// namespace vs
