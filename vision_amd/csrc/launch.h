// Host-callable launchers for the kernels in kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "dev_scene.h"

namespace vs {

hipError_t launch_trace(const SceneView &sc, const RayRec *rays, HitRec *hits, uint8_t *occ, uint32_t n, bool any,
                        unsigned long long *counters, hipStream_t stream);

hipError_t launch_bsdf(const SceneView &sc, uint32_t spectrum_dim, uint32_t mat, const float *wo, const float *wi,
                       const float *lambda, uint32_t n, int sample, uint32_t frame, float *wi_out, float *f,
                       float *pdf, hipStream_t stream);
// diffuse_only: the scene holds only diffuse / Oren-Nayar / black materials
// (the kernel is compiled without the other material families)
hipError_t launch_pt(const SceneView &sc, const CameraView &cam, const RenderView &rv, float4 *rt, float4 *out,
                     unsigned long long *counters, bool diffuse_only, hipStream_t stream);
// the saved picture: tone map (again) + optional sRGB gamma over the output buffer (film.hip)
hipError_t launch_final_picture(const float4 *in, float4 *out, uint32_t n, uint32_t tone_mapper, bool gamma,
                                hipStream_t stream);

}// namespace vs
