// G-buffer and SVGF denoiser state shared by the host driver (capi.cpp) and
// the kernels (svgf.hip).
//
// G-buffer: FrameBuffer's rt_geom / rt_gradient kernels
// (src/base/frame_buffer.cpp:72-177), PixelGeometry (frame_buffer.h:15-20),
// the albedo and emission buffers.  SVGF: src/render_core/denoisers/SVGF/*.
//
// HBM layout, float4 field-major:
//   geo: 2 slots (previous / current by frame parity, cur = (frame + 1) & 1,
//        frame_buffer.h:44-47) x 2 fields x pixels:
//        (normal.xyz, normal fwidth), (p_film.xy, depth gradient, linear depth);
//        a miss keeps depth -1
//   alb, emi: albedo and emission in linear sRGB (w = 1), pixels each
//   rec: 3 buffers x 2 fields x pixels: (illumination.rgb, variance),
//        (history, moment 1, moment 2, 0).  Buffers 0 / 1 are the SVGFData of
//        the previous / current frame by parity; buffer 2 is the second half
//        of the ping-pong the filter passes run on.  The frame's last pass
//        lands in the current buffer, which the next frame reprojects from.
// The reference's passes run in place (each a-trous pass writes the records
// its neighbours read; reproject reads the history it overwrites at a
// neighbouring pixel).  That is a race whose outcome depends on thread
// order; these kernels take the race-free reading: a pass reads only what
// the previous pass or the previous frame wrote.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "dev_scene.h"
#include "restir.h"

namespace vs {

struct SvgfParams {
    uint32_t N;// a-trous passes (step 1, 2, 4, ...)
    float alpha, moments_alpha;
    uint32_t history_limit;
    int moments_radius;
    float sigma_rt, sigma_normal;
    uint32_t reproject, filter_moments;
};

struct SvgfView {
    float4 *geo;
    float4 *alb, *emi;
    float4 *rec;
    uint32_t width, height, pixels;
    uint32_t frame;
};

// the G-buffer of the frame (the current slot); counts its camera rays
hipError_t gbuffer_compute(const SceneView &sc, const CameraView &cam, const RestirView &rv, const SvgfView &sv,
                           uint32_t spectrum_dim, bool rgb, int material_set, unsigned long long *counters,
                           hipStream_t stream);
// SVGF::dispatch for one frame (svgf.cpp:89-106): demodulate + reproject,
// filter moments, N a-trous passes, modulate into `out`.  `motion` is the
// motion vector field (float4 .xy per pixel), `rad` the film's running mean.
hipError_t svgf_denoise(const SvgfView &sv, const SvgfParams &p, const float4 *motion, const float4 *rad,
                        float4 *out, hipStream_t stream);
hipError_t svgf_clear(const SvgfView &sv, hipStream_t stream);

}// namespace vs
