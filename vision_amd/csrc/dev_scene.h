// Kernel-argument views of the device-resident scene, film and render state.
// Plain pointers into HBM owned by the host-side DeviceScene (capi.cpp).
#pragma once

#include <cstdint>

#include "records.h"
#include "vmath.h"

namespace vs {

struct SceneView {
    const uint32_t *nodes;   // Bvh8Node[], 80 B each
    const float *leaf_tris;  // LeafTri[], 48 B each
    const f3 *P;             // world positions per vertex
    const f3 *N;             // world (unnormalized) normals per vertex
    const f2 *UV;
    const u3 *tri;           // vertex indices per global triangle
    const uint32_t *tri_inst;// instance per global triangle
    const InstanceRecord *inst;
    const MaterialRecord *mat;
    const LightRecord *lights;
    const AliasEntry *alias; // all alias tables, concatenated
    const float *alias_func; // their weights
    uint32_t num_lights;
    uint32_t light_sampler;  // 0 = uniform, 1 = power
    uint32_t select_offset;  // power: offset of the light-selection table
    float select_integral;   // power: sum(weights)/n
    uint32_t stack_size;     // traversal stack entries per lane
    float ray_offset_factor;
    uint32_t num_tris;
    uint32_t pad;
    // spectra: CIE X, Y, Z and D65 tables (94 samples each, at offsets 0, 94,
    // 188, 282) followed by the materials' ior / eta / k curves
    const float *spd;
    const float *lut;// energy-compensation tables (offsets in dev_bsdf.h)
    float cie_interval;
    float cie_y_integral;
    uint32_t spectrum_rgb;// 1: sRGB mode, 0: hero-wavelength mode
    uint32_t env_light;   // light index of the environment map, 0xFFFFFFFF if none
    // environment map (spherical light): texels hold linear sRGB (sRGB mode)
    // or sigmoid coefficients + unbound scale (hero mode)
    const float *env_tex;
    const AliasEntry *env_cond;// per-row alias tables, W*H
    const float *env_cond_func;
    const AliasEntry *env_marg;// row-marginal alias table, H
    const float *env_marg_func;
    uint32_t env_w, env_h;
    float env_marg_integral;
    float world_diameter;
    float env_w2o[9];// row-major 3x3
    float env_o2w[9];
    float env_scale;
    float pad2;
    // images of textured colour slots: float4 texels, per texture {first
    // texel, width, height, 0}
    const float4 *tex;
    const uint4 *tex_info;
};

struct CameraView {
    float c2w[12];// 3x4 row-major: rows of [R | t]
    float lower_x, upper_y;
    float raster_sx, raster_sy;// span_x / W, span_y / H
    float tan_half_fov;
    float lens_radius, focal_distance;
    uint32_t width, height;
    uint32_t filter;// 0 box, 1 triangle, 2 fitted curve (filter importance sampling)
    float filter_radius;
    uint32_t thin_lens;
    // fitted-curve filters (FilterSampler): normalised 20x20 value table and
    // the 2D alias table over |value|
    const float *fis_lut;
    const AliasEntry *fis_cond;
    const float *fis_cond_func;
    const AliasEntry *fis_marg;
    const float *fis_marg_func;
    float fis_marg_integral;
};

struct RenderView {
    uint32_t max_depth, min_depth;
    float rr_threshold;
    int mis_mode;
    uint32_t frame_start;// frame index of the first sample this launch adds
    uint32_t spp;        // samples (frames) this launch adds per pixel
    uint32_t accumulation;
    float exposure;
    uint32_t tone_mapper;// 0 linear, 1 aces, 2 reinhard
    uint32_t write_output;
    uint32_t count_rays;
    uint32_t spectrum_dim;// hero wavelengths per path (3 in sRGB mode)
    // ConvergenceInspector (adaptive/inspector.h): VarianceStats (avg, var,
    // N bits, 0) per pixel, null when adaptive sampling is off
    float4 *vstats;
    float adapt_threshold;
    uint32_t adapt_min_samples;
};

// device ray / hit records of the trace entry points (32 B / 24 B)
struct RayRec {
    float o[3];
    float tmin;
    float d[3];
    float tmax;
};
struct HitRec {
    uint32_t inst;// 0xFFFFFFFF on a miss
    uint32_t prim;
    float bary[2];
    float t;
    uint32_t tri;// global triangle id (scene order)
};

}// namespace vs
