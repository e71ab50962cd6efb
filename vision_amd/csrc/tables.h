// Host construction of sampling tables and device records.
#pragma once

#include <vector>

#include "records.h"
#include "scene.h"

namespace vs {

struct AliasTableHost {
    std::vector<AliasEntry> table;
    std::vector<float> func;
    float integral{0.f};
};

// Vose alias table with the reference's construction order
// (AliasTable::build, src/render_core/warpers/alias.h:86-122)
AliasTableHost build_alias(const std::vector<float> &weights);

// light power per LightType (area.cpp:87-89, point.cpp:32-34, spot.cpp:44-46,
// spherical.cpp:57-60)
f3 light_power(const LightHost &l, float world_radius);

// scene records in device layout
struct PackedScene {
    bool rgb{true};
    std::vector<MaterialRecord> mats;
    std::vector<float> spd;// CIE X, Y, Z, D65 (4 x 94) then material curves
    float cie_interval{0.f}, cie_y_integral{0.f};
    std::vector<LightRecord> lights;
    std::vector<AliasEntry> alias;
    std::vector<float> func;
    uint32_t select_offset{0};
    float select_integral{0.f};
    // environment map
    uint32_t env_light{kInvalid};
    uint32_t env_w{0}, env_h{0};
    std::vector<float> env_tex;// W*H*4
    std::vector<AliasEntry> env_cond, env_marg;
    std::vector<float> env_cond_func, env_marg_func;
    float env_marg_integral{0.f};
    float env_w2o[9]{}, env_o2w[9]{};
    float env_scale{1.f};
    float world_radius{10.f};
    // textures of the colour slots: float4 texels (linear sRGB, or the
    // albedo sigmoid coefficients of each texel in hero mode), and per
    // texture {first texel, width, height, 0}
    std::vector<float> tex;
    std::vector<uint32_t> tex_info;
};

PackedScene pack_scene(const SceneHost &s);

// FilterSampler (src/render_core/filters/fitted_curve.h:16-73): the 20x20
// table of a fitted-curve filter (gaussian, mitchell, sinc) over the positive
// quadrant of its radius, normalised by its mean, and the 2D alias table of
// its absolute values
struct FilterTable {
    std::vector<float> lut;
    std::vector<AliasEntry> cond, marg;
    std::vector<float> cond_func, marg_func;
    float marg_integral{0.f};
};
constexpr int kFilterTableSize = 20;
// 0 box, 1 triangle, 2 gaussian, 3 mitchell, 4 sinc
int filter_kind(const CameraHost &c);
float filter_evaluate(const CameraHost &c, float x, float y);
FilterTable build_filter_table(const CameraHost &c);

}// namespace vs
