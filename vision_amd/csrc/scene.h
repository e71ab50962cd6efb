// Host-side scene model: Vision scene JSON -> resolved descriptors ->
// world-space triangle soup plus material / light / camera records.
//
// Descriptor keys and defaults follow the reference's *Desc::init functions
// (src/core/node_desc.cpp:31-369) and plugin constructors cited per field.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "json.h"
#include "records.h"
#include "vmath.h"

namespace vs {

struct MeshHost {
    std::vector<f3> P, N;
    std::vector<f2> UV;
    std::vector<u3> tris;
};

struct InstanceHost {
    uint32_t mesh{0};
    m4 o2w{m4_identity()};
    int32_t material{-1};// index into SceneHost::materials, -1 = none
    int32_t light{-1};   // index into SceneHost::lights, -1 = not emissive
    std::string name;
    std::string material_name;
};

// A material with its resolved slot values; packed into MaterialRecord.
struct MaterialHost {
    std::string type;// diffuse | black_body | metal | mirror | glass | principled_bsdf | mix | add
    std::string name;
    Json param;
    // why this material cannot be rendered (empty: it can); reported when a
    // shape uses it, so a scene file listing materials no shape uses loads
    std::string unsupported;
};

// an image used as a colour parameter (ImageNode, src/render_core/shadernodes/image.cpp):
// linear RGB rows, top first
struct TextureHost {
    std::string key;// resolved path + "|" + colour space (+ "|illum")
    uint32_t width{0}, height{0};
    bool illum{false};// an emission image (illumination spectra), not an albedo
    std::vector<float> rgb;
};

struct LightHost {
    uint32_t type{LIGHT_AREA};
    std::string name;
    f3 color{1, 1, 1};// normalized by its max component (light.cpp:10-17); an image's average
    int32_t tex{-1};  // area: emission image, index into SceneHost::textures
    float scale{1.f};
    uint32_t two_sided{0};
    uint32_t inst{0xFFFFFFFFu};
    f3 position{0, 0, 0};
    f3 direction{0, 0, 1};
    float angle{0.f}, falloff{0.f};
    // environment map (spherical): linear RGB texels, top row first
    std::string env_fn;
    std::vector<float> env_rgb;
    uint32_t env_w{0}, env_h{0};
    bool env_constant{false};// a constant colour, not an image
    m4 env_o2w{m4_identity()};
    bool flip_u{false};
    // per-emitter triangle area weights (area lights) filled at flatten time
    std::vector<float> tri_areas;
    uint32_t tri_base{0};
};

struct CameraHost {
    std::string type{"thin_lens"};
    float fov_y{20.f};
    f3 position{0, 0, 0};
    float yaw{0.f}, pitch{0.f};// degrees, re-derived from the transform
    m4 c2w{m4_identity()};
    float lens_radius{0.f}, focal_distance{5.f};
    uint32_t width{1280}, height{720};
    std::string filter_type{"gaussian"};
    float filter_radius{0.5f};
    // fitted-curve filter parameters: gaussian sigma, mitchell b / c, sinc tau
    float filter_sigma{1.f}, filter_b{1.f / 3.f}, filter_c{1.f / 3.f}, filter_tau{3.f};
    float exposure{1.f};
    uint32_t accumulation{1};
    std::string tone_mapper{"linear"};
};

// ReSTIR DI settings of the "rt" integrator's "direct" block: ReSTIRDI
// (direct.cpp:10-17), ReSTIR (ReSTIR/common.h:84-99), SpatialResamplingParam
// and TemporalResamplingParam (common.h:19-67)
struct RestirHost {
    bool open{true};
    uint32_t M_light{10}, M_bsdf{1}, max_age{30}, max_recursion{5};
    float diff_factor{0.3f};
    bool pairwise{false};
    bool s_open{true};
    float s_theta{35.f}, s_depth{0.3f}, s_radius{30.f};
    uint32_t s_num{1};
    bool t_open{true}, t_mis{false};
    uint32_t t_limit{5}, t_N{9};
    float t_radius{8.f}, t_theta{35.f}, t_depth{0.3f};
    // spatial passes run before the reference's one (0 = the reference)
    uint32_t extra_spatial_passes{0};
};

// the "indirect" block: ReSTIR GI (ReSTIR/indirect.h:10-28, the ReSTIR base
// and resampling parameters of ReSTIR/common.h:19-113)
struct RestirGiHost {
    bool open{true};
    uint32_t max_age{30};
    float diff_factor{0.3f};
    bool s_open{true};
    float s_theta{35.f}, s_depth{0.3f}, s_radius{30.f};
    uint32_t s_num{1};
    bool t_open{true};
    uint32_t t_limit{5}, t_N{9};
    float t_radius{8.f}, t_theta{35.f}, t_depth{0.3f};
};

// the integrator's "denoiser" block (IntegratorDesc::init, node_desc.cpp:102-111:
// type "svgf" by default) with SVGF's parameters (SVGF::SVGF, SVGF/svgf.h:44-53).
// Like the reference's GUI switches, `on`, `reproject` and `filter_moments`
// are set from the host (vs_scene_set_denoiser); SVGF runs after each frame
// of the rt integrator (rt.cpp:157-158).
struct SvgfHost {
    bool svgf{true};// false for another denoiser type (not run)
    bool on{false}, reproject{true}, filter_moments{true};
    uint32_t N{3};
    float alpha{0.05f}, moments_alpha{0.2f};
    uint32_t history_limit{32};
    int moments_radius{3};
    float sigma_rt{10.f}, sigma_normal{30.f};
};

// the pt integrator's "adaptive" block: ConvergenceInspector's threshold and
// start_index (adaptive/inspector.cpp:9-12); switched on from the host
// (vs_scene_set_adaptive), as the reference's GUI does
struct AdaptiveHost {
    bool on{false};
    float threshold{0.001f};
    uint32_t min_samples{128};
};

struct IntegratorHost {
    std::string type{"pt"};
    uint32_t max_depth{16}, min_depth{5};
    float rr_threshold{1.f};
    int mis_mode{0};
    Json param;
    RestirHost restir;
    RestirGiHost gi;
    SvgfHost svgf;
    AdaptiveHost adaptive;
};

struct SceneHost {
    std::string base_dir;
    std::vector<MeshHost> meshes;
    std::vector<InstanceHost> instances;
    std::vector<MaterialHost> materials;
    std::vector<LightHost> lights;
    CameraHost camera;
    IntegratorHost integrator;
    std::string spectrum_type{"srgb"};
    uint32_t spectrum_dim{3};
    std::string light_sampler_type{"uniform"};
    float env_prob{0.5f};
    bool env_separate{false};
    std::string sampler_type{"independent"};
    float ray_offset_factor{1.f};
    float min_world_radius{10.f};
    f3 aabb_lo{0, 0, 0}, aabb_hi{0, 0, 0};// world bounds of all triangles (after flatten)
    std::vector<TextureHost> textures;     // images the materials' colour slots name
    uint32_t output_spp{0};
    std::string output_fn{"output.png"};

    // flattened world-space geometry (filled by flatten())
    std::vector<f3> wP, wN;
    std::vector<f2> wUV;
    std::vector<u3> wTri;         // global vertex indices
    std::vector<uint32_t> triInst;// instance id per global triangle
    std::vector<uint32_t> triPrim;// mesh-local primitive id per global triangle
    std::vector<uint32_t> instTriBase;
    std::vector<MaterialRecord> matRecords;
    std::vector<LightRecord> lightRecords;
    std::vector<InstanceRecord> instRecords;

    void flatten();
};

// Read a PFM image (linear RGB float) into top-row-first RGB texels.
// an environment map (EXR, HDR, PFM: linear; PNG: sRGB-encoded) as linear RGB rows, top first
void load_env_image(const std::string &path, std::vector<float> &rgb, uint32_t &w, uint32_t &h);

// fn relative to dir unless absolute or dir is empty
std::string join_path(const std::string &dir, const std::string &fn);

// the index of the image an image shader node names in s.textures (loading
// it on first use); colour_space "srgb" decodes 8/16-bit values, "linear"
// keeps them, unset means linear for .exr / .hdr / .pfm and sRGB otherwise
// (ImagePool::load_texture, src/base/mgr/image_pool.cpp:13-34)
uint32_t texture_index(SceneHost &s, const Json &node, bool illum = false);
// the index of an already loaded image node
uint32_t find_texture(const SceneHost &s, const Json &node);
// the image node of a parameter slot, or nullptr for constants
const Json *slot_image_node(const Json &param, const char *key);

// Parse a scene from JSON text; base_dir resolves relative mesh/texture paths.
SceneHost load_scene(const std::string &text, const std::string &base_dir);
std::string read_text_file(const std::string &path);

// c2w of a camera from the reference's yaw/pitch/position re-derivation
// (src/base/sensor/sensor.cpp:78-83,163-172).
m4 camera_to_world(f3 position, float yaw_deg, float pitch_deg);

}// namespace vs
