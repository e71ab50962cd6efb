#include "tables.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <stdexcept>

#include "data_tables.h"

namespace vs {

namespace {
constexpr float kPiF = 3.14159265358979323846f;
constexpr float kPeak[3] = {602.785f, 539.285f, 445.772f};// rgb_spectrum_peak_wavelengths (spectrum.h:14)

float luminance(f3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }

const Json *slot_node(const Json &p, const char *key) {
    const Json &j = p[key];
    if (j.is_null()) return nullptr;
    if (j.is_object() && j.contains("channels")) return &j["node"];
    return &j;
}
std::vector<double> slot_numbers(const Json &p, const char *key) {
    const Json *node = slot_node(p, key);
    if (!node) return {};
    if (node->is_number() || node->is_array()) return node->as_numbers();
    if (node->is_object()) {
        if (node->contains("fn") || (*node)["type"].as_string("number") == "image")
            throw std::runtime_error(std::string("textured slot '") + key + "' is not supported");
        if (node->contains("param")) return (*node)["param"]["value"].as_numbers();
        return (*node)["value"].as_numbers();
    }
    return {};
}
float slot1(const Json &p, const char *key, float def) {
    auto v = slot_numbers(p, key);
    return v.empty() ? def : float(v[0]);
}
f3 slot3(const Json &p, const char *key, f3 def) {
    auto v = slot_numbers(p, key);
    if (v.size() >= 3) return mk3(float(v[0]), float(v[1]), float(v[2]));
    if (v.size() == 1) return mk3(float(v[0]));
    return def;
}

uint32_t fbits_u(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
float ubits_f(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// calculate_alpha (src/base/scattering/microfacet.h:44-58)
void calculate_alpha(float alpha, float aniso, float &ax, float &ay) {
    if (std::fabs(aniso) <= 1e-4f) {
        ax = ay = alpha;
        return;
    }
    ax = aniso < 0 ? alpha / (1.f + aniso) : alpha * (1.f - aniso);
    ay = aniso < 0 ? alpha * (1.f + aniso) : alpha / (1.f - aniso);
}

// glass ior curves (glass.cpp:104-134) plus the Cauchy curve n = A + B / lambda_um^2
float ior_curve(const std::string &name, const Json &p, float lambda_nm) {
    float l = lambda_nm / 1000.f;
    float l2 = l * l;
    if (name == "LASF9") {
        float f = 2.00029547f * l2 / (l2 - 0.0121426017f) + 0.298926886f * l2 / (l2 - 0.0538736236f) +
                  1.80691843f * l2 / (l2 - 156.530829f);
        return std::sqrt(f + 1);
    }
    if (name == "Cauchy") {
        float A = float(p["cauchy_a"].as_float(1.5)), B = float(p["cauchy_b"].as_float(0.004));
        return A + B / l2;
    }
    float f = 1.03961212f * l2 / (l2 - 0.00600069867f) + 0.231792344f * l2 / (l2 - 0.0200179144f) +
              1.01046945f * l2 / (l2 - 103.560653f);
    return std::sqrt(f + 1);
}

// SPD::to_list (spd.h:40-47): samples every 5 nm from 360 while < 830
std::vector<float> curve_to_list(const std::string &name, const Json &p) {
    std::vector<float> r;
    for (float lambda = 360.f; lambda < 830.f; lambda += 5.f) r.push_back(ior_curve(name, p, lambda));
    return r;
}

void put_slot(MaterialRecord &r, int off, f3 rgb, bool is_rgb) {
    if (is_rgb) {
        r.p[off] = rgb.x;
        r.p[off + 1] = rgb.y;
        r.p[off + 2] = rgb.z;
        r.p[off + 3] = 0.f;
    } else {
        f4 c = rgb_to_albedo_coeffs(rgb);
        r.p[off] = c.x;
        r.p[off + 1] = c.y;
        r.p[off + 2] = c.z;
        r.p[off + 3] = c.w;
    }
}

uint32_t append_curve(std::vector<float> &spd, const std::vector<float> &c) {
    uint32_t off = uint32_t(spd.size());
    spd.insert(spd.end(), c.begin(), c.end());
    return off;
}

// a colour slot at p[off]: a constant, or the index of the image it names
void put_color_slot(MaterialRecord &r, int off, const SceneHost &s, const Json &p, const char *key, f3 def,
                    bool is_rgb) {
    if (const Json *node = slot_image_node(p, key)) {
        r.p[off] = ubits_f(find_texture(s, *node));
        r.p[off + 1] = r.p[off + 2] = r.p[off + 3] = 0.f;
        r.flags |= MATF_TEXTURE_SLOT0 << (off / 4);
        return;
    }
    put_slot(r, off, slot3(p, key, def), is_rgb);
}

MaterialRecord pack_material(const MaterialHost &m, bool rgb, std::vector<float> &spd, const SceneHost &s) {
    MaterialRecord r{};
    const Json &p = m.param;
    if (!m.unsupported.empty()) {
        // listed but used by no shape (the loader refuses a shape that uses
        // it): a black body keeps the material indices in place
        r.type = MAT_BLACK;
    } else if (m.type == "diffuse") {
        // DiffuseMaterial::initialize_slots (diffuse.cpp:38-46)
        r.type = MAT_DIFFUSE;
        put_color_slot(r, 0, s, p, "color", mk3(0.5f), rgb);
        if (p.contains("sigma")) {
            r.type = MAT_OREN_NAYAR;
            r.p[4] = std::clamp(slot1(p, "sigma", 0.5f), 0.f, 1.f);
        }
    } else if (m.type == "black_body") {
        r.type = MAT_BLACK;
    } else if (m.type == "metal") {
        // MetalMaterial (metal.cpp:39-157)
        r.type = MAT_METAL;
        float rough = std::clamp(slot1(p, "roughness", 0.01f), 0.0001f, 1.f);
        float aniso = std::clamp(slot1(p, "anisotropic", 0.f), -0.9f, 0.9f);
        if (p["remapping_roughness"].as_bool(true)) rough = rough * rough;
        calculate_alpha(rough, aniso, r.p[6], r.p[7]);
        std::vector<float> eta, k;
        metal_ior(p["material_name"].as_string(""), eta, k, nullptr);
        for (int c = 0; c < 3; ++c) {
            r.p[c] = spd_eval_host(eta, kPeak[c]);
            r.p[3 + c] = spd_eval_host(k, kPeak[c]);
        }
        if (!rgb) {
            r.p[8] = ubits_f(append_curve(spd, eta));
            r.p[9] = ubits_f(append_curve(spd, k));
            r.p[10] = 471.f / float(eta.size());
        }
    } else if (m.type == "mirror") {
        // MirrorMaterial (src/render_core/materials/mirror.cpp:26-72): colour
        // slot, roughness 0.001 remapped to alpha, anisotropy
        r.type = MAT_MIRROR;
        put_color_slot(r, 0, s, p, "color", mk3(1.f), rgb);
        float rough = std::clamp(slot1(p, "roughness", 0.001f), 0.0001f, 1.f);
        float aniso = std::clamp(slot1(p, "anisotropic", 0.f), -0.9f, 0.9f);
        if (p["remapping_roughness"].as_bool(true)) rough = rough * rough;
        calculate_alpha(rough, aniso, r.p[6], r.p[7]);
    } else if (m.type == "glass") {
        // GlassMaterial (glass.cpp:171-258)
        r.type = MAT_GLASS;
        put_color_slot(r, 0, s, p, "color", mk3(1.f), rgb);
        float rough = std::clamp(slot1(p, "roughness", 0.5f), 0.01f, 1.f);
        float aniso = std::clamp(slot1(p, "anisotropic", 0.f), -0.9f, 0.9f);
        if (p["remapping_roughness"].as_bool(true)) rough = rough * rough;
        calculate_alpha(rough, aniso, r.p[5], r.p[6]);
        std::string name = p["material_name"].as_string("");
        if (name.empty()) {
            r.p[4] = slot1(p, "ior", 1.5f);
        } else {
            if (name != "BK7" && name != "LASF9" && name != "Cauchy") name = "BK7";
            r.p[4] = ior_curve(name, p, kPeak[0]);
            if (!rgb) {
                std::vector<float> curve = curve_to_list(name, p);
                r.p[7] = ubits_f(append_curve(spd, curve));
                r.p[8] = 471.f / float(curve.size());
                r.flags |= 1u;// dispersive
            }
        }
    } else if (m.type == "principled_bsdf") {
        // PrincipledBSDF::initialize_slots (principled_bsdf.cpp:281-307)
        r.type = MAT_PRINCIPLED;
        put_color_slot(r, 0, s, p, "color", mk3(1.f), rgb);
        put_color_slot(r, 4, s, p, "spec_tint", mk3(1.f), rgb);
        put_color_slot(r, 8, s, p, "sheen_tint", mk3(1.f), rgb);
        put_color_slot(r, 12, s, p, "coat_tint", mk3(1.f), rgb);
        r.p[16] = slot1(p, "metallic", 0.f);
        r.p[17] = slot1(p, "ior", 1.5f);
        r.p[18] = slot1(p, "roughness", 0.5f);
        r.p[19] = slot1(p, "anisotropic", 0.f);
        r.p[20] = slot1(p, "sheen_weight", 0.f);
        r.p[21] = slot1(p, "sheen_roughness", 0.5f);
        r.p[22] = slot1(p, "coat_weight", 0.f);
        r.p[23] = slot1(p, "coat_roughness", 0.2f);
        r.p[24] = slot1(p, "coat_ior", 1.5f);
        r.p[25] = slot1(p, "transmission_weight", 0.f);
    } else if (m.type == "mix" || m.type == "add") {
        // MixMaterial / AddMaterial (mix.cpp, add.cpp): the component records
        // are packed after the scene's own materials (pack_scene fills their
        // indices into p[0], p[1]); the mix fraction slot defaults to 0.5
        r.type = m.type == "mix" ? MAT_MIX : MAT_ADD;
        r.p[2] = m.type == "mix" ? slot1(p, "frac", 0.5f) : 0.f;
    } else {
        throw std::runtime_error("unsupported material type: " + m.type);
    }
    return r;
}

}// namespace

AliasTableHost build_alias(const std::vector<float> &weights) {
    AliasTableHost out;
    const size_t n = weights.size();
    if (n == 0) return out;
    double sum = std::accumulate(weights.begin(), weights.end(), 0.0);
    double ratio = double(n) / sum;
    std::vector<uint32_t> over, under;
    out.table.resize(n);
    for (size_t i = 0; i < n; ++i) {
        float p = static_cast<float>(weights[i] * ratio);
        out.table[i] = {p, uint32_t(i)};
        (p > 1.0f ? over : under).push_back(uint32_t(i));
    }
    while (!over.empty() && !under.empty()) {
        uint32_t o = over.back(), u = under.back();
        over.pop_back();
        under.pop_back();
        out.table[o].prob -= 1.0f - out.table[u].prob;
        out.table[u].alias = o;
        if (out.table[o].prob > 1.0f) over.push_back(o);
        else if (out.table[o].prob < 1.0f) under.push_back(o);
    }
    for (uint32_t i : over) out.table[i] = {1.0f, i};
    for (uint32_t i : under) out.table[i] = {1.0f, i};
    out.integral = float(sum / double(n));
    out.func = weights;
    return out;
}

f3 light_power(const LightHost &l, float world_radius) {
    f3 avg = l.color * l.scale;
    switch (l.type) {
        case LIGHT_AREA: {
            float area = 0.f;
            for (float a : l.tri_areas) area += a;
            return avg * ((l.two_sided ? 2.f : 1.f) * area * kPiF);
        }
        case LIGHT_POINT: return avg * (4.f * kPiF);
        case LIGHT_SPOT: return avg * (2.f * kPiF * (1.f - .5f * (l.angle * 2.f + l.falloff)));
        case LIGHT_ENV: {
            double acc[3] = {0, 0, 0};
            size_t n = size_t(l.env_w) * l.env_h;
            for (size_t i = 0; i < n; ++i)
                for (int c = 0; c < 3; ++c) acc[c] += l.env_rgb[3 * i + c];
            // Pi * sqr(world_radius) * average(), average = mean texel * scale
            f3 mean = n ? mk3(float(acc[0] / double(n)), float(acc[1] / double(n)), float(acc[2] / double(n))) : mk3(0.f);
            return (mean * l.scale) * (kPiF * (world_radius * world_radius));
        }
        default: return mk3(0.f);
    }
}

PackedScene pack_scene(const SceneHost &s) {
    PackedScene out;
    out.rgb = s.spectrum_type == "srgb";
    if (!out.rgb) {
        if (s.spectrum_dim != 3 && s.spectrum_dim != 4)
            throw std::runtime_error("unsupported hero spectrum dimension (3 or 4 are supported)");
        if (!rgb2spec_available()) throw std::runtime_error("rgb2spec_srgb.f32 is missing: run vision_amd/build.py");
    }
    const CieSpd &cie = cie_spd();
    out.spd.insert(out.spd.end(), cie.x.begin(), cie.x.end());
    out.spd.insert(out.spd.end(), cie.y.begin(), cie.y.end());
    out.spd.insert(out.spd.end(), cie.z.begin(), cie.z.end());
    out.spd.insert(out.spd.end(), cie.d65.begin(), cie.d65.end());
    out.cie_interval = cie.interval;
    out.cie_y_integral = cie.y_integral;
    for (const MaterialHost &m : s.materials) out.mats.push_back(pack_material(m, out.rgb, out.spd, s));
    // components of mix / add materials follow the scene's own materials
    for (size_t i = 0; i < s.materials.size(); ++i) {
        const MaterialHost &m = s.materials[i];
        if ((m.type != "mix" && m.type != "add") || !m.unsupported.empty()) continue;
        for (int k = 0; k < 2; ++k) {
            const Json &c = m.param[k == 0 ? "mat0" : "mat1"];
            MaterialHost ch;
            ch.type = c["type"].as_string("diffuse");
            ch.name = c["name"].as_string("");
            ch.param = c["param"];
            out.mats[i].p[k] = ubits_f(uint32_t(out.mats.size()));
            out.mats.push_back(pack_material(ch, out.rgb, out.spd, s));
        }
    }
    for (const TextureHost &t : s.textures) {
        out.tex_info.insert(out.tex_info.end(), {uint32_t(out.tex.size() / 4), t.width, t.height, t.illum ? 1u : 0u});
        for (size_t i = 0; i < size_t(t.width) * t.height; ++i) {
            f3 c = mk3(t.rgb[3 * i], t.rgb[3 * i + 1], t.rgb[3 * i + 2]);
            if (out.rgb) {
                out.tex.insert(out.tex.end(), {c.x, c.y, c.z, 0.f});
            } else if (t.illum) {
                // emission texels: RGBIlluminationSpectrum coefficients
                f4 q = rgb_to_unbound_coeffs(c);
                out.tex.insert(out.tex.end(), {q.x, q.y, q.z, q.w});
            } else {
                f4 q = rgb_to_albedo_coeffs(c);
                out.tex.insert(out.tex.end(), {q.x, q.y, q.z, q.w});
            }
        }
    }

    f3 diag = s.aabb_hi - s.aabb_lo;
    out.world_radius = std::max(0.5f * length(diag), s.min_world_radius);
    if (s.lights.empty()) throw std::runtime_error("scene has no lights");
    for (size_t li = 0; li < s.lights.size(); ++li) {
        const LightHost &l = s.lights[li];
        LightRecord r{};
        r.type = l.type;
        r.two_sided = l.two_sided;
        r.inst = l.inst;
        r.tex = l.tex >= 0 ? uint32_t(l.tex) : 0xFFFFFFFFu;
        r.color[0] = l.color.x;
        r.color[1] = l.color.y;
        r.color[2] = l.color.z;
        r.scale = l.scale;
        r.position[0] = l.position.x;
        r.position[1] = l.position.y;
        r.position[2] = l.position.z;
        r.direction[0] = l.direction.x;
        r.direction[1] = l.direction.y;
        r.direction[2] = l.direction.z;
        r.angle = l.angle;
        r.falloff = l.falloff;
        r.cos_angle = std::cos(l.angle);
        r.cos_falloff_start = std::cos(std::max(0.f, l.angle - l.falloff));
        if (!out.rgb) {
            f4 c = rgb_to_unbound_coeffs(l.color);
            r.illum[0] = c.x;
            r.illum[1] = c.y;
            r.illum[2] = c.z;
            r.illum[3] = c.w;
        }
        if (l.type == LIGHT_AREA) {
            AliasTableHost a = build_alias(l.tri_areas);
            r.tri_base = l.tri_base;
            r.tri_count = uint32_t(l.tri_areas.size());
            r.alias_offset = uint32_t(out.alias.size());
            r.alias_integral = a.integral;
            out.alias.insert(out.alias.end(), a.table.begin(), a.table.end());
            out.func.insert(out.func.end(), a.func.begin(), a.func.end());
        } else if (l.type == LIGHT_ENV) {
            if (out.env_light != kInvalid) throw std::runtime_error("only one environment light is supported");
            out.env_light = uint32_t(li);
            out.env_w = l.env_w;
            out.env_h = l.env_h;
            out.env_scale = l.scale;
            size_t n = size_t(l.env_w) * l.env_h;
            out.env_tex.resize(4 * n);
            std::vector<float> weights(n);
            for (size_t i = 0; i < n; ++i) {
                f3 c = mk3(l.env_rgb[3 * i], l.env_rgb[3 * i + 1], l.env_rgb[3 * i + 2]);
                if (out.rgb) {
                    out.env_tex[4 * i] = c.x;
                    out.env_tex[4 * i + 1] = c.y;
                    out.env_tex[4 * i + 2] = c.z;
                    out.env_tex[4 * i + 3] = 0.f;
                } else {
                    f4 k = rgb_to_unbound_coeffs(c);
                    out.env_tex[4 * i] = k.x;
                    out.env_tex[4 * i + 1] = k.y;
                    out.env_tex[4 * i + 2] = k.z;
                    out.env_tex[4 * i + 3] = k.w;
                }
                // SphericalMap::calculate_weights (spherical.cpp:170-196)
                float v = float(int(i) / int(l.env_h)) + 0.5f;
                float theta = v / float(l.env_w);
                weights[i] = l.env_constant ? 1.f : luminance(c) * std::sin(kPiF * theta);
            }
            // AliasTable2D::build (alias2d.cpp:32-71)
            std::vector<float> marg(l.env_h);
            out.env_cond.reserve(n);
            for (uint32_t y = 0; y < l.env_h; ++y) {
                std::vector<float> row(weights.begin() + size_t(y) * l.env_w, weights.begin() + size_t(y + 1) * l.env_w);
                AliasTableHost a = build_alias(row);
                marg[y] = a.integral;
                out.env_cond.insert(out.env_cond.end(), a.table.begin(), a.table.end());
                out.env_cond_func.insert(out.env_cond_func.end(), a.func.begin(), a.func.end());
            }
            AliasTableHost m = build_alias(marg);
            out.env_marg = m.table;
            out.env_marg_func = m.func;
            out.env_marg_integral = m.integral;
            // o2w = o2w_desc * rotation_x(-90) * scale(1, flip_u ? 1 : -1, 1)
            double R[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) R[i][j] = m4_at(l.env_o2w, i, j);
            double fx = l.flip_u ? 1.0 : -1.0;
            // rotation_x(-90): y -> -z, z -> y ; then scale y by fx
            double B[3][3] = {{1, 0, 0}, {0, 0, 1 * 1.0}, {0, -1, 0}};
            double M[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double acc = 0;
                    for (int k = 0; k < 3; ++k) acc += R[i][k] * B[k][j];
                    M[i][j] = acc * (j == 1 ? fx : 1.0);
                }
            double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                         M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                         M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
            double inv[3][3] = {{(M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det, (M[0][2] * M[2][1] - M[0][1] * M[2][2]) / det,
                                 (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / det},
                                {(M[1][2] * M[2][0] - M[1][0] * M[2][2]) / det, (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / det,
                                 (M[0][2] * M[1][0] - M[0][0] * M[1][2]) / det},
                                {(M[1][0] * M[2][1] - M[1][1] * M[2][0]) / det, (M[0][1] * M[2][0] - M[0][0] * M[2][1]) / det,
                                 (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / det}};
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    out.env_o2w[3 * i + j] = float(M[i][j]);
                    out.env_w2o[3 * i + j] = float(inv[i][j]);
                }
        }
        out.lights.push_back(r);
    }
    if (s.light_sampler_type == "power") {
        std::vector<float> w;
        for (const LightHost &l : s.lights) w.push_back(luminance(light_power(l, out.world_radius)));
        AliasTableHost a = build_alias(w);
        out.select_offset = uint32_t(out.alias.size());
        out.select_integral = a.integral;
        out.alias.insert(out.alias.end(), a.table.begin(), a.table.end());
        out.func.insert(out.func.end(), a.func.begin(), a.func.end());
    }
    (void)fbits_u;
    return out;
}

int filter_kind(const CameraHost &c) {
    if (c.filter_type == "box") return 0;
    if (c.filter_type == "triangle") return 1;
    if (c.filter_type == "gaussian") return 2;
    if (c.filter_type == "mitchell") return 3;
    if (c.filter_type == "sinc") return 4;
    throw std::runtime_error("unsupported filter: " + c.filter_type);
}

namespace {

// gaussian (src/math/util.h:36-43)
float gaussian(float x, float mu, float sigma) {
    return 1.f / std::sqrt(2 * kPiF * (sigma * sigma)) * std::exp(-((x - mu) * (x - mu)) / (2 * (sigma * sigma)));
}
// MitchellFilter::mitchell_1d (mitchell.cpp:22-35)
float mitchell_1d(float x, float b, float c) {
    x = std::fabs(x);
    if (x <= 1)
        return ((12 - 9 * b - 6 * c) * (x * x * x) + (-18 + 12 * b + 6 * c) * (x * x) + (6 - 2 * b)) * (1.f / 6.f);
    if (x <= 2)
        return ((-b - 6 * c) * (x * x * x) + (6 * b + 30 * c) * (x * x) + (-12 * b - 48 * c) * x + (8 * b + 24 * c)) *
               (1.f / 6.f);
    return 0;
}
// windowed_sinc (src/math/util.h:51-66)
float sin_x_over_x(float x) { return 1 + x * x == 1 ? 1.f : std::sin(x) / x; }
float sinc(float x) { return sin_x_over_x(kPiF * x); }
float windowed_sinc(float x, float radius, float tau) {
    return std::fabs(x) > radius ? 0.f : sinc(x) * sinc(x / tau);
}

}// namespace

// Filter::evaluate of gaussian.cpp:30-34, mitchell.cpp:37-40, sinc.cpp:21-24
float filter_evaluate(const CameraHost &c, float x, float y) {
    const float r = c.filter_radius;
    switch (filter_kind(c)) {
        case 2: {
            float ex = gaussian(r, 0, c.filter_sigma);
            float vx = gaussian(x, 0, c.filter_sigma) - ex;
            float vy = gaussian(y, 0, c.filter_sigma) - ex;
            return std::max(0.f, vx) * std::max(0.f, vy);
        }
        case 3: return mitchell_1d(2 * x / r, c.filter_b, c.filter_c) * mitchell_1d(2 * y / r, c.filter_b, c.filter_c);
        case 4: return windowed_sinc(x, r, c.filter_tau) * windowed_sinc(y, r, c.filter_tau) * 4;
        default: return 1.f;
    }
}

// FilterSampler::build (fitted_curve.h:33-53)
FilterTable build_filter_table(const CameraHost &c) {
    FilterTable t;
    const int n = kFilterTableSize, len = n * n;
    t.lut.resize(len);
    std::vector<float> func(len);
    const float r = c.filter_radius;
    for (int i = 0; i < len; ++i) {
        int x = i % n, y = i / n;
        float px = (x + 0.5f) / n * r, py = (y + 0.5f) / n * r;
        float val = filter_evaluate(c, px, py);
        func[i] = std::fabs(val);
        t.lut[i] = val;
    }
    float sum = 0.f;
    for (float v : t.lut) sum += v;
    float integral = sum / len;
    for (float &v : t.lut) v = v / integral;
    // AliasTable2D::build (alias2d.cpp:32-71)
    std::vector<float> marg(n);
    for (int y = 0; y < n; ++y) {
        std::vector<float> row(func.begin() + y * n, func.begin() + (y + 1) * n);
        AliasTableHost a = build_alias(row);
        marg[y] = a.integral;
        t.cond.insert(t.cond.end(), a.table.begin(), a.table.end());
        t.cond_func.insert(t.cond_func.end(), a.func.begin(), a.func.end());
    }
    AliasTableHost m = build_alias(marg);
    t.marg = m.table;
    t.marg_func = m.func;
    t.marg_integral = m.integral;
    return t;
}

}// namespace vs
