#include "tables.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace vs {

namespace {
constexpr float kPiF = 3.14159265358979323846f;
float luminance(f3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }

float slot1(const Json &p, const char *key, float def) {
    const Json &j = p[key];
    if (j.is_null()) return def;
    const Json *node = &j;
    if (j.is_object() && j.contains("channels")) node = &j["node"];
    std::vector<double> v;
    if (node->is_number() || node->is_array()) v = node->as_numbers();
    else if (node->is_object() && node->contains("param")) v = (*node)["param"]["value"].as_numbers();
    else if (node->is_object()) v = (*node)["value"].as_numbers();
    return v.empty() ? def : float(v[0]);
}
f3 slot3(const Json &p, const char *key, f3 def) {
    const Json &j = p[key];
    if (j.is_null()) return def;
    const Json *node = &j;
    if (j.is_object() && j.contains("channels")) node = &j["node"];
    std::vector<double> v;
    if (node->is_number() || node->is_array()) v = node->as_numbers();
    else if (node->is_object() && node->contains("param")) v = (*node)["param"]["value"].as_numbers();
    else if (node->is_object()) {
        if (node->contains("fn")) throw std::runtime_error(std::string("textured slot '") + key + "' is not supported");
        v = (*node)["value"].as_numbers();
    }
    if (v.size() >= 3) return mk3(float(v[0]), float(v[1]), float(v[2]));
    if (v.size() == 1) return mk3(float(v[0]));
    return def;
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

f3 light_power(const LightHost &l) {
    f3 avg = l.color * l.scale;
    switch (l.type) {
        case LIGHT_AREA: {
            float area = 0.f;
            for (float a : l.tri_areas) area += a;
            return avg * ((l.two_sided ? 2.f : 1.f) * area * kPiF);
        }
        case LIGHT_POINT: return avg * (4.f * kPiF);
        case LIGHT_SPOT: return avg * (2.f * kPiF * (1.f - .5f * (l.angle * 2.f + l.falloff)));
        default: return mk3(0.f);
    }
}

std::vector<MaterialRecord> pack_materials(const SceneHost &s) {
    std::vector<MaterialRecord> out;
    for (const MaterialHost &m : s.materials) {
        MaterialRecord r{};
        const Json &p = m.param;
        if (m.type == "diffuse") {
            // DiffuseMaterial::initialize_slots (diffuse.cpp:38-46): color default 0.5
            f3 c = slot3(p, "color", mk3(0.5f));
            r.type = MAT_DIFFUSE;
            r.p[0] = c.x;
            r.p[1] = c.y;
            r.p[2] = c.z;
            if (p.contains("sigma")) {
                r.type = MAT_OREN_NAYAR;
                r.p[3] = std::clamp(slot1(p, "sigma", 0.5f), 0.f, 1.f);
            }
        } else if (m.type == "black_body") {
            r.type = MAT_BLACK;
        } else {
            throw std::runtime_error("unsupported material type: " + m.type);
        }
        out.push_back(r);
    }
    return out;
}

LightTables pack_lights(const SceneHost &s) {
    LightTables t;
    if (s.lights.empty()) throw std::runtime_error("scene has no lights");
    for (const LightHost &l : s.lights) {
        LightRecord r{};
        r.type = l.type;
        r.two_sided = l.two_sided;
        r.inst = l.inst;
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
        if (l.type == LIGHT_AREA) {
            AliasTableHost a = build_alias(l.tri_areas);
            r.tri_base = l.tri_base;
            r.tri_count = uint32_t(l.tri_areas.size());
            r.alias_offset = uint32_t(t.alias.size());
            r.alias_integral = a.integral;
            t.alias.insert(t.alias.end(), a.table.begin(), a.table.end());
            t.func.insert(t.func.end(), a.func.begin(), a.func.end());
        } else if (l.type == LIGHT_ENV) {
            throw std::runtime_error("environment lights are not supported yet");
        }
        t.records.push_back(r);
    }
    if (s.light_sampler_type == "power") {
        std::vector<float> w;
        for (const LightHost &l : s.lights) w.push_back(luminance(light_power(l)));
        AliasTableHost a = build_alias(w);
        t.select_offset = uint32_t(t.alias.size());
        t.select_integral = a.integral;
        t.alias.insert(t.alias.end(), a.table.begin(), a.table.end());
        t.func.insert(t.func.end(), a.func.begin(), a.func.end());
    }
    return t;
}

}// namespace vs
