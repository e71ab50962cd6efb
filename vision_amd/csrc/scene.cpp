#include "scene.h"

#include "image_io.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace vs {

std::string join_path(const std::string &dir, const std::string &fn) {
    if (fn.empty() || fn[0] == '/' || dir.empty()) return fn;
    return dir + "/" + fn;
}

namespace {

constexpr float kPi = 3.14159265358979323846f;

float radians(float deg) { return deg * (kPi / 180.f); }
float degrees(float rad) { return rad * (180.f / kPi); }

m4 mat_mul(const m4 &a, const m4 &b) {
    m4 r;
    for (int col = 0; col < 4; ++col) {
        float v[4];
        for (int row = 0; row < 4; ++row) {
            float s = 0.f;
            for (int k = 0; k < 4; ++k) s += m4_at(a, row, k) * m4_at(b, k, col);
            v[row] = s;
        }
        r.c[col] = mk4(v[0], v[1], v[2], v[3]);
    }
    return r;
}

m4 translation(f3 t) {
    m4 m = m4_identity();
    m.c[3] = mk4(t.x, t.y, t.z, 1.f);
    return m;
}
m4 scale_m(f3 s) {
    m4 m = m4_identity();
    m.c[0].x = s.x;
    m.c[1].y = s.y;
    m.c[2].z = s.z;
    return m;
}
// right-handed rotations about the principal axes, angle in degrees
m4 rotation_x(float deg) {
    float c = std::cos(radians(deg)), s = std::sin(radians(deg));
    m4 m = m4_identity();
    m.c[1] = mk4(0, c, s, 0);
    m.c[2] = mk4(0, -s, c, 0);
    return m;
}
m4 rotation_y(float deg) {
    float c = std::cos(radians(deg)), s = std::sin(radians(deg));
    m4 m = m4_identity();
    m.c[0] = mk4(c, 0, -s, 0);
    m.c[2] = mk4(s, 0, c, 0);
    return m;
}
m4 rotation_z(float deg) {
    float c = std::cos(radians(deg)), s = std::sin(radians(deg));
    m4 m = m4_identity();
    m.c[0] = mk4(c, s, 0, 0);
    m.c[1] = mk4(-s, c, 0, 0);
    return m;
}
m4 rotation_axis(f3 axis, float deg) {
    f3 a = normalize(axis);
    float c = std::cos(radians(deg)), s = std::sin(radians(deg)), t = 1 - c;
    m4 m = m4_identity();
    m.c[0] = mk4(t * a.x * a.x + c, t * a.x * a.y + s * a.z, t * a.x * a.z - s * a.y, 0);
    m.c[1] = mk4(t * a.x * a.y - s * a.z, t * a.y * a.y + c, t * a.y * a.z + s * a.x, 0);
    m.c[2] = mk4(t * a.x * a.z + s * a.y, t * a.y * a.z - s * a.x, t * a.z * a.z + c, 0);
    return m;
}

f3 as_f3(const Json &j, f3 def) {
    auto v = j.as_numbers();
    if (v.size() >= 3) return mk3(float(v[0]), float(v[1]), float(v[2]));
    if (v.size() == 1) return mk3(float(v[0]));
    return def;
}

// look_at: camera columns (right, up, forward, position) with
// forward = normalize(target - position), right = normalize(forward x up).
m4 look_at(f3 pos, f3 target, f3 up) {
    f3 fwd = normalize(target - pos);
    f3 right = normalize(cross(fwd, up));
    f3 up2 = cross(right, fwd);
    m4 m;
    m.c[0] = mk4(right.x, right.y, right.z, 0);
    m.c[1] = mk4(up2.x, up2.y, up2.z, 0);
    m.c[2] = mk4(fwd.x, fwd.y, fwd.z, 0);
    m.c[3] = mk4(pos.x, pos.y, pos.z, 1);
    return m;
}

// TransformDesc::init (src/core/node_desc.cpp:31-58)
m4 parse_transform(const Json &t) {
    if (t.is_null()) return m4_identity();
    std::string type = t["type"].as_string("matrix4x4");
    const Json &p = t["param"];
    if (type == "look_at") {
        return look_at(as_f3(p["position"], mk3(0.f)), as_f3(p["target_pos"], mk3(0, 0, 1)),
                       as_f3(p["up"], mk3(0, 1, 0)));
    }
    if (type == "Euler") {
        m4 yaw = rotation_y(float(p["yaw"].as_float(0)));
        m4 roll = rotation_z(float(p["roll"].as_float(0)));
        m4 pitch = rotation_x(float(p["pitch"].as_float(0)));
        m4 tt = translation(as_f3(p["position"], mk3(0.f)));
        return mat_mul(mat_mul(mat_mul(tt, pitch), roll), yaw);
    }
    if (type == "trs") {
        f3 tr = as_f3(p["t"], mk3(0.f));
        auto r = p["r"].as_numbers();
        f3 s = as_f3(p["s"], mk3(1.f));
        m4 rot = m4_identity();
        if (r.size() == 4 && (r[0] != 0 || r[1] != 0 || r[2] != 0))
            rot = rotation_axis(mk3(float(r[0]), float(r[1]), float(r[2])), float(r[3]));
        return mat_mul(mat_mul(translation(tr), rot), scale_m(s));
    }
    if (type == "matrix4x4") {
        const Json &mj = p["matrix4x4"];
        m4 m = m4_identity();
        if (mj.is_array() && mj.size() == 4) {
            for (int col = 0; col < 4; ++col) {
                auto v = mj.at(col).as_numbers();
                if (v.size() != 4) throw std::runtime_error("matrix4x4 column must have 4 entries");
                m.c[col] = mk4(float(v[0]), float(v[1]), float(v[2]), float(v[3]));
            }
        }
        return m;
    }
    throw std::runtime_error("transform type error " + type);
}

// Slot value resolution (SlotDesc::init, ShaderNodeDesc::init,
// src/core/node_desc.cpp:126-147,318-345).  Only constant ("number") nodes
// are supported; image/texture nodes raise.
bool slot_values(const Json &slot, std::vector<float> &out) {
    if (slot.is_null()) return false;
    const Json *node = &slot;
    if (slot.is_object() && slot.contains("channels")) node = &slot["node"];
    std::vector<double> v;
    if (node->is_array() || node->is_number()) {
        v = node->as_numbers();
    } else if (node->is_object()) {
        if (node->contains("param")) {
            std::string t = (*node)["type"].as_string("");
            if (t != "number") throw std::runtime_error("unsupported shader node type: " + t);
            v = (*node)["param"]["value"].as_numbers();
        } else {
            std::string t = (*node)["type"].as_string("image");
            if (t == "number") v = (*node)["value"].as_numbers();
            else throw std::runtime_error("unsupported shader node (textures are not supported): " + t);
        }
    } else {
        return false;
    }
    out.clear();
    for (double d : v) out.push_back(float(d));
    return !out.empty();
}

f3 slot_f3(const Json &slot, f3 def) {
    std::vector<float> v;
    if (!slot_values(slot, v)) return def;
    if (v.size() >= 3) return mk3(v[0], v[1], v[2]);
    return mk3(v[0]);
}

float slot_f1(const Json &slot, float def) {
    std::vector<float> v;
    if (!slot_values(slot, v)) return def;
    return v[0];
}

// NumberArray::normalize (src/render_core/shadernodes/number.cpp:40-49):
// divides by the max component when it is >= 1 and returns the factor.
float normalize_color(f3 &c) {
    float m = std::max(c.x, std::max(c.y, c.z));
    if (m < 1.f) return 1.f;
    c = mk3(c.x / m, c.y / m, c.z / m);
    return m;
}

void add_vertex(MeshHost &m, f3 p, f3 n, f2 uv) {
    m.P.push_back(p);
    m.N.push_back(n);
    m.UV.push_back(uv);
}

// Quad (src/render_core/shapes/quad.cpp:21-50)
MeshHost make_quad(float width, float height) {
    MeshHost m;
    float w = width / 2, h = height / 2;
    f3 P[4] = {mk3(w, 0, h), mk3(w, 0, -h), mk3(-w, 0, h), mk3(-w, 0, -h)};
    f2 UV[4] = {mk2(1, 1), mk2(1, 0), mk2(0, 1), mk2(0, 0)};
    f3 ng_un = cross(P[0] - P[2], P[1] - P[2]);
    for (int i = 0; i < 4; ++i) add_vertex(m, P[i], ng_un, UV[i]);
    m.tris = {{0, 1, 2}, {2, 1, 3}};
    return m;
}

// Sphere (src/render_core/shapes/sphere.cpp:19-86): a latitude-longitude
// mesh of theta_div rings and 2 theta_div segments; the vertex and triangle
// order is the reference's, so triangle ids and uvs agree
MeshHost make_sphere(float radius, uint32_t theta_div) {
    if (theta_div < 2 || theta_div > 4096) throw std::runtime_error("sphere sub_div must be in 2..4096");
    const uint32_t phi_div = 2 * theta_div;
    MeshHost m;
    add_vertex(m, mk3(0, radius, 0), mk3(0, 1, 0), mk2(0, 0));
    for (uint32_t i = 1; i < theta_div; ++i) {
        const float v = float(i) / float(theta_div);
        const float theta = float(kPi) * v;
        const float y = radius * std::cos(theta), r = radius * std::sin(theta);
        for (uint32_t j = 0; j < phi_div; ++j) {
            const float u = float(j) / float(phi_div);
            const float phi = u * (2.f * float(kPi));
            const f3 p = j == 0 ? mk3(r, y, 0.f) : mk3(std::cos(phi) * r, y, std::sin(phi) * r);
            add_vertex(m, p, normalize(p), mk2(u, v));
        }
    }
    add_vertex(m, mk3(0, -radius, 0), mk3(0, -1, 0), mk2(0, 1));
    for (uint32_t i = 0; i < phi_div; ++i) m.tris.push_back({0u, (i + 1) % phi_div + 1, i + 1});
    for (uint32_t i = 0; i + 2 < theta_div; ++i) {
        const uint32_t row = 1 + i * phi_div;
        for (uint32_t j = 0; j < phi_div; ++j) {
            const uint32_t a = row + j;
            // the last segment wraps to the ring's first vertex
            const uint32_t b = j + 1 < phi_div ? a + 1 : a + 1 - phi_div;
            if (j + 1 < phi_div) {
                m.tris.push_back({a, b, a + phi_div});
                m.tris.push_back({b, a + phi_div + 1, a + phi_div});
            } else {
                m.tris.push_back({a, b, a + phi_div});
                m.tris.push_back({b, a + 1, a + phi_div});
            }
        }
    }
    const uint32_t last = uint32_t(m.P.size()) - 1;
    for (uint32_t i = 0; i < phi_div; ++i) m.tris.push_back({last, last - ((1 + i) % phi_div + 1), last - (i + 1)});
    return m;
}

// Cube (src/render_core/shapes/cube.cpp:20-71)
MeshHost make_cube(float x, float y, float z) {
    y = y == 0 ? x : y;
    z = z == 0 ? y : z;
    x /= 2.f;
    y /= 2.f;
    z /= 2.f;
    f3 P[24] = {
        mk3(-x, -y, z), mk3(x, -y, z), mk3(-x, y, z), mk3(x, y, z),
        mk3(-x, y, -z), mk3(x, y, -z), mk3(-x, -y, -z), mk3(x, -y, -z),
        mk3(-x, y, z), mk3(x, y, z), mk3(-x, y, -z), mk3(x, y, -z),
        mk3(-x, -y, z), mk3(x, -y, z), mk3(-x, -y, -z), mk3(x, -y, -z),
        mk3(x, -y, z), mk3(x, y, z), mk3(x, y, -z), mk3(x, -y, -z),
        mk3(-x, -y, z), mk3(-x, y, z), mk3(-x, y, -z), mk3(-x, -y, -z)};
    f3 N[6] = {mk3(0, 0, 1), mk3(0, 0, -1), mk3(0, 1, 0), mk3(0, -1, 0), mk3(1, 0, 0), mk3(-1, 0, 0)};
    f2 UV[24] = {
        mk2(0, 0), mk2(1, 0), mk2(0, 1), mk2(1, 1),
        mk2(0, 1), mk2(1, 1), mk2(0, 0), mk2(1, 0),
        mk2(0, 1), mk2(1, 1), mk2(0, 0), mk2(1, 0),
        mk2(0, 1), mk2(1, 1), mk2(0, 0), mk2(1, 0),
        mk2(0, 1), mk2(1, 1), mk2(1, 0), mk2(0, 0),
        mk2(0, 1), mk2(1, 1), mk2(1, 0), mk2(0, 0)};
    MeshHost m;
    for (int i = 0; i < 24; ++i) add_vertex(m, P[i], N[i / 4], UV[i]);
    m.tris = {{0, 1, 3}, {0, 3, 2}, {6, 5, 7}, {4, 5, 6}, {10, 9, 11}, {8, 9, 10},
              {13, 14, 15}, {13, 12, 14}, {18, 17, 19}, {17, 16, 19}, {21, 22, 23}, {20, 21, 23}};
    return m;
}

MeshHost load_obj(const std::string &path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open mesh " + path);
    std::vector<f3> v, vn;
    std::vector<f2> vt;
    MeshHost m;
    struct Key {
        int v, t, n;
        bool operator==(const Key &o) const { return v == o.v && t == o.t && n == o.n; }
    };
    struct KeyHash {
        size_t operator()(const Key &k) const {
            return (size_t(k.v) * 73856093u) ^ (size_t(k.t) * 19349663u) ^ (size_t(k.n) * 83492791u);
        }
    };
    std::unordered_map<Key, uint32_t, KeyHash> remap;
    std::string line;
    auto resolve = [](int idx, size_t n) -> int { return idx < 0 ? int(n) + idx : idx - 1; };
    while (std::getline(in, line)) {
        if (line.size() < 2) continue;
        std::istringstream ss(line);
        std::string tag;
        ss >> tag;
        if (tag == "v") {
            f3 p;
            ss >> p.x >> p.y >> p.z;
            v.push_back(p);
        } else if (tag == "vn") {
            f3 n;
            ss >> n.x >> n.y >> n.z;
            vn.push_back(n);
        } else if (tag == "vt") {
            f2 t;
            ss >> t.x >> t.y;
            vt.push_back(t);
        } else if (tag == "f") {
            std::vector<uint32_t> poly;
            std::string tok;
            while (ss >> tok) {
                int vi = 0, ti = 0, ni = 0;
                const char *c = tok.c_str();
                vi = std::atoi(c);
                const char *s1 = std::strchr(c, '/');
                if (s1) {
                    if (s1[1] != '/') ti = std::atoi(s1 + 1);
                    const char *s2 = std::strchr(s1 + 1, '/');
                    if (s2) ni = std::atoi(s2 + 1);
                }
                Key k{resolve(vi, v.size()), ti ? resolve(ti, vt.size()) : -1, ni ? resolve(ni, vn.size()) : -1};
                auto it = remap.find(k);
                uint32_t id;
                if (it == remap.end()) {
                    id = uint32_t(m.P.size());
                    if (k.v < 0 || size_t(k.v) >= v.size()) throw std::runtime_error("bad obj vertex index in " + path);
                    add_vertex(m, v[k.v], k.n >= 0 ? vn.at(k.n) : mk3(0.f), k.t >= 0 ? vt.at(k.t) : mk2(0, 0));
                    remap.emplace(k, id);
                } else {
                    id = it->second;
                }
                poly.push_back(id);
            }
            for (size_t i = 2; i < poly.size(); ++i) m.tris.push_back({poly[0], poly[i - 1], poly[i]});
        }
    }
    return m;
}

// Binary mesh: "VSMESH01", u32 nv, u32 nt, u32 flags (1: normals, 2: uvs),
// f32 P[nv*3], [f32 N[nv*3]], [f32 UV[nv*2]], u32 tri[nt*3]
MeshHost load_vsmesh(const std::string &path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open mesh " + path);
    char magic[8];
    in.read(magic, 8);
    if (std::memcmp(magic, "VSMESH01", 8) != 0) throw std::runtime_error("bad vsmesh magic in " + path);
    uint32_t hdr[3];
    in.read(reinterpret_cast<char *>(hdr), sizeof(hdr));
    uint32_t nv = hdr[0], nt = hdr[1], flags = hdr[2];
    MeshHost m;
    m.P.resize(nv);
    in.read(reinterpret_cast<char *>(m.P.data()), std::streamsize(nv) * 12);
    m.N.assign(nv, mk3(0.f));
    if (flags & 1) in.read(reinterpret_cast<char *>(m.N.data()), std::streamsize(nv) * 12);
    m.UV.assign(nv, mk2(0, 0));
    if (flags & 2) in.read(reinterpret_cast<char *>(m.UV.data()), std::streamsize(nv) * 8);
    m.tris.resize(nt);
    in.read(reinterpret_cast<char *>(m.tris.data()), std::streamsize(nt) * 12);
    if (!in) throw std::runtime_error("truncated vsmesh " + path);
    for (const u3 &t : m.tris)
        if (t.x >= nv || t.y >= nv || t.z >= nv) throw std::runtime_error("vsmesh index out of range in " + path);
    return m;
}

LightHost parse_light(const Json &desc, const std::string &base_dir, SceneHost &s) {
    LightHost l;
    std::string type = desc["type"].as_string("area");
    const Json &p = desc["param"];
    l.name = desc["name"].as_string("");
    l.scale = float(p["scale"].as_float(1.0));
    if (type == "area" && slot_image_node(p, "color")) {
        // an emission image (ImageNode): radiance = texel at the light
        // surface's uv times scale (AreaLight::L, area.cpp:90-94); an image
        // node does not normalise (ShaderNode::normalize returns 1) and the
        // light's power takes the image's average
        l.tex = int32_t(texture_index(s, *slot_image_node(p, "color"), true));
        const TextureHost &t = s.textures[size_t(l.tex)];
        double acc[3] = {0, 0, 0};
        const size_t n = size_t(t.width) * t.height;
        for (size_t i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) acc[c] += t.rgb[3 * i + c];
        l.color = mk3(float(acc[0] / double(n)), float(acc[1] / double(n)), float(acc[2] / double(n)));
    } else if (type != "spherical") {
        // Light::Light folds the colour normalization factor into scale
        l.color = slot_f3(p["color"], mk3(1.f));
        l.scale *= normalize_color(l.color);
    }
    if (type == "area") {
        l.type = LIGHT_AREA;
        l.two_sided = p["two_sided"].as_bool(false) ? 1u : 0u;
    } else if (type == "point") {
        l.type = LIGHT_POINT;
        l.position = as_f3(p["position"], mk3(0.f));
    } else if (type == "spot") {
        // SpotLight ctor (src/render_core/lights/spot.cpp:34-40); the falloff
        // clamp bound is the angle already converted to radians
        l.type = LIGHT_SPOT;
        l.position = as_f3(p["position"], mk3(0.f));
        l.angle = radians(std::clamp(float(p["angle"].as_float(45.0)), 1.f, 89.f));
        l.falloff = radians(std::clamp(float(p["falloff"].as_float(10.0)), 0.f, l.angle));
        l.direction = normalize(as_f3(p["direction"], mk3(0, 0, 1)));
    } else if (type == "spherical") {
        // SphericalMap (src/render_core/lights/environments/spherical.cpp:33-45)
        l.type = LIGHT_ENV;
        const Json &c = p["color"];
        const Json &node = c.contains("node") ? c["node"] : c;
        std::string fn = node["fn"].as_string("");
        if (fn.empty() && node.contains("param")) fn = node["param"]["fn"].as_string("");
        l.scale = float(p["scale"].as_float(1.0));
        if (fn.empty()) {
            // a constant colour: a 1 x 1 map sampled with weight 1
            // (SphericalMap::prepare, spherical.cpp:198-211), the colour
            // normalised into the scale as for every light (light.cpp:10-17)
            f3 c = slot_f3(p["color"], mk3(1.f));
            l.scale *= normalize_color(c);
            l.env_w = l.env_h = 1;
            l.env_rgb = {c.x, c.y, c.z};
            l.env_constant = true;
        } else {
            l.env_fn = join_path(base_dir, fn);
            load_env_image(l.env_fn, l.env_rgb, l.env_w, l.env_h);
        }
        l.color = mk3(1.f);
        l.flip_u = p["flip_u"].as_bool(false);
        l.env_o2w = parse_transform(p["o2w"]);
    } else {
        throw std::runtime_error("unsupported light type: " + type);
    }
    return l;
}

}// namespace

const Json *slot_image_node(const Json &param, const char *key) {
    const Json &j = param[key];
    if (j.is_null()) return nullptr;
    const Json *node = (j.is_object() && (j.contains("channels") || j.contains("node"))) ? &j["node"] : &j;
    if (!node->is_object()) return nullptr;
    if ((*node)["type"].as_string("") == "image" || (*node)["param"].contains("fn") || node->contains("fn"))
        return node;
    return nullptr;
}

namespace {

// resolved path, colour space and the registry key of an image node
void texture_key(const SceneHost &s, const Json &node, std::string &path, std::string &cs, std::string &key,
                 bool illum = false) {
    const Json &p = node.contains("param") ? node["param"] : node;
    std::string fn = p["fn"].as_string("");
    if (fn.empty()) throw std::runtime_error("image node without \"fn\"");
    path = join_path(s.base_dir, fn);
    cs = p["color_space"].as_string("");
    std::string e = image_extension(path);
    if (cs.empty()) cs = (e == "exr" || e == "hdr" || e == "pfm") ? "linear" : "srgb";
    if (cs != "linear" && cs != "srgb") throw std::runtime_error("unsupported colour space " + cs + " for " + fn);
    key = path + "|" + cs + (illum ? "|illum" : "");
}

}// namespace

uint32_t find_texture(const SceneHost &s, const Json &node) {
    std::string path, cs, key;
    texture_key(s, node, path, cs, key);
    for (size_t i = 0; i < s.textures.size(); ++i)
        if (s.textures[i].key == key) return uint32_t(i);
    throw std::logic_error("texture not loaded: " + path);
}

uint32_t texture_index(SceneHost &s, const Json &node, bool illum) {
    std::string path, cs, key;
    texture_key(s, node, path, cs, key, illum);
    for (size_t i = 0; i < s.textures.size(); ++i)
        if (s.textures[i].key == key) return uint32_t(i);
    ImageRGBA img = read_image(path);
    TextureHost t;
    t.key = key;
    t.illum = illum;
    t.width = img.width;
    t.height = img.height;
    t.rgb.resize(size_t(img.width) * img.height * 3);
    const bool srgb = cs == "srgb";
    for (size_t i = 0; i < size_t(img.width) * img.height; ++i)
        for (int c = 0; c < 3; ++c) {
            float v = img.rgba[i * 4 + c];
            t.rgb[i * 3 + c] = srgb ? srgb_to_linear(v) : v;
        }
    s.textures.push_back(std::move(t));
    return uint32_t(s.textures.size() - 1);
}

void load_env_image(const std::string &path, std::vector<float> &rgb, uint32_t &w, uint32_t &h) {
    ImageRGBA img = read_image(path);
    // 8/16-bit files hold sRGB-encoded values unless the scene says otherwise
    // (ImagePool::load_texture, src/base/mgr/image_pool.cpp:17-22)
    std::string e = image_extension(path);
    bool srgb = e == "png" || e == "jpg" || e == "jpeg";
    w = img.width;
    h = img.height;
    rgb.resize(size_t(w) * h * 3);
    for (size_t i = 0; i < size_t(w) * h; ++i)
        for (int c = 0; c < 3; ++c) {
            float v = img.rgba[i * 4 + c];
            rgb[i * 3 + c] = srgb ? srgb_to_linear(v) : v;
        }
}

m4 camera_to_world(f3 position, float yaw, float pitch) {
    m4 rot = mat_mul(mat_mul(scale_m(mk3(1, 1, -1)), rotation_y(yaw)), rotation_x(-pitch));
    return mat_mul(translation(position), rot);
}

std::string read_text_file(const std::string &path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::ostringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

SceneHost load_scene(const std::string &text, const std::string &base_dir) {
    Json root = Json::parse(text);
    SceneHost s;
    s.base_dir = base_dir;

    // integrator (IntegratorDesc::init; IlluminationIntegrator ctor
    // src/base/integrator.cpp:60-67)
    {
        const Json &d = root["integrator"];
        s.integrator.type = d["type"].as_string("pt");
        const Json &p = d["param"];
        s.integrator.param = p;
        s.integrator.max_depth = p["max_depth"].as_uint(16);
        s.integrator.min_depth = p["min_depth"].as_uint(5);
        s.integrator.rr_threshold = float(p["rr_threshold"].as_float(1.0));
        s.integrator.mis_mode = p["mis_mode"].as_int(0);
        if (s.integrator.type != "pt" && s.integrator.type != "rt")
            throw std::runtime_error("unsupported integrator type: " + s.integrator.type);
        {
            SvgfHost &v = s.integrator.svgf;
            const Json &dn = p["denoiser"];
            v.svgf = dn["type"].as_string("svgf") == "svgf";
            const Json &dp = dn["param"];
            v.N = dp["N"].as_uint(3);
            v.alpha = float(dp["alpha"].as_float(0.05));
            v.moments_alpha = float(dp["moments_alpha"].as_float(0.2));
            v.history_limit = dp["history_limit"].as_uint(32);
            v.moments_radius = dp["moments_filter_radius"].as_int(3);
            v.sigma_rt = float(dp["sigma_rt"].as_float(10.0));
            v.sigma_normal = float(dp["sigma_normal"].as_float(30.0));
            if (v.N > 16 || v.moments_radius < 0 || v.moments_radius > 16 || v.history_limit == 0)
                throw std::runtime_error("SVGF settings out of range");
            AdaptiveHost &a = s.integrator.adaptive;
            const Json &ad = p["adaptive"];
            a.threshold = float(ad["threshold"].as_float(0.001));
            a.min_samples = ad["start_index"].as_uint(128);
        }
        if (s.integrator.type == "rt") {
            RestirHost &r = s.integrator.restir;
            const Json &di = p["direct"];
            r.open = di["open"].as_bool(true);
            r.M_light = di["M_light"].as_uint(10);
            r.M_bsdf = di["M_bsdf"].as_uint(1);
            r.max_age = di["max_age"].as_uint(30);
            r.max_recursion = di["max_recursion"].as_uint(5);
            r.diff_factor = float(di["diffuse_factor"].as_float(0.3));
            r.pairwise = di["pairwise"].as_bool(false);
            const Json &sp = di["spatial"];
            r.s_theta = float(sp["theta"].as_float(35.0));
            r.s_depth = float(sp["depth"].as_float(0.3));
            r.s_radius = float(sp["radius"].as_float(30.0));
            r.s_num = sp["sample_num"].as_uint(1);
            r.s_open = sp["open"].as_bool(true);
            if (!r.s_open) r.s_num = 1;
            r.extra_spatial_passes = sp["extra_passes"].as_uint(0);
            const Json &tp = di["temporal"];
            r.t_limit = tp["history_limit"].as_uint(5);
            r.t_radius = float(tp["radius"].as_float(8.0));
            r.t_theta = float(tp["theta"].as_float(35.0));
            r.t_depth = float(tp["depth"].as_float(0.3));
            r.t_open = tp["open"].as_bool(true);
            r.t_N = tp["N"].as_uint(9);
            r.t_mis = tp["mis"].as_bool(false);
            if (!r.open)
                throw std::runtime_error("unsupported: rt integrator with ReSTIR DI closed");
            // ReSTIR GI (ReSTIR(integrator, desc), ReSTIR/common.h:98-105)
            RestirGiHost &g = s.integrator.gi;
            const Json &ii = p["indirect"];
            g.open = ii["open"].as_bool(true);
            g.max_age = ii["max_age"].as_uint(30);
            g.diff_factor = float(ii["diffuse_factor"].as_float(0.3));
            const Json &gs = ii["spatial"];
            g.s_theta = float(gs["theta"].as_float(35.0));
            g.s_depth = float(gs["depth"].as_float(0.3));
            g.s_radius = float(gs["radius"].as_float(30.0));
            g.s_num = gs["sample_num"].as_uint(1);
            g.s_open = gs["open"].as_bool(true);
            if (!g.s_open) g.s_num = 1;
            const Json &gt = ii["temporal"];
            g.t_limit = gt["history_limit"].as_uint(5);
            g.t_radius = float(gt["radius"].as_float(8.0));
            g.t_theta = float(gt["theta"].as_float(35.0));
            g.t_depth = float(gt["depth"].as_float(0.3));
            g.t_open = gt["open"].as_bool(true);
            g.t_N = gt["N"].as_uint(9);
            if (g.s_num > 64 || g.t_N > 64) throw std::runtime_error("ReSTIR GI settings out of range");
            if (r.s_num > 64 || r.t_N > 64 || r.M_light > 4096 || r.M_bsdf > 64 || r.max_recursion > 64 ||
                r.extra_spatial_passes > 8)
                throw std::runtime_error("ReSTIR DI settings out of range");
        }
    }
    // spectrum (SpectrumDesc::init; hero dimension default 3, hero.cpp:239)
    {
        const Json &d = root["spectrum"];
        s.spectrum_type = d["type"].as_string("srgb");
        if (s.spectrum_type == "hero") {
            s.spectrum_dim = d["param"]["dimension"].as_uint(3);
            if (s.spectrum_dim != 3 && s.spectrum_dim != 4)
                throw std::runtime_error("unsupported hero spectrum dimension (3 or 4 supported): " +
                                         std::to_string(s.spectrum_dim));
        }
        else if (s.spectrum_type == "srgb") s.spectrum_dim = 3;
        else throw std::runtime_error("unsupported spectrum type: " + s.spectrum_type);
        if (s.integrator.type == "rt" && s.integrator.gi.open && s.spectrum_type != "srgb")
            throw std::runtime_error("unsupported: ReSTIR GI (the rt integrator's open \"indirect\" block) with "
                                     "hero-wavelength spectra; close it or render in sRGB mode");
    }
    // sampler
    {
        const Json &d = root["sampler"];
        s.sampler_type = d["type"].as_string("independent");
        if (s.sampler_type != "independent") throw std::runtime_error("unsupported sampler: " + s.sampler_type);
    }
    // render_setting (RenderSettingDesc::init)
    s.ray_offset_factor = float(root["render_setting"]["ray_offset_factor"].as_float(1.0));
    s.min_world_radius = float(root["render_setting"]["min_world_radius"].as_float(10.0));
    // output (OutputDesc::init)
    s.output_spp = root["output"]["spp"].as_uint(0);
    s.output_fn = root["output"]["fn"].as_string("output.png");

    // light sampler + its light list (LightSamplerDesc::init, LightSampler ctor
    // src/base/illumination/lightsampler.cpp:11-26)
    struct LightQuad {
        size_t light;
        float width, height;
        m4 o2w;
    };
    std::vector<LightQuad> light_quads;
    {
        const Json &d = root["light_sampler"];
        s.light_sampler_type = d["type"].as_string("uniform");
        if (s.light_sampler_type != "uniform" && s.light_sampler_type != "power")
            throw std::runtime_error("unsupported light sampler: " + s.light_sampler_type);
        const Json &p = d["param"];
        s.env_separate = p["env_separate"].as_bool(false);
        s.env_prob = std::clamp(float(p["env_prob"].as_float(0.5)), 0.01f, 0.99f);
        const Json &lights = p["lights"];
        if (lights.is_array())
            for (const Json &l : lights.elements()) {
                LightHost lh = parse_light(l, base_dir, s);
                // an area light of its own brings its emitter: a width x
                // height quad placed by "o2w" (AreaLight::init_geometry,
                // src/render_core/lights/area.cpp:57-68)
                if (lh.type == LIGHT_AREA) {
                    const Json &lp = l["param"];
                    light_quads.push_back({s.lights.size(), float(lp["width"].as_float(1.0)),
                                           float(lp["height"].as_float(1.0)), parse_transform(lp["o2w"])});
                }
                s.lights.push_back(lh);
            }
    }

    // materials (MaterialDesc::init)
    {
        const Json &mats = root["materials"];
        if (mats.is_array()) {
            for (const Json &m : mats.elements()) {
                MaterialHost mh;
                mh.type = m["type"].as_string("diffuse");
                mh.name = m["name"].as_string("");
                mh.param = m["param"];
                // mix / add (MaterialDesc::init, node_desc.cpp:156-161) hold
                // two component materials, param.mat0 and param.mat1
                auto leaf = [](const std::string &t) {
                    return t == "diffuse" || t == "metal" || t == "mirror" || t == "glass" || t == "principled_bsdf";
                };
                std::vector<const Json *> with_slots{&mh.param};
                if (mh.type == "mix" || mh.type == "add") {
                    for (const char *k : {"mat0", "mat1"}) {
                        const Json &c = mh.param[k];
                        std::string ct = c.is_object() ? c["type"].as_string("diffuse") : std::string("(missing)");
                        if (!c.is_object() || !leaf(ct)) {
                            mh.unsupported = "unsupported " + mh.type + " component " + k + ": " + ct +
                                             " (components are diffuse, metal, mirror, glass or principled_bsdf)";
                            continue;
                        }
                        with_slots.push_back(&c["param"]);
                    }
                    if (s.integrator.type == "rt")
                        mh.unsupported = "unsupported: " + mh.type + " materials with the rt integrator (path "
                                                                     "tracing renders them)";
                } else if (!leaf(mh.type) && mh.type != "black_body") {
                    mh.unsupported = "unsupported material type: " + mh.type;
                }
                // colour slots may name images; load them once here
                if (mh.unsupported.empty())
                    for (const Json *sp : with_slots)
                        for (const char *key : {"color", "spec_tint", "sheen_tint", "coat_tint"})
                            if (const Json *node = slot_image_node(*sp, key)) texture_index(s, *node);
                s.materials.push_back(mh);
            }
        }
    }

    // camera (SensorDesc::init, Sensor::init, RadianceCollector ctor)
    {
        const Json &d = root["camera"];
        CameraHost &c = s.camera;
        c.type = d["type"].as_string("thin_lens");
        if (c.type != "pinhole" && c.type != "thin_lens") throw std::runtime_error("unsupported camera: " + c.type);
        const Json &p = d["param"];
        c.fov_y = float(p["fov_y"].as_float(20.0));
        m4 m = parse_transform(p["transform"]);
        // Sensor::update_mat re-derives pitch/yaw/position (sensor.cpp:78-83)
        c.pitch = degrees(std::atan2(m.c[1].z, m.c[1].y));
        c.yaw = degrees(std::atan2(m.c[2].x, m.c[0].x));
        c.position = mk3(m.c[3].x, m.c[3].y, m.c[3].z);
        c.c2w = camera_to_world(c.position, c.yaw, c.pitch);
        if (c.type == "thin_lens") {
            c.focal_distance = float(p["focal_distance"].as_float(5.0));
            c.lens_radius = float(p["lens_radius"].as_float(0.0));
        }
        const Json &rc = p["radiance_collector"]["param"];
        auto res = rc["resolution"].as_numbers();
        if (res.size() >= 2) {
            c.width = uint32_t(res[0]);
            c.height = uint32_t(res[1]);
        }
        c.exposure = float(rc["exposure"].as_float(1.0));
        c.accumulation = rc["accumulation"].as_uint(1);
        c.tone_mapper = rc["tone_mapper"]["type"].as_string("linear");
        const Json &f = p["filter"];
        c.filter_type = f["type"].as_string("gaussian");
        const Json &fr = f["param"]["radius"];
        auto frv = fr.as_numbers();
        c.filter_radius = frv.empty() ? 0.5f : float(frv[0]);
        // GaussianFilter / MitchellFilter / LanczosSincFilter parameters
        // (gaussian.cpp:17-20, mitchell.cpp:16-18, sinc.cpp:16-17)
        c.filter_sigma = float(f["param"]["sigma"].as_float(1.0));
        c.filter_b = float(f["param"]["b"].as_float(1.0 / 3.0));
        c.filter_c = float(f["param"]["c"].as_float(1.0 / 3.0));
        c.filter_tau = float(f["param"]["tau"].as_float(3.0));
        if (c.filter_type != "box" && c.filter_type != "triangle" && c.filter_type != "gaussian" &&
            c.filter_type != "mitchell" && c.filter_type != "sinc")
            throw std::runtime_error("unsupported filter: " + c.filter_type);
        if (c.width == 0 || c.height == 0) throw std::runtime_error("resolution must be positive");
    }

    // shapes (ShapeDesc::init, Scene::add_shape src/base/mgr/scene.cpp:118-154)
    {
        const Json &shapes = root["shapes"];
        if (shapes.is_array()) {
            for (const Json &sh : shapes.elements()) {
                std::string type = sh["type"].as_string("");
                const Json &p = sh["param"];
                MeshHost mesh;
                if (type == "quad") {
                    mesh = make_quad(float(p["width"].as_float(1.0)), float(p["height"].as_float(1.0)));
                } else if (type == "sphere") {
                    mesh = make_sphere(float(p["radius"].as_float(1.0)), p["sub_div"].as_uint(60));
                } else if (type == "cube") {
                    mesh = make_cube(float(p["x"].as_float(1.0)), float(p["y"].as_float(1.0)), float(p["z"].as_float(1.0)));
                } else if (type == "model") {
                    std::string fn = join_path(base_dir, p["fn"].as_string(""));
                    if (fn.size() >= 7 && fn.compare(fn.size() - 7, 7, ".vsmesh") == 0) mesh = load_vsmesh(fn);
                    else mesh = load_obj(fn);
                } else {
                    throw std::runtime_error("unsupported shape type: " + type);
                }
                InstanceHost inst;
                inst.mesh = uint32_t(s.meshes.size());
                inst.o2w = parse_transform(p["transform"]);
                inst.name = sh["name"].as_string("");
                inst.material_name = p["material"].as_string("");
                for (size_t i = 0; i < s.materials.size(); ++i) {
                    if (s.materials[i].name == inst.material_name) {
                        if (!s.materials[i].unsupported.empty())
                            throw std::runtime_error(s.materials[i].unsupported + " (material \"" +
                                                     inst.material_name + "\" of shape \"" + inst.name + "\")");
                        inst.material = int32_t(i);
                        break;
                    }
                }
                if (p.contains("emission")) {
                    LightHost l = parse_light(p["emission"], base_dir, s);
                    if (l.type != LIGHT_AREA) throw std::runtime_error("shape emission must be an area light");
                    l.inst = uint32_t(s.instances.size());
                    inst.light = int32_t(s.lights.size());
                    s.lights.push_back(l);
                }
                s.meshes.push_back(std::move(mesh));
                s.instances.push_back(inst);
            }
        }
    }
    // the light list's own area emitters, after the scene's shapes, with the
    // black-body material (Scene::obtain_black_body: no scattering)
    if (!light_quads.empty()) {
        MaterialHost black;
        black.type = "black_body";
        const int32_t black_index = int32_t(s.materials.size());
        s.materials.push_back(black);
        for (const LightQuad &q : light_quads) {
            InstanceHost inst;
            inst.mesh = uint32_t(s.meshes.size());
            inst.o2w = q.o2w;
            inst.material = black_index;
            inst.light = int32_t(q.light);
            s.lights[q.light].inst = uint32_t(s.instances.size());
            s.meshes.push_back(make_quad(q.width, q.height));
            s.instances.push_back(inst);
        }
    }
    s.flatten();
    return s;
}

namespace {

// inverse-transpose of the upper 3x3 (normal transform)
void normal_matrix(const m4 &m, float out[3][3]) {
    double a[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) a[r][c] = m4_at(m, r, c);
    double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                 a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    double inv[3][3];
    inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
    inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
    inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
    inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
    inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
    inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
    inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
    inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
    inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
    // out = transpose(inv)
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) out[r][c] = float(inv[c][r]);
}

}// namespace

void SceneHost::flatten() {
    wP.clear();
    wN.clear();
    wUV.clear();
    wTri.clear();
    triInst.clear();
    triPrim.clear();
    instTriBase.clear();
    instRecords.clear();
    for (size_t i = 0; i < instances.size(); ++i) {
        const InstanceHost &inst = instances[i];
        const MeshHost &mesh = meshes[inst.mesh];
        uint32_t vbase = uint32_t(wP.size());
        instTriBase.push_back(uint32_t(wTri.size()));
        float nm[3][3];
        normal_matrix(inst.o2w, nm);
        for (size_t v = 0; v < mesh.P.size(); ++v) {
            wP.push_back(xform_point(inst.o2w, mesh.P[v]));
            f3 n = mesh.N[v];
            wN.push_back(mk3(nm[0][0] * n.x + nm[0][1] * n.y + nm[0][2] * n.z,
                             nm[1][0] * n.x + nm[1][1] * n.y + nm[1][2] * n.z,
                             nm[2][0] * n.x + nm[2][1] * n.y + nm[2][2] * n.z));
            wUV.push_back(mesh.UV[v]);
        }
        for (size_t t = 0; t < mesh.tris.size(); ++t) {
            u3 tr = mesh.tris[t];
            wTri.push_back({tr.x + vbase, tr.y + vbase, tr.z + vbase});
            triInst.push_back(uint32_t(i));
            triPrim.push_back(uint32_t(t));
        }
        InstanceRecord ir{};
        ir.mat = inst.material >= 0 ? uint32_t(inst.material) : kInvalid;
        ir.light = inst.light >= 0 ? uint32_t(inst.light) : kInvalid;
        ir.tri_base = instTriBase.back();
        instRecords.push_back(ir);
    }
    aabb_lo = mk3(3.4e38f);
    aabb_hi = mk3(-3.4e38f);
    for (const f3 &p : wP) {
        aabb_lo = vmin(aabb_lo, p);
        aabb_hi = vmax(aabb_hi, p);
    }
    // per-triangle areas of emitters (ShapeInstance::surface_areas)
    for (LightHost &l : lights) {
        if (l.type != LIGHT_AREA) continue;
        l.tri_base = instTriBase[l.inst];
        size_t n = meshes[instances[l.inst].mesh].tris.size();
        l.tri_areas.resize(n);
        for (size_t t = 0; t < n; ++t) {
            u3 tr = wTri[l.tri_base + t];
            f3 v0 = wP[tr.x], v1 = wP[tr.y], v2 = wP[tr.z];
            l.tri_areas[t] = 0.5f * length(cross(v1 - v0, v2 - v0));
        }
    }
}

}// namespace vs
