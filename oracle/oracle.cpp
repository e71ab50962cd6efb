// CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// A scalar C++ restatement of the reference's megakernel path-tracing path
// (FaithZL/Vision), used by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg as the checker.  The product (libvision_amd.so) never
// links, loads or calls this code.
//
// Every function cites the reference file:line it restates.  Pinning (see
// tests/test_oracle.py): TEA against the published TEA known answer (32
// rounds, zero key), LCG against its recurrence's known first outputs,
// warps/BSDF sampling by chi-square against their pdfs, alias tables by exact
// reconstruction of their distribution, BVH2 traversal by brute force,
// the Li loop by the furnace closed form L = Le / (1 - albedo) and a
// point-light single-bounce closed form.  Functions only an import of the
// absent ocarina library could pin (ray-origin offset constants, frame
// orientation, perspective arithmetic) follow the published methods named
// inline; for those the oracle defines the behaviour, parity unpinned.
//
// Built with -ffp-contract=off so the watertight triangle test matches the
// HIP kernel bit for bit on identical rays.

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kOneMinusEps = 0x1.fffffep-1f;
constexpr float kShadowEps = 1e-4f;
constexpr uint32_t kInv = 0xFFFFFFFFu;

struct V3 {
    float x, y, z;
};
inline V3 v3(float x, float y, float z) { return {x, y, z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }
inline float dotp(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 crossp(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float len(V3 a) { return std::sqrt(dotp(a, a)); }
inline V3 unit(V3 a) { return mul(a, 1.f / len(a)); }
inline float at(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// ---------------------------------------------------------------- sampler
// tea<N> (src/math/util.h:13-25) generalised over key and round count so the
// 32-round form can be checked against TEA's published known answer.
uint32_t tea_rounds(uint32_t v0, uint32_t v1, const uint32_t k[4], uint32_t rounds) {
    uint32_t s = 0;
    for (uint32_t n = 0; n < rounds; ++n) {
        s += 0x9e3779b9u;
        v0 += ((v1 << 4) + k[0]) ^ (v1 + s) ^ ((v1 >> 5) + k[1]);
        v1 += ((v0 << 4) + k[2]) ^ (v0 + s) ^ ((v0 >> 5) + k[3]);
    }
    return v0;
}
const uint32_t kVisionKey[4] = {0xa341316cu, 0xc8013ea4u, 0xad90777du, 0x7e95761eu};
uint32_t tea4(uint32_t a, uint32_t b) { return tea_rounds(a, b, kVisionKey, 4); }

// IndependentSampler (src/render_core/samplers/independent.cpp:24-42), lcg (util.h:27-34)
struct Sampler {
    uint32_t state{0};
    void start(uint32_t px, uint32_t py, uint32_t index, uint32_t dim) { state = tea4(tea4(px, py), tea4(index, dim)); }
    float next1d() {
        state = 1664525u * state + 1013904223u;
        return float(state & 0x00ffffffu) * (1.f / 16777216.f);
    }
};

// ---------------------------------------------------------------- warps (src/math/warp.h)
void square_to_disk(float ux, float uy, float &x, float &y) {
    float r = std::sqrt(ux);
    float th = 2.f * kPi * uy;
    x = r * std::cos(th);
    y = r * std::sin(th);
}
V3 cosine_hemisphere(float ux, float uy) {
    float x, y;
    square_to_disk(ux, uy, x, y);
    return v3(x, y, std::sqrt(std::max(0.f, 1.f - x * x - y * y)));
}
float sample_linear(float u, float a, float b) {
    float x = u * (a + b) / (a + std::sqrt(a * a + u * (b * b - a * a)));
    float r = std::min(x, kOneMinusEps);
    return (u == 0.f && a == 0.f) ? 0.f : r;
}
float sample_tent(float u, float r) {
    return u < 0.5f ? -r * sample_linear((0.5f - u) * 2.f, 1.f, 0.f) : r * sample_linear((u - 0.5f) * 2.f, 1.f, 0.f);
}

// ---------------------------------------------------------------- alias table
// AliasTable::build and offset_u_remapped (src/render_core/warpers/alias.h:86-158)
struct Alias {
    std::vector<float> prob, func;
    std::vector<uint32_t> alias;
    float integral{0.f};
};
Alias alias_build(const float *w, uint32_t n) {
    Alias a;
    double sum = 0.0;
    for (uint32_t i = 0; i < n; ++i) sum += w[i];
    double ratio = double(n) / sum;
    a.prob.resize(n);
    a.alias.resize(n);
    std::vector<uint32_t> over, under;
    for (uint32_t i = 0; i < n; ++i) {
        float p = float(w[i] * ratio);
        a.prob[i] = p;
        a.alias[i] = i;
        if (p > 1.f) over.push_back(i);
        else under.push_back(i);
    }
    while (!over.empty() && !under.empty()) {
        uint32_t o = over.back(), u = under.back();
        over.pop_back();
        under.pop_back();
        a.prob[o] -= 1.f - a.prob[u];
        a.alias[u] = o;
        if (a.prob[o] > 1.f) over.push_back(o);
        else if (a.prob[o] < 1.f) under.push_back(o);
    }
    for (uint32_t i : over) a.prob[i] = 1.f, a.alias[i] = i;
    for (uint32_t i : under) a.prob[i] = 1.f, a.alias[i] = i;
    a.integral = float(sum / double(n));
    a.func.assign(w, w + n);
    return a;
}
uint32_t alias_sample(const Alias &a, float u, float *u_rem) {
    uint32_t n = uint32_t(a.prob.size());
    u = u * float(n);
    uint32_t idx = std::min(uint32_t(u), n - 1u);
    u = std::min(u - float(idx), kOneMinusEps);
    float p = a.prob[idx];
    bool keep = u < p;
    if (u_rem) *u_rem = keep ? std::min(u / p, kOneMinusEps) : std::min((1.f - u) / (1.f - p), kOneMinusEps);
    return keep ? idx : a.alias[idx];
}
float alias_pmf(const Alias &a, uint32_t i) {
    return a.integral > 0.f ? a.func[i] / (a.integral * float(a.prob.size())) : 0.f;
}

// ---------------------------------------------------------------- scene view
struct Scene {
    const float *P, *N, *UV;
    const uint32_t *tri, *tri_inst, *inst, *mat, *light;
    const float *cam, *render;
    uint32_t nv, nt, ni, nm, nl;
    V3 p(uint32_t v) const { return v3(P[3 * v], P[3 * v + 1], P[3 * v + 2]); }
    V3 n(uint32_t v) const { return v3(N[3 * v], N[3 * v + 1], N[3 * v + 2]); }
    float matp(uint32_t m, int k) const {
        float f;
        std::memcpy(&f, &mat[32 * m + 2 + k], 4);
        return f;
    }
    uint32_t mat_type(uint32_t m) const { return mat[32 * m]; }
    float lightf(uint32_t l, int word) const {
        float f;
        std::memcpy(&f, &light[28 * l + word], 4);
        return f;
    }
    uint32_t lightu(uint32_t l, int word) const { return light[28 * l + word]; }
};

// ---------------------------------------------------------------- intersection
// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013).
struct Ray {
    V3 o, d;
    float tmin, tmax;
};
struct RayPre {
    V3 o;
    int kx, ky, kz;
    float Sx, Sy, Sz;
};
RayPre ray_pre(V3 o, V3 d) {
    RayPre r;
    r.o = o;
    float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    r.kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    r.kx = (r.kz + 1) % 3;
    r.ky = (r.kx + 1) % 3;
    if (at(d, r.kz) < 0.f) std::swap(r.kx, r.ky);
    r.Sx = at(d, r.kx) / at(d, r.kz);
    r.Sy = at(d, r.ky) / at(d, r.kz);
    r.Sz = 1.f / at(d, r.kz);
    return r;
}
bool hit_triangle(const RayPre &r, V3 v0, V3 v1, V3 v2, float &t, float &b1, float &b2) {
    V3 A = sub(v0, r.o), B = sub(v1, r.o), C = sub(v2, r.o);
    float Ax = at(A, r.kx) - r.Sx * at(A, r.kz), Ay = at(A, r.ky) - r.Sy * at(A, r.kz);
    float Bx = at(B, r.kx) - r.Sx * at(B, r.kz), By = at(B, r.ky) - r.Sy * at(B, r.kz);
    float Cx = at(C, r.kx) - r.Sx * at(C, r.kz), Cy = at(C, r.ky) - r.Sy * at(C, r.kz);
    float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if (U == 0.f || V == 0.f || W == 0.f) {
        U = float(double(Cx) * double(By) - double(Cy) * double(Bx));
        V = float(double(Ax) * double(Cy) - double(Ay) * double(Cx));
        W = float(double(Bx) * double(Ay) - double(By) * double(Ax));
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
    float det = U + V + W;
    if (det == 0.f) return false;
    float Az = r.Sz * at(A, r.kz), Bz = r.Sz * at(B, r.kz), Cz = r.Sz * at(C, r.kz);
    float T = U * Az + V * Bz + W * Cz;
    t = T / det;
    b1 = V / det;
    b2 = W / det;
    return true;
}

struct Hit {
    float t;
    uint32_t tri;
    float b1, b2;
};

// simple median-split BVH2 with double-precision slab tests (conservative),
// checked against brute force in the tests
struct Bvh2 {
    struct Node {
        double lo[3], hi[3];
        uint32_t left, right, first, count;
    };
    std::vector<Node> nodes;
    std::vector<uint32_t> idx;
};

void bvh2_build(const Scene &s, Bvh2 &b) {
    b.idx.resize(s.nt);
    for (uint32_t i = 0; i < s.nt; ++i) b.idx[i] = i;
    std::vector<V3> cen(s.nt);
    for (uint32_t i = 0; i < s.nt; ++i) {
        V3 a = s.p(s.tri[3 * i]), c = s.p(s.tri[3 * i + 1]), d = s.p(s.tri[3 * i + 2]);
        cen[i] = mul(add(add(a, c), d), 1.f / 3.f);
    }
    struct Task {
        uint32_t node, first, count;
    };
    std::vector<Task> stack;
    b.nodes.push_back({});
    stack.push_back({0, 0, s.nt});
    while (!stack.empty()) {
        Task t = stack.back();
        stack.pop_back();
        Bvh2::Node nd{};
        for (int k = 0; k < 3; ++k) nd.lo[k] = DBL_MAX, nd.hi[k] = -DBL_MAX;
        for (uint32_t i = t.first; i < t.first + t.count; ++i)
            for (int c = 0; c < 3; ++c) {
                V3 p = s.p(s.tri[3 * b.idx[i] + c]);
                for (int k = 0; k < 3; ++k) {
                    nd.lo[k] = std::min(nd.lo[k], double(at(p, k)));
                    nd.hi[k] = std::max(nd.hi[k], double(at(p, k)));
                }
            }
        if (t.count <= 4) {
            nd.first = t.first;
            nd.count = t.count;
            b.nodes[t.node] = nd;
            continue;
        }
        int axis = 0;
        double ext = -1;
        for (int k = 0; k < 3; ++k)
            if (nd.hi[k] - nd.lo[k] > ext) ext = nd.hi[k] - nd.lo[k], axis = k;
        uint32_t mid = t.first + t.count / 2;
        std::nth_element(b.idx.begin() + t.first, b.idx.begin() + mid, b.idx.begin() + t.first + t.count,
                         [&](uint32_t a, uint32_t c) { return at(cen[a], axis) < at(cen[c], axis); });
        nd.left = uint32_t(b.nodes.size());
        nd.right = nd.left + 1;
        nd.count = 0;
        b.nodes[t.node] = nd;
        b.nodes.push_back({});
        b.nodes.push_back({});
        stack.push_back({nd.left, t.first, mid - t.first});
        stack.push_back({nd.right, mid, t.first + t.count - mid});
    }
}

bool box_hit(const Bvh2::Node &n, V3 o, V3 d, double tmin, double tmax) {
    for (int k = 0; k < 3; ++k) {
        double ok = at(o, k), dk = at(d, k);
        if (dk == 0.0) {
            if (ok < n.lo[k] || ok > n.hi[k]) return false;
            continue;
        }
        double inv = 1.0 / dk;
        double t0 = (n.lo[k] - ok) * inv, t1 = (n.hi[k] - ok) * inv;
        if (t0 > t1) std::swap(t0, t1);
        tmin = std::max(tmin, t0 - 1e-9 * std::fabs(t0));
        tmax = std::min(tmax, t1 + 1e-9 * std::fabs(t1));
        if (tmin > tmax) return false;
    }
    return true;
}

void consider(const Scene &s, const RayPre &rp, uint32_t i, float tmin, float tmax, Hit &h) {
    float t, b1, b2;
    if (!hit_triangle(rp, s.p(s.tri[3 * i]), s.p(s.tri[3 * i + 1]), s.p(s.tri[3 * i + 2]), t, b1, b2)) return;
    if (t > tmin && t < tmax && (t < h.t || (t == h.t && i < h.tri))) h = {t, i, b1, b2};
}

Hit trace_brute(const Scene &s, const Ray &r, bool any) {
    RayPre rp = ray_pre(r.o, r.d);
    Hit h{r.tmax, kInv, 0.f, 0.f};
    for (uint32_t i = 0; i < s.nt; ++i) {
        consider(s, rp, i, r.tmin, r.tmax, h);
        if (any && h.tri != kInv) return h;
    }
    return h;
}

Hit trace_bvh(const Scene &s, const Bvh2 &b, const Ray &r, bool any) {
    RayPre rp = ray_pre(r.o, r.d);
    Hit h{r.tmax, kInv, 0.f, 0.f};
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Bvh2::Node &n = b.nodes[stack[--sp]];
        if (!box_hit(n, r.o, r.d, r.tmin, h.t)) continue;
        if (n.count) {
            for (uint32_t k = n.first; k < n.first + n.count; ++k) {
                consider(s, rp, b.idx[k], r.tmin, r.tmax, h);
                if (any && h.tri != kInv) return h;
            }
        } else {
            stack[sp++] = n.left;
            stack[sp++] = n.right;
        }
    }
    return h;
}

// ---------------------------------------------------------------- geometry
// offset_ray_origin: Waechter & Binder, "A Fast and Robust Method for
// Avoiding Self-Intersection" (Ray Tracing Gems, ch. 6)
float off1(float p, float n) {
    int of = int(256.f * n);
    int pi;
    std::memcpy(&pi, &p, 4);
    pi += p < 0.f ? -of : of;
    float pf;
    std::memcpy(&pf, &pi, 4);
    return std::fabs(p) < 1.f / 32.f ? p + (1.f / 65536.f) * n : pf;
}
V3 offset_origin(V3 p, V3 n) { return v3(off1(p.x, n.x), off1(p.y, n.y), off1(p.z, n.z)); }

struct Frame {
    V3 x, y, z;
    V3 local(V3 v) const { return v3(dotp(v, x), dotp(v, y), dotp(v, z)); }
    V3 world(V3 v) const { return add(add(mul(x, v.x), mul(y, v.y)), mul(z, v.z)); }
};

void coordinate_system(V3 a, V3 &b, V3 &c) {
    if (std::fabs(a.x) > std::fabs(a.y)) {
        float inv = 1.f / std::sqrt(a.x * a.x + a.z * a.z);
        b = v3(-a.z * inv, 0.f, a.x * inv);
    } else {
        float inv = 1.f / std::sqrt(a.y * a.y + a.z * a.z);
        b = v3(0.f, a.z * inv, -a.y * inv);
    }
    c = crossp(a, b);
}

// Geometry::compute_surface_interaction (src/base/mgr/geometry.cpp:78-165);
// the shading frame is the orthonormal form of PartialDerivative::update
// (src/base/scattering/interaction.h:96-100): z = ns, x = cross(ns, dp_dv)
struct Surf {
    V3 pos, ng, wo;
    Frame fr;
    float area;
    uint32_t inst;
};
Surf surface(const Scene &s, uint32_t t, float b1, float b2, bool complete) {
    Surf o{};
    uint32_t i0 = s.tri[3 * t], i1 = s.tri[3 * t + 1], i2 = s.tri[3 * t + 2];
    V3 p0 = s.p(i0), p1 = s.p(i1), p2 = s.p(i2);
    float b0 = 1.f - b1 - b2;
    o.pos = add(add(mul(p0, b0), mul(p1, b1)), mul(p2, b2));
    V3 dp02 = sub(p0, p2), dp12 = sub(p1, p2);
    V3 ngu = crossp(dp02, dp12);
    float l = len(ngu);
    o.area = 0.5f * l;
    o.ng = mul(ngu, 1.f / l);
    o.inst = s.tri_inst[t];
    if (complete) {
        float u0 = s.UV[2 * i0], v0 = s.UV[2 * i0 + 1], u1 = s.UV[2 * i1], v1 = s.UV[2 * i1 + 1];
        float u2 = s.UV[2 * i2], v2 = s.UV[2 * i2 + 1];
        float du02 = u0 - u2, dv02 = v0 - v2, du12 = u1 - u2, dv12 = v1 - v2;
        float det = du02 * dv12 - dv02 * du12;
        V3 dpdv;
        if (std::fabs(det) < 1e-8f) dpdv = unit(sub(p2, p0));
        else dpdv = mul(add(mul(dp02, -du12), mul(dp12, du02)), 1.f / det);
        V3 nn = add(add(mul(s.n(i0), b0), mul(s.n(i1), b1)), mul(s.n(i2), b2));
        V3 z = (nn.x != 0.f || nn.y != 0.f || nn.z != 0.f) ? unit(nn) : o.ng;
        V3 x = crossp(z, dpdv);
        float lx = len(x);
        if (!(lx > 1e-20f)) {
            V3 a, c;
            coordinate_system(z, a, c);
            x = a;
        } else {
            x = mul(x, 1.f / lx);
        }
        o.fr = {x, crossp(z, x), z};
    }
    return o;
}

// ---------------------------------------------------------------- lights
struct LightTables {
    std::vector<Alias> tri_alias;// per light (area)
    Alias select;
    bool power{false};
};

float luminance(V3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }

// light power (area.cpp:87-89, point.cpp:32-34, spot.cpp:44-46)
V3 light_power(const Scene &s, uint32_t l, const LightTables &lt) {
    V3 avg = mul(v3(s.lightf(l, 8), s.lightf(l, 9), s.lightf(l, 10)), s.lightf(l, 11));
    uint32_t type = s.lightu(l, 0);
    if (type == 0) {
        float area = 0.f;
        for (float a : lt.tri_alias[l].func) area += a;
        return mul(avg, (s.lightu(l, 1) ? 2.f : 1.f) * area * kPi);
    }
    if (type == 1) return mul(avg, 4.f * kPi);
    if (type == 2) return mul(avg, 2.f * kPi * (1.f - .5f * (s.lightf(l, 15) * 2.f + s.lightf(l, 19))));
    return v3(0, 0, 0);
}

LightTables build_lights(const Scene &s) {
    LightTables lt;
    lt.tri_alias.resize(s.nl);
    for (uint32_t l = 0; l < s.nl; ++l) {
        if (s.lightu(l, 0) != 0) continue;
        uint32_t base = s.lightu(l, 3), cnt = s.lightu(l, 4);
        std::vector<float> areas(cnt);
        for (uint32_t k = 0; k < cnt; ++k) {
            uint32_t t = base + k;
            V3 v0 = s.p(s.tri[3 * t]), v1 = s.p(s.tri[3 * t + 1]), v2 = s.p(s.tri[3 * t + 2]);
            areas[k] = 0.5f * len(crossp(sub(v1, v0), sub(v2, v0)));
        }
        lt.tri_alias[l] = alias_build(areas.data(), cnt);
    }
    lt.power = s.render[4] != 0.f;
    if (lt.power) {
        std::vector<float> w(s.nl);
        for (uint32_t l = 0; l < s.nl; ++l) w[l] = luminance(light_power(s, l, lt));
        lt.select = alias_build(w.data(), s.nl);
    }
    return lt;
}

float select_pmf(const Scene &s, const LightTables &lt, uint32_t l) {
    return lt.power ? alias_pmf(lt.select, l) : 1.f / float(s.nl);
}

struct LSample {
    V3 L;
    float pdf;
    V3 p;
};

// LightSampler::sample_wi (src/base/illumination/lightsampler.cpp:199-216),
// AreaLight::sample_wi (src/render_core/lights/area.cpp:120-149),
// IPointLight::sample_wi (src/base/illumination/light.cpp:42-51),
// PointLight::Le (point.cpp:39-43), SpotLight::Le/falloff (spot.cpp:51-70)
LSample sample_light(const Scene &s, const LightTables &lt, const Surf &it, Sampler &sp) {
    LSample ls{};
    float ul = sp.next1d();
    float ux = sp.next1d(), uy = sp.next1d();
    uint32_t l;
    float pmf_sel;
    if (!lt.power) {
        l = uint32_t(std::min(ul * float(s.nl), float(s.nl) - 1.f));
        pmf_sel = 1.f / float(s.nl);
    } else {
        l = alias_sample(lt.select, ul, nullptr);
        pmf_sel = alias_pmf(lt.select, l);
    }
    V3 color = mul(v3(s.lightf(l, 8), s.lightf(l, 9), s.lightf(l, 10)), s.lightf(l, 11));
    uint32_t type = s.lightu(l, 0);
    if (type == 0) {
        const Alias &a = lt.tri_alias[l];
        float urem;
        uint32_t prim = alias_sample(a, ux, &urem);
        float pmf = alias_pmf(a, prim);
        float su = std::sqrt(urem);
        float b1 = 1.f - su, b2 = uy * su;
        Surf p = surface(s, s.lightu(l, 3) + prim, b1, b2, false);
        float pdf_pos = (1.f / p.area) * pmf;
        V3 w = sub(it.pos, p.pos);
        bool front = dotp(w, p.ng) > 0.f || s.lightu(l, 1);
        ls.L = front ? color : v3(0, 0, 0);
        float pdf = pdf_pos * dotp(w, w) / std::fabs(dotp(p.ng, unit(w)));
        ls.pdf = (std::isinf(pdf) || std::isnan(pdf)) ? 0.f : pdf;
        float f = dotp(p.ng, w) > 0.f ? 1.f : -1.f;
        ls.p = offset_origin(p.pos, mul(p.ng, f * s.render[5]));
    } else {
        V3 pos = v3(s.lightf(l, 12), s.lightf(l, 13), s.lightf(l, 14));
        V3 wu = sub(it.pos, pos);
        V3 val = mul(color, 1.f / dotp(wu, wu));
        if (type == 2) {
            V3 dir = v3(s.lightf(l, 16), s.lightf(l, 17), s.lightf(l, 18));
            float angle = s.lightf(l, 15), fall = s.lightf(l, 19);
            float start = std::max(0.f, angle - fall);
            float ca = std::cos(angle), cs = std::cos(start);// spot.cpp:51-59
            float ct = std::min(std::max(dotp(dir, unit(wu)), ca), cs);
            float f = (ct - ca) / (cs - ca);
            val = mul(val, f * f * f * f);
        }
        ls.L = val;
        ls.pdf = -1.f;
        ls.p = pos;
    }
    ls.pdf *= pmf_sel;
    return ls;
}

// LightSampler::evaluate_hit_wi (lightsampler.cpp:252-267)
void eval_hit(const Scene &s, const LightTables &lt, const Surf &it, uint32_t l, uint32_t tri, V3 ref, V3 &L, float &pdf) {
    uint32_t prim = tri - s.lightu(l, 3);
    float pdf_pos = (1.f / it.area) * alias_pmf(lt.tri_alias[l], prim);
    V3 w = sub(ref, it.pos);
    bool front = dotp(w, it.ng) > 0.f || s.lightu(l, 1);
    L = front ? mul(v3(s.lightf(l, 8), s.lightf(l, 9), s.lightf(l, 10)), s.lightf(l, 11)) : v3(0, 0, 0);
    float p = pdf_pos * dotp(w, w) / std::fabs(dotp(it.ng, unit(w)));
    p = (std::isinf(p) || std::isnan(p)) ? 0.f : p;
    pdf = p * select_pmf(s, lt, l);
}

// ---------------------------------------------------------------- materials
struct Eval {
    V3 f;
    float pdf;
};

float sin_theta(V3 w) { return std::sqrt(std::max(0.f, 1.f - w.z * w.z)); }
float cos_phi(V3 w) {
    float st = sin_theta(w);
    return st == 0.f ? 1.f : std::min(std::max(w.x / st, -1.f), 1.f);
}
float sin_phi(V3 w) {
    float st = sin_theta(w);
    return st == 0.f ? 0.f : std::min(std::max(w.y / st, -1.f), 1.f);
}

// MaterialEvaluator::evaluate (src/base/scattering/material.cpp:132-148);
// Lambert (bxdf.h:82-96), Oren-Nayar (bxdf.cpp:94-121), black body
// (src/render_core/materials/black_body.cpp:36-43)
Eval eval_mat(const Scene &s, uint32_t m, const Surf &it, V3 wi_w) {
    Eval e{};
    uint32_t type = s.mat_type(m);
    V3 wo = it.fr.local(it.wo), wi = it.fr.local(wi_w);
    if (type == 2) {
        e.f = v3(0, 0, 0);
        e.pdf = 1.f;
    } else {
        bool same = wo.z * wi.z > 0.f;
        V3 R = v3(s.matp(m, 0), s.matp(m, 1), s.matp(m, 2));
        float k = kInvPi;
        if (type == 1) {
            float sig = s.matp(m, 3) * (kPi * 0.5f);
            float s2 = (sig * sig) * (sig * sig);
            float A = 1.f - (s2 / (2.f * (s2 + 0.33f)));
            float B = 0.45f * s2 / (s2 + 0.09f);
            float sti = sin_theta(wi), sto = sin_theta(wo);
            float dcos = cos_phi(wi) * cos_phi(wo) + sin_phi(wi) * sin_phi(wo);
            float maxc = std::max(0.f, dcos);
            bool cond = std::fabs(wi.z) > std::fabs(wo.z);
            float sin_a = cond ? sto : sti;
            float tan_b = cond ? sti / std::fabs(wi.z) : sto / std::fabs(wo.z);
            k = kInvPi * (A + B * maxc * sin_a * tan_b);
        }
        e.f = same ? mul(R, k) : v3(0, 0, 0);
        e.pdf = same ? std::fabs(wi.z) * kInvPi : 0.f;
        e.f = mul(e.f, std::fabs(wi.z));
    }
    if (!(dotp(it.wo, it.ng) * dotp(wi_w, it.ng) > 0.f)) e.pdf = 0.f;
    return e;
}

struct BSample {
    Eval e;
    V3 wi;
    float eta;
};

// MaterialEvaluator::sample (material.cpp:162-184); BxDF::sample_wi (bxdf.cpp:42-46)
BSample sample_mat(const Scene &s, uint32_t m, const Surf &it, Sampler &sp) {
    BSample b{};
    b.eta = 1.f;
    if (s.mat_type(m) == 2) {
        b.wi = it.wo;
        b.e = {v3(0, 0, 0), 1.f};
        if (!(dotp(it.wo, it.ng) * dotp(b.wi, it.ng) > 0.f)) b.e.pdf = 0.f;
        return b;
    }
    V3 wo = it.fr.local(it.wo);
    float u0 = sp.next1d(), u1 = sp.next1d();
    V3 wi = cosine_hemisphere(u0, u1);
    if (wo.z < 0.f) wi.z = -wi.z;
    b.wi = it.fr.world(wi);
    b.e = eval_mat(s, m, it, b.wi);
    return b;
}

// ---------------------------------------------------------------- integrator
struct Ctx {
    const Scene *s;
    const Bvh2 *bvh;
    const LightTables *lt;
    bool brute;
    Hit trace(const Ray &r, bool any) const { return brute ? trace_brute(*s, r, any) : trace_bvh(*s, *bvh, r, any); }
};

// IlluminationIntegrator::Li (src/base/integrator.cpp:134-283) in sRGB mode
V3 Li(const Ctx &c, V3 ro, V3 rd, Sampler &sp) {
    const Scene &s = *c.s;
    uint32_t max_depth = uint32_t(s.render[0]), min_depth = uint32_t(s.render[1]);
    float rr_th = s.render[2];
    int mis_mode = int(s.render[3]);
    V3 ret = v3(0, 0, 0), thr = v3(1, 1, 1);
    float scatter_pdf = 1e16f, eta_scale = 1.f;
    uint32_t skips = 0;
    for (uint32_t b = 0; b < max_depth; ++b) {
        Hit h = c.trace({ro, rd, 0.f, FLT_MAX}, false);
        if (h.tri == kInv) break;
        Surf it = surface(s, h.tri, h.b1, h.b2, true);
        it.wo = unit(neg(rd));
        uint32_t inst = it.inst;
        uint32_t mat = s.inst[4 * inst], light = s.inst[4 * inst + 1];
        if (mat == kInv) {
            V3 n = dotp(it.ng, rd) > 0.f ? it.ng : neg(it.ng);
            ro = offset_origin(it.pos, n);
            if (++skips > 64u) break;
            --b;
            continue;
        }
        if (light != kInv) {
            V3 Le;
            float lpdf;
            eval_hit(s, *c.lt, it, light, h.tri, ro, Le, lpdf);
            float w = scatter_pdf / (scatter_pdf + lpdf);
            if (mis_mode == 2) w = 1.f;
            else if (mis_mode == 1) w = b == 0 ? w : 0.f;
            ret = add(ret, v3(Le.x * thr.x * w, Le.y * thr.y * w, Le.z * thr.z * w));
        }
        LSample ls = sample_light(s, *c.lt, it, sp);
        Eval se = eval_mat(s, mat, it, unit(sub(ls.p, it.pos)));
        BSample bs = sample_mat(s, mat, it, sp);
        bool delta = ls.pdf < 0.f;
        float w = delta ? 1.f : ls.pdf / (ls.pdf + se.pdf);
        if (mis_mode == 1) w = 1.f;
        float lpdf = delta ? -ls.pdf : ls.pdf;
        if (se.pdf > 0.f && lpdf > 0.f && mis_mode != 2) {
            V3 dir = sub(ls.p, it.pos);
            V3 n = dotp(it.ng, dir) > 0.f ? it.ng : neg(it.ng);
            Hit sh = c.trace({offset_origin(it.pos, n), dir, 0.f, 1.f - kShadowEps}, true);
            if (sh.tri == kInv) {
                V3 Ld = v3(ls.L.x * se.f.x * w / lpdf, ls.L.y * se.f.y * w / lpdf, ls.L.z * se.f.z * w / lpdf);
                ret = add(ret, v3(thr.x * Ld.x, thr.y * Ld.y, thr.z * Ld.z));
            }
        }
        eta_scale *= bs.eta * bs.eta;
        float lum = std::max(thr.x, std::max(thr.y, thr.z));
        if (!(bs.e.pdf > 0.f) || lum == 0.f) break;
        thr = v3(thr.x * (bs.e.f.x / bs.e.pdf), thr.y * (bs.e.f.y / bs.e.pdf), thr.z * (bs.e.f.z / bs.e.pdf));
        if (eta_scale * lum < rr_th && b >= min_depth) {
            float q = std::min(0.95f, lum);
            float rr = sp.next1d();
            if (q < rr) break;
            thr = v3(thr.x / q, thr.y / q, thr.z / q);
        }
        scatter_pdf = bs.e.pdf;
        V3 n = dotp(it.ng, bs.wi) > 0.f ? it.ng : neg(it.ng);
        ro = offset_origin(it.pos, n);
        rd = bs.wi;
    }
    return ret;
}

// Sensor::generate_ray (src/base/sensor/sensor.cpp:44-56), thin lens
// (src/render_core/sensors/thin_lens.cpp:34-42), Sampler::sensor_sample
// (src/base/sampler.h:65-73), box / tent filters (box.cpp:16-20, triangle.cpp:16-18)
void camera_ray(const Scene &s, uint32_t px, uint32_t py, Sampler &sp, V3 &o, V3 &d) {
    const float *c = s.cam;
    float ux = sp.next1d(), uy = sp.next1d();
    float r = c[22];
    float fx, fy;
    if (c[21] == 0.f) {
        fx = -r + ux * (r + r);
        fy = -r + uy * (r + r);
    } else {
        fx = sample_tent(ux, r);
        fy = sample_tent(uy, r);
    }
    float lx = sp.next1d(), ly = sp.next1d();
    (void)sp.next1d();
    float ffx = float(px) + 0.5f + fx, ffy = float(py) + 0.5f + fy;
    float sx = c[12] + ffx * c[14], sy = c[13] - ffy * c[15];
    V3 dc = unit(v3(sx * c[16], sy * c[16], 1.f));
    V3 oc = v3(0, 0, 0);
    if (c[23] != 0.f && c[17] > 0.f) {
        float dx, dy;
        square_to_disk(lx, ly, dx, dy);
        dx *= c[17];
        dy *= c[17];
        float ft = c[18] / dc.z;
        V3 pf = mul(dc, ft);
        oc = v3(dx, dy, 0.f);
        dc = unit(sub(pf, oc));
    }
    o = v3(c[0] * oc.x + c[1] * oc.y + c[2] * oc.z + c[3], c[4] * oc.x + c[5] * oc.y + c[6] * oc.z + c[7],
           c[8] * oc.x + c[9] * oc.y + c[10] * oc.z + c[11]);
    d = v3(c[0] * dc.x + c[1] * dc.y + c[2] * dc.z, c[4] * dc.x + c[5] * dc.y + c[6] * dc.z,
           c[8] * dc.x + c[9] * dc.y + c[10] * dc.z);
}

float zero_nan_inf(float v) { return (std::isnan(v) || std::isinf(v)) ? 0.f : v; }

Scene make_scene(const float *P, const float *N, const float *UV, const uint32_t *tri, const uint32_t *tri_inst,
                 const uint32_t *inst, const uint32_t *mat, const uint32_t *light, const float *cam, const float *render,
                 uint32_t nv, uint32_t nt, uint32_t ni, uint32_t nm, uint32_t nl) {
    Scene s{P, N, UV, tri, tri_inst, inst, mat, light, cam, render, nv, nt, ni, nm, nl};
    return s;
}

}// namespace

extern "C" {

uint32_t or_tea(uint32_t v0, uint32_t v1, const uint32_t *key, uint32_t rounds) {
    return tea_rounds(v0, v1, key ? key : kVisionKey, rounds);
}
uint32_t or_tea_v1(uint32_t v0, uint32_t v1, const uint32_t *key, uint32_t rounds) {
    // second word of the TEA block, for the full-cipher known answer
    const uint32_t *k = key ? key : kVisionKey;
    uint32_t s = 0;
    for (uint32_t n = 0; n < rounds; ++n) {
        s += 0x9e3779b9u;
        v0 += ((v1 << 4) + k[0]) ^ (v1 + s) ^ ((v1 >> 5) + k[1]);
        v1 += ((v0 << 4) + k[2]) ^ (v0 + s) ^ ((v0 >> 5) + k[3]);
    }
    return v1;
}

// the LCG recurrence from an arbitrary state (util.h:27-34)
void or_lcg(uint32_t state, uint32_t count, uint32_t *states, float *values) {
    Sampler s;
    s.state = state;
    for (uint32_t i = 0; i < count; ++i) {
        values[i] = s.next1d();
        states[i] = s.state;
    }
}

// `count` draws of the sampler stream (pixel, index, dim)
void or_sampler_stream(uint32_t px, uint32_t py, uint32_t index, uint32_t dim, uint32_t count, float *out,
                       uint32_t *states) {
    Sampler s;
    s.start(px, py, index, dim);
    if (states) states[0] = s.state;
    for (uint32_t i = 0; i < count; ++i) {
        out[i] = s.next1d();
        if (states) states[i + 1] = s.state;
    }
}

void or_cosine_hemisphere(const float *u, uint32_t n, float *out) {
    for (uint32_t i = 0; i < n; ++i) {
        V3 w = cosine_hemisphere(u[2 * i], u[2 * i + 1]);
        out[3 * i] = w.x, out[3 * i + 1] = w.y, out[3 * i + 2] = w.z;
    }
}

void or_square_to_triangle(const float *u, uint32_t n, float *out) {
    for (uint32_t i = 0; i < n; ++i) {
        float su = std::sqrt(u[2 * i]);
        out[2 * i] = 1.f - su;
        out[2 * i + 1] = u[2 * i + 1] * su;
    }
}

void or_sample_tent(const float *u, uint32_t n, float r, float *out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = sample_tent(u[i], r);
}

void or_alias_build(const float *w, uint32_t n, float *prob, uint32_t *alias, float *integral) {
    Alias a = alias_build(w, n);
    std::memcpy(prob, a.prob.data(), n * 4);
    std::memcpy(alias, a.alias.data(), n * 4);
    *integral = a.integral;
}

void or_alias_sample(const float *w, uint32_t nw, const float *u, uint32_t n, uint32_t *idx, float *u_rem, float *pmf) {
    Alias a = alias_build(w, nw);
    for (uint32_t i = 0; i < n; ++i) {
        idx[i] = alias_sample(a, u[i], u_rem ? &u_rem[i] : nullptr);
        if (pmf) pmf[i] = alias_pmf(a, idx[i]);
    }
}

void or_offset_ray_origin(const float *p, const float *n, uint32_t count, float *out) {
    for (uint32_t i = 0; i < count; ++i) {
        V3 r = offset_origin(v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]), v3(n[3 * i], n[3 * i + 1], n[3 * i + 2]));
        out[3 * i] = r.x, out[3 * i + 1] = r.y, out[3 * i + 2] = r.z;
    }
}

// Oren-Nayar / Lambert evaluation in a given local frame (for closed-form checks)
void or_eval_diffuse(uint32_t type, const float *color, float sigma, const float *wo, const float *wi, uint32_t n,
                     float *f, float *pdf) {
    uint32_t mat[32] = {};
    mat[0] = type;
    std::memcpy(&mat[2], color, 12);
    std::memcpy(&mat[5], &sigma, 4);
    uint32_t inst[4] = {0, kInv, 0, 0};
    Scene s{};
    s.mat = mat;
    s.inst = inst;
    for (uint32_t i = 0; i < n; ++i) {
        Surf it{};
        it.fr = {v3(1, 0, 0), v3(0, 1, 0), v3(0, 0, 1)};
        it.ng = v3(0, 0, 1);
        it.wo = v3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
        Eval e = eval_mat(s, 0, it, v3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]));
        f[3 * i] = e.f.x, f[3 * i + 1] = e.f.y, f[3 * i + 2] = e.f.z;
        pdf[i] = e.pdf;
    }
}

// closest / any hit; mode 0 brute force, 1 BVH2
void or_trace(const float *P, const uint32_t *tri, uint32_t nv, uint32_t nt, const float *rays, uint32_t n, int any,
              int mode, uint32_t *hit_tri, float *hit_t, float *hit_bary) {
    Scene s{};
    s.P = P;
    s.tri = tri;
    s.nv = nv;
    s.nt = nt;
    Bvh2 b;
    if (mode == 1) bvh2_build(s, b);
    auto work = [&](uint32_t lo, uint32_t hi) {
        for (uint32_t i = lo; i < hi; ++i) {
            const float *r = rays + 8 * i;
            Ray ray{v3(r[0], r[1], r[2]), v3(r[4], r[5], r[6]), r[3], r[7]};
            Hit h = mode == 1 ? trace_bvh(s, b, ray, any != 0) : trace_brute(s, ray, any != 0);
            hit_tri[i] = h.tri;
            hit_t[i] = h.tri == kInv ? ray.tmax : h.t;
            hit_bary[2 * i] = h.b1;
            hit_bary[2 * i + 1] = h.b2;
        }
    };
    uint32_t nth = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    if (n < 4096) nth = 1;
    std::vector<std::thread> th;
    uint32_t chunk = (n + nth - 1) / nth;
    for (uint32_t t = 0; t < nth; ++t) {
        uint32_t lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo < hi) th.emplace_back(work, lo, hi);
    }
    for (auto &t : th) t.join();
}

// Render pixels of the window [x0,x1)x[y0,y1) for frames
// [frame_start, frame_start+spp) into rt (W*H*4, running mean in place,
// NormalRadianceCollector::add_sample, normal.cpp:87-101).  mode 0 brute
// force, 1 BVH2.  threads = 0 uses all hardware threads.
// Returns the number of rays cast.
uint64_t or_render(const float *P, const float *N, const float *UV, const uint32_t *tri, const uint32_t *tri_inst,
                   const uint32_t *inst, const uint32_t *mat, const uint32_t *light, const float *cam,
                   const float *render, uint32_t nv, uint32_t nt, uint32_t ni, uint32_t nm, uint32_t nl,
                   uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1, uint32_t frame_start, uint32_t spp, int mode,
                   uint32_t threads, float *rt) {
    Scene s = make_scene(P, N, UV, tri, tri_inst, inst, mat, light, cam, render, nv, nt, ni, nm, nl);
    Bvh2 b;
    if (mode == 1) bvh2_build(s, b);
    LightTables lt = build_lights(s);
    Ctx c{&s, &b, &lt, mode == 0};
    uint32_t W = uint32_t(cam[19]);
    bool accumulate = render[6] != 0.f;
    std::atomic<uint32_t> next_row{y0};
    auto work = [&]() {
        while (true) {
            uint32_t y = next_row.fetch_add(1);
            if (y >= y1) break;
            for (uint32_t x = x0; x < x1; ++x) {
                float *acc = rt + 4 * (size_t(y) * W + x);
                for (uint32_t f = frame_start; f < frame_start + spp; ++f) {
                    Sampler sp;
                    sp.start(x, y, f, 0u);
                    V3 o, d;
                    camera_ray(s, x, y, sp, o, d);
                    V3 L = Li(c, o, d, sp);
                    float v[4] = {zero_nan_inf(L.x), zero_nan_inf(L.y), zero_nan_inf(L.z), 1.f};
                    if (accumulate) {
                        float a = 1.f / float(f + 1u);
                        for (int k = 0; k < 4; ++k) acc[k] = acc[k] + a * (v[k] - acc[k]);
                    } else {
                        for (int k = 0; k < 4; ++k) acc[k] = v[k];
                    }
                }
            }
        }
    };
    uint32_t nth = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::thread> th;
    for (uint32_t t = 0; t < nth; ++t) th.emplace_back(work);
    for (auto &t : th) t.join();
    return 0;
}

}// extern "C"
