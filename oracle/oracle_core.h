// CPU oracle, part 1 (TEST INFRASTRUCTURE ONLY): sampler, warps, alias
// tables, scene view, watertight intersection, BVH2, ray offset and shading
// frames.  Included once, by oracle.cpp.
#pragma once

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

// The elementary functions shared with the HIP kernels: sampling decisions
// depend on their last bits, so both sides evaluate the same operation
// sequence (tests/test_elementary.py pins them against mpmath).
#include "../vision_amd/csrc/elementary.h"

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
    float sn, cs;
    vs::em::sincos(th, &sn, &cs);
    x = r * cs;
    y = r * sn;
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
    float u, v;// texture coordinates (barycentric blend of the vertices')
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
    {
        float u0 = s.UV[2 * i0], v0 = s.UV[2 * i0 + 1], u1 = s.UV[2 * i1], v1 = s.UV[2 * i1 + 1];
        float u2 = s.UV[2 * i2], v2 = s.UV[2 * i2 + 1];
        o.u = u0 * b0 + u1 * b1 + u2 * b2;
        o.v = v0 * b0 + v1 * b1 + v2 * b2;
    }
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

}// namespace
