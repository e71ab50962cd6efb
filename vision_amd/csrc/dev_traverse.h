// BVH8 closest-hit / any-hit traversal for gfx950 (replaces the reference's
// OptiX Accel::trace_closest / trace_occlusion, src/base/mgr/geometry.cpp:167-184).
//
// Per lane: one stack of node groups in LDS (entry = child_base, remaining
// hit bits in near-to-far visiting order, internal-child mask), at most one
// entry per tree level.  Leaf triangles of a node are tested before the next
// node is fetched, so closest-hit culling uses the tightest t.
//
// The ray/triangle test is the watertight test of Woop, Benthin and Wald
// (JCGT 2013) with FP contraction disabled, so the oracle's scalar restatement
// produces bit-identical (t, barycentrics, triangle) for identical rays.
// Valid hits satisfy tmin < t < tmax; equal t resolves to the lower triangle id.
#pragma once

#include "dev_common.h"
#include "dev_scene.h"

namespace vs {
namespace dev {

struct RayPre {
    f3 o, d;
    float tmin;
    f3 idir;
    uint32_t oct;
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

__device__ __forceinline__ float sel3(f3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

__device__ __forceinline__ float safe_rcp(float d) {
    const float eps = 1e-12f;
    return 1.f / (fabsf(d) > eps ? d : copysignf(eps, d));
}

__device__ __forceinline__ RayPre make_ray_pre(f3 o, f3 d, float tmin) {
    RayPre r;
    r.o = o;
    r.d = d;
    r.tmin = tmin;
    r.idir = mk3(safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z));
    r.oct = (d.x < 0.f ? 1u : 0u) | (d.y < 0.f ? 2u : 0u) | (d.z < 0.f ? 4u : 0u);
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (sel3(d, kz) < 0.f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    float dz = sel3(d, kz);
    r.Sx = sel3(d, kx) / dz;
    r.Sy = sel3(d, ky) / dz;
    r.Sz = 1.f / dz;
    return r;
}

// watertight ray/triangle test; returns false on a miss
__device__ __forceinline__ bool intersect_triangle(const RayPre &r, f3 v0, f3 v1, f3 v2,
                                                   float &t, float &b1, float &b2) {
#pragma clang fp contract(off)
    f3 A = v0 - r.o, B = v1 - r.o, C = v2 - r.o;
    float Akz = sel3(A, r.kz), Bkz = sel3(B, r.kz), Ckz = sel3(C, r.kz);
    float Ax = sel3(A, r.kx) - r.Sx * Akz;
    float Ay = sel3(A, r.ky) - r.Sy * Akz;
    float Bx = sel3(B, r.kx) - r.Sx * Bkz;
    float By = sel3(B, r.ky) - r.Sy * Bkz;
    float Cx = sel3(C, r.kx) - r.Sx * Ckz;
    float Cy = sel3(C, r.ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.f || V == 0.f || W == 0.f) {
        double CxBy = double(Cx) * double(By), CyBx = double(Cy) * double(Bx);
        U = float(CxBy - CyBx);
        double AxCy = double(Ax) * double(Cy), AyCx = double(Ay) * double(Cx);
        V = float(AxCy - AyCx);
        double BxAy = double(Bx) * double(Ay), ByAx = double(By) * double(Ax);
        W = float(BxAy - ByAx);
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
    float det = U + V + W;
    if (det == 0.f) return false;
    float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    t = T / det;
    b1 = V / det;
    b2 = W / det;
    return true;
}

struct TraceHit {
    float t;
    uint32_t tri;// global triangle id, kInvalid on a miss
    float b1, b2;
};

__device__ __forceinline__ uint32_t permute_oct(uint32_t m, uint32_t oct) {
    if (oct & 1u) m = ((m & 0x55u) << 1) | ((m & 0xAAu) >> 1);
    if (oct & 2u) m = ((m & 0x33u) << 2) | ((m & 0xCCu) >> 2);
    if (oct & 4u) m = ((m & 0x0Fu) << 4) | ((m & 0xF0u) >> 4);
    return m;
}

__device__ __forceinline__ float qbyte(uint32_t lo_word, uint32_t hi_word, int s) {
    uint32_t w = s < 4 ? lo_word : hi_word;
    return float((w >> ((s & 3) * 8)) & 0xFFu);
}

// intersect the 8 child boxes of node `idx`; outputs the internal-child hit
// mask in visiting order and the leaf-triangle mask relative to tri_base
__device__ __forceinline__ void process_node(const uint4 *__restrict__ nodes, uint32_t idx, const RayPre &r,
                                             float tmax, uint32_t &cbase, uint32_t &chits, uint32_t &cimask,
                                             uint32_t &tbase, uint32_t &tmask, uint32_t &nfetch) {
    const uint4 *np = nodes + size_t(idx) * 5;
    uint4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3], n4 = np[4];
    ++nfetch;
    float sx = __uint_as_float((n0.w & 0xFFu) << 23) * r.idir.x;
    float sy = __uint_as_float(((n0.w >> 8) & 0xFFu) << 23) * r.idir.y;
    float sz = __uint_as_float(((n0.w >> 16) & 0xFFu) << 23) * r.idir.z;
    float ax = (__uint_as_float(n0.x) - r.o.x) * r.idir.x;
    float ay = (__uint_as_float(n0.y) - r.o.y) * r.idir.y;
    float az = (__uint_as_float(n0.z) - r.o.z) * r.idir.z;
    bool px = r.idir.x >= 0.f, py = r.idir.y >= 0.f, pz = r.idir.z >= 0.f;
    // per axis the quantised planes the ray enters and leaves by: the low
    // ones first where its direction is positive (chosen once per node, not
    // per child)
    const uint32_t nx0 = px ? n2.x : n3.z, nx1 = px ? n2.y : n3.w, fx0 = px ? n3.z : n2.x, fx1 = px ? n3.w : n2.y;
    const uint32_t ny0 = py ? n2.z : n4.x, ny1 = py ? n2.w : n4.y, fy0 = py ? n4.x : n2.z, fy1 = py ? n4.y : n2.w;
    const uint32_t nz0 = pz ? n3.x : n4.z, nz1 = pz ? n3.y : n4.w, fz0 = pz ? n4.z : n3.x, fz1 = pz ? n4.w : n3.y;
    uint32_t imask = n0.w >> 24;
    uint32_t hitmask = 0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        float nx = fmaf(qbyte(nx0, nx1, s), sx, ax), fx = fmaf(qbyte(fx0, fx1, s), sx, ax);
        float ny = fmaf(qbyte(ny0, ny1, s), sy, ay), fy = fmaf(qbyte(fy0, fy1, s), sy, ay);
        float nz = fmaf(qbyte(nz0, nz1, s), sz, az), fz = fmaf(qbyte(fz0, fz1, s), sz, az);
        float tn = fmaxf(fmaxf(nx, ny), fmaxf(nz, r.tmin));
        float tf = fminf(fminf(fx, fy), fminf(fz, tmax));
        hitmask |= (tn <= tf ? 1u : 0u) << s;
    }
    uint32_t leafhits = hitmask & ~imask;
    uint32_t tm = 0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        uint32_t meta = ((s < 4 ? n1.z : n1.w) >> ((s & 3) * 8)) & 0xFFu;
        uint32_t cnt = meta >> 5, off = meta & 31u;
        uint32_t bits = ((1u << cnt) - 1u) << off;
        tm |= (leafhits >> s) & 1u ? bits : 0u;
    }
    tmask = tm;
    tbase = n1.y;
    cbase = n1.x;
    cimask = imask;
    chits = permute_oct(hitmask & imask, r.oct);
}

template <bool ANY>
__device__ __forceinline__ TraceHit traverse(const SceneView &sc, f3 o, f3 d, float tmin, float tmax,
                                             uint2 *stack, uint32_t stride, uint32_t &nfetch, uint32_t &tfetch) {
    const uint4 *__restrict__ nodes = reinterpret_cast<const uint4 *>(sc.nodes);
    const float4 *__restrict__ ltri = reinterpret_cast<const float4 *>(sc.leaf_tris);
    RayPre r = make_ray_pre(o, d, tmin);
    TraceHit h;
    h.t = tmax;
    h.tri = kInvalid;
    h.b1 = h.b2 = 0.f;
    uint32_t sp = 0;
    uint32_t gbase, ghits, gimask, tbase, tmask;
    process_node(nodes, 0, r, tmax, gbase, ghits, gimask, tbase, tmask, nfetch);
    while (true) {
        while (tmask) {
            uint32_t j = __builtin_ctz(tmask);
            tmask &= tmask - 1u;
            const float4 *tp = ltri + size_t(tbase + j) * 3;
            float4 a = tp[0], b = tp[1], c = tp[2];
            ++tfetch;
            float t, b1, b2;
            if (intersect_triangle(r, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(c.x, c.y, c.z), t, b1, b2)) {
                uint32_t id = __float_as_uint(a.w);
                if (t > tmin && t < tmax && (t < h.t || (t == h.t && id < h.tri))) {
                    h.t = t;
                    h.tri = id;
                    h.b1 = b1;
                    h.b2 = b2;
                    if (ANY) return h;
                }
            }
        }
        if (ghits == 0) {
            if (sp == 0) break;
            --sp;
            uint2 e = stack[sp * stride];
            gbase = e.x;
            ghits = e.y & 0xFFu;
            gimask = e.y >> 8;
        }
        uint32_t i = __builtin_ctz(ghits);
        ghits &= ghits - 1u;
        uint32_t slot = i ^ r.oct;
        uint32_t child = gbase + __popc(gimask & ((1u << slot) - 1u));
        uint32_t nb, nh, ni;
        process_node(nodes, child, r, ANY ? tmax : h.t, nb, nh, ni, tbase, tmask, nfetch);
        if (nh) {
            if (ghits) {
                if (sp < sc.stack_size) stack[sp * stride] = make_uint2(gbase, ghits | (gimask << 8));
                ++sp;
                if (sp > sc.stack_size) sp = sc.stack_size;// unreachable: stack sized to tree depth
            }
            gbase = nb;
            ghits = nh;
            gimask = ni;
        }
    }
    return h;
}

}// namespace dev
}// namespace vs
