// G-buffer and SVGF denoiser kernels for gfx950.
//
// The G-buffer kernel traces one camera ray per pixel (8x8 pixel tiles, one
// wave each, traversal stack in LDS).  The SVGF passes are 2D stencils over
// 32-byte records (16x16 pixels per 256-thread block): each reads its
// neighbourhood through the L1/L2 caches and writes one record, so every
// pass is bound by HBM bandwidth (algorithmic bytes per pixel: reproject
// 32 B G-buffer x 2 + 32 B of the previous frame's record x 4 taps + 48 B of
// radiance / albedo / emission in, 32 B out; moments and a-trous 32 B in and
// 32 B out once neighbourhoods are cache-resident).
#include <hip/hip_runtime.h>

#include "dev_pt.h"
#include "dev_restir.h"
#include "svgf.h"
#include "xcd.h"

namespace vs {

using namespace dev;

namespace {

struct Geo {
    f3 n;
    float fw;
    f2 pf;
    float dg, depth;
};
struct Rec {
    f3 il;
    float var, hist, m1, m2;
};

__device__ __forceinline__ float4 *geo_field(const SvgfView &v, uint32_t slot, uint32_t f) {
    return v.geo + (size_t(slot) * 2u + f) * v.pixels;
}
__device__ __forceinline__ float4 *rec_field(const SvgfView &v, uint32_t buf, uint32_t f) {
    return v.rec + (size_t(buf) * 2u + f) * v.pixels;
}
__device__ __forceinline__ Geo load_geo(const SvgfView &v, uint32_t slot, uint32_t i) {
    const float4 a = geo_field(v, slot, 0)[i], b = geo_field(v, slot, 1)[i];
    Geo g;
    g.n = mk3(a.x, a.y, a.z);
    g.fw = a.w;
    g.pf = mk2(b.x, b.y);
    g.dg = b.z;
    g.depth = b.w;
    return g;
}
__device__ __forceinline__ Rec load_rec(const SvgfView &v, uint32_t buf, uint32_t i) {
    const float4 a = rec_field(v, buf, 0)[i], b = rec_field(v, buf, 1)[i];
    Rec r;
    r.il = mk3(a.x, a.y, a.z);
    r.var = a.w;
    r.hist = b.x;
    r.m1 = b.y;
    r.m2 = b.z;
    return r;
}
__device__ __forceinline__ void store_rec(const SvgfView &v, uint32_t buf, uint32_t i, const Rec &r) {
    rec_field(v, buf, 0)[i] = make_float4(r.il.x, r.il.y, r.il.z, r.var);
    rec_field(v, buf, 1)[i] = make_float4(r.hist, r.m1, r.m2, 0.f);
}

__device__ __forceinline__ float lum3(f3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }
__device__ __forceinline__ float sat(float x) { return fminf(fmaxf(x, 0.f), 1.f); }
__device__ __forceinline__ bool in_screen(int x, int y, int W, int H) { return x >= 0 && y >= 0 && x < W && y < H; }

// SVGF::cal_weight (svgf.cpp:69-78)
__device__ __forceinline__ float cal_weight(float cd, float nd, float sd, f3 cn, f3 nn, float sn, float ci, float ni,
                                            float si) {
    const float dn = cn.x * nn.x + cn.y * nn.y + cn.z * nn.z;
    const float w_normal = em::pow(sat(dn), sn);
    const float w_depth = sd == 0.f ? 0.f : fabsf(cd - nd) / sd;
    const float w_illum = fabsf(ci - ni) / si;
    return em::exp(-(fmaxf(w_illum, 0.f) + fmaxf(w_depth, 0.f))) * w_normal;
}

// Reproject::is_valid_reproject (reproject.cpp:14-27)
__device__ __forceinline__ bool valid_reproject(const Geo &cur, const Geo &prev, int W, int H) {
    const bool inside = in_screen(int(prev.pf.x), int(prev.pf.y), W, H);
    const bool z_ok = (fabsf(cur.depth - prev.depth) / (cur.dg + 1e-2f)) < 10.f;
    const float dx = cur.n.x - prev.n.x, dy = cur.n.y - prev.n.y, dz = cur.n.z - prev.n.z;
    const bool n_ok = (sqrtf(dx * dx + dy * dy + dz * dz) / (cur.fw + 1e-2f)) < 16.f;
    return inside && z_ok && n_ok;
}

// ---- G-buffer (FrameBuffer::compile_compute_geom, frame_buffer.cpp:72-122)
template <int NS, bool RGB, int MS>
__global__ void __launch_bounds__(64) gbuffer_kernel(SceneView sc, CameraView cam, RestirView rv, SvgfView sv,
                                                     unsigned long long *__restrict__ counters) {
    extern __shared__ uint2 lds_stack[];
    const uint2 tile = xcd_tile_2d();
    const uint32_t px = tile.x * 8u + (threadIdx.x & 7u);
    const uint32_t py = tile.y * 8u + (threadIdx.x >> 3);
    if (px >= sv.width || py >= sv.height) return;
    const uint32_t pix = py * sv.width + px;
    Swl<NS> swl = sample_wavelengths<NS, RGB>(px, py, sv.frame);
    Sampler sp;
    sp.start(px, py, sv.frame, 0u);
    f3 o, d;
    float fw;
    f2 p_film;
    camera_ray(cam, px, py, sp, o, d, fw, &p_film);
    uint32_t nodes = 0, tris = 0;
    const TraceHit h = traverse<false>(sc, o, d, 0.f, 3.402823466e38f, lds_stack + threadIdx.x, 64u, nodes, tris);
    float4 g0 = make_float4(0.f, 0.f, 0.f, 0.f), g1 = make_float4(p_film.x, p_film.y, 0.f, -1.f);
    f3 a = mk3(0.f, 0.f, 0.f), e = mk3(0.f, 0.f, 0.f);
    if (h.tri != kInvalid) {
        Surface it = surface_at(sc, h.tri, h.b1, h.b2, true);
        it.wo = normalize(-d);
        g0 = make_float4(it.frame.z.x, it.frame.z.y, it.frame.z.z, 0.f);
        g1.w = linear_depth(rv, it.pos);
        const InstanceRecord inst = sc.inst[it.inst];
        if (inst.mat != kInvalid) {
            bool dispersive;
            MatCtx<NS> mc = make_material<NS, RGB, MS>(sc, sc.mat[inst.mat], it.frame, it.ng, it.wo, swl, dispersive,
                                                       it.uv);
            a = to_linear_srgb<NS, RGB>(sc, mat_albedo<NS, RGB, MS>(sc, mc, swl), swl);
        }
        if (inst.light != kInvalid) {
            Spec<NS> Le;
            float lpdf;
            eval_hit_light<NS, RGB>(sc, it, inst.light, h.tri, o, swl, Le, lpdf);
            e = to_linear_srgb<NS, RGB>(sc, Le, swl);
        }
    }
    const uint32_t cur = (sv.frame + 1u) & 1u;
    geo_field(sv, cur, 0)[pix] = g0;
    geo_field(sv, cur, 1)[pix] = g1;
    sv.alb[pix] = make_float4(a.x, a.y, a.z, 1.f);
    sv.emi[pix] = make_float4(e.x, e.y, e.z, 1.f);
    if (counters) {
        atomicAdd(counters + 0, 1ull);
        atomicAdd(counters + 1, (unsigned long long)nodes);
        atomicAdd(counters + 2, (unsigned long long)tris);
    }
}

// FrameBuffer::compute_gradient (frame_buffer.cpp:124-168): writes only the
// fwidth and depth-gradient words, which no thread of the pass reads
__global__ void __launch_bounds__(256) gradient_kernel(SvgfView sv) {
    const uint2 tile = xcd_tile_2d();
    const int x = int(tile.x * 16u + (threadIdx.x & 15u)), y = int(tile.y * 16u + (threadIdx.x >> 4));
    const int W = int(sv.width), H = int(sv.height);
    if (x >= W || y >= H) return;
    const uint32_t cur = (sv.frame + 1u) & 1u;
    const float4 *g0 = geo_field(sv, cur, 0), *g1 = geo_field(sv, cur, 1);
    const size_t i = size_t(y) * W + x;
    const float4 ca = g0[i];
    const float cd = g1[i].w;
    uint32_t xn = 0, yn = 0;
    f3 ndx = mk3(0.f, 0.f, 0.f), ndy = mk3(0.f, 0.f, 0.f);
    float ddx = 0.f, ddy = 0.f;
    const int x0 = max(0, x - 1), x1 = min(x + 1, W - 1), y0 = max(0, y - 1), y1 = min(y + 1, H - 1);
    for (int nx = x0; nx <= x1; ++nx) {
        for (int ny = y0; ny <= y1; ++ny) {
            const size_t j = size_t(ny) * W + nx;
            const float4 qa = g0[j];
            const float qd = g1[j].w;
            if (x > nx) {
                ++xn;
                ndx = mk3(ndx.x + (ca.x - qa.x), ndx.y + (ca.y - qa.y), ndx.z + (ca.z - qa.z));
                ddx += cd - qd;
            } else if (nx > x) {
                ++xn;
                ndx = mk3(ndx.x + (qa.x - ca.x), ndx.y + (qa.y - ca.y), ndx.z + (qa.z - ca.z));
                ddx += qd - cd;
            }
            if (y > ny) {
                ++yn;
                ndy = mk3(ndy.x + (ca.x - qa.x), ndy.y + (ca.y - qa.y), ndy.z + (ca.z - qa.z));
                ddy += cd - qd;
            } else if (ny > y) {
                ++yn;
                ndy = mk3(ndy.x + (qa.x - ca.x), ndy.y + (qa.y - ca.y), ndy.z + (qa.z - ca.z));
                ddy += qd - cd;
            }
        }
    }
    const float fxn = float(xn), fyn = float(yn);
    const float fx = fabsf(ndx.x / fxn) + fabsf(ndy.x / fyn);
    const float fy = fabsf(ndx.y / fxn) + fabsf(ndy.y / fyn);
    const float fz = fabsf(ndx.z / fxn) + fabsf(ndy.z / fyn);
    reinterpret_cast<float *>(geo_field(sv, cur, 0) + i)[3] = sqrtf(fx * fx + fy * fy + fz * fz);
    reinterpret_cast<float *>(geo_field(sv, cur, 1) + i)[2] = fabsf(ddx / fxn) + fabsf(ddy / fyn);
}

// ---- SVGF passes

// Modulator's demodulate (modulator.cpp:29-37) fused with the reproject
// kernel (reproject.cpp:29-162); writes buffer `out`
__global__ void __launch_bounds__(256) reproject_kernel(SvgfView sv, SvgfParams p, const float4 *__restrict__ motion,
                                                        const float4 *__restrict__ rad, uint32_t out) {
    const uint2 tile = xcd_tile_2d();
    const int x = int(tile.x * 16u + (threadIdx.x & 15u)), y = int(tile.y * 16u + (threadIdx.x >> 4));
    const int W = int(sv.width), H = int(sv.height);
    if (x >= W || y >= H) return;
    const size_t i = size_t(y) * W + x;
    const uint32_t cur = (sv.frame + 1u) & 1u, prev = sv.frame & 1u;
    const float4 L = rad[i], al = sv.alb[i], em = sv.emi[i];
    Rec r;
    r.il = mk3((L.x - em.x) / fmaxf(al.x, 0.001f), (L.y - em.y) / fmaxf(al.y, 0.001f),
               (L.z - em.z) / fmaxf(al.z, 0.001f));
    r.var = r.hist = r.m1 = r.m2 = 0.f;
    if (p.reproject) {
        const Geo cg = load_geo(sv, cur, uint32_t(i));
        const float4 mv = motion[i];
        // Reproject::load_prev_data
        const float ppx = cg.pf.x - mv.x, ppy = cg.pf.y - mv.y;
        const int qx = int(ppx - 0.5f), qy = int(ppy - 0.5f);
        const bool inside = in_screen(qx, qy, W, H);
        f3 pil = mk3(0.f, 0.f, 0.f);
        float pm1 = 0.f, pm2 = 0.f;
        bool valid = false;
        bool v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int tx = qx + (k & 1), ty = qy + (k >> 1);
            v[k] = false;
            if (in_screen(tx, ty, W, H)) {
                v[k] = valid_reproject(cg, load_geo(sv, prev, uint32_t(ty) * sv.width + uint32_t(tx)), W, H);
                valid = valid || v[k];
            }
        }
        if (valid) {
            float wsum = 0.f;
            const float fx = ppx - floorf(ppx), fy = ppy - floorf(ppy);
            const float w[4] = {(1.f - fx) * (1.f - fy), fx * (1.f - fy), (1.f - fx) * fy, fx * fy};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int tx = qx + (k & 1), ty = qy + (k >> 1);
                if (v[k] && in_screen(tx, ty, W, H)) {
                    const Rec q = load_rec(sv, prev, uint32_t(ty) * sv.width + uint32_t(tx));
                    pil = mk3(pil.x + w[k] * q.il.x, pil.y + w[k] * q.il.y, pil.z + w[k] * q.il.z);
                    pm1 += w[k] * q.m1;
                    pm2 += w[k] * q.m2;
                    wsum += w[k];
                }
            }
            valid = wsum >= 0.01f;
            pil = valid ? mk3(pil.x / wsum, pil.y / wsum, pil.z / wsum) : mk3(0.f, 0.f, 0.f);
            pm1 = valid ? pm1 / wsum : 0.f;
            pm2 = valid ? pm2 / wsum : 0.f;
        }
        if (!valid) {
            float n = 0.f;
            const int x0 = max(0, qx - 1), x1 = min(qx + 1, W - 1), y0 = max(0, qy - 1), y1 = min(qy + 1, H - 1);
            for (int nx = x0; nx <= x1; ++nx) {
                for (int ny = y0; ny <= y1; ++ny) {
                    const uint32_t j = uint32_t(ny) * sv.width + uint32_t(nx);
                    if (valid_reproject(cg, load_geo(sv, prev, j), W, H)) {
                        const Rec q = load_rec(sv, prev, j);
                        pil = mk3(pil.x + q.il.x, pil.y + q.il.y, pil.z + q.il.z);
                        pm1 += q.m1;
                        pm2 += q.m2;
                        n += 1.f;
                    }
                }
            }
            if (n > 0.f) {
                valid = true;
                pil = mk3(pil.x / n, pil.y / n, pil.z / n);
                pm1 /= n;
                pm2 /= n;
            }
        }
        float history;
        if (valid && inside) {
            history = load_rec(sv, prev, uint32_t(qy) * sv.width + uint32_t(qx)).hist;
        } else {
            history = 0.f;
            pil = mk3(0.f, 0.f, 0.f);
            pm1 = pm2 = 0.f;
        }
        history = fminf(float(p.history_limit), valid ? history + 1.0f : 1.0f);
        const float alpha = valid ? fmaxf(p.alpha, 1.f / history) : 1.f;
        const float malpha = valid ? fmaxf(p.moments_alpha, 1.f / history) : 1.f;
        float m1 = lum3(r.il), m2 = m1 * m1;
        m1 = pm1 + malpha * (m1 - pm1);
        m2 = pm2 + malpha * (m2 - pm2);
        r.var = fmaxf(0.f, m2 - m1 * m1);
        r.il = mk3(pil.x + alpha * (r.il.x - pil.x), pil.y + alpha * (r.il.y - pil.y), pil.z + alpha * (r.il.z - pil.z));
        r.m1 = m1;
        r.m2 = m2;
        r.hist = history;
    }
    store_rec(sv, out, uint32_t(i), r);
}

// FilterMoments (filter_moments.cpp:13-60): buffer `in` -> `out`
__global__ void __launch_bounds__(256) moments_kernel(SvgfView sv, SvgfParams p, uint32_t in, uint32_t out) {
    const uint2 tile = xcd_tile_2d();
    const int x = int(tile.x * 16u + (threadIdx.x & 15u)), y = int(tile.y * 16u + (threadIdx.x >> 4));
    const int W = int(sv.width), H = int(sv.height);
    if (x >= W || y >= H) return;
    const uint32_t i = uint32_t(y) * sv.width + uint32_t(x);
    const uint32_t cur = (sv.frame + 1u) & 1u;
    Rec r = load_rec(sv, in, i);
    const Geo cg = load_geo(sv, cur, i);
    if (r.hist >= 4.f && cg.depth >= 0.f) {
        const float sigma_depth = fmaxf(cg.dg, 1e-8f) * 3.f;
        const float cl = lum3(r.il);
        float wsum = 0.f, m1 = 0.f, m2 = 0.f;
        f3 s = mk3(0.f, 0.f, 0.f);
        const int rad = p.moments_radius;
        const int x0 = max(0, x - rad), x1 = min(x + rad, W - 1), y0 = max(0, y - rad), y1 = min(y + rad, H - 1);
        for (int nx = x0; nx <= x1; ++nx) {
            for (int ny = y0; ny <= y1; ++ny) {
                const uint32_t j = uint32_t(ny) * sv.width + uint32_t(nx);
                const Rec q = load_rec(sv, in, j);
                const float4 ga = geo_field(sv, cur, 0)[j];
                const float gd = geo_field(sv, cur, 1)[j].w;
                const float w = cal_weight(cg.depth, gd, sigma_depth, cg.n, mk3(ga.x, ga.y, ga.z), p.sigma_normal, cl,
                                           lum3(q.il), p.sigma_rt);
                wsum += w;
                s = mk3(s.x + q.il.x * w, s.y + q.il.y * w, s.z + q.il.z * w);
                m1 += q.m1 * w;
                m2 += q.m2 * w;
            }
        }
        wsum = fmaxf(wsum, 1e-6f);
        r.il = mk3(s.x / wsum, s.y / wsum, s.z / wsum);
        m1 /= wsum;
        m2 /= wsum;
        float var = m2 - m1 * m1;
        var *= 4.f / r.hist;
        r.var = var;
    }
    store_rec(sv, out, i, r);
}

// one a-trous pass (atrous.cpp:13-93): buffer `in` -> `out`; the last pass
// also runs the Modulator's modulate kernel (modulator.cpp:19-27) into `pic`
__global__ void __launch_bounds__(256) atrous_kernel(SvgfView sv, SvgfParams p, uint32_t in, uint32_t out, int step,
                                                     float4 *__restrict__ pic) {
    const uint2 tile = xcd_tile_2d();
    const int x = int(tile.x * 16u + (threadIdx.x & 15u)), y = int(tile.y * 16u + (threadIdx.x >> 4));
    const int W = int(sv.width), H = int(sv.height);
    if (x >= W || y >= H) return;
    const uint32_t i = uint32_t(y) * sv.width + uint32_t(x);
    const uint32_t cur = (sv.frame + 1u) & 1u;
    Rec r = load_rec(sv, in, i);
    const float cl = lum3(r.il);
    float variance = 0.f;
    {
        const int x0 = max(0, x - 1), x1 = min(x + 1, W - 1), y0 = max(0, y - 1), y1 = min(y + 1, H - 1);
        for (int nx = x0; nx <= x1; ++nx) {
            for (int ny = y0; ny <= y1; ++ny) {
                const int dx = abs(nx - x), dy = abs(ny - y);
                const float t = dx == 0 ? (dy == 0 ? 1.f / 4 : 1.f / 8) : (dy == 0 ? 1.f / 8 : 1.f / 16);
                variance += rec_field(sv, in, 0)[uint32_t(ny) * sv.width + uint32_t(nx)].w * t;
            }
        }
    }
    const Geo cg = load_geo(sv, cur, i);
    if (cg.depth >= 0.f) {
        const float sigma_depth = fmaxf(cg.dg, 1e-8f) * float(step);
        const float sigma_il = p.sigma_rt * sqrtf(fmaxf(0.f, 1e-10f + variance));
        float wsum = 1.f;
        float s0 = r.il.x, s1 = r.il.y, s2 = r.il.z, s3 = r.var;
        const int x0 = max(0, x - 2), x1 = min(x + 2, W - 1), y0 = max(0, y - 2), y1 = min(y + 2, H - 1);
        for (int nx = x0; nx <= x1; ++nx) {
            for (int ny = y0; ny <= y1; ++ny) {
                const int ox = x - nx, oy = y - ny;
                if (ox == 0 && oy == 0) continue;
                const int ax = abs(ox), ay = abs(oy);
                const float kx = ax == 0 ? 1.f : (ax == 1 ? 2.f / 3 : 1.f / 6);
                const float ky = ay == 0 ? 1.f : (ay == 1 ? 2.f / 3 : 1.f / 6);
                const float kern = kx * ky;
                const int tx = x + ox * step, ty = y + oy * step;
                if (!in_screen(tx, ty, W, H)) continue;
                const uint32_t j = uint32_t(ty) * sv.width + uint32_t(tx);
                const float4 qa = rec_field(sv, in, 0)[j];
                const float4 ga = geo_field(sv, cur, 0)[j];
                const float gd = geo_field(sv, cur, 1)[j].w;
                const float dist = sqrtf(float(ox) * float(ox) + float(oy) * float(oy));
                const float w = cal_weight(cg.depth, gd, sigma_depth * dist, cg.n, mk3(ga.x, ga.y, ga.z),
                                           p.sigma_normal, cl, lum3(mk3(qa.x, qa.y, qa.z)), sigma_il);
                const float iw = w * kern;
                wsum += iw;
                s0 += iw * qa.x;
                s1 += iw * qa.y;
                s2 += iw * qa.z;
                s3 += (iw * iw) * qa.w;
            }
        }
        r.il = mk3(s0 / wsum, s1 / wsum, s2 / wsum);
        r.var = s3 / (wsum * wsum);
    }
    store_rec(sv, out, i, r);
    if (pic) {
        const float4 al = sv.alb[i], em = sv.emi[i];
        pic[i] = make_float4(r.il.x * al.x + em.x, r.il.y * al.y + em.y, r.il.z * al.z + em.z, 1.f);
    }
}

__global__ void __launch_bounds__(256) modulate_kernel(SvgfView sv, uint32_t in, float4 *__restrict__ pic) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sv.pixels) return;
    const float4 il = rec_field(sv, in, 0)[i], al = sv.alb[i], em = sv.emi[i];
    pic[i] = make_float4(il.x * al.x + em.x, il.y * al.y + em.y, il.z * al.z + em.z, 1.f);
}

// zero records, G-buffer slots at PixelGeometry's defaults (depth -1)
__global__ void __launch_bounds__(256) svgf_clear_kernel(SvgfView sv) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sv.pixels) return;
    for (uint32_t b = 0; b < 3u; ++b) {
        rec_field(sv, b, 0)[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        rec_field(sv, b, 1)[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    for (uint32_t s = 0; s < 2u; ++s) {
        geo_field(sv, s, 0)[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        geo_field(sv, s, 1)[i] = make_float4(0.f, 0.f, 0.f, -1.f);
    }
    sv.alb[i] = make_float4(0.f, 0.f, 0.f, 1.f);
    sv.emi[i] = make_float4(0.f, 0.f, 0.f, 1.f);
}

template <int NS, bool RGB, int MS>
hipError_t launch_gbuffer(const SceneView &sc, const CameraView &cam, const RestirView &rv, const SvgfView &sv,
                          unsigned long long *counters, hipStream_t stream) {
    const size_t lds = size_t(sc.stack_size) * 64 * sizeof(uint2);
    hipLaunchKernelGGL((gbuffer_kernel<NS, RGB, MS>), dim3((sv.width + 7u) / 8u, (sv.height + 7u) / 8u), dim3(64), lds,
                       stream, sc, cam, rv, sv, counters);
    return hipGetLastError();
}

dim3 tiles16(const SvgfView &sv) { return dim3((sv.width + 15u) / 16u, (sv.height + 15u) / 16u); }

}// namespace

hipError_t gbuffer_compute(const SceneView &sc, const CameraView &cam, const RestirView &rv, const SvgfView &sv,
                           uint32_t spectrum_dim, bool rgb, int material_set, unsigned long long *counters,
                           hipStream_t stream) {
    if (sv.pixels == 0) return hipSuccess;
    if (material_set != MS_ALL && material_set != MS_DIFFUSE) return hipErrorInvalidValue;
    const bool all = material_set == MS_ALL;
    hipError_t e;
    if (rgb) e = all ? launch_gbuffer<3, true, MS_ALL>(sc, cam, rv, sv, counters, stream)
                     : launch_gbuffer<3, true, MS_DIFFUSE>(sc, cam, rv, sv, counters, stream);
    else if (spectrum_dim == 4) e = all ? launch_gbuffer<4, false, MS_ALL>(sc, cam, rv, sv, counters, stream)
                                        : launch_gbuffer<4, false, MS_DIFFUSE>(sc, cam, rv, sv, counters, stream);
    else if (spectrum_dim == 3) e = all ? launch_gbuffer<3, false, MS_ALL>(sc, cam, rv, sv, counters, stream)
                                        : launch_gbuffer<3, false, MS_DIFFUSE>(sc, cam, rv, sv, counters, stream);
    else return hipErrorInvalidValue;
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(gradient_kernel, tiles16(sv), dim3(256), 0, stream, sv);
    return hipGetLastError();
}

hipError_t svgf_denoise(const SvgfView &sv, const SvgfParams &p, const float4 *motion, const float4 *rad, float4 *out,
                        hipStream_t stream) {
    if (sv.pixels == 0) return hipSuccess;
    if (p.N > 16 || p.moments_radius < 0 || p.moments_radius > 16) return hipErrorInvalidValue;
    // ping-pong between the current buffer and the spare one so that the
    // frame's last pass lands in the current buffer
    const uint32_t cur = (sv.frame + 1u) & 1u, spare = 2u;
    const uint32_t passes = (p.filter_moments ? 1u : 0u) + p.N;
    uint32_t b = (passes & 1u) ? spare : cur;
    hipLaunchKernelGGL(reproject_kernel, tiles16(sv), dim3(256), 0, stream, sv, p, motion, rad, b);
    if (p.filter_moments) {
        const uint32_t nb = b == cur ? spare : cur;
        hipLaunchKernelGGL(moments_kernel, tiles16(sv), dim3(256), 0, stream, sv, p, b, nb);
        b = nb;
    }
    for (uint32_t k = 0; k < p.N; ++k) {
        const uint32_t nb = b == cur ? spare : cur;
        hipLaunchKernelGGL(atrous_kernel, tiles16(sv), dim3(256), 0, stream, sv, p, b, nb, int(1u << k),
                           k + 1 == p.N ? out : nullptr);
        b = nb;
    }
    if (p.N == 0)
        hipLaunchKernelGGL(modulate_kernel, dim3((sv.pixels + 255u) / 256u), dim3(256), 0, stream, sv, b, out);
    return hipGetLastError();
}

hipError_t svgf_clear(const SvgfView &sv, hipStream_t stream) {
    if (sv.pixels == 0) return hipSuccess;
    hipLaunchKernelGGL(svgf_clear_kernel, dim3((sv.pixels + 255u) / 256u), dim3(256), 0, stream, sv);
    return hipGetLastError();
}

}// namespace vs
