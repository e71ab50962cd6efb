// ReSTIR GI: per-pixel reservoirs of indirect-light samples (a surface point
// seen from the primary surface and the radiance leaving it towards that
// surface), filled from one path per pixel, from the previous frame and from
// screen-space neighbours, then shaded and added to ReSTIR DI's direct light.
//
// Restates the indirect half of the reference's real-time integrator
// (src/render_core/integrators/ReSTIR/indirect.cpp): init_sample and the
// initial-samples kernel (:64-111), eval_bsdf / compute_p_hat (:113-130),
// combine_temporal and temporal_reuse (:132-236), constant_combine /
// spatial_reuse / shading (:238-333), Jacobian_det (:14-28); the reservoir
// is GIReservoir with GISample (ReSTIR/reservoir.h:120-222), and the combine
// is RealTimeIntegrator's (rt.cpp:119-126).  sRGB mode only: the reference
// divides a linear-sRGB path radiance by the first three channels of the
// BSDF value, which only means something when those are RGB.
//
// The reference's initial-sample and temporal kernels are fused here (the
// temporal pass reads only its own pixel's new sample), so the sample
// buffer never goes through HBM; the combine is fused into the spatial /
// shading kernel.  Sampler dimensions are the reference's: 3 for the path,
// 4 for temporal reuse (after a temporary camera sample at 0), 5 for spatial
// reuse.
#pragma once

#include "dev_restir.h"

namespace vs {
namespace dev {

struct GiSample {
    f3 pos, n, Lo;
    float age;
};

struct GiReservoir {
    float weight_sum, C, W;
    GiSample s;
    // GIReservoir::update / truncation / cal_W (reservoir.h:196-219)
    __device__ __forceinline__ void update(float u, const GiSample &v, float weight, float new_C) {
        weight_sum += weight;
        C += new_C;
        if (u * weight_sum < weight || weight_sum == 0.f) s = v;
    }
    __device__ __forceinline__ void truncation(float limit) {
        float factor = limit / C;
        C = C < limit ? C : limit;
        weight_sum = factor < 1.f ? weight_sum * factor : weight_sum;
    }
    __device__ __forceinline__ void update_W(float p_hat) { W = p_hat == 0.f ? 0.f : weight_sum / (p_hat * C); }
};

__device__ __forceinline__ GiSample empty_gi_sample() {
    GiSample s;
    s.pos = mk3(0.f, 0.f, 0.f);
    s.n = mk3(0.f, 0.f, 0.f);
    s.Lo = mk3(0.f, 0.f, 0.f);
    s.age = 0.f;
    return s;
}
__device__ __forceinline__ GiReservoir empty_gi_reservoir() {
    GiReservoir r;
    r.weight_sum = 0.f;
    r.C = 0.f;
    r.W = 0.f;
    r.s = empty_gi_sample();
    return r;
}

// ---- HBM records: 3 buffers x 4 fields x pixels
__device__ __forceinline__ GiReservoir load_gi_reservoir(const RestirGiView &g, uint32_t pixels, uint32_t buf,
                                                         uint32_t i) {
    const float4 *b = g.rsv + size_t(buf) * 4u * pixels;
    float4 a = b[i], p = b[size_t(pixels) + i], n = b[2u * size_t(pixels) + i], l = b[3u * size_t(pixels) + i];
    GiReservoir r;
    r.weight_sum = a.x;
    r.C = a.y;
    r.W = a.z;
    r.s.age = a.w;
    r.s.pos = mk3(p.x, p.y, p.z);
    r.s.n = mk3(n.x, n.y, n.z);
    r.s.Lo = mk3(l.x, l.y, l.z);
    return r;
}
__device__ __forceinline__ void store_gi_reservoir(const RestirGiView &g, uint32_t pixels, uint32_t buf, uint32_t i,
                                                   const GiReservoir &r) {
    float4 *b = g.rsv + size_t(buf) * 4u * pixels;
    b[i] = make_float4(r.weight_sum, r.C, r.W, r.s.age);
    b[size_t(pixels) + i] = make_float4(r.s.pos.x, r.s.pos.y, r.s.pos.z, 0.f);
    b[2u * size_t(pixels) + i] = make_float4(r.s.n.x, r.s.n.y, r.s.n.z, 0.f);
    b[3u * size_t(pixels) + i] = make_float4(r.s.Lo.x, r.s.Lo.y, r.s.Lo.z, 0.f);
}

// the primary surface of a pixel as the reference rebuilds it
// (Pipeline::compute_surface_interaction(hit, view_pos))
__device__ __forceinline__ Surface surface_seen_from(const SceneView &sc, const SurfaceRec &cs, f3 view) {
    Surface it = surface_at(sc, cs.tri, cs.b1, cs.b2, true);
    it.wo = normalize(view - it.pos);
    return it;
}

// ReSTIRGI::eval_bsdf (indirect.cpp:113-124): f |cos| of the surface's
// material towards the sample point, first three channels; a surface
// without material gives 0.  `mc` is the surface's material context.
template <int MS>
__device__ __forceinline__ f3 gi_bsdf(const MatCtx<3> &mc, bool has_material, f3 pos, const GiSample &g) {
    if (!has_material) return mk3(0.f, 0.f, 0.f);
    EvalN<3> e = mat_eval<3, MS>(mc, normalize(g.pos - pos), nullptr);
    return mk3(e.f.v[0], e.f.v[1], e.f.v[2]);
}
// GISample::p_hat (reservoir.h:166-168): luminance of Lo f
template <int MS>
__device__ __forceinline__ float gi_p_hat(const MatCtx<3> &mc, bool has_material, f3 pos, const GiSample &g) {
    f3 f = gi_bsdf<MS>(mc, has_material, pos, g);
    return luminance(mk3(g.Lo.x * f.x, g.Lo.y * f.y, g.Lo.z * f.z));
}

// ReSTIRGI::Jacobian_det (indirect.cpp:14-28): the solid-angle change of a
// sample point re-used from a neighbour's surface, clamped to [0.6, 1/0.6]
__device__ __forceinline__ float gi_jacobian(f3 cur, f3 neighbor, const GiSample &g) {
    f3 cv = cur - g.pos, nv = neighbor - g.pos;
    float cos_c = fabsf(dot(normalize(cv), g.n)), cos_n = fabsf(dot(normalize(nv), g.n));
    float r = (cos_c * dot(nv, nv)) / (cos_n * dot(cv, cv));
    r = (isnan(r) || isinf(r)) ? 0.f : r;
    const float lower = 0.6f;
    return fminf(fmaxf(r, lower), 1.f / lower);
}

// the initial sample of a pixel (init_sample, indirect.cpp:64-81, and the
// initial-samples kernel, :83-111): a path from the primary surface along
// ReSTIR DI's BSDF candidate direction with that candidate's throughput
// f / pdf, integrated by IlluminationIntegrator::Li at depth g.depth; the
// sample is the path's first surface (its bounce-0 hit, HitContext::it,
// integrator.cpp:181-183) and its radiance divided by the throughput, times
// the near-specular chain's throughput of a replaced primary surface.
template <int MS>
__device__ __forceinline__ GiSample gi_initial_sample(const SceneView &sc, const RenderView &rv0,
                                                      const RestirView &rv, const RestirGiView &g,
                                                      const SurfaceRec &cs, uint32_t px, uint32_t py, uint32_t pix,
                                                      Swl<3> &swl, uint2 *stack, uint32_t stride, PathCounters &pc) {
    const uint32_t cslot = cur_slot(rv.frame);
    SurfaceExt ext;
    ext.throughput = mk3(1.f, 1.f, 1.f);
    f3 view = mk3(rv.cam_pos[0], rv.cam_pos[1], rv.cam_pos[2]);
    if (cs.replaced) {
        ext = load_ext(rv, cslot, pix);
        view = ext.view_pos;
    }
    Surface it = surface_seen_from(sc, cs, view);
    const float4 h0 = g.aux[pix], h1 = g.aux[size_t(rv.pixels) + pix];
    const f3 wi = mk3(h0.x, h0.y, h0.z);
    const float pdf = h0.w;
    // HitBSDF::safe_throughput: f / pdf, 0 where not finite
    const f3 t = mk3(zero_if_nan_inf(h1.x / pdf), zero_if_nan_inf(h1.y / pdf), zero_if_nan_inf(h1.z / pdf));
    RenderView prv = rv0;
    prv.max_depth = g.depth;
    PathState<3> ps;
    ps.swl = swl;
    ps.sampler.start(px, py, rv.frame, 3u);
    f3 n = dot(it.ng, wi) > 0.f ? it.ng : -it.ng;
    ps.o = offset_ray_origin(it.pos, n);
    ps.d = wi;
    ps.thr.v[0] = t.x;
    ps.thr.v[1] = t.y;
    ps.thr.v[2] = t.z;
    ps.L = mk3(0.f, 0.f, 0.f);
    ps.scatter_pdf = pdf;
    ps.eta_scale = 1.f;
    ps.bounce = 0u;
    ps.skips = 0u;
    ps.tail = 0u;
    ps.fw = 1.f;
    ps.alive = prv.max_depth > 0u ? 1u : 0u;
    GiSample smp = empty_gi_sample();
    while (ps.alive) {
        TraceHit h = traverse<false>(sc, ps.o, ps.d, 0.f, 3.402823466e38f, stack, stride, pc.nodes, pc.tris);
        ++pc.rays;
        // the bounce-0 surface with a material (also the one the extra
        // BSDF ray of max_depth < 2 reaches, which runs as bounce 0)
        if (h.tri != kInvalid && (ps.tail || ps.bounce == 0u) && sc.inst[sc.tri_inst[h.tri]].mat != kInvalid) {
            Surface fh = surface_at(sc, h.tri, h.b1, h.b2, false);
            smp.pos = fh.pos;
            smp.n = fh.ng;
        }
        ShadowRay sr;
        path_shade<3, true, MS>(sc, prv, ps, h, sr);
        if (sr.valid) {
            uint32_t dn = 0, dt = 0;
            TraceHit sh = traverse<true>(sc, sr.o, sr.d, 0.f, 1.f - kShadowEpsilon, stack, stride, dn, dt);
            pc.nodes += dn;
            pc.tris += dt;
            ++pc.rays;
            if (sh.tri == kInvalid) ps.L += sr.c;
        }
    }
    swl = ps.swl;
    smp.Lo = mk3(zero_if_nan_inf(ps.L.x / t.x), zero_if_nan_inf(ps.L.y / t.y), zero_if_nan_inf(ps.L.z / t.z));
    if (cs.replaced) smp.Lo = mk3(smp.Lo.x * ext.throughput.x, smp.Lo.y * ext.throughput.y, smp.Lo.z * ext.throughput.z);
    return smp;
}

// ReSTIRGI::combine_temporal (indirect.cpp:132-151): the current sample and
// the previous frame's, re-weighted at the current surface; the Jacobian
// when the previous sample comes from another pixel than the reprojected
// one
template <int MS>
__device__ __forceinline__ GiReservoir gi_combine_temporal(const SceneView &sc, const GiReservoir &cur,
                                                           const Surface &it, const MatCtx<3> &mc, bool has_mat,
                                                           GiReservoir other, const SurfaceRec *nb, Sampler &sp) {
    other.s.age += 1.f;
    GiReservoir ret = empty_gi_reservoir();
    float cur_p = gi_p_hat<MS>(mc, has_mat, it.pos, cur.s);
    ret.update(sp.next1d(), cur.s, safe_weight(cur.C, cur_p, cur.W), 1.f);
    float other_p = gi_p_hat<MS>(mc, has_mat, it.pos, other.s);
    ret.update(sp.next1d(), other.s, safe_weight(other.C, other_p, other.W), other.C);
    float p = gi_p_hat<MS>(mc, has_mat, it.pos, ret.s);
    if (nb) p = p * gi_jacobian(it.pos, surface_at(sc, nb->tri, nb->b1, nb->b2, false).pos, other.s);
    ret.update_W(p);
    return ret;
}

// the fused initial-sample and temporal-reuse pass for one pixel (the
// initial-samples kernel and compile_temporal_reuse, indirect.cpp:83-111,
// 153-236); the reservoir goes to the pass-through buffer.  Pixels that see
// nothing get zero indirect radiance and keep their reservoirs.
template <int MS>
__device__ __forceinline__ void restir_gi_initial_temporal(const SceneView &sc, const CameraView &cam, const RenderView &rv0,
                                           const RestirView &rv, const RestirGiView &g, uint32_t px, uint32_t py,
                                           uint2 *stack, uint32_t stride, PathCounters &pc) {
    const uint32_t pix = py * rv.width + px;
    const uint32_t W = rv.width, H = rv.height;
    const uint32_t cslot = cur_slot(rv.frame), pslot = prev_slot(rv.frame);
    const SurfaceRec cs = load_surface(rv, cslot, pix);
    if (!cs.hit()) {
        g.rad[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    Swl<3> swl = sample_wavelengths<3, true>(px, py, rv.frame);
    GiSample smp = gi_initial_sample<MS>(sc, rv0, rv, g, cs, px, py, pix, swl, stack, stride, pc);
    // temporal reuse: the film position of the pixel's camera sample
    f2 p_film;
    {
        Sampler tmp;
        tmp.start(px, py, rv.frame, 0u);
        f3 o, d;
        float fw;
        camera_ray(cam, px, py, tmp, o, d, fw, &p_film);
    }
    Sampler sp;
    sp.start(px, py, rv.frame, 4u);
    const float4 h0 = g.aux[pix], h1 = g.aux[size_t(rv.pixels) + pix], mv = g.aux[2u * size_t(rv.pixels) + pix];
    GiReservoir rsv = empty_gi_reservoir();
    {
        float p_hat = luminance(mk3(smp.Lo.x * h1.x, smp.Lo.y * h1.y, smp.Lo.z * h1.z));
        rsv.update(0.5f, smp, safe_weight(1.f, p_hat, 1.f / h0.w), 1.f);
        rsv.update_W(p_hat);
    }
    const float prev_x = p_film.x - mv.x, prev_y = p_film.y - mv.y;
    const float limit = rsv.C * float(g.t_limit);
    const int ix = to_i32(prev_x), iy = to_i32(prev_y);
    if (ix >= 0 && iy >= 0 && ix < int(W) && iy < int(H) && g.temporal) {
        f3 view = mk3(rv.cam_pos[0], rv.cam_pos[1], rv.cam_pos[2]);
        if (cs.replaced) view = load_ext(rv, cslot, pix).view_pos;
        const Surface it = surface_seen_from(sc, cs, view);
        const bool has_mat = sc.inst[it.inst].mat != kInvalid;
        MatCtx<3> mc = material_at<3, true, MS>(sc, it, swl);
        uint32_t j = min(to_u32(prev_y), H - 1u) * W + min(to_u32(prev_x), W - 1u);
        GiReservoir pr = load_gi_reservoir(g, rv.pixels, pslot, j);
        pr.truncation(limit);
        SurfaceRec ps = load_surface(rv, pslot, j);
        if (pr.s.age < float(g.max_age) && valid_neighbor(cs, ps, g.t_dot, g.t_depth, g.diff_factor)) {
            rsv = gi_combine_temporal<MS>(sc, rsv, it, mc, has_mat, pr, nullptr, sp);
        } else {
            for (uint32_t i = 0; i < g.t_N; ++i) {
                f2 dk = square_to_disk(sp.next2d());
                f2 q = mk2(dk.x * g.t_radius + prev_x, dk.y * g.t_radius + prev_y);
                j = min(to_u32(q.y), H - 1u) * W + min(to_u32(q.x), W - 1u);
                GiReservoir ar = load_gi_reservoir(g, rv.pixels, pslot, j);
                ar.truncation(limit);
                SurfaceRec as = load_surface(rv, pslot, j);
                if (ar.s.age < float(g.max_age) && valid_neighbor(cs, as, g.t_dot, g.t_depth, g.diff_factor)) {
                    rsv = gi_combine_temporal<MS>(sc, rsv, it, mc, has_mat, ar, &as, sp);
                    break;
                }
            }
        }
    }
    store_gi_reservoir(g, rv.pixels, kPassThrough, pix, rsv);
}

// a neighbour pixel of a spatial draw: a disk offset of `radius`, rounded,
// clamped to the image (spatial_reuse, indirect.cpp:278-283); 2 draws
__device__ __forceinline__ uint32_t gi_neighbor(const RestirView &rv, float radius, uint32_t px, uint32_t py,
                                                Sampler &sp) {
    f2 dk = square_to_disk(sp.next2d());
    f2 off = mk2(dk.x * radius, dk.y * radius);
    int ax = min(max(int(px) + int(roundf(off.x)), 0), int(rv.width) - 1);
    int ay = min(max(int(py) + int(roundf(off.y)), 0), int(rv.height) - 1);
    return uint32_t(ay) * rv.width + uint32_t(ax);
}

// spatial reuse with visibility and shading for one pixel (the spatial /
// shading kernel, indirect.cpp:311-333, with spatial_reuse and
// constant_combine, :238-294, and shading, :296-309): all neighbour draws
// first (2 each), then one draw per valid neighbour whose sample carries
// light.  Returns the pixel's indirect radiance; the reservoir is kept for
// the next frame.
template <int MS>
__device__ __forceinline__ f3 restir_gi_spatial_shade(const SceneView &sc, const RestirView &rv, const RestirGiView &g,
                                      const SurfaceRec &cs, uint32_t px, uint32_t py, uint2 *stack, uint32_t stride,
                                      PathCounters &pc) {
    const uint32_t pix = py * rv.width + px;
    const uint32_t cslot = cur_slot(rv.frame);
    Swl<3> swl = sample_wavelengths<3, true>(px, py, rv.frame);
    Sampler sp;
    sp.start(px, py, rv.frame, 5u);
    GiReservoir rsv = load_gi_reservoir(g, rv.pixels, kPassThrough, pix);
    f3 view = mk3(rv.cam_pos[0], rv.cam_pos[1], rv.cam_pos[2]);
    if (cs.replaced) view = load_ext(rv, cslot, pix).view_pos;
    Surface it = surface_seen_from(sc, cs, view);
    const bool has_mat = sc.inst[it.inst].mat != kInvalid;
    MatCtx<3> mc = material_at<3, true, MS>(sc, it, swl);
    if (g.spatial) {
        // the draws of the neighbour positions are replayed from the saved
        // state while the combination draws continue from sp
        Sampler replay = sp;
        for (uint32_t i = 0; i < g.s_num; ++i) gi_neighbor(rv, g.s_radius, px, py, sp);
        GiReservoir ret = rsv;
        for (uint32_t i = 0; i < g.s_num; ++i) {
            const uint32_t j = gi_neighbor(rv, g.s_radius, px, py, replay);
            const SurfaceRec ns = load_surface(rv, cslot, j);
            if (!valid_neighbor(cs, ns, g.s_dot, g.s_depth, g.diff_factor)) continue;
            GiReservoir nr = load_gi_reservoir(g, rv.pixels, kPassThrough, j);
            if (!(luminance(nr.s.Lo) > 0.f)) continue;
            f3 npos = surface_at(sc, ns.tri, ns.b1, ns.b2, false).pos;
            float p = gi_p_hat<MS>(mc, has_mat, it.pos, nr.s) * gi_jacobian(it.pos, npos, nr.s);
            float v = occluded(sc, it, nr.s.pos, stack, stride, pc) ? 0.f : 1.f;
            float w = safe_weight(nr.C, p, nr.W);
            ret.update(sp.next1d(), nr.s, w * v, nr.C * v);
        }
        ret.update_W(gi_p_hat<MS>(mc, has_mat, it.pos, ret.s));
        rsv = ret;
    }
    f3 f = gi_bsdf<MS>(mc, has_mat, it.pos, rsv.s);
    store_gi_reservoir(g, rv.pixels, cur_slot(rv.frame), pix, rsv);
    return mk3(rsv.s.Lo.x * f.x * rsv.W, rsv.s.Lo.y * f.y * rsv.W, rsv.s.Lo.z * f.z * rsv.W);
}

}// namespace dev
}// namespace vs
