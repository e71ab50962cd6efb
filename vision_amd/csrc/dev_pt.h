// Megakernel path tracer (sRGB mode): one thread per pixel, all samples of a
// launch processed in-thread with the film's running mean kept in registers.
//
// Restates PathTracingIntegrator::compile (src/render_core/integrators/pt.cpp:50-75)
// and IlluminationIntegrator::Li (src/base/integrator.cpp:134-283), with the
// sampler-dimension order of the reference (see Sampler draws noted inline).
#pragma once

#include "dev_common.h"
#include "dev_traverse.h"

namespace vs {
namespace dev {

struct Spec3 {
    float v[3];
};
__device__ __forceinline__ Spec3 spec(float s) { return {{s, s, s}}; }
__device__ __forceinline__ Spec3 spec(f3 c) { return {{c.x, c.y, c.z}}; }
__device__ __forceinline__ Spec3 operator*(Spec3 a, Spec3 b) { return {{a.v[0] * b.v[0], a.v[1] * b.v[1], a.v[2] * b.v[2]}}; }
__device__ __forceinline__ Spec3 operator*(Spec3 a, float s) { return {{a.v[0] * s, a.v[1] * s, a.v[2] * s}}; }
__device__ __forceinline__ Spec3 operator/(Spec3 a, float s) { return {{a.v[0] / s, a.v[1] / s, a.v[2] / s}}; }
__device__ __forceinline__ Spec3 operator+(Spec3 a, Spec3 b) { return {{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]}}; }
__device__ __forceinline__ float smax(Spec3 a) { return fmaxf(a.v[0], fmaxf(a.v[1], a.v[2])); }

// ---- surface interaction (Geometry::compute_surface_interaction,
//      src/base/mgr/geometry.cpp:78-165); barycentrics weight (v1, v2)
struct Surface {
    f3 pos, ng, wo;
    Frame frame;
    f2 uv;
    float prim_area;
    uint32_t inst;
};

__device__ __forceinline__ Surface surface_at(const SceneView &sc, uint32_t tri, float b1, float b2, bool complete) {
    Surface s;
    u3 t = sc.tri[tri];
    f3 p0 = sc.P[t.x], p1 = sc.P[t.y], p2 = sc.P[t.z];
    float b0 = 1.f - b1 - b2;
    s.pos = p0 * b0 + p1 * b1 + p2 * b2;
    f3 dp02 = p0 - p2, dp12 = p1 - p2;
    f3 ng_un = cross(dp02, dp12);
    float len = length(ng_un);
    s.prim_area = 0.5f * len;
    s.ng = ng_un * (1.f / len);
    f2 uv0 = sc.UV[t.x], uv1 = sc.UV[t.y], uv2 = sc.UV[t.z];
    s.uv = mk2(uv0.x * b0 + uv1.x * b1 + uv2.x * b2, uv0.y * b0 + uv1.y * b1 + uv2.y * b2);
    s.inst = sc.tri_inst[tri];
    if (complete) {
        f2 duv02 = uv0 - uv2, duv12 = uv1 - uv2;
        float det = duv02.x * duv12.y - duv02.y * duv12.x;
        f3 dp_dv;
        if (fabsf(det) < 1e-8f) {
            dp_dv = normalize(p2 - p0);
        } else {
            float inv_det = 1.f / det;
            dp_dv = (dp02 * (-duv12.x) + dp12 * duv02.x) * inv_det;
        }
        f3 n = sc.N[t.x] * b0 + sc.N[t.y] * b1 + sc.N[t.z] * b2;
        f3 z = (n.x != 0.f || n.y != 0.f || n.z != 0.f) ? normalize(n) : s.ng;
        f3 x = cross(z, dp_dv);
        float lx = length(x);
        if (!(lx > 1e-20f)) {
            f3 a, b;
            coordinate_system(z, a, b);
            x = a;
        } else {
            x = x * (1.f / lx);
        }
        s.frame.x = x;
        s.frame.y = cross(z, x);
        s.frame.z = z;
    }
    return s;
}

// ---- alias table (src/render_core/warpers/alias.h:141-158)
__device__ __forceinline__ uint32_t alias_sample(const SceneView &sc, uint32_t off, uint32_t n, float u, float *u_remapped) {
    u = u * float(n);
    uint32_t idx = min(uint32_t(u), n - 1u);
    u = fminf(u - float(idx), kOneMinusEpsilon);
    AliasEntry e = sc.alias[off + idx];
    bool keep = u < e.prob;
    if (u_remapped) *u_remapped = keep ? fminf(u / e.prob, kOneMinusEpsilon) : fminf((1.f - u) / (1.f - e.prob), kOneMinusEpsilon);
    return keep ? idx : e.alias;
}
__device__ __forceinline__ float alias_pmf(const SceneView &sc, uint32_t off, uint32_t n, float integral, uint32_t i) {
    return integral > 0.f ? sc.alias_func[off + i] / (integral * float(n)) : 0.f;
}

// light-selection PMF (LightSampler::PMF, uniform.cpp:12-20 / power.cpp:17-19)
__device__ __forceinline__ float light_select_pmf(const SceneView &sc, uint32_t li) {
    if (sc.light_sampler == 0) return 1.f / float(sc.num_lights);
    return alias_pmf(sc, sc.select_offset, sc.num_lights, sc.select_integral, li);
}

struct LightSampleD {
    Spec3 L;
    float pdf;
    f3 p_light;
};

// LightSampler::sample_wi (src/base/illumination/lightsampler.cpp:199-216)
// draws: select 1, surface 2
__device__ __forceinline__ LightSampleD sample_light(const SceneView &sc, const Surface &it, Sampler &sampler) {
    LightSampleD ls;
    float u_light = sampler.next1d();
    f2 u = sampler.next2d();
    uint32_t li;
    float sel_pmf;
    if (sc.light_sampler == 0) {
        li = uint32_t(fminf(u_light * float(sc.num_lights), float(sc.num_lights) - 1.f));
        sel_pmf = 1.f / float(sc.num_lights);
    } else {
        li = alias_sample(sc, sc.select_offset, sc.num_lights, u_light, nullptr);
        sel_pmf = alias_pmf(sc, sc.select_offset, sc.num_lights, sc.select_integral, li);
    }
    const LightRecord &L = sc.lights[li];
    Spec3 color = spec(mk3(L.color[0], L.color[1], L.color[2])) * L.scale;
    if (L.type == LIGHT_AREA) {
        // AreaLight::sample_wi (src/render_core/lights/area.cpp:120-149)
        float ux;
        uint32_t prim = alias_sample(sc, L.alias_offset, L.tri_count, u.x, &ux);
        float pmf = alias_pmf(sc, L.alias_offset, L.tri_count, L.alias_integral, prim);
        f2 bary = square_to_triangle(mk2(ux, u.y));
        Surface p = surface_at(sc, L.tri_base + prim, bary.x, bary.y, false);
        float pdf_pos = (1.f / p.prim_area) * pmf;
        f3 w = it.pos - p.pos;
        bool front = dot(w, p.ng) > 0.f || L.two_sided;
        ls.L = front ? color : spec(0.f);
        float cos_t = fabsf(dot(p.ng, normalize(w)));
        float pdf = pdf_pos * length_squared(w) / cos_t;
        ls.pdf = (isinf(pdf) || isnan(pdf)) ? 0.f : pdf;
        float factor = dot(p.ng, w) > 0.f ? 1.f : -1.f;
        ls.p_light = offset_ray_origin(p.pos, p.ng * factor * sc.ray_offset_factor);
    } else {
        // IPointLight::sample_wi (src/base/illumination/light.cpp:42-51);
        // point.cpp:39-43, spot.cpp:51-70
        f3 pos = mk3(L.position[0], L.position[1], L.position[2]);
        f3 w_un = it.pos - pos;
        Spec3 val = color / length_squared(w_un);
        if (L.type == LIGHT_SPOT) {
            f3 dir = mk3(L.direction[0], L.direction[1], L.direction[2]);
            float cos_angle = L.cos_angle;
            float cos_start = L.cos_falloff_start;
            float ct = fminf(fmaxf(dot(dir, normalize(w_un)), cos_angle), cos_start);
            float f = (ct - cos_angle) / (cos_start - cos_angle);
            float f2v = f * f;
            val = val * (f2v * f2v);
        }
        ls.L = val;
        ls.pdf = -1.f;
        ls.p_light = pos;
    }
    ls.pdf *= sel_pmf;
    return ls;
}

// LightSampler::evaluate_hit_wi for an area emitter hit by a BSDF ray
// (src/base/illumination/lightsampler.cpp:252-267)
__device__ __forceinline__ void eval_hit_light(const SceneView &sc, const Surface &it, uint32_t li, uint32_t tri,
                                               f3 ref_pos, Spec3 &L, float &pdf) {
    const LightRecord &Lr = sc.lights[li];
    uint32_t prim = tri - Lr.tri_base;
    float pdf_pos = (1.f / it.prim_area) * alias_pmf(sc, Lr.alias_offset, Lr.tri_count, Lr.alias_integral, prim);
    f3 w = ref_pos - it.pos;
    bool front = dot(w, it.ng) > 0.f || Lr.two_sided;
    L = front ? spec(mk3(Lr.color[0], Lr.color[1], Lr.color[2])) * Lr.scale : spec(0.f);
    float cos_t = fabsf(dot(it.ng, normalize(w)));
    float p = pdf_pos * length_squared(w) / cos_t;
    p = (isinf(p) || isnan(p)) ? 0.f : p;
    pdf = p * light_select_pmf(sc, li);
}

// ---- materials
struct BsdfEval {
    Spec3 f;// BSDF * |cos wi|
    float pdf;
};

__device__ __forceinline__ bool same_hemisphere(f3 a, f3 b, f3 n) { return dot(a, n) * dot(b, n) > 0.f; }

// MaterialEvaluator::evaluate with individual shading normals
// (src/base/scattering/material.cpp:132-148): lobe eval in the shading frame,
// f *= |cos|, pdf zeroed when wo/wi straddle the geometric surface
__device__ __forceinline__ BsdfEval eval_material(const MaterialRecord &m, const Surface &it, f3 wi_w) {
    BsdfEval e;
    f3 wo = it.frame.to_local(it.wo), wi = it.frame.to_local(wi_w);
    if (m.type == MAT_BLACK) {
        e.f = spec(0.f);
        e.pdf = 1.f;
    } else {
        // Lambert (bxdf.h:82-96, bxdf.cpp:13-52); Oren-Nayar (bxdf.cpp:94-121),
        // whose sigma2 is sigma^4 as written there
        bool s = wo.z * wi.z > 0.f;
        Spec3 kr = spec(mk3(m.p[0], m.p[1], m.p[2]));
        float k = kInvPi;
        if (m.type == MAT_OREN_NAYAR) {
            float sig = m.p[3] * (kPi * 0.5f);
            float s2 = (sig * sig) * (sig * sig);
            float A = 1.f - (s2 / (2.f * (s2 + 0.33f)));
            float B = 0.45f * s2 / (s2 + 0.09f);
            float sti = sqrtf(fmaxf(0.f, 1.f - wi.z * wi.z)), sto = sqrtf(fmaxf(0.f, 1.f - wo.z * wo.z));
            float cpi = sti == 0.f ? 1.f : fminf(fmaxf(wi.x / sti, -1.f), 1.f);
            float spi = sti == 0.f ? 0.f : fminf(fmaxf(wi.y / sti, -1.f), 1.f);
            float cpo = sto == 0.f ? 1.f : fminf(fmaxf(wo.x / sto, -1.f), 1.f);
            float spo = sto == 0.f ? 0.f : fminf(fmaxf(wo.y / sto, -1.f), 1.f);
            float maxc = fmaxf(0.f, cpi * cpo + spi * spo);
            bool cond = fabsf(wi.z) > fabsf(wo.z);
            float sin_a = cond ? sto : sti;
            float tan_b = cond ? sti / fabsf(wi.z) : sto / fabsf(wo.z);
            k = kInvPi * (A + B * maxc * sin_a * tan_b);
        }
        e.f = s ? kr * k : spec(0.f);
        e.pdf = s ? fabsf(wi.z) * kInvPi : 0.f;
        e.f = e.f * fabsf(wi.z);
    }
    bool discard = !same_hemisphere(it.wo, wi_w, it.ng);
    if (discard) e.pdf = 0.f;
    return e;
}

struct BsdfSample {
    BsdfEval eval;
    f3 wi;
    float eta;
};

// MaterialEvaluator::sample (material.cpp:162-184); Lambert draws 2
__device__ __forceinline__ BsdfSample sample_material(const MaterialRecord &m, const Surface &it, Sampler &sampler) {
    BsdfSample bs;
    bs.eta = 1.f;
    if (m.type == MAT_BLACK) {
        // BlackBodyLobe::sample_local: wi = wo, pdf 1, no draws
        bs.wi = it.wo;
        bs.eval.f = spec(0.f);
        bs.eval.pdf = 1.f;
    } else {
        f3 wo = it.frame.to_local(it.wo);
        f3 wi = square_to_cosine_hemisphere(sampler.next2d());
        if (wo.z < 0.f) wi.z = -wi.z;
        bs.wi = it.frame.to_world(wi);
        bs.eval = eval_material(m, it, bs.wi);
        return bs;
    }
    bool discard = !same_hemisphere(it.wo, bs.wi, it.ng);
    if (discard) bs.eval.pdf = 0.f;
    return bs;
}

struct PathCounters {
    uint32_t rays;
    uint32_t nodes;
    uint32_t tris;
};

// one camera path; returns linear sRGB radiance (before the filter weight)
__device__ f3 trace_path(const SceneView &sc, const RenderView &rv, f3 ray_o, f3 ray_d, Sampler &sampler,
                         uint2 *stack, uint32_t stride, PathCounters &pc) {
    Spec3 ret = spec(0.f);
    Spec3 throughput = spec(1.f);
    float scatter_pdf = 1e16f;
    float eta_scale = 1.f;
    f3 prev_ng = ray_d;
    uint32_t skips = 0;
    for (uint32_t bounces = 0; bounces < rv.max_depth; ++bounces) {
        TraceHit h = traverse<false>(sc, ray_o, ray_d, 0.f, 3.402823466e38f, stack, stride, pc.nodes, pc.tris);
        ++pc.rays;
        if (h.tri == kInvalid) break;// no environment light in this mode
        Surface it = surface_at(sc, h.tri, h.b1, h.b2, true);
        it.wo = normalize(-ray_d);
        const InstanceRecord inst = sc.inst[it.inst];
        if (inst.mat == kInvalid) {
            // material-less surface: continue straight through (integrator.cpp:167-173)
            f3 n = dot(it.ng, ray_d) > 0.f ? it.ng : -it.ng;
            ray_o = offset_ray_origin(it.pos, n);
            if (++skips > 64u) break;
            --bounces;
            continue;
        }
        if (inst.light != kInvalid) {
            Spec3 Le;
            float lpdf;
            eval_hit_light(sc, it, inst.light, h.tri, ray_o, Le, lpdf);
            float w = mis_weight(scatter_pdf, lpdf);
            if (rv.mis_mode == 2) w = 1.f;
            else if (rv.mis_mode == 1) w = bounces == 0 ? w : 0.f;
            ret = ret + Le * throughput * w;
        }
        prev_ng = it.ng;
        (void)prev_ng;
        // next-event estimation
        LightSampleD ls = sample_light(sc, it, sampler);
        const MaterialRecord &m = sc.mat[inst.mat];
        f3 wi_l = normalize(ls.p_light - it.pos);
        BsdfEval se = eval_material(m, it, wi_l);
        BsdfSample bs = sample_material(m, it, sampler);
        bool delta = ls.pdf < 0.f;
        float w = delta ? 1.f : mis_weight(ls.pdf, se.pdf);
        if (rv.mis_mode == 1) w = 1.f;
        float lpdf = delta ? -ls.pdf : ls.pdf;
        if (se.pdf > 0.f && lpdf > 0.f) {
            f3 dir = ls.p_light - it.pos;
            f3 n = dot(it.ng, dir) > 0.f ? it.ng : -it.ng;
            uint32_t dn = 0, dt = 0;
            TraceHit sh = traverse<true>(sc, offset_ray_origin(it.pos, n), dir, 0.f, 1.f - kShadowEpsilon, stack,
                                         stride, dn, dt);
            pc.nodes += dn;
            pc.tris += dt;
            ++pc.rays;
            if (sh.tri == kInvalid && rv.mis_mode != 2) {
                Spec3 Ld = ls.L * se.f * w / lpdf;
                ret = ret + throughput * Ld;
            }
        }
        eta_scale *= bs.eta * bs.eta;
        float lum = smax(throughput);
        if (!(bs.eval.pdf > 0.f) || lum == 0.f) break;
        throughput = throughput * (bs.eval.f / bs.eval.pdf);
        if (eta_scale * lum < rv.rr_threshold && bounces >= rv.min_depth) {
            float q = fminf(0.95f, lum);
            float rr = sampler.next1d();
            if (q < rr) break;
            throughput = throughput / q;
        }
        scatter_pdf = bs.eval.pdf;
        f3 n = dot(it.ng, bs.wi) > 0.f ? it.ng : -it.ng;
        ray_o = offset_ray_origin(it.pos, n);
        ray_d = bs.wi;
    }
    return mk3(ret.v[0], ret.v[1], ret.v[2]);
}

// Sensor::generate_ray (src/base/sensor/sensor.cpp:44-56) + filter sample
// (Sampler::sensor_sample, src/base/sampler.h:65-73): draws filter 2, lens 2, time 1
__device__ __forceinline__ void camera_ray(const CameraView &cam, uint32_t px, uint32_t py, Sampler &sampler,
                                           f3 &o, f3 &d, float &weight) {
    f2 uf = sampler.next2d();
    f2 fs;
    if (cam.filter == 0) {
        float r = cam.filter_radius;
        fs = mk2(-r + uf.x * (r + r), -r + uf.y * (r + r));
    } else {
        fs = mk2(sample_tent(uf.x, cam.filter_radius), sample_tent(uf.y, cam.filter_radius));
    }
    weight = 1.f;
    f2 p_lens = sampler.next2d();
    (void)sampler.next1d();// time
    float fx = float(px) + 0.5f + fs.x;
    float fy = float(py) + 0.5f + fs.y;
    float sx = cam.lower_x + fx * cam.raster_sx;
    float sy = cam.upper_y - fy * cam.raster_sy;
    f3 dc = normalize(mk3(sx * cam.tan_half_fov, sy * cam.tan_half_fov, 1.f));
    f3 oc = mk3(0.f, 0.f, 0.f);
    if (cam.thin_lens && cam.lens_radius > 0.f) {
        f2 pl = square_to_disk(p_lens);
        pl = pl * cam.lens_radius;
        float ft = cam.focal_distance / dc.z;
        f3 p_focus = dc * ft;
        oc = mk3(pl.x, pl.y, 0.f);
        dc = normalize(p_focus - oc);
    }
    const float *m = cam.c2w;
    o = mk3(m[0] * oc.x + m[1] * oc.y + m[2] * oc.z + m[3], m[4] * oc.x + m[5] * oc.y + m[6] * oc.z + m[7],
            m[8] * oc.x + m[9] * oc.y + m[10] * oc.z + m[11]);
    d = mk3(m[0] * dc.x + m[1] * dc.y + m[2] * dc.z, m[4] * dc.x + m[5] * dc.y + m[6] * dc.z,
            m[8] * dc.x + m[9] * dc.y + m[10] * dc.z);
}

__device__ __forceinline__ float zero_if_nan_inf(float v) { return (isnan(v) || isinf(v)) ? 0.f : v; }

__device__ __forceinline__ float tone_map(float v, uint32_t tm) {
    if (tm == 1) {
        float r = (v * (2.51f * v + 0.03f)) / (v * (2.43f * v + 0.59f) + 0.14f);
        return fminf(fmaxf(r, 0.f), 1.f);
    }
    if (tm == 2) return v / (v + 1.f);
    return v;
}

}// namespace dev
}// namespace vs
