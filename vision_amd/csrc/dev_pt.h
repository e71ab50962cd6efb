// Megakernel path tracer: one thread per pixel, all samples of a launch
// processed in-thread with the film's running mean kept in registers.
//
// Restates PathTracingIntegrator::compile (src/render_core/integrators/pt.cpp:50-75)
// and IlluminationIntegrator::Li (src/base/integrator.cpp:134-283) with the
// reference's sampler-dimension order (draw counts noted inline).  Template
// parameters: NS spectral channels, RGB = sRGB mode (else hero wavelengths).
#pragma once

#include "dev_common.h"
#include "dev_material.h"
#include "dev_traverse.h"

namespace vs {
namespace dev {

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

// ---- alias tables (src/render_core/warpers/alias.h:141-158)
__device__ __forceinline__ uint32_t alias_pick(const AliasEntry *tab, uint32_t n, float u, float *u_remapped) {
    u = u * float(n);
    uint32_t idx = min(uint32_t(u), n - 1u);
    u = fminf(u - float(idx), kOneMinusEpsilon);
    AliasEntry e = tab[idx];
    bool keep = u < e.prob;
    if (u_remapped)
        *u_remapped = keep ? fminf(u / e.prob, kOneMinusEpsilon) : fminf((1.f - u) / (1.f - e.prob), kOneMinusEpsilon);
    return keep ? idx : e.alias;
}
__device__ __forceinline__ uint32_t alias_sample(const SceneView &sc, uint32_t off, uint32_t n, float u, float *u_remapped) {
    return alias_pick(sc.alias + off, n, u, u_remapped);
}
__device__ __forceinline__ float alias_pmf(const SceneView &sc, uint32_t off, uint32_t n, float integral, uint32_t i) {
    return integral > 0.f ? sc.alias_func[off + i] / (integral * float(n)) : 0.f;
}

// light-selection PMF (LightSampler::PMF, uniform.cpp:12-20 / power.cpp:17-19)
__device__ __forceinline__ float light_select_pmf(const SceneView &sc, uint32_t li) {
    if (sc.light_sampler == 0) return 1.f / float(sc.num_lights);
    return alias_pmf(sc, sc.select_offset, sc.num_lights, sc.select_integral, li);
}

// light selection (LightSampler::select_light, lightsampler.cpp:178-197;
// uniform.cpp:22-31 / power.cpp:21-26)
__device__ __forceinline__ uint32_t select_light(const SceneView &sc, float u) {
    if (sc.light_sampler == 0) return uint32_t(fminf(u * float(sc.num_lights), float(sc.num_lights) - 1.f));
    return alias_sample(sc, sc.select_offset, sc.num_lights, u, nullptr);
}

// ---- spectra of lights
template <int NS, bool RGB>
__device__ __forceinline__ Spec<NS> illum_spectrum(const SceneView &sc, const float *rgb, const float *coef,
                                                   float scale, const Swl<NS> &swl) {
    Spec<NS> r;
    if (RGB) {
#pragma unroll
        for (int i = 0; i < NS; ++i) r.v[i] = rgb[i < 3 ? i : 2] * scale;
    } else {
        // RGBIlluminationSpectrum: sigmoid(lambda) * unbound scale * D65(lambda)
#pragma unroll
        for (int i = 0; i < NS; ++i)
            r.v[i] = sigmoid_poly(coef, swl.lambda[i]) * coef[3] * spd_eval(sc.spd + 282, sc.cie_interval, swl.lambda[i]) *
                     scale;
    }
    return r;
}

// a light's radiance spectrum; an area light with an emission image takes
// the nearest texel at the light surface's uv (wrapped, v = 0 the first row)
template <int NS, bool RGB>
__device__ __forceinline__ Spec<NS> light_color(const SceneView &sc, const LightRecord &L, const Swl<NS> &swl,
                                                f2 uv) {
    if (L.tex != kInvalid) {
        const uint4 info = sc.tex_info[L.tex];
        const float u = uv.x - floorf(uv.x), v = uv.y - floorf(uv.y);
        const uint32_t x = min(uint32_t(u * float(info.y)), info.y - 1u);
        const uint32_t y = min(uint32_t(v * float(info.z)), info.z - 1u);
        const float4 t = sc.tex[size_t(info.x) + size_t(y) * info.y + x];
        const float q[4] = {t.x, t.y, t.z, t.w};
        return illum_spectrum<NS, RGB>(sc, q, q, L.scale, swl);
    }
    return illum_spectrum<NS, RGB>(sc, L.color, L.illum, L.scale, swl);
}

// ---- environment map (SphericalMap, src/render_core/lights/environments/spherical.cpp)
__device__ __forceinline__ f3 mat3_apply(const float *m, f3 v) {
    return mk3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
               m[6] * v.x + m[7] * v.y + m[8] * v.z);
}
__device__ __forceinline__ f2 spherical_uv(f3 local) {
    float theta = em::acos(fminf(fmaxf(local.z, -1.f), 1.f));
    float phi = em::atan2(local.y, local.x);
    if (phi < 0.f) phi += 2.f * kPi;
    return mk2(phi * (0.5f * kInvPi), theta * kInvPi);
}
template <int NS, bool RGB>
__device__ __forceinline__ Spec<NS> env_texel(const SceneView &sc, f2 uv, const Swl<NS> &swl) {
    uint32_t x = min(uint32_t(fmaxf(uv.x, 0.f) * float(sc.env_w)), sc.env_w - 1u);
    uint32_t y = min(uint32_t(fmaxf(uv.y, 0.f) * float(sc.env_h)), sc.env_h - 1u);
    const float *t = sc.env_tex + 4 * (size_t(y) * sc.env_w + x);
    return illum_spectrum<NS, RGB>(sc, t, t, sc.env_scale, swl);
}
// AliasTable2D::PDF (alias2d.cpp:101-105)
__device__ __forceinline__ float env_map_pdf(const SceneView &sc, f2 uv) {
    uint32_t iu = min(uint32_t(fmaxf(uv.x * float(sc.env_w), 0.f)), sc.env_w - 1u);
    uint32_t iv = min(uint32_t(fmaxf(uv.y * float(sc.env_h), 0.f)), sc.env_h - 1u);
    return sc.env_marg_integral > 0.f ? sc.env_cond_func[size_t(iv) * sc.env_w + iu] / sc.env_marg_integral : 0.f;
}
// SphericalMap::evaluate_wi for a direction (pdf in solid angle, before light selection)
template <int NS, bool RGB>
__device__ __forceinline__ void env_eval(const SceneView &sc, f3 dir, const Swl<NS> &swl, Spec<NS> &L, float &pdf) {
    f3 local = mat3_apply(sc.env_w2o, dir);
    float theta = em::acos(fminf(fmaxf(local.z, -1.f), 1.f));
    float sin_t = em::sin(theta);
    f2 uv = spherical_uv(local);
    L = env_texel<NS, RGB>(sc, uv, swl);
    float p = env_map_pdf(sc, uv) / (2.f * kPi * kPi * sin_t);
    pdf = sin_t == 0.f ? 0.f : p;
}

// AliasTable2D::sample_continuous (alias2d.cpp:106-129) over the environment
// map: (u, v) in [0, 1)^2 and the map pdf
__device__ __forceinline__ f2 env_sample_uv(const SceneView &sc, f2 u, float &pdf_map) {
    float uvr;
    uint32_t iv = alias_pick(sc.env_marg, sc.env_h, u.y, &uvr);
    float fv = (float(iv) + uvr) / float(sc.env_h);
    float pdf_v = sc.env_marg_integral > 0.f ? sc.env_marg_func[iv] / sc.env_marg_integral : 0.f;
    size_t row = size_t(sc.env_w) * iv;
    float uur;
    uint32_t iu = alias_pick(sc.env_cond + row, sc.env_w, u.x, &uur);
    float fu = (float(iu) + uur) / float(sc.env_w);
    float integral_u = sc.env_marg_func[iv];
    float pdf_u = integral_u > 0.f ? sc.env_cond_func[row + iu] / integral_u : 0.f;
    pdf_map = pdf_u * pdf_v;
    return mk2(fu, fv);
}

template <int NS>
struct LightSampleN {
    Spec<NS> L;
    float pdf;
    f3 p_light;
};

// LightSampler::sample_wi (src/base/illumination/lightsampler.cpp:199-216)
// draws: select 1, surface 2
template <int NS, bool RGB>
__device__ __forceinline__ LightSampleN<NS> sample_light(const SceneView &sc, f3 ref_pos, Sampler &sampler,
                                                         const Swl<NS> &swl) {
    LightSampleN<NS> ls;
    float u_light = sampler.next1d();
    f2 u = sampler.next2d();
    uint32_t li = select_light(sc, u_light);
    float sel_pmf = light_select_pmf(sc, li);
    const LightRecord &L = sc.lights[li];
    if (L.type == LIGHT_AREA) {
        // AreaLight::sample_wi (src/render_core/lights/area.cpp:120-149)
        float ux;
        uint32_t prim = alias_sample(sc, L.alias_offset, L.tri_count, u.x, &ux);
        float pmf = alias_pmf(sc, L.alias_offset, L.tri_count, L.alias_integral, prim);
        f2 bary = square_to_triangle(mk2(ux, u.y));
        Surface p = surface_at(sc, L.tri_base + prim, bary.x, bary.y, false);
        float pdf_pos = (1.f / p.prim_area) * pmf;
        f3 w = ref_pos - p.pos;
        bool front = dot(w, p.ng) > 0.f || L.two_sided;
        ls.L = front ? light_color<NS, RGB>(sc, L, swl, p.uv) : Spec<NS>::fill(0.f);
        float cos_t = fabsf(dot(p.ng, normalize(w)));
        float pdf = pdf_pos * length_squared(w) / cos_t;
        ls.pdf = (isinf(pdf) || isnan(pdf)) ? 0.f : pdf;
        float factor = dot(p.ng, w) > 0.f ? 1.f : -1.f;
        ls.p_light = offset_ray_origin(p.pos, p.ng * factor * sc.ray_offset_factor);
    } else if (L.type == LIGHT_ENV) {
        // SphericalMap::sample_wi / evaluate (spherical.cpp:94-124, 157-163);
        // AliasTable2D::sample_continuous (alias2d.cpp:109-129)
        float pdf_map;
        f2 fuv = env_sample_uv(sc, u, pdf_map);
        float theta = fuv.y * kPi, phi = fuv.x * 2.f * kPi;
        float sin_t, cos_t;
        em::sincos(theta, &sin_t, &cos_t);
        float sp, cp;
        em::sincos(phi, &sp, &cp);
        f3 local = mk3(sin_t * cp, sin_t * sp, cos_t);
        f3 world = normalize(mat3_apply(sc.env_o2w, local));
        float pdf_dir = pdf_map / (2.f * kPi * kPi * sin_t);
        ls.pdf = isinf(pdf_dir) ? 0.f : pdf_dir;
        ls.L = env_texel<NS, RGB>(sc, spherical_uv(local), swl);
        ls.p_light = ref_pos + world * sc.world_diameter;
    } else {
        // IPointLight::sample_wi (src/base/illumination/light.cpp:42-51);
        // point.cpp:39-43, spot.cpp:51-70
        f3 pos = mk3(L.position[0], L.position[1], L.position[2]);
        f3 w_un = ref_pos - pos;
        Spec<NS> val = light_color<NS, RGB>(sc, L, swl, mk2(0.f, 0.f)) / length_squared(w_un);
        if (L.type == LIGHT_SPOT) {
            f3 dir = mk3(L.direction[0], L.direction[1], L.direction[2]);
            float ct = fminf(fmaxf(dot(dir, normalize(w_un)), L.cos_angle), L.cos_falloff_start);
            float f = (ct - L.cos_angle) / (L.cos_falloff_start - L.cos_angle);
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
template <int NS, bool RGB>
__device__ __forceinline__ void eval_hit_light(const SceneView &sc, const Surface &it, uint32_t li, uint32_t tri,
                                               f3 ref_pos, const Swl<NS> &swl, Spec<NS> &L, float &pdf) {
    const LightRecord &Lr = sc.lights[li];
    uint32_t prim = tri - Lr.tri_base;
    float pdf_pos = (1.f / it.prim_area) * alias_pmf(sc, Lr.alias_offset, Lr.tri_count, Lr.alias_integral, prim);
    f3 w = ref_pos - it.pos;
    bool front = dot(w, it.ng) > 0.f || Lr.two_sided;
    L = front ? light_color<NS, RGB>(sc, Lr, swl, it.uv) : Spec<NS>::fill(0.f);
    float cos_t = fabsf(dot(it.ng, normalize(w)));
    float p = pdf_pos * length_squared(w) / cos_t;
    p = (isinf(p) || isnan(p)) ? 0.f : p;
    pdf = p * light_select_pmf(sc, li);
}

// Spectrum::linear_srgb: identity in sRGB mode; in hero mode the CIE XYZ
// estimate of HeroWavelengthSpectrum::cie_xyz (hero.cpp:262-274) mapped to
// linear sRGB (cie.h:414-420)
template <int NS, bool RGB>
__device__ __forceinline__ f3 to_linear_srgb(const SceneView &sc, const Spec<NS> &s, const Swl<NS> &swl) {
    if (RGB) return mk3(s.v[0], s.v[1], s.v[2]);
    float X = 0.f, Y = 0.f, Z = 0.f;
    uint32_t valid = 0;
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        float l = swl.lambda[i], p = swl.pdf[i];
        valid += p > 0.f ? 1u : 0u;
        float x = spd_eval(sc.spd + 0, sc.cie_interval, l) * s.v[i];
        float y = spd_eval(sc.spd + 94, sc.cie_interval, l) * s.v[i];
        float z = spd_eval(sc.spd + 188, sc.cie_interval, l) * s.v[i];
        X += p == 0.f ? 0.f : x / p;
        Y += p == 0.f ? 0.f : y / p;
        Z += p == 0.f ? 0.f : z / p;
    }
    float factor = 1.f / (float(valid) * sc.cie_y_integral);
    X *= factor;
    Y *= factor;
    Z *= factor;
    return mk3(3.240479f * X - 1.537150f * Y - 0.498535f * Z, -0.969256f * X + 1.875991f * Y + 0.041556f * Z,
               0.055648f * X - 0.204043f * Y + 1.057311f * Z);
}

struct PathCounters {
    uint32_t rays;
    uint32_t nodes;
    uint32_t tris;
};

// State of one camera path between bounces.  The bounce loop of
// IlluminationIntegrator::Li is split at its traversals: path_shade runs one
// loop iteration after the closest hit is known and hands back the NEE
// shadow ray, whose (already weighted) contribution the caller adds when the
// ray is unoccluded.  The megakernel runs begin / trace / shade / shadow in
// one thread; the wavefront kernels run each step over a queue of paths.
// Both execute the same arithmetic in the same order.
template <int NS>
struct PathState {
    f3 o, d;          // next ray (t in (0, inf))
    Spec<NS> thr;     // throughput
    f3 L;             // linear sRGB radiance gathered so far
    float scatter_pdf;// pdf of the last BSDF sample (1e16 before the first, pt.cpp:69)
    float eta_scale;
    uint32_t bounce;  // scattering events so far
    uint32_t skips;   // material-less surfaces passed through
    uint32_t alive;
    uint32_t tail;    // the BSDF-only emitter bounce of max_depth < 2 (integrator.cpp:276-281)
    float fw;         // film weight of the sensor sample
    Sampler sampler;
    Swl<NS> swl;
};

struct ShadowRay {
    f3 o, d;// t in (0, 1 - ShadowEpsilon)
    f3 c;   // linear sRGB contribution if unoccluded
    uint32_t valid;
};

// one bounce of IlluminationIntegrator::Li (src/base/integrator.cpp:150-273)
// given the closest hit of ps.o + t ps.d
template <int NS, bool RGB, int MS = MS_ALL>
__device__ __forceinline__ void path_shade(const SceneView &sc, const RenderView &rv, PathState<NS> &ps, const TraceHit &h,
                           ShadowRay &sr) {
    sr.valid = 0u;
    const uint32_t bounces = ps.tail ? 0u : ps.bounce;
    if (h.tri == kInvalid) {
        // IlluminationIntegrator::evaluate_miss (integrator.cpp:111-132)
        if (sc.env_light != kInvalid) {
            f3 wdir = normalize((ps.o + ps.d) - ps.o);
            Spec<NS> L;
            float lpdf;
            env_eval<NS, RGB>(sc, wdir, ps.swl, L, lpdf);
            lpdf *= light_select_pmf(sc, sc.env_light);
            float w = mis_weight(ps.scatter_pdf, lpdf);
            if (rv.mis_mode == 2) w = 1.f;
            else if (rv.mis_mode == 1) w = bounces == 0 ? w : 0.f;
            ps.L += to_linear_srgb<NS, RGB>(sc, (L * w) * ps.thr, ps.swl);
        }
        ps.alive = 0u;
        return;
    }
    Surface it = surface_at(sc, h.tri, h.b1, h.b2, true);
    it.wo = normalize(-ps.d);
    const InstanceRecord inst = sc.inst[it.inst];
    if (inst.mat == kInvalid) {
        // material-less surface: continue straight through (integrator.cpp:180-186)
        f3 n = dot(it.ng, ps.d) > 0.f ? it.ng : -it.ng;
        ps.o = offset_ray_origin(it.pos, n);
        if (++ps.skips > 64u) ps.alive = 0u;
        return;
    }
    if (inst.light != kInvalid) {
        Spec<NS> Le;
        float lpdf;
        eval_hit_light<NS, RGB>(sc, it, inst.light, h.tri, ps.o, ps.swl, Le, lpdf);
        float w = mis_weight(ps.scatter_pdf, lpdf);
        if (rv.mis_mode == 2) w = 1.f;
        else if (rv.mis_mode == 1) w = bounces == 0 ? w : 0.f;
        ps.L += to_linear_srgb<NS, RGB>(sc, Le * ps.thr * w, ps.swl);
    }
    if (ps.tail) {
        ps.alive = 0u;
        return;
    }
    // next-event estimation (direct_lighting, integrator.cpp:22-38)
    LightSampleN<NS> ls = sample_light<NS, RGB>(sc, it.pos, ps.sampler, ps.swl);
    bool dispersive;
    MatCtx<NS> mc = make_material<NS, RGB, MS>(sc, sc.mat[inst.mat], it.frame, it.ng, it.wo, ps.swl, dispersive,
                                                   it.uv);
    if (!RGB && dispersive) {
        // SampledWavelengths::check_dispersive (spectrum.cpp:21-39)
        bool secondary_valid = NS > 1;
#pragma unroll
        for (int i = 1; i < NS; ++i) secondary_valid = secondary_valid && ps.swl.pdf[i] != 0.f;
        if (secondary_valid) {
#pragma unroll
            for (int i = 1; i < NS; ++i) ps.swl.pdf[i] = 0.f;
        }
    }
    f3 wi_l = normalize(ls.p_light - it.pos);
    EvalN<NS> se = mat_eval_any<NS, RGB, MS>(sc, mc, wi_l, nullptr, ps.swl);
    SampleN<NS> bs = mat_sample_any<NS, RGB, MS>(sc, mc, ps.sampler, ps.swl);
    bool delta = ls.pdf < 0.f;
    float w = delta ? 1.f : mis_weight(ls.pdf, se.pdf);
    if (rv.mis_mode == 1) w = 1.f;
    float lpdf = delta ? -ls.pdf : ls.pdf;
    if (se.pdf > 0.f && lpdf > 0.f && rv.mis_mode != 2) {
        f3 dir = ls.p_light - it.pos;
        f3 n = dot(it.ng, dir) > 0.f ? it.ng : -it.ng;
        sr.o = offset_ray_origin(it.pos, n);
        sr.d = dir;
        Spec<NS> Ld = ls.L * se.f * w / lpdf;
        sr.c = to_linear_srgb<NS, RGB>(sc, ps.thr * Ld, ps.swl);
        sr.valid = 1u;
    }
    ps.eta_scale *= bs.eta * bs.eta;
    float lum = ps.thr.maxc();
    if (!(bs.e.pdf > 0.f) || lum == 0.f) {
        ps.alive = 0u;
        return;
    }
    ps.thr = ps.thr * (bs.e.f / bs.e.pdf);
    if (ps.eta_scale * lum < rv.rr_threshold && bounces >= rv.min_depth) {
        float q = fminf(0.95f, lum);
        float rr = ps.sampler.next1d();
        if (q < rr) {
            ps.alive = 0u;
            return;
        }
        ps.thr = ps.thr / q;
    }
    ps.scatter_pdf = bs.e.pdf;
    f3 n = dot(it.ng, bs.wi) > 0.f ? it.ng : -it.ng;
    // with max_depth < 2 the previous ray is the primary one; the extra
    // bounce below is skipped when the new direction equals it (:154-158)
    const bool same_dir = bs.wi.x == ps.d.x && bs.wi.y == ps.d.y && bs.wi.z == ps.d.z;
    ps.o = offset_ray_origin(it.pos, n);
    ps.d = bs.wi;
    ps.bounce = bounces + 1u;
    if (ps.bounce >= rv.max_depth) {
        // with max_depth < 2 and full MIS, one more BSDF-sampled ray adds
        // emitter / environment light only (integrator.cpp:276-281)
        if (rv.max_depth < 2u && rv.mis_mode == 0 && !same_dir) ps.tail = 1u;
        else ps.alive = 0u;
    }
}

// Sensor::generate_ray (src/base/sensor/sensor.cpp:44-56) + filter sample
// (Sampler::sensor_sample, src/base/sampler.h:65-73): draws filter 2, lens 2, time 1
__device__ __forceinline__ void camera_ray(const CameraView &cam, uint32_t px, uint32_t py, Sampler &sampler,
                                           f3 &o, f3 &d, float &weight, f2 *p_film = nullptr) {
    f2 uf = sampler.next2d();
    f2 fs;
    weight = 1.f;
    if (cam.filter == 0) {
        float r = cam.filter_radius;
        fs = mk2(-r + uf.x * (r + r), -r + uf.y * (r + r));
    } else if (cam.filter == 1) {
        fs = mk2(sample_tent(uf.x, cam.filter_radius), sample_tent(uf.y, cam.filter_radius));
    } else {
        // FilterSampler::sample (src/render_core/filters/fitted_curve.h:67-73) and
        // FittedCurveFilter::sample (:90-94): the 2D alias table over |f| on the
        // positive quadrant, mirrored by the signs of u; weight = f / pdf
        f2 u = mk2(uf.x * 2.f - 1.f, uf.y * 2.f - 1.f);
        float uvr;
        uint32_t iv = alias_pick(cam.fis_marg, 20u, fabsf(u.y), &uvr);
        float fv = (float(iv) + uvr) / 20.f;
        float pdf_v = cam.fis_marg_integral > 0.f ? cam.fis_marg_func[iv] / cam.fis_marg_integral : 0.f;
        uint32_t row = 20u * iv;
        float uur;
        uint32_t iu = alias_pick(cam.fis_cond + row, 20u, fabsf(u.x), &uur);
        float fu = (float(iu) + uur) / 20.f;
        float integral_u = cam.fis_marg_func[iv];
        float pdf_u = integral_u > 0.f ? cam.fis_cond_func[row + iu] / integral_u : 0.f;
        float pdf = pdf_u * pdf_v;
        f2 p = mk2(u.x < 0.f ? -fu : fu, u.y < 0.f ? -fv : fv);
        weight = cam.fis_lut[row + iu] / pdf;
        fs = mk2(p.x * cam.filter_radius, p.y * cam.filter_radius);
    }
    f2 p_lens = sampler.next2d();
    (void)sampler.next1d();// time
    float fx = float(px) + 0.5f + fs.x;
    float fy = float(py) + 0.5f + fs.y;
    if (p_film) *p_film = mk2(fx, fy);
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

// RenderEnv::initial (src/base/integrator.cpp:49-58): wavelengths come from
// the stream (pixel, frame, 0xFFFFFFFF) without advancing the main stream
template <int NS, bool RGB>
__device__ __forceinline__ Swl<NS> sample_wavelengths(uint32_t px, uint32_t py, uint32_t frame) {
    Swl<NS> swl;
    if (RGB) {
        const float peaks[3] = {602.785f, 539.285f, 445.772f};
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            swl.lambda[i] = peaks[i < 3 ? i : 2];
            swl.pdf[i] = 1.f;
        }
    } else {
        Sampler t;
        t.start(px, py, frame, 0xFFFFFFFFu);
        float u = t.next1d();
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            float offset = float(i) * (1.f / float(NS));
            float up = u + offset;
            up = up - floorf(up);
            float l = sample_visible_wavelength(up);
            swl.lambda[i] = l;
            swl.pdf[i] = visible_wavelength_pdf(l);
        }
    }
    return swl;
}

// start of a camera path (pt.cpp:53-73): wavelengths from their own stream,
// sensor sample, camera ray; the film weight of box and tent filters is 1
template <int NS, bool RGB>
__device__ __forceinline__ void path_begin(const CameraView &cam, const RenderView &rv, uint32_t px, uint32_t py,
                                           uint32_t frame, PathState<NS> &ps) {
    ps.swl = sample_wavelengths<NS, RGB>(px, py, frame);
    ps.sampler.start(px, py, frame, 0u);
    camera_ray(cam, px, py, ps.sampler, ps.o, ps.d, ps.fw);
    ps.thr = Spec<NS>::fill(1.f);
    ps.L = mk3(0.f, 0.f, 0.f);
    ps.scatter_pdf = 1e16f;
    ps.eta_scale = 1.f;
    ps.bounce = 0u;
    ps.skips = 0u;
    ps.tail = 0u;
    ps.alive = rv.max_depth > 0u ? 1u : 0u;
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

// sRGB transfer function (IEC 61966-2-1): linear segment below 0.0031308,
// 1.055 x^(1/2.4) - 0.055 above
__device__ __forceinline__ float linear_to_srgb(float x) {
    return x <= 0.0031308f ? 12.92f * x : 1.055f * em::pow(x, 1.f / 2.4f) - 0.055f;
}

// one sample into the film's running mean (NormalRadianceCollector::add_sample,
// src/render_core/radiancecollectors/normal.cpp:87-101)
__device__ __forceinline__ float4 film_add(const RenderView &rv, float4 acc, f3 L, uint32_t frame) {
    float4 v = make_float4(zero_if_nan_inf(L.x), zero_if_nan_inf(L.y), zero_if_nan_inf(L.z), 1.f);
    if (!rv.accumulation) return v;
    float a = 1.f / float(frame + 1u);
    return make_float4(acc.x + a * (v.x - acc.x), acc.y + a * (v.y - acc.y), acc.z + a * (v.z - acc.z),
                       acc.w + a * (v.w - acc.w));
}

// the output buffer: tone_map(1 - exp(-v * exposure)) (radiance_collector.cpp:24-26)
__device__ __forceinline__ float4 film_output(const RenderView &rv, float4 acc) {
    float4 o4;
    o4.x = tone_map(1.f - em::exp(-acc.x * rv.exposure), rv.tone_mapper);
    o4.y = tone_map(1.f - em::exp(-acc.y * rv.exposure), rv.tone_mapper);
    o4.z = tone_map(1.f - em::exp(-acc.z * rv.exposure), rv.tone_mapper);
    o4.w = 1.f;
    return o4;
}

// The samples [frame_start, frame_start + frames) of up to n_px pixels of one
// column (rows y0, y0 + dy, ...) into the running mean rt (and the output
// buffer when out is given), back to back: a lane whose path ends starts its
// next sample at once (the next frame of its pixel, then its next pixel)
// instead of waiting at the end of each frame for the longest path of its
// wave.  Every pixel's samples still run in frame order in one lane with the
// same arithmetic, so the film equals the frame-by-frame loop's bit for bit.
// VarianceStats::add_sample_impl (adaptive/inspector.h:22-31) on (avg, var,
// N bits, 0)
__device__ __forceinline__ float4 vstats_add(float4 st, float x) {
    const uint32_t N = __float_as_uint(st.z);
    const float fN = float(N), fN1 = float(N + 1u);
    const float new_avg = (st.x * fN + x) / fN1;
    const float d = st.x - new_avg, e = x - st.x;
    const float new_var = (fN * st.y + fN * (d * d) + e * e + 2.f * e * d + d * d) / fN1;
    return make_float4(new_avg, new_var, __uint_as_float(N + 1u), 0.f);
}

// ConvergenceInspector::is_convergence (inspector.cpp:33-36)
__device__ __forceinline__ bool vstats_converged(const RenderView &rv, float4 st) {
    return __float_as_uint(st.z) > rv.adapt_min_samples && st.y / st.x < rv.adapt_threshold;
}

// COUNT: tally rays and BVH fetches into pc (the counting runs).
// ADAPT: adaptive sampling (pt.cpp:43-64): a pixel whose statistics have
// converged adds no more samples (a skipped sample adds nothing, so the rest
// of its frames are skipped too); each added sample feeds the statistics
// the output buffer's value x + y + z
template <int NS, bool RGB, int MS = MS_ALL, bool ADAPT = false, bool COUNT = true>
__device__ __forceinline__ void render_column(const SceneView &sc, const CameraView &cam, const RenderView &rv,
                              float4 *__restrict__ rt, float4 *__restrict__ out, uint32_t px, uint32_t y0,
                              uint32_t dy, uint32_t n_px, uint32_t frame_start, uint32_t frames, uint2 *stack,
                              uint32_t stride, PathCounters &pc) {
    if (px >= cam.width || y0 >= cam.height || frames == 0u || n_px == 0u) return;
    const uint32_t n = min(n_px, (cam.height - 1u - y0) / dy + 1u);
    uint32_t j = 0, s = 0, py = y0;
    uint32_t idx = py * cam.width + px;
    float4 acc = rt[idx];
    float4 vst = ADAPT ? rv.vstats[idx] : make_float4(0.f, 0.f, 0.f, 0.f);
    PathState<NS> ps;
    bool fresh = true, skip = false;
    while (true) {
        if (fresh) {
            if (ADAPT && vstats_converged(rv, vst)) {
                skip = true;
                ps.alive = 0u;
            } else {
                path_begin<NS, RGB>(cam, rv, px, py, frame_start + s, ps);
            }
            fresh = false;
        }
        if (ps.alive) {
            // COUNT = false: the fetch counts are dead and their increments
            // in the traversal loops compile away
            uint32_t dn = 0, dt = 0;
            TraceHit h = traverse<false>(sc, ps.o, ps.d, 0.f, 3.402823466e38f, stack, stride, dn, dt);
            ShadowRay sr;
            path_shade<NS, RGB, MS>(sc, rv, ps, h, sr);
            if (sr.valid) {
                TraceHit sh = traverse<true>(sc, sr.o, sr.d, 0.f, 1.f - kShadowEpsilon, stack, stride, dn, dt);
                if (COUNT) ++pc.rays;
                if (sh.tri == kInvalid) ps.L += sr.c;
            }
            if (COUNT) {
                pc.nodes += dn;
                pc.tris += dt;
                ++pc.rays;
            }
        }
        if (!ps.alive) {
            if (!ADAPT || !skip) {
                acc = film_add(rv, acc, ps.L * ps.fw, frame_start + s);
                if (ADAPT) {
                    const float4 o = film_output(rv, acc);
                    vst = vstats_add(vst, o.x + o.y + o.z);
                }
            }
            fresh = true;
            if (++s == frames || (ADAPT && skip)) {
                rt[idx] = acc;
                if (out) out[idx] = film_output(rv, acc);
                if (ADAPT) rv.vstats[idx] = vst;
                s = 0;
                skip = false;
                if (++j == n) break;
                py += dy;
                idx = py * cam.width + px;
                acc = rt[idx];
                if (ADAPT) vst = rv.vstats[idx];
            }
        }
    }
}

}// namespace dev
}// namespace vs
