// Per-hit material evaluator: MaterialEvaluator with individual shading
// normals (src/base/scattering/material.cpp:132-184) over the lobes of
// diffuse.cpp, metal.cpp, glass.cpp and principled_bsdf.cpp.
#pragma once

#include "dev_bsdf.h"
#include "dev_spectrum.h"

namespace vs {
namespace dev {

// material record layouts (packed by the host, tables.cpp)
//  diffuse / oren-nayar: p[0..3] colour slot, p[4] sigma
//  metal:  p[0..2] eta rgb, p[3..5] k rgb, p[6] ax, p[7] ay, p[8] spd eta offset,
//          p[9] spd k offset, p[10] spd interval
//  glass:  p[0..3] colour slot, p[4] ior, p[5] ax, p[6] ay, p[7] spd offset,
//          p[8] spd interval; flags bit 0: dispersive ior curve
//  principled: p[0..3] color, p[4..7] spec_tint, p[8..11] sheen_tint,
//          p[12..15] coat_tint, p[16] metallic, p[17] ior, p[18] roughness,
//          p[19] anisotropic, p[20] sheen_weight, p[21] sheen_roughness,
//          p[22] coat_weight, p[23] coat_roughness, p[24] coat_ior,
//          p[25] transmission_weight
//  mix / add: p[0], p[1] indices (uint bits) of the two component records
//          (never mix / add / black body themselves), p[2] the mix fraction
// A colour slot holds linear sRGB in sRGB mode and the sigmoid-polynomial
// coefficients of the RGB albedo spectrum in hero mode.
constexpr uint32_t MATF_DISPERSIVE = 1u;
// material sets a kernel is compiled for: every material, or one family
// (the diffuse family: Lambert, Oren-Nayar, black body; metal; glass;
// principled) for scenes or queues that hold nothing else, which keeps the
// evaluator small.  The black body is handled in every set.
constexpr int MS_ALL = 0;
constexpr int MS_DIFFUSE = 1;
constexpr int MS_METAL = 2;
constexpr int MS_GLASS = 3;
constexpr int MS_PRINCIPLED = 4;
template <int MS>
__device__ __forceinline__ constexpr bool ms_has(int family) { return MS == MS_ALL || MS == family; }
constexpr uint32_t FLAG_TRANSMISSION = 4u;// BxDFFlag::Transmission

template <int NS, bool RGB>
__device__ __forceinline__ Spec<NS> eval_albedo(const float *slot, const Swl<NS> &swl) {
    Spec<NS> r;
    if (RGB) {
#pragma unroll
        for (int i = 0; i < NS; ++i) r.v[i] = slot[i < 3 ? i : 2];
    } else {
#pragma unroll
        for (int i = 0; i < NS; ++i) r.v[i] = sigmoid_poly(slot, swl.lambda[i]);
    }
    return r;
}

// a colour slot of material m at p[off]: its constant, or the texel of the
// image it names at uv (nearest texel, uv wrapped to [0, 1), v = 0 the first
// row)
template <int NS, bool RGB>
__device__ __forceinline__ Spec<NS> slot_albedo(const SceneView &sc, const MaterialRecord &m, int off, f2 uv,
                                                const Swl<NS> &swl) {
    if (m.flags & (MATF_TEXTURE_SLOT0 << (off >> 2))) {
        const uint4 info = sc.tex_info[__float_as_uint(m.p[off])];
        const float u = uv.x - floorf(uv.x), v = uv.y - floorf(uv.y);
        const uint32_t x = min(uint32_t(u * float(info.y)), info.y - 1u);
        const uint32_t y = min(uint32_t(v * float(info.z)), info.z - 1u);
        const float4 t = sc.tex[size_t(info.x) + size_t(y) * info.y + x];
        const float q[4] = {t.x, t.y, t.z, t.w};
        return eval_albedo<NS, RGB>(q, swl);
    }
    return eval_albedo<NS, RGB>(m.p + off, swl);
}

template <int NS>
struct MatCtx {
    uint32_t type;
    Frame fr;
    f3 ng, wo_w, wo;
    Ggx ggx;
    // diffuse / metal / glass
    Spec<NS> a0;// kr (diffuse), eta (metal), kt (glass)
    Spec<NS> a1;// k (metal), per-channel eta (glass)
    float onA, onB;
    float comp;
    float eta0;
    // principled
    float sh_a, sh_b;
    Spec<NS> sh_tint, cc_kr, mt_kr, mt_F0, mt_B, tr_kt, tr_F0, sp_kr, sp_F0, df_kr;
    float cc_eta, cc_alpha, mt_comp, tr_eta, tr_weight, tr_comp, sp_eta;
    float sw[6];// normalized sample weights: sheen, coat, metal, trans, spec, diffuse
    // mix / add: the component records, their evaluation and sample weights
    // in the flattened lobe set, the surface uv (components are rebuilt on use)
    uint32_t sub[2];
    float sub_e[2], sub_s[2];
    f2 uv;
};

template <int NS>
struct EvalN {
    Spec<NS> f;// BSDF * |cos wi|
    float pdf;
    uint32_t flags;
};

__device__ __forceinline__ float layering_max(float a, float w) { return w != 0.f ? a / w : 0.f; }
template <int NS>
__device__ __forceinline__ Spec<NS> layering_weight(const Spec<NS> &albedo, const Spec<NS> &weight) {
    float m = layering_max(albedo.v[0], weight.v[0]);
#pragma unroll
    for (int i = 1; i < NS; ++i) m = fmaxf(m, layering_max(albedo.v[i], weight.v[i]));
    return weight * fminf(fmaxf(1.f - m, 0.f), 1.f);
}

__device__ __forceinline__ float dielectric_comp(const float *lut, float ax, float ay, float abs_cos_o, float eta) {
    // DielectricLobe::sample_lut / refl_compensate (lobe.cpp:263-306)
    const float *t = lut + (eta > 1.f ? LUT_DIEL : LUT_DIEL_INV);
    float x = sqrtf(sqrtf(ax * ay));
    float e = eta > 1.f ? eta : 1.f / eta;
    float z = (e - 1.003f) / (5.f - 1.003f);
    float v[2];
    lut3<2>(t, 32, x, abs_cos_o, z, v);
    return 1.f / v[0];
}

__device__ __forceinline__ float pure_reflection_comp(const float *lut, float ax, float ay, float cos_o) {
    // PureReflectionLobe::compensate_factor (lobe.cpp:716-721)
    float v;
    lut2<1>(lut + LUT_PURE, 32, sqrtf(ax * ay), cos_o, &v);
    return 1.f / v;
}

// Build the evaluator for material `m` at `it` (create_lobe_set of each
// material).  `dispersive` reports a wavelength-dependent ior (glass with an
// ior curve), on which the caller drops the secondary wavelengths.
template <int NS, bool RGB, int MS = MS_ALL>
__device__ __forceinline__ MatCtx<NS> make_material(const SceneView &sc, const MaterialRecord &m, const Frame &fr,
                                                    f3 ng, f3 wo_w, const Swl<NS> &swl, bool &dispersive, f2 uv) {
    MatCtx<NS> c;
    c.type = m.type;
    c.fr = fr;
    c.ng = ng;
    c.wo_w = wo_w;
    c.wo = fr.to_local(wo_w);
    dispersive = false;
    const float *p = m.p;
    if (MS == MS_ALL && (m.type == MAT_MIX || m.type == MAT_ADD)) {
        // LobeSet::create_mix / create_add (lobe.cpp:495-521) then flatten
        // (lobe.cpp:533-562): component k carries sample weight s_k (mix:
        // 1 - frac or frac; add: 1/2 after normalisation) and evaluation
        // weight e_k (mix: as s_k; add: 1, or 1/2 when the component is a
        // lobe set, whose lobes flatten with the parent's sample weight)
        const bool add = m.type == MAT_ADD;
        c.uv = uv;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            c.sub[k] = __float_as_uint(p[k]);
            const MaterialRecord &cm = sc.mat[c.sub[k]];
            const float wk = k == 0 ? 1.f - p[2] : p[2];
            c.sub_s[k] = add ? 0.5f : wk;
            c.sub_e[k] = add ? (cm.type == MAT_PRINCIPLED ? 0.5f : 1.f) : wk;
            if (!RGB && cm.type == MAT_GLASS && (cm.flags & MATF_DISPERSIVE)) dispersive = true;
        }
        return c;
    }
    if (ms_has<MS>(MS_DIFFUSE) && (m.type == MAT_DIFFUSE || m.type == MAT_OREN_NAYAR)) {
        c.a0 = slot_albedo<NS, RGB>(sc, m, 0, uv, swl);
        if (m.type == MAT_OREN_NAYAR) {
            float sig = p[4] * (kPi * 0.5f);
            float s2 = (sig * sig) * (sig * sig);
            c.onA = 1.f - (s2 / (2.f * (s2 + 0.33f)));
            c.onB = 0.45f * s2 / (s2 + 0.09f);
        }
    } else if (ms_has<MS>(MS_METAL) && m.type == MAT_METAL) {
        c.ggx = {p[6], p[7]};
        if (RGB) {
#pragma unroll
            for (int i = 0; i < NS; ++i) {
                c.a0.v[i] = p[i < 3 ? i : 2];
                c.a1.v[i] = p[3 + (i < 3 ? i : 2)];
            }
        } else {
            const float *se = sc.spd + __float_as_uint(p[8]);
            const float *sk = sc.spd + __float_as_uint(p[9]);
#pragma unroll
            for (int i = 0; i < NS; ++i) {
                c.a0.v[i] = spd_eval(se, p[10], swl.lambda[i]);
                c.a1.v[i] = spd_eval(sk, p[10], swl.lambda[i]);
            }
        }
        c.comp = pure_reflection_comp(sc.lut, c.ggx.ax, c.ggx.ay, c.wo.z);
    } else if (ms_has<MS>(MS_METAL) && m.type == MAT_MIRROR) {
        // MirrorLobe: the tint as kr, constant Fresnel, compensated
        c.ggx = {p[6], p[7]};
        c.a0 = slot_albedo<NS, RGB>(sc, m, 0, uv, swl);
        c.comp = pure_reflection_comp(sc.lut, c.ggx.ax, c.ggx.ay, c.wo.z);
    } else if (ms_has<MS>(MS_GLASS) && m.type == MAT_GLASS) {
        c.ggx = {p[5], p[6]};
        c.a0 = slot_albedo<NS, RGB>(sc, m, 0, uv, swl);
        bool front = dot(wo_w, ng) > 0.f;// Interaction::correct_eta
        if (!RGB && (m.flags & MATF_DISPERSIVE)) {
            dispersive = true;
            const float *s = sc.spd + __float_as_uint(p[7]);
#pragma unroll
            for (int i = 0; i < NS; ++i) {
                float ior = spd_eval(s, p[8], swl.lambda[i]);
                c.a1.v[i] = front ? ior : 1.f / ior;
            }
        } else {
            float e = front ? p[4] : 1.f / p[4];
            c.a1 = Spec<NS>::fill(e);
        }
        c.eta0 = c.a1.v[0];
        c.comp = dielectric_comp(sc.lut, c.ggx.ax, c.ggx.ay, fabsf(c.wo.z), c.eta0);
    } else if (ms_has<MS>(MS_PRINCIPLED) && m.type == MAT_PRINCIPLED) {
        // PrincipledBSDF::create_lobe_set (principled_bsdf.cpp:352-461)
        Spec<NS> color = slot_albedo<NS, RGB>(sc, m, 0, uv, swl);
        Spec<NS> spec_tint = slot_albedo<NS, RGB>(sc, m, 4, uv, swl);
        float ior = p[17];
        float rough = fminf(fmaxf(p[18], 0.0001f), 1.f);
        float aniso = p[19];
        float aspect = sqrtf(1.f - aniso * 0.9f);
        float r2 = rough * rough;
        c.ggx = {fmaxf(0.001f, r2 / aspect), fmaxf(0.001f, r2 * aspect)};
        Spec<NS> weight = Spec<NS>::fill(1.f);
        float cos_theta = dot(wo_w, ng);
        float front = cos_theta > 0.f ? 1.f : 0.f;
        float sw[6];
        {   // sheen (SheenLTC, approximate mode)
            Spec<NS> tint = slot_albedo<NS, RGB>(sc, m, 8, uv, swl) * (p[20] * front) * weight;
            float ltc[4];
            lut2<4>(sc.lut + LUT_SHEEN, 32, cos_theta, p[21], ltc);
            c.sh_a = ltc[0];
            c.sh_b = ltc[1];
            c.sh_tint = tint * ltc[2];
            sw[0] = c.sh_tint.average();
            weight = layering_weight(c.sh_tint, weight);
        }
        {   // coat
            float cw = p[22] * front;
            float cr = fminf(fmaxf(p[23], 0.0001f), 1.f);
            cr = cr * cr;
            c.cc_alpha = cr;
            c.cc_eta = p[24];
            c.cc_kr = (cw * weight) * slot_albedo<NS, RGB>(sc, m, 12, uv, swl);
            float s;
            lut3<1>(sc.lut + LUT_COAT, 32, sqrtf(sqrtf(cr * cr)), cos_theta, (c.cc_eta - 1.003f) / (4.f - 1.003f), &s);
            Spec<NS> cc_albedo = c.cc_kr * s;
            sw[1] = cc_albedo.average();
            weight = layering_weight(cc_albedo, weight);
        }
        {   // metallic (F82-tint conductor with energy compensation)
            float metallic = p[16] * front;
            c.mt_F0 = color;
            const float f = 6.f / 7.f;
            const float f5 = f * f * f * f * f;
            Spec<NS> fs = color + f5 * (1.f - color);
            c.mt_B = fs * (7.f / (f5 * f)) * (1.f - spec_tint);
            c.mt_kr = weight * metallic;
            c.mt_comp = pure_reflection_comp(sc.lut, c.ggx.ax, c.ggx.ay, c.wo.z);
            sw[2] = metallic * weight.average();
            weight = weight * (1.f - metallic);
        }
        {   // transmission (dielectric with Schlick Fresnel)
            float tw = p[25];
            float eta = cos_theta > 0.f ? ior : 1.f / ior;
            c.tr_eta = eta;
            c.tr_kt = color;
            c.tr_F0 = spec_tint * schlick_f0_from_ior(eta);
            Spec<NS> t_weight = weight * tw;
            c.tr_weight = t_weight.average();
            c.tr_comp = dielectric_comp(sc.lut, c.ggx.ax, c.ggx.ay, fabsf(c.wo.z), eta);
            sw[3] = c.tr_weight;
            weight = weight * (1.f - tw);
        }
        {   // specular
            c.sp_eta = ior;
            c.sp_F0 = spec_tint * schlick_f0_from_ior(ior);
            c.sp_kr = weight;
            float s;
            float z = sqrtf(fabsf((ior - 1.f) / (ior + 1.f)));
            lut3<1>(sc.lut + LUT_SPEC, 32, sqrtf(sqrtf(c.ggx.ax * c.ggx.ay)), cos_theta, z, &s);
            Spec<NS> alb = (c.sp_F0 + s * (1.f - c.sp_F0)) * c.sp_kr;
            sw[4] = alb.average();
            weight = layering_weight(alb, weight);
        }
        {   // diffuse
            c.df_kr = color * weight * front;
            sw[5] = c.df_kr.average();
        }
        float sum = 0.f;
#pragma unroll
        for (int i = 0; i < 6; ++i) sum += sw[i];
#pragma unroll
        for (int i = 0; i < 6; ++i) c.sw[i] = sw[i] / sum;
    }
    return c;
}

// ---- local lobe evaluations (f without the cosine factor applied)
template <int NS>
__device__ __forceinline__ void eval_conductor_like(const Ggx &g, f3 wo, f3 wi, bool f82, const Spec<NS> &a,
                                                    const Spec<NS> &b, const Spec<NS> &kr, float comp,
                                                    EvalN<NS> &e) {
    // MicrofacetReflection::f / PDF through BxDF::safe_evaluate, then
    // PureReflectionLobe compensation (bxdf.cpp:64-79, lobe.cpp:723-730)
    bool s = wo.z * wi.z > 0.f;
    f3 wh = normalize(wo + wi);
    f3 whf = face_forward(wh, mk3(0.f, 0.f, 1.f));
    float c = fabsf(dot(wo, whf));
    Spec<NS> F;
    if (f82) {
        F = fresnel_f82(a, b, c);
    } else {
#pragma unroll
        for (int i = 0; i < NS; ++i) F.v[i] = fresnel_conductor(c, a.v[i], b.v[i]);
    }
    e.f = s ? (g.brdf(wo, whf, wi) * F) * kr : Spec<NS>::fill(0.f);
    e.pdf = s ? g.pdf_reflection(wo, wh) : 0.f;
    e.f = e.f * comp;
    e.flags = 0;
}

// MicrofacetReflection with FresnelConstant (fresnel.h:132-137): f = D G kr
// / (4 |cos o cos i|), then the pure-reflection compensation
template <int NS>
__device__ __forceinline__ void eval_mirror(const Ggx &g, f3 wo, f3 wi, const Spec<NS> &kr, float comp,
                                            EvalN<NS> &e) {
    bool s = wo.z * wi.z > 0.f;
    f3 wh = normalize(wo + wi);
    f3 whf = face_forward(wh, mk3(0.f, 0.f, 1.f));
    e.f = s ? kr * g.brdf(wo, whf, wi) : Spec<NS>::fill(0.f);
    e.pdf = s ? g.pdf_reflection(wo, wh) : 0.f;
    e.f = e.f * comp;
    e.flags = 0;
}

template <int NS>
__device__ __forceinline__ void eval_reflection_dielectric(const Ggx &g, f3 wo, f3 wi, float eta, bool schlick,
                                                           const Spec<NS> &F0, const Spec<NS> &kr, EvalN<NS> &e) {
    // MicrofacetReflection with a dielectric / Schlick Fresnel (coat, specular)
    bool s = wo.z * wi.z > 0.f;
    f3 wh = normalize(wo + wi);
    f3 whf = face_forward(wh, mk3(0.f, 0.f, 1.f));
    float c = fabsf(dot(wo, whf));
    Spec<NS> F = schlick ? fresnel_schlick(F0, eta, c) : Spec<NS>::fill(fresnel_dielectric(c, eta));
    e.f = s ? (g.brdf(wo, whf, wi) * F) * kr : Spec<NS>::fill(0.f);
    e.pdf = s ? g.pdf_reflection(wo, wh) : 0.f;
    e.flags = 0;
}

// DielectricLobe::evaluate_local_impl with eta (lobe.cpp:318-415)
template <int NS>
__device__ __forceinline__ void eval_dielectric(const Ggx &g, f3 wo, f3 wi, const Spec<NS> &eta_s, float eta0,
                                                bool schlick, const Spec<NS> &F0, const Spec<NS> &kt, float comp,
                                                EvalN<NS> &e, float *eta_out) {
    bool reflect = wo.z * wi.z > 0.f;
    float eta_p = reflect ? 1.f : eta0;
    if (eta_out) *eta_out = eta_p;
    f3 wh = face_forward(normalize(wo + wi * eta_p), wo);
    float c = fabsf(dot(wh, wo));
    Spec<NS> F;
    if (schlick) {
        F = fresnel_schlick(F0, eta0, c);
    } else {
#pragma unroll
        for (int i = 0; i < NS; ++i) F.v[i] = fresnel_dielectric(c, eta_s.v[i]);
    }
    Spec<NS> T = 1.f - F;
    float rp = F.average() / (T * kt + F).average();
    if (reflect) {
        e.f = F * g.brdf(wo, wh, wi);
        e.pdf = g.pdf_reflection(wo, wh) * rp;
        e.flags = 0;
    } else {
        f3 wh2 = normalize(wo + wi * eta0);
        e.f = (g.btdf(wo, wh2, wi, eta0) * T) * kt;
        e.pdf = g.pdf_transmission(wo, wh, wi, eta0) * (1.f - rp);
        e.flags = FLAG_TRANSMISSION;
    }
    e.f = e.f * comp;
}

// SheenLTC::evaluate_local_impl (principled_bsdf.cpp:59-71)
template <int NS>
__device__ __forceinline__ void eval_sheen(float a, float b, const Spec<NS> &tint, f3 wo, f3 wi, EvalN<NS> &e) {
    f3 w = mk3(a * wi.x + b * wi.z, a * wi.y, wi.z);
    float l = length(w);
    w = w / l;
    float jac = (a * a) / (l * l * l);
    float ltc = w.z * kInvPi * jac;// cosine_hemisphere_PDF(cos_theta(w)) * jacobian
    e.f = (wi.z < 0.f || wo.z < 0.f) ? Spec<NS>::fill(0.f) : (ltc * tint) / wi.z;
    e.pdf = ltc;
    e.flags = 0;
}

template <int NS>
__device__ __forceinline__ void eval_lambert(const Spec<NS> &kr, f3 wo, f3 wi, float A, float B, bool oren,
                                             EvalN<NS> &e) {
    bool s = wo.z * wi.z > 0.f;
    Spec<NS> f = kr * kInvPi;
    if (oren) {
        float sti = sqrtf(fmaxf(0.f, 1.f - wi.z * wi.z)), sto = sqrtf(fmaxf(0.f, 1.f - wo.z * wo.z));
        float cpi = sti == 0.f ? 1.f : fminf(fmaxf(wi.x / sti, -1.f), 1.f);
        float spi = sti == 0.f ? 0.f : fminf(fmaxf(wi.y / sti, -1.f), 1.f);
        float cpo = sto == 0.f ? 1.f : fminf(fmaxf(wo.x / sto, -1.f), 1.f);
        float spo = sto == 0.f ? 0.f : fminf(fmaxf(wo.y / sto, -1.f), 1.f);
        float maxc = fmaxf(0.f, cpi * cpo + spi * spo);
        bool cond = fabsf(wi.z) > fabsf(wo.z);
        float sin_a = cond ? sto : sti;
        float tan_b = cond ? sti / fabsf(wi.z) : sto / fabsf(wo.z);
        f = f * (A + B * maxc * sin_a * tan_b);
    }
    e.f = s ? f : Spec<NS>::fill(0.f);
    e.pdf = s ? fabsf(wi.z) * kInvPi : 0.f;
    e.flags = 0;
}

// evaluation of a leaf material's lobe(s), added into e: f * |cos wi| and
// pdf before the geometric hemisphere check, each lobe's evaluation weight
// times ew and sample weight times sw (1 and 1 for the material alone; a
// component's weights inside a mix / add).  Lobe by lobe as
// LobeSet::evaluate_impl (lobe.cpp:668-688) sums them: f |cos| times (weight
// x factor), pdf times (sample weight x factor), the factor masking lobes
// that only reflect to the side of wo.
template <int NS, int MS = MS_ALL>
__device__ __forceinline__ void leaf_eval_into(const MatCtx<NS> &c, f3 wi_w, float *eta_out, float ew, float sw,
                                               EvalN<NS> &e) {
    f3 wo = c.wo, wi = c.fr.to_local(wi_w);
    EvalN<NS> t;
    if (ms_has<MS>(MS_PRINCIPLED) && c.type == MAT_PRINCIPLED) {
        float acos = fabsf(wi.z);
        float same = wo.z * wi.z > 0.f ? 1.f : 0.f;
        eval_sheen(c.sh_a, c.sh_b, c.sh_tint, wo, wi, t);
        e.f = e.f + (t.f * acos) * (ew * same);
        e.pdf += t.pdf * ((c.sw[0] * sw) * same);
        eval_reflection_dielectric(Ggx{c.cc_alpha, c.cc_alpha}, wo, wi, c.cc_eta, false, c.cc_kr, c.cc_kr, t);
        e.f = e.f + (t.f * acos) * (ew * same);
        e.pdf += t.pdf * ((c.sw[1] * sw) * same);
        eval_conductor_like(c.ggx, wo, wi, true, c.mt_F0, c.mt_B, c.mt_kr, c.mt_comp, t);
        e.f = e.f + (t.f * acos) * (ew * same);
        e.pdf += t.pdf * ((c.sw[2] * sw) * same);
        eval_dielectric(c.ggx, wo, wi, c.tr_F0, c.tr_eta, true, c.tr_F0, c.tr_kt, c.tr_comp, t, eta_out);
        e.f = e.f + (t.f * acos) * (c.tr_weight * ew);
        e.pdf += t.pdf * (c.sw[3] * sw);
        e.flags |= t.flags;
        eval_reflection_dielectric(c.ggx, wo, wi, c.sp_eta, true, c.sp_F0, c.sp_kr, t);
        e.f = e.f + (t.f * acos) * (ew * same);
        e.pdf += t.pdf * ((c.sw[4] * sw) * same);
        eval_lambert(c.df_kr, wo, wi, 0.f, 0.f, false, t);
        e.f = e.f + (t.f * acos) * (ew * same);
        e.pdf += t.pdf * ((c.sw[5] * sw) * same);
        return;
    }
    if (ms_has<MS>(MS_DIFFUSE) && (c.type == MAT_DIFFUSE || c.type == MAT_OREN_NAYAR)) {
        eval_lambert(c.a0, wo, wi, c.onA, c.onB, c.type == MAT_OREN_NAYAR, t);
    } else if (ms_has<MS>(MS_METAL) && c.type == MAT_METAL) {
        eval_conductor_like(c.ggx, wo, wi, false, c.a0, c.a1, Spec<NS>::fill(1.f), c.comp, t);
    } else if (ms_has<MS>(MS_METAL) && c.type == MAT_MIRROR) {
        eval_mirror(c.ggx, wo, wi, c.a0, c.comp, t);
    } else if (ms_has<MS>(MS_GLASS) && c.type == MAT_GLASS) {
        eval_dielectric(c.ggx, wo, wi, c.a1, c.eta0, false, c.a1, c.a0, c.comp, t, eta_out);
    } else {
        // black body: f = 0, pdf = 1 (black_body.cpp:36-43)
        t.f = Spec<NS>::fill(0.f);
        t.pdf = 1.f;
        t.flags = 0;
        e.f = e.f + t.f;
        e.pdf += t.pdf * sw;
        return;
    }
    // one lobe: its hemisphere factor is already in f and pdf (zero where
    // a reflection lobe sees wi on the other side; a dielectric's is 1)
    t.f = t.f * fabsf(wi.z);
    e.f = e.f + t.f * ew;
    e.pdf += t.pdf * sw;
    e.flags |= t.flags;
}

template <int NS, int MS = MS_ALL>
__device__ __forceinline__ EvalN<NS> leaf_eval(const MatCtx<NS> &c, f3 wi_w, float *eta_out) {
    EvalN<NS> e;
    e.f = Spec<NS>::fill(0.f);
    e.pdf = 0.f;
    e.flags = 0;
    leaf_eval_into<NS, MS>(c, wi_w, eta_out, 1.f, 1.f, e);
    return e;
}

// the geometric hemisphere check of MaterialEvaluator::evaluate
// (material.cpp:132-148): no pdf for a reflection across the geometric
// surface or a transmission that stays on one side
template <int NS>
__device__ __forceinline__ void geometric_check(const MatCtx<NS> &c, f3 wi_w, EvalN<NS> &e) {
    bool transmission = (e.flags & FLAG_TRANSMISSION) != 0;
    bool same_geo = dot(c.wo_w, c.ng) * dot(wi_w, c.ng) > 0.f;
    if (same_geo == transmission) e.pdf = 0.f;
}

// full evaluation of a leaf material in world space: f * |cos wi|, pdf
template <int NS, int MS = MS_ALL>
__device__ __forceinline__ EvalN<NS> mat_eval(const MatCtx<NS> &c, f3 wi_w, float *eta_out) {
    EvalN<NS> e = leaf_eval<NS, MS>(c, wi_w, eta_out);
    geometric_check(c, wi_w, e);
    return e;
}

template <int NS>
struct SampleN {
    EvalN<NS> e;
    f3 wi;
    float eta;
};

// local direction sampling of one lobe; returns validity
template <int NS>
__device__ __forceinline__ bool sample_dielectric_dir(const Ggx &g, f3 wo, const Spec<NS> &eta_s, float eta0,
                                                      bool schlick, const Spec<NS> &F0, const Spec<NS> &kt,
                                                      Sampler &sp, f3 &wi) {
    // DielectricLobe::sample_wi_local_impl (lobe.cpp:452-470): wh 2, uc 1
    f3 wh = g.sample_wh(wo, sp.next2d());
    float d = dot(wo, wh);
    Spec<NS> F;
    if (schlick) {
        F = fresnel_schlick(F0, eta0, fabsf(d));
    } else {
#pragma unroll
        for (int i = 0; i < NS; ++i) F.v[i] = fresnel_dielectric(fabsf(d), eta_s.v[i]);
    }
    float rp = F.average() / ((1.f - F) * kt + F).average();
    float uc = sp.next1d();
    if (uc < rp) {
        wi = reflect_dir(wo, wh);
        return wo.z * wi.z > 0.f;
    }
    bool ok = refract_dir(wo, wh, eta0, wi);
    return ok && !(wo.z * wi.z > 0.f);
}

// direction of lobe k of a leaf material (k: the principled stack's lobe,
// 0 for single-lobe materials), in local coordinates; returns validity
template <int NS, int MS = MS_ALL>
__device__ __forceinline__ bool leaf_sample_dir(const MatCtx<NS> &c, int k, Sampler &sp, f3 &wi) {
    f3 wo = c.wo;
    bool valid = true;
    wi = mk3(0.f, 0.f, 1.f);
    if (ms_has<MS>(MS_DIFFUSE) && (c.type == MAT_DIFFUSE || c.type == MAT_OREN_NAYAR)) {
        wi = square_to_cosine_hemisphere(sp.next2d());
        if (wo.z < 0.f) wi.z = -wi.z;
    } else if (ms_has<MS>(MS_METAL) && (c.type == MAT_METAL || c.type == MAT_MIRROR)) {
        f3 wh = c.ggx.sample_wh(wo, sp.next2d());
        wi = reflect_dir(wo, wh);
        valid = wo.z * wi.z > 0.f;
    } else if (ms_has<MS>(MS_GLASS) && c.type == MAT_GLASS) {
        valid = sample_dielectric_dir(c.ggx, wo, c.a1, c.eta0, false, c.a1, c.a0, sp, wi);
    } else if (ms_has<MS>(MS_PRINCIPLED) && c.type == MAT_PRINCIPLED) {
        if (k == 0) {
            f3 w = square_to_cosine_hemisphere(sp.next2d());
            w = mk3(w.x / c.sh_a - w.z * c.sh_b / c.sh_a, w.y / c.sh_a, w.z);
            wi = normalize(w);
        } else if (k == 1 || k == 2 || k == 4) {
            Ggx g = k == 1 ? Ggx{c.cc_alpha, c.cc_alpha} : c.ggx;
            f3 wh = g.sample_wh(wo, sp.next2d());
            wi = reflect_dir(wo, wh);
            valid = wo.z * wi.z > 0.f;
        } else if (k == 3) {
            valid = sample_dielectric_dir(c.ggx, wo, c.tr_F0, c.tr_eta, true, c.tr_F0, c.tr_kt, sp, wi);
        } else {
            wi = square_to_cosine_hemisphere(sp.next2d());
            if (wo.z < 0.f) wi.z = -wi.z;
        }
    }
    return valid;
}

template <int NS, int MS = MS_ALL>
__device__ __forceinline__ SampleN<NS> mat_sample(const MatCtx<NS> &c, Sampler &sp) {
    SampleN<NS> s;
    s.eta = 1.f;
    if (c.type == MAT_BLACK) {
        // BlackBodyLobe::sample_local: wi = wo, no draws
        s.wi = c.wo_w;
        s.e.f = Spec<NS>::fill(0.f);
        s.e.pdf = 1.f;
        s.e.flags = 0;
        bool same_geo = dot(c.wo_w, c.ng) * dot(s.wi, c.ng) > 0.f;
        if (!same_geo) s.e.pdf = 0.f;
        return s;
    }
    int k = 0;
    if (ms_has<MS>(MS_PRINCIPLED) && c.type == MAT_PRINCIPLED) {
        // LobeSet::sample_wi_impl (lobe.cpp:597-627): uc 1 + an unused 2D draw,
        // then the chosen lobe's draws
        float uc = sp.next1d();
        (void)sp.next2d();
        float sum = 0.f;
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            k = uc > sum ? i : k;
            sum += c.sw[i];
        }
    }
    f3 wi;
    bool valid = leaf_sample_dir<NS, MS>(c, k, sp, wi);
    s.wi = c.fr.to_world(wi);
    s.e = mat_eval<NS, MS>(c, s.wi, &s.eta);
    if (!valid) s.e.pdf = 0.f;
    return s;
}

// ---- mix / add (MixMaterial, AddMaterial; mix.cpp, add.cpp)
// A mix holds only its components' record indices and weights; each use
// rebuilds the component evaluators one at a time, so a kernel's register
// budget is that of one leaf evaluator.
template <int NS, bool RGB>
__device__ __forceinline__ MatCtx<NS> mix_component(const SceneView &sc, const MatCtx<NS> &c, int k,
                                                    const Swl<NS> &swl) {
    bool d;
    return make_material<NS, RGB, MS_ALL>(sc, sc.mat[c.sub[k]], c.fr, c.ng, c.wo_w, swl, d, c.uv);
}

// LobeSet::evaluate_impl over the flattened lobes (lobe.cpp:668-688): each
// component's lobes (already masked by their hemisphere factors) weighted by
// the component's evaluation and sample weights; the dielectric lobes report
// eta in order, the last one's kept
template <int NS, bool RGB>
__device__ __forceinline__ EvalN<NS> mix_eval(const SceneView &sc, const MatCtx<NS> &c, f3 wi_w, float *eta_out,
                                              const Swl<NS> &swl) {
    EvalN<NS> e;
    e.f = Spec<NS>::fill(0.f);
    e.pdf = 0.f;
    e.flags = 0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        MatCtx<NS> m = mix_component<NS, RGB>(sc, c, k, swl);
        leaf_eval_into<NS, MS_ALL>(m, wi_w, eta_out, c.sub_e[k], c.sub_s[k], e);
    }
    geometric_check(c, wi_w, e);
    return e;
}

// LobeSet::sample_wi_impl over the flattened lobes: uc 1 + an unused 2D
// draw pick a lobe by its sample weight (the principled stack's six in
// order, scaled by its component's weight; a single lobe has the
// component's weight), then that lobe's draws, then the whole set is
// evaluated at the sampled direction
template <int NS, bool RGB>
__device__ __forceinline__ SampleN<NS> mix_sample(const SceneView &sc, const MatCtx<NS> &c, Sampler &sp,
                                                  const Swl<NS> &swl) {
    float uc = sp.next1d();
    (void)sp.next2d();
    int pick_c = 0, pick_l = 0;
    float sum = 0.f;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        if (sc.mat[c.sub[k]].type == MAT_PRINCIPLED) {
            MatCtx<NS> m = mix_component<NS, RGB>(sc, c, k, swl);
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                if (uc > sum) {
                    pick_c = k;
                    pick_l = i;
                }
                sum += m.sw[i] * c.sub_s[k];
            }
        } else {
            if (uc > sum) {
                pick_c = k;
                pick_l = 0;
            }
            sum += c.sub_s[k];
        }
    }
    SampleN<NS> s;
    s.eta = 1.f;
    f3 wi;
    bool valid;
    {
        MatCtx<NS> m = mix_component<NS, RGB>(sc, c, pick_c, swl);
        valid = leaf_sample_dir<NS, MS_ALL>(m, pick_l, sp, wi);
    }
    s.wi = c.fr.to_world(wi);
    s.e = mix_eval<NS, RGB>(sc, c, s.wi, &s.eta, swl);
    if (!valid) s.e.pdf = 0.f;
    return s;
}

// MaterialEvaluator::evaluate / ::sample for any material of the set MS
template <int NS, bool RGB, int MS = MS_ALL>
__device__ __forceinline__ EvalN<NS> mat_eval_any(const SceneView &sc, const MatCtx<NS> &c, f3 wi_w,
                                                  float *eta_out, const Swl<NS> &swl) {
    if (MS == MS_ALL && (c.type == MAT_MIX || c.type == MAT_ADD)) return mix_eval<NS, RGB>(sc, c, wi_w, eta_out, swl);
    return mat_eval<NS, MS>(c, wi_w, eta_out);
}
template <int NS, bool RGB, int MS = MS_ALL>
__device__ __forceinline__ SampleN<NS> mat_sample_any(const SceneView &sc, const MatCtx<NS> &c, Sampler &sp,
                                                      const Swl<NS> &swl) {
    if (MS == MS_ALL && (c.type == MAT_MIX || c.type == MAT_ADD)) return mix_sample<NS, RGB>(sc, c, sp, swl);
    return mat_sample<NS, MS>(c, sp);
}

// ---- albedo (MaterialEvaluator::albedo, material.cpp:91-98) at the cosine
// of wo against the shading normal.  Lobe::albedo of each kind:
// MicrofacetLobe kr F(cos) (lobe.cpp:208-210), CoatLobe s kr and
// SpecularLobe (F0 + s (1 - F0)) kr with s from their tables
// (principled_bsdf.cpp:154-160, :198-205), DielectricLobe kt (1 - F(|cos|))
// + F (lobe.cpp:308-313), DiffuseLobe / SheenLTC their R / tint; a lobe set
// adds its lobes' albedos times their evaluation weights (LobeSet::albedo,
// lobe.cpp:564-570).  The black body's albedo is zero.

// the principled stack's lobes added to r, each weight times `scale` (a mix
// component's sample weight when flattened, 1 on its own)
template <int NS>
__device__ __forceinline__ void principled_albedo_into(const SceneView &sc, const MatCtx<NS> &c, float ct,
                                                       float scale, Spec<NS> &r) {
    r = r + c.sh_tint * scale;
    {
        float s;
        lut3<1>(sc.lut + LUT_COAT, 32, sqrtf(sqrtf(c.cc_alpha * c.cc_alpha)), ct, (c.cc_eta - 1.003f) / (4.f - 1.003f),
                &s);
        r = r + (c.cc_kr * s) * scale;
    }
    r = r + (c.mt_kr * fresnel_f82(c.mt_F0, c.mt_B, ct)) * scale;
    {
        Spec<NS> F = fresnel_schlick(c.tr_F0, c.tr_eta, fabsf(ct));
        r = r + (c.tr_kt * (1.f - F) + F) * (c.tr_weight * scale);
    }
    {
        float s;
        lut3<1>(sc.lut + LUT_SPEC, 32, sqrtf(sqrtf(c.ggx.ax * c.ggx.ay)), ct,
                sqrtf(fabsf((c.sp_eta - 1.f) / (c.sp_eta + 1.f))), &s);
        r = r + ((c.sp_F0 + s * (1.f - c.sp_F0)) * c.sp_kr) * scale;
    }
    r = r + c.df_kr * scale;
}

// the albedo of a single-lobe material
template <int NS, int MS>
__device__ __forceinline__ Spec<NS> single_lobe_albedo(const MatCtx<NS> &c, float ct) {
    if (ms_has<MS>(MS_METAL) && c.type == MAT_METAL) {
        Spec<NS> F;
#pragma unroll
        for (int i = 0; i < NS; ++i) F.v[i] = fresnel_conductor(ct, c.a0.v[i], c.a1.v[i]);
        return Spec<NS>::fill(1.f) * F;
    }
    if (ms_has<MS>(MS_GLASS) && c.type == MAT_GLASS) {
        Spec<NS> F;
#pragma unroll
        for (int i = 0; i < NS; ++i) F.v[i] = fresnel_dielectric(fabsf(ct), c.a1.v[i]);
        return c.a0 * (1.f - F) + F;
    }
    if (c.type == MAT_BLACK) return Spec<NS>::fill(0.f);
    return c.a0;// Lambert, Oren-Nayar, mirror (constant Fresnel)
}

template <int NS, bool RGB, int MS = MS_ALL>
__device__ __forceinline__ Spec<NS> mat_albedo(const SceneView &sc, const MatCtx<NS> &c, const Swl<NS> &swl) {
    const float ct = c.wo.z;
    if (MS == MS_ALL && (c.type == MAT_MIX || c.type == MAT_ADD)) {
        Spec<NS> r = Spec<NS>::fill(0.f);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            MatCtx<NS> m = mix_component<NS, RGB>(sc, c, k, swl);
            if (m.type == MAT_PRINCIPLED) principled_albedo_into(sc, m, ct, c.sub_s[k], r);
            else r = r + single_lobe_albedo<NS, MS_ALL>(m, ct) * c.sub_e[k];
        }
        return r;
    }
    if (ms_has<MS>(MS_PRINCIPLED) && c.type == MAT_PRINCIPLED) {
        Spec<NS> r = Spec<NS>::fill(0.f);
        principled_albedo_into(sc, c, ct, 1.f, r);
        return r;
    }
    return single_lobe_albedo<NS, MS>(c, ct);
}

}// namespace dev
}// namespace vs
