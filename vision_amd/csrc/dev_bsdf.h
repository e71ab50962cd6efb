// Device BSDFs: Lambert / Oren-Nayar, GGX conductor ("metal"), rough
// dielectric ("glass") and the principled lobe stack, each restating the
// reference's lobe as cited; evaluation returns f * |cos wi| and the pdf.
#pragma once

#include "dev_common.h"
#include "dev_spectrum.h"

namespace vs {
namespace dev {

// offsets (in floats) of the energy-compensation tables in SceneView::lut
constexpr uint32_t LUT_DIEL = 0;         // float2[32][32][32]
constexpr uint32_t LUT_DIEL_INV = 65536; // float2[32][32][32]
constexpr uint32_t LUT_SPEC = 131072;    // float[32][32][32]
constexpr uint32_t LUT_COAT = 163840;    // float[32][32][32]
constexpr uint32_t LUT_PURE = 196608;    // float[32][32]
constexpr uint32_t LUT_SHEEN = 197632;   // float4[32][32] (approximate mode)
constexpr uint32_t LUT_TOTAL = 201728;

// linear filtering with clamp-to-edge addressing of a table whose texel i
// sits at (i + 0.5) / res
__device__ __forceinline__ void lut_coord(float u, int res, int &i0, int &i1, float &f) {
    float x = fminf(fmaxf(u * float(res) - 0.5f, 0.f), float(res - 1));
    float fl = floorf(x);
    i0 = int(fl);
    i1 = min(i0 + 1, res - 1);
    f = x - fl;
}
template <int C>
__device__ __forceinline__ void lut2(const float *t, int res, float u, float v, float *out) {
    int x0, x1, y0, y1;
    float fx, fy;
    lut_coord(u, res, x0, x1, fx);
    lut_coord(v, res, y0, y1, fy);
#pragma unroll
    for (int c = 0; c < C; ++c) {
        float a = t[(y0 * res + x0) * C + c], b = t[(y0 * res + x1) * C + c];
        float d = t[(y1 * res + x0) * C + c], e = t[(y1 * res + x1) * C + c];
        float r0 = a + fx * (b - a), r1 = d + fx * (e - d);
        out[c] = r0 + fy * (r1 - r0);
    }
}
template <int C>
__device__ __forceinline__ void lut3(const float *t, int res, float u, float v, float w, float *out) {
    int x0, x1, y0, y1, z0, z1;
    float fx, fy, fz;
    lut_coord(u, res, x0, x1, fx);
    lut_coord(v, res, y0, y1, fy);
    lut_coord(w, res, z0, z1, fz);
#pragma unroll
    for (int c = 0; c < C; ++c) {
        auto at = [&](int x, int y, int z) { return t[((z * res + y) * res + x) * C + c]; };
        float c00 = at(x0, y0, z0) + fx * (at(x1, y0, z0) - at(x0, y0, z0));
        float c10 = at(x0, y1, z0) + fx * (at(x1, y1, z0) - at(x0, y1, z0));
        float c01 = at(x0, y0, z1) + fx * (at(x1, y0, z1) - at(x0, y0, z1));
        float c11 = at(x0, y1, z1) + fx * (at(x1, y1, z1) - at(x0, y1, z1));
        float c0 = c00 + fy * (c10 - c00), c1 = c01 + fy * (c11 - c01);
        out[c] = c0 + fz * (c1 - c0);
    }
}

// ---- GGX (src/base/scattering/microfacet.cpp:11-174), visible-normal sampling
struct Ggx {
    float ax, ay;
    __device__ __forceinline__ float D(f3 wh) const {
        f3 H = mk3(wh.x / ax, wh.y / ay, wh.z);
        float l2 = dot(H, H);
        return kInvPi / (ax * ay * (l2 * l2));
    }
    __device__ __forceinline__ float lambda(f3 w) const {
        float a = ax * w.x, b = ay * w.y;
        float r = (a * a + b * b) / (w.z * w.z);
        return w.z == 0.f ? 0.f : 0.5f * (sqrtf(1.f + r) - 1.f);
    }
    __device__ __forceinline__ float G1(f3 w) const { return 1.f / (1.f + lambda(w)); }
    __device__ __forceinline__ float G(f3 wo, f3 wi) const { return 1.f / (1.f + lambda(wo) + lambda(wi)); }
    __device__ __forceinline__ f3 sample_wh(f3 wo, f2 u) const {
        bool flip = wo.z < 0.f;
        f3 Ve = flip ? -wo : wo;
        f3 Vh = normalize(mk3(ax * Ve.x, ay * Ve.y, Ve.z));
        float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
        f3 T1 = lensq > 1e-7f ? mk3(-Vh.y, Vh.x, 0.f) / sqrtf(lensq) : mk3(1.f, 0.f, 0.f);
        f3 T2 = lensq > 1e-7f ? cross(Vh, T1) : mk3(0.f, 1.f, 0.f);
        f2 t = square_to_disk(u);
        float s = 0.5f * (1.f + Vh.z);
        float a = sqrtf(fmaxf(0.f, 1.f - t.x * t.x));
        t.y = a + s * (t.y - a);
        f3 Nh = T1 * t.x + T2 * t.y + Vh * sqrtf(fmaxf(0.f, 1.f - (t.x * t.x + t.y * t.y)));
        f3 Ne = normalize(mk3(ax * Nh.x, ay * Nh.y, fmaxf(0.f, Nh.z)));
        return flip ? -Ne : Ne;
    }
    __device__ __forceinline__ float pdf_wh(f3 wo, f3 wh) const {
        return D(wh) * G1(wo) * fabsf(dot(wo, wh)) / fabsf(wo.z);
    }
    __device__ __forceinline__ float pdf_reflection(f3 wo, f3 wh) const {
        return pdf_wh(wo, wh) / (4.f * fabsf(dot(wo, wh)));
    }
    __device__ __forceinline__ float pdf_transmission(f3 wo, f3 wh, f3 wi, float eta) const {
        float d = dot(wi, wh) * eta + dot(wo, wh);
        return pdf_wh(wo, wh) * (fabsf(dot(wi, wh)) / (d * d));
    }
    __device__ __forceinline__ float brdf(f3 wo, f3 wh, f3 wi) const {
        return D(wh) * G(wo, wi) / fabsf(4.f * wo.z * wi.z);
    }
    // BTDF_div_ft in radiance transport (microfacet.cpp:160-174)
    __device__ __forceinline__ float btdf(f3 wo, f3 wh, f3 wi, float eta) const {
        float num = D(wh) * G(wo, wi) * fabsf(dot(wi, wh) * dot(wo, wh));
        float d = dot(wi, wh) * eta + dot(wo, wh);
        float denom = (d * d) * fabsf(wi.z * wo.z);
        float ft = denom == 0.f ? 0.f : num / denom;
        return ft * (1.f / (eta * eta));
    }
};

__device__ __forceinline__ f3 face_forward(f3 v, f3 n) { return dot(v, n) < 0.f ? -v : v; }
__device__ __forceinline__ f3 reflect_dir(f3 wo, f3 n) { return -wo + n * (2.f * dot(wo, n)); }
// optics.h:27-39
__device__ __forceinline__ bool refract_dir(f3 wi, f3 n, float eta, f3 &wt) {
    float cos_i = dot(n, wi);
    float sin2_i = fmaxf(0.f, 1.f - cos_i * cos_i);
    float sin2_t = sin2_i / (eta * eta);
    float cos_t = sqrtf(fmaxf(0.f, 1.f - sin2_t));
    wt = -wi / eta + n * (cos_i / eta - cos_t);
    return sin2_t < 1.f;
}
// optics.h:70-79
__device__ __forceinline__ float fresnel_dielectric(float abs_cos_i, float eta) {
    float sin2_i = 1.f - abs_cos_i * abs_cos_i;
    float sin2_t = sin2_i / (eta * eta);
    float cos_t = sqrtf(fmaxf(0.f, 1.f - sin2_t));
    float r_parl = (eta * abs_cos_i - cos_t) / (eta * abs_cos_i + cos_t);
    float r_perp = (abs_cos_i - eta * cos_t) / (abs_cos_i + eta * cos_t);
    return sin2_t >= 1.f ? 1.f : (r_parl * r_parl + r_perp * r_perp) * 0.5f;
}

// complex Fresnel of a conductor (optics.h:92-115) in explicit complex arithmetic
struct Cpx {
    float re, im;
};
__device__ __forceinline__ Cpx cadd(Cpx a, Cpx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cpx csub(Cpx a, Cpx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ Cpx cmul(Cpx a, Cpx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ Cpx cdiv(Cpx a, Cpx b) {
    float s = 1.f / (b.re * b.re + b.im * b.im);
    return {(a.re * b.re + a.im * b.im) * s, (a.im * b.re - a.re * b.im) * s};
}
__device__ __forceinline__ float cnorm(Cpx a) { return a.re * a.re + a.im * a.im; }
// principal square root (complex_sqrt, src/math/complex.h:53-61)
__device__ __forceinline__ Cpx csqrt(Cpx z) {
    float n = sqrtf(cnorm(z));
    if (n == 0.f) return {0.f, 0.f};
    float t1 = sqrtf(0.5f * (n + fabsf(z.re)));
    float t2 = 0.5f * z.im / t1;
    if (z.re >= 0.f) return {t1, t2};
    return {fabsf(t2), copysignf(t1, z.im)};
}
__device__ __forceinline__ float fresnel_conductor(float cos_i, float eta, float k) {
    Cpx e{eta, k};
    float sin2 = 1.f - cos_i * cos_i;
    Cpx sin2_t = cdiv({sin2, 0.f}, cmul(e, e));
    Cpx cos_t = csqrt(csub({1.f, 0.f}, sin2_t));
    Cpx ec = {e.re * cos_i, e.im * cos_i};
    Cpx r_parl = cdiv(csub(ec, cos_t), cadd(ec, cos_t));
    Cpx ect = cmul(e, cos_t);
    Cpx r_perp = cdiv(csub({cos_i, 0.f}, ect), cadd({cos_i, 0.f}, ect));
    return (cnorm(r_parl) + cnorm(r_perp)) * 0.5f;
}

__device__ __forceinline__ float schlick_f0_from_ior(float ior) {
    float r = (ior - 1.f) / (ior + 1.f);
    return r * r;
}

// FresnelSchlick::evaluate (fresnel.h:56-63)
template <int NS>
__device__ __forceinline__ Spec<NS> fresnel_schlick(const Spec<NS> &F0, float eta, float cos_t) {
    float F_real = fresnel_dielectric(cos_t, eta);
    float F0_real = schlick_f0_from_ior(eta);
    float t = fminf(fmaxf((F_real - F0_real) / (1.f - F0_real), 0.f), 1.f);
    return F0 + t * (1.f - F0);
}

// FresnelF82Tint (fresnel.h:98-130)
template <int NS>
__device__ __forceinline__ Spec<NS> fresnel_f82(const Spec<NS> &F0, const Spec<NS> &B, float cos_t) {
    float mu = fminf(fmaxf(1.f - cos_t, 0.f), 1.f);
    float mu2 = mu * mu;
    float mu5 = mu2 * mu2 * mu;
    Spec<NS> fs = F0 + mu5 * (1.f - F0);
    return saturate(fs - ((B * cos_t) * mu5) * mu);
}

}// namespace dev
}// namespace vs
