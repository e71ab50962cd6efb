// Device spectra: NS-channel sampled spectra, hero-wavelength sampling, SPD
// tables and RGB -> sigmoid-polynomial spectra.
//
// NS = 3 with rgb = true is the reference's sRGB mode (srgb.cpp:10-59: the
// three channels are R, G, B at the sRGB peak wavelengths); NS = 4 with
// rgb = false is hero-wavelength mode (hero.cpp:227-348).
#pragma once

#include "dev_common.h"
#include "dev_scene.h"

namespace vs {
namespace dev {

template <int NS>
struct Spec {
    float v[NS];
    __device__ __forceinline__ float &operator[](int i) { return v[i]; }
    __device__ __forceinline__ float operator[](int i) const { return v[i]; }
    __device__ __forceinline__ static Spec fill(float s) {
        Spec r;
#pragma unroll
        for (int i = 0; i < NS; ++i) r.v[i] = s;
        return r;
    }
    __device__ __forceinline__ float average() const {
        float s = 0.f;
#pragma unroll
        for (int i = 0; i < NS; ++i) s += v[i];
        return s / float(NS);
    }
    __device__ __forceinline__ float maxc() const {
        float m = v[0];
#pragma unroll
        for (int i = 1; i < NS; ++i) m = fmaxf(m, v[i]);
        return m;
    }
};

#define VS_SPEC_BINOP(op)                                                                         \
    template <int NS>                                                                             \
    __device__ __forceinline__ Spec<NS> operator op(const Spec<NS> &a, const Spec<NS> &b) {       \
        Spec<NS> r;                                                                               \
        _Pragma("unroll") for (int i = 0; i < NS; ++i) r.v[i] = a.v[i] op b.v[i];                 \
        return r;                                                                                 \
    }                                                                                             \
    template <int NS>                                                                             \
    __device__ __forceinline__ Spec<NS> operator op(const Spec<NS> &a, float b) {                 \
        Spec<NS> r;                                                                               \
        _Pragma("unroll") for (int i = 0; i < NS; ++i) r.v[i] = a.v[i] op b;                      \
        return r;                                                                                 \
    }                                                                                             \
    template <int NS>                                                                             \
    __device__ __forceinline__ Spec<NS> operator op(float a, const Spec<NS> &b) {                 \
        Spec<NS> r;                                                                               \
        _Pragma("unroll") for (int i = 0; i < NS; ++i) r.v[i] = a op b.v[i];                      \
        return r;                                                                                 \
    }
VS_SPEC_BINOP(+)
VS_SPEC_BINOP(-)
VS_SPEC_BINOP(*)
VS_SPEC_BINOP(/)
#undef VS_SPEC_BINOP

template <int NS>
__device__ __forceinline__ Spec<NS> saturate(const Spec<NS> &a) {
    Spec<NS> r;
#pragma unroll
    for (int i = 0; i < NS; ++i) r.v[i] = fminf(fmaxf(a.v[i], 0.f), 1.f);
    return r;
}

// sampled wavelengths (SampledWavelengths, src/base/color/spectrum.h:18-56)
template <int NS>
struct Swl {
    float lambda[NS];
    float pdf[NS];
};

// SPD::eval (src/base/color/spd.cpp:64-86): piecewise-linear table over
// [360, 830] nm with `interval` = 471 / table size
__device__ __forceinline__ float spd_eval(const float *tab, float interval, float lambda) {
    float t = (fminf(fmaxf(lambda, 360.f), 830.f) - 360.f) / interval;
    uint32_t count = uint32_t((830.f - 360.f) / interval) + 1u;
    uint32_t i = uint32_t(fminf(t, float(count - 2u)));
    float l = tab[i], r = tab[i + 1];
    float f = t - floorf(t);
    return l + f * (r - l);
}

// hero.cpp:15-25 (visible-wavelength importance sampling of pbrt-v4)
__device__ __forceinline__ float sample_visible_wavelength(float u) {
    return 538.f - 138.888889f * em::atanh(0.85691062f - 1.82750197f * u);
}
__device__ __forceinline__ float visible_wavelength_pdf(float lambda) {
    float c = em::cosh(0.0072f * (lambda - 538.f));
    return 0.0039398042f / (c * c);
}

// RGBSigmoidPolynomial (hero.cpp:27-48)
__device__ __forceinline__ float sigmoid_poly(const float *c, float lambda) {
    float x = fmaf(fmaf(c[0], lambda, c[1]), lambda, c[2]);
    if (isinf(x)) return x > 0.f ? 1.f : 0.f;
    return 0.5f * fmaf(x, 1.f / sqrtf(fmaf(x, x, 1.f)), 1.f);
}

}// namespace dev
}// namespace vs
