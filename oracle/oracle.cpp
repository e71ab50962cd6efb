// CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// A scalar C++ restatement of the reference's megakernel path-tracing path
// (FaithZL/Vision), used by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg as the checker.  The product (libvision_amd.so) never
// links, loads or calls this code.
//
// Every function cites the reference file:line it restates.  Pinning (see
// tests/test_oracle.py and tests/test_oracle_spectral.py):
//  - TEA against the published TEA known answer (32 rounds), the LCG against
//    its recurrence's first outputs;
//  - warps, alias tables and every lobe's sampling by chi-square / exact
//    reconstruction against their pdfs;
//  - BVH2 traversal by brute force; the Li loop by the furnace closed form
//    L = Le / (1 - albedo), point/spot/area single-bounce closed forms;
//  - GGX by the normalisation of D (int D cos = 1) and of the visible-normal
//    pdf, the weak white furnace (int D G1 <wo,wh>/cos = 1), Fresnel by its
//    normal-incidence and grazing closed forms and the conductor/dielectric
//    limit (k = 0 equals the dielectric formula);
//  - energy compensation: a white conductor with its LUT factor reflects
//    close to 1 at every roughness;
//  - hero wavelengths: the wavelength pdf integrates to 1 over 360-830 nm
//    and matches the histogram of its sampler, the CIE Y integral equals its
//    published value, a constant spectrum converts to the white point;
//  - RGB -> spectrum: the decoded sigmoid spectrum integrated against the
//    CIE curves reproduces the input colour.
// Functions only an import of the absent ocarina library could pin (ray
// origin offset, frame orientation, texture filtering of the LUTs, the
// lerp argument order) follow the published methods named inline; for those
// the oracle defines the behaviour: parity unpinned.
//
// Built with -ffp-contract=off so the watertight triangle test matches the
// HIP kernel bit for bit on identical rays.

#include "oracle_core.h"

namespace {

// ---------------------------------------------------------------- spectral data
constexpr int kMaxNS = 4;

// Tables the spectral path reads: the SPD buffer exported with the scene
// (CIE X, Y, Z, D65 at 94 samples each, then material curves), the energy
// compensation tables (src/base/scattering/precomputed_table.h) and the
// sRGB -> sigmoid coefficient table [3][64][64][64][3].
struct Spectral {
    bool rgb{true};
    int ns{3};
    const float *spd{nullptr};
    float cie_interval{471.f / 94.f};
    float cie_y_integral{1.f};
    const float *lut_diel{nullptr}, *lut_diel_inv{nullptr}, *lut_spec{nullptr}, *lut_coat{nullptr};
    const float *lut_pure{nullptr}, *lut_sheen{nullptr};
    const float *rgb2spec{nullptr};
    // images of textured colour slots (library export): float4 texels and
    // {first texel, width, height, 0} per texture
    const float *tex{nullptr};
    const uint32_t *tex_info{nullptr};
    uint32_t ntex{0};
};

struct Spec {
    float v[kMaxNS]{0.f, 0.f, 0.f, 0.f};
};
inline Spec sfill(float x) {
    Spec s;
    for (float &v : s.v) v = x;
    return s;
}
inline Spec operator+(const Spec &a, const Spec &b) {
    Spec r;
    for (int i = 0; i < kMaxNS; ++i) r.v[i] = a.v[i] + b.v[i];
    return r;
}
inline Spec operator-(const Spec &a, const Spec &b) {
    Spec r;
    for (int i = 0; i < kMaxNS; ++i) r.v[i] = a.v[i] - b.v[i];
    return r;
}
inline Spec operator*(const Spec &a, const Spec &b) {
    Spec r;
    for (int i = 0; i < kMaxNS; ++i) r.v[i] = a.v[i] * b.v[i];
    return r;
}
inline Spec operator*(const Spec &a, float b) {
    Spec r;
    for (int i = 0; i < kMaxNS; ++i) r.v[i] = a.v[i] * b;
    return r;
}
inline Spec operator*(float a, const Spec &b) {
    Spec r;
    for (int i = 0; i < kMaxNS; ++i) r.v[i] = a * b.v[i];
    return r;
}
inline Spec operator/(const Spec &a, float b) {
    Spec r;
    for (int i = 0; i < kMaxNS; ++i) r.v[i] = a.v[i] / b;
    return r;
}
inline Spec operator-(float a, const Spec &b) {
    Spec r;
    for (int i = 0; i < kMaxNS; ++i) r.v[i] = a - b.v[i];
    return r;
}
// SampledSpectrum::average / max over the sampled dimension (spectrum.h:58-200)
inline float savg(const Spec &s, int n) {
    float a = 0.f;
    for (int i = 0; i < n; ++i) a += s.v[i];
    return a / float(n);
}
inline float smax(const Spec &s, int n) {
    float m = s.v[0];
    for (int i = 1; i < n; ++i) m = std::max(m, s.v[i]);
    return m;
}

struct Swl {
    int n{3};
    float lambda[kMaxNS]{};
    float pdf[kMaxNS]{};
};

// SPD::eval (src/base/color/spd.cpp:60-70): table over 360-830 nm with the
// sample spacing 471 / size; fract-weighted interpolation of neighbours
float spd_eval(const float *tab, float interval, float lambda) {
    float t = (std::min(std::max(lambda, 360.f), 830.f) - 360.f) / interval;
    uint32_t count = uint32_t((830.f - 360.f) / interval) + 1u;
    uint32_t i = uint32_t(std::min(t, float(count - 2u)));
    float l = tab[i], r = tab[i + 1];
    float f = t - std::floor(t);
    return l + f * (r - l);
}

// hero.cpp:15-25
float sample_visible_wavelength(float u) { return 538.f - 138.888889f * vs::em::atanh(0.85691062f - 1.82750197f * u); }
float visible_wavelength_pdf(float lambda) {
    float c = vs::em::cosh(0.0072f * (lambda - 538.f));
    return 0.0039398042f / (c * c);
}

// RGBSigmoidPolynomial (hero.cpp:27-48)
float sigmoid_poly(const float *c, float lambda) {
    float x = std::fma(std::fma(c[0], lambda, c[1]), lambda, c[2]);
    if (std::isinf(x)) return x > 0.f ? 1.f : 0.f;
    return 0.5f * std::fma(x, 1.f / std::sqrt(std::fma(x, x, 1.f)), 1.f);
}

float srgb_to_y(V3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }

// RGBToSpectrumTable::decode_albedo / decode_unbound, host form (hero.cpp:91-139)
float inverse_smooth_step(float x) { return 0.5f - std::sin(std::asin(1.0f - 2.0f * x) * (1.0f / 3.0f)); }
void decode_albedo(const float *table, V3 rgb_in, float out[4]) {
    constexpr uint32_t res = 64;
    float rgb[3] = {std::min(std::max(rgb_in.x, 0.f), 1.f), std::min(std::max(rgb_in.y, 0.f), 1.f),
                    std::min(std::max(rgb_in.z, 0.f), 1.f)};
    float lum = srgb_to_y(v3(rgb[0], rgb[1], rgb[2]));
    if (rgb[0] == rgb[1] && rgb[1] == rgb[2]) {
        out[0] = 0.f;
        out[1] = 0.f;
        out[2] = (rgb[0] - 0.5f) / std::sqrt(rgb[0] * (1.0f - rgb[0]));
        out[3] = lum;
        return;
    }
    uint32_t maxc = (rgb[0] > rgb[1]) ? ((rgb[0] > rgb[2]) ? 0u : 2u) : ((rgb[1] > rgb[2]) ? 1u : 2u);
    float z = rgb[maxc];
    float x = rgb[(maxc + 1u) % 3u] * float(res - 1u) / z;
    float y = rgb[(maxc + 2u) % 3u] * float(res - 1u) / z;
    float zz = inverse_smooth_step(inverse_smooth_step(z)) * float(res - 1u);
    uint32_t xi = std::min(uint32_t(x), res - 2u), yi = std::min(uint32_t(y), res - 2u);
    uint32_t zi = std::min(uint32_t(zz), res - 2u);
    float dx = x - float(xi), dy = y - float(yi), dz = zz - float(zi);
    auto co = [&](uint32_t ox, uint32_t oy, uint32_t oz, int i) {
        size_t idx = (((size_t(maxc) * res + (zi + oz)) * res + (yi + oy)) * res + (xi + ox)) * 3 + i;
        return table[idx];
    };
    auto lerp = [](float t, float a, float b) { return a + t * (b - a); };
    for (int i = 0; i < 3; ++i) {
        out[i] = lerp(dz, lerp(dy, lerp(dx, co(0, 0, 0, i), co(1, 0, 0, i)), lerp(dx, co(0, 1, 0, i), co(1, 1, 0, i))),
                      lerp(dy, lerp(dx, co(0, 0, 1, i), co(1, 0, 1, i)), lerp(dx, co(0, 1, 1, i), co(1, 1, 1, i))));
    }
    out[3] = lum;
}
void decode_unbound(const float *table, V3 rgb_in, float out[4]) {
    V3 rgb = v3(std::max(rgb_in.x, 0.f), std::max(rgb_in.y, 0.f), std::max(rgb_in.z, 0.f));
    float m = std::max(rgb.x, std::max(rgb.y, rgb.z));
    float scale = 2.f * m;
    V3 a = scale == 0.f ? v3(0.f, 0.f, 0.f) : v3(rgb.x / scale, rgb.y / scale, rgb.z / scale);
    decode_albedo(table, a, out);
    out[3] = scale;
}

// albedo of a colour slot (ShaderNodeSlot::eval_albedo_spectrum): linear
// sRGB in sRGB mode, the sigmoid spectrum of its coefficients in hero mode
Spec albedo_slot(const Spectral &sd, const float *slot, const Swl &w) {
    Spec r;
    if (sd.rgb) {
        for (int i = 0; i < 3; ++i) r.v[i] = slot[i];
    } else {
        for (int i = 0; i < w.n; ++i) r.v[i] = sigmoid_poly(slot, w.lambda[i]);
    }
    return r;
}

// a colour slot that may name an image (ImageNode::evaluate,
// src/render_core/shadernodes/image.cpp:84-95): flag bit 16 + off / 4 marks
// it, p[off] holds the texture index; the texel is the nearest one at uv
// wrapped to [0, 1) with v = 0 the first row (the filtering is ocarina's,
// absent here: parity unpinned)
Spec albedo_at(const Spectral &sd, uint32_t flags, const float *p, int off, float u, float v, const Swl &w) {
    if (flags & (1u << (16 + off / 4))) {
        uint32_t t;
        std::memcpy(&t, p + off, 4);
        if (t >= sd.ntex) return sfill(0.f);
        const uint32_t *info = sd.tex_info + 4 * t;
        float uu = u - std::floor(u), vv = v - std::floor(v);
        uint32_t x = std::min(uint32_t(uu * float(info[1])), info[1] - 1u);
        uint32_t y = std::min(uint32_t(vv * float(info[2])), info[2] - 1u);
        const float *q = sd.tex + 4 * (size_t(info[0]) + size_t(y) * info[1] + x);
        return albedo_slot(sd, q, w);
    }
    return albedo_slot(sd, p + off, w);
}

// RGBIlluminationSpectrum (hero.cpp:171-186) times a light scale; `coef`
// holds decode_unbound of the colour
Spec illumination(const Spectral &sd, const float *rgb, const float *coef, float scale, const Swl &w) {
    Spec r;
    if (sd.rgb) {
        for (int i = 0; i < 3; ++i) r.v[i] = rgb[i] * scale;
    } else {
        for (int i = 0; i < w.n; ++i)
            r.v[i] = sigmoid_poly(coef, w.lambda[i]) * coef[3] * spd_eval(sd.spd + 282, sd.cie_interval, w.lambda[i]) * scale;
    }
    return r;
}

// Spectrum::linear_srgb: identity in sRGB mode (srgb.cpp:47-49); hero mode:
// HeroWavelengthSpectrum::cie_xyz (hero.cpp:262-274) then cie.h:414-420
V3 linear_srgb(const Spectral &sd, const Spec &s, const Swl &w) {
    if (sd.rgb) return v3(s.v[0], s.v[1], s.v[2]);
    float X = 0.f, Y = 0.f, Z = 0.f;
    uint32_t valid = 0;
    for (int i = 0; i < w.n; ++i) {
        float l = w.lambda[i], p = w.pdf[i];
        valid += p > 0.f ? 1u : 0u;
        float x = spd_eval(sd.spd, sd.cie_interval, l) * s.v[i];
        float y = spd_eval(sd.spd + 94, sd.cie_interval, l) * s.v[i];
        float z = spd_eval(sd.spd + 188, sd.cie_interval, l) * s.v[i];
        X += p == 0.f ? 0.f : x / p;
        Y += p == 0.f ? 0.f : y / p;
        Z += p == 0.f ? 0.f : z / p;
    }
    float factor = 1.f / (float(valid) * sd.cie_y_integral);
    X *= factor;
    Y *= factor;
    Z *= factor;
    return v3(3.240479f * X + -1.537150f * Y + -0.498535f * Z, -0.969256f * X + 1.875991f * Y + 0.041556f * Z,
              0.055648f * X + -0.204043f * Y + 1.057311f * Z);
}

// RenderEnv::initial (src/base/integrator.cpp:49-58) with
// HeroWavelengthSpectrum::sample_wavelength (hero.cpp:276-290): the
// wavelengths come from the stream (pixel, frame, 0xFFFFFFFF)
Swl sample_wavelengths(const Spectral &sd, uint32_t px, uint32_t py, uint32_t frame) {
    Swl w;
    w.n = sd.ns;
    if (sd.rgb) {
        const float peaks[3] = {602.785f, 539.285f, 445.772f};
        for (int i = 0; i < 3; ++i) {
            w.lambda[i] = peaks[i];
            w.pdf[i] = 1.f;
        }
        return w;
    }
    Sampler t;
    t.start(px, py, frame, 0xFFFFFFFFu);
    float u = t.next1d();
    for (int i = 0; i < w.n; ++i) {
        float offset = float(i) * (1.f / float(w.n));
        float up = u + offset;
        up = up - std::floor(up);
        float l = sample_visible_wavelength(up);
        w.lambda[i] = l;
        w.pdf[i] = visible_wavelength_pdf(l);
    }
    return w;
}

// ---------------------------------------------------------------- LUTs
// MaterialLut::sample: a texture fetch with linear filtering and
// clamp-to-edge addressing, texel i centred at (i + 0.5) / res
void lut_axis(float u, int res, int &i0, int &i1, float &f) {
    float x = std::min(std::max(u * float(res) - 0.5f, 0.f), float(res - 1));
    float fl = std::floor(x);
    i0 = int(fl);
    i1 = std::min(i0 + 1, res - 1);
    f = x - fl;
}
void lut2(const float *t, int res, int C, float u, float v, float *out) {
    int x0, x1, y0, y1;
    float fx, fy;
    lut_axis(u, res, x0, x1, fx);
    lut_axis(v, res, y0, y1, fy);
    for (int c = 0; c < C; ++c) {
        float a = t[(y0 * res + x0) * C + c], b = t[(y0 * res + x1) * C + c];
        float d = t[(y1 * res + x0) * C + c], e = t[(y1 * res + x1) * C + c];
        float r0 = a + fx * (b - a), r1 = d + fx * (e - d);
        out[c] = r0 + fy * (r1 - r0);
    }
}
void lut3(const float *t, int res, int C, float u, float v, float w, float *out) {
    int x0, x1, y0, y1, z0, z1;
    float fx, fy, fz;
    lut_axis(u, res, x0, x1, fx);
    lut_axis(v, res, y0, y1, fy);
    lut_axis(w, res, z0, z1, fz);
    for (int c = 0; c < C; ++c) {
        auto at3 = [&](int x, int y, int z) { return t[((z * res + y) * res + x) * C + c]; };
        float c00 = at3(x0, y0, z0) + fx * (at3(x1, y0, z0) - at3(x0, y0, z0));
        float c10 = at3(x0, y1, z0) + fx * (at3(x1, y1, z0) - at3(x0, y1, z0));
        float c01 = at3(x0, y0, z1) + fx * (at3(x1, y0, z1) - at3(x0, y0, z1));
        float c11 = at3(x0, y1, z1) + fx * (at3(x1, y1, z1) - at3(x0, y1, z1));
        float c0 = c00 + fy * (c10 - c00), c1 = c01 + fy * (c11 - c01);
        out[c] = c0 + fz * (c1 - c0);
    }
}

// ---------------------------------------------------------------- microfacet
// GGX with visible-normal sampling (src/base/scattering/microfacet.cpp:11-174,
// microfacet.h:128-187)
struct Ggx {
    float ax{1.f}, ay{1.f};
    float D(V3 wh) const {
        V3 H = v3(wh.x / ax, wh.y / ay, wh.z / 1.f);
        float l2 = dotp(H, H);
        float alpha2 = ax * ay;
        return kInvPi / (alpha2 * (l2 * l2));
    }
    float lambda(V3 w) const {
        float a = ax * w.x, b = ay * w.y;
        float r = (a * a + b * b) / (w.z * w.z);
        float ret = 0.5f * (std::sqrt(1.0f + r) - 1.0f);
        return w.z == 0.f ? 0.f : ret;
    }
    float G1(V3 w) const { return 1.f / (1.f + lambda(w)); }
    float G(V3 wo, V3 wi) const { return 1.f / (1.f + lambda(wo) + lambda(wi)); }
    // sample_GGX_VNDF (microfacet.cpp:66-86), Heitz 2018
    V3 sample_vndf(V3 Ve, float u0, float u1) const {
        V3 Vh = unit(v3(ax * Ve.x, ay * Ve.y, Ve.z));
        float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
        V3 T1 = v3(1.f, 0.f, 0.f), T2 = v3(0.f, 1.f, 0.f);
        if (lensq > 1e-7f) {
            float s = std::sqrt(lensq);
            T1 = v3(-Vh.y / s, Vh.x / s, 0.f / s);
            T2 = crossp(Vh, T1);
        }
        float tx, ty;
        square_to_disk(u0, u1, tx, ty);
        float a = std::sqrt(std::max(0.f, 1.0f - tx * tx));
        float s = 0.5f * (1.0f + Vh.z);
        ty = a + s * (ty - a);
        V3 Nh = add(add(mul(T1, tx), mul(T2, ty)), mul(Vh, std::sqrt(std::max(0.f, 1.0f - (tx * tx + ty * ty)))));
        return unit(v3(ax * Nh.x, ay * Nh.y, std::max(0.0f, Nh.z)));
    }
    // sample_wh (microfacet.cpp:88-104)
    V3 sample_wh(V3 wo, float u0, float u1) const {
        bool flip = wo.z < 0.f;
        V3 w = flip ? neg(wo) : wo;
        V3 wh = sample_vndf(w, u0, u1);
        return flip ? neg(wh) : wh;
    }
    float pdf_wh(V3 wo, V3 wh) const { return D(wh) * G1(wo) * std::fabs(dotp(wo, wh)) / std::fabs(wo.z); }
    float pdf_reflection(V3 wo, V3 wh) const { return pdf_wh(wo, wh) / (4.f * std::fabs(dotp(wo, wh))); }
    float pdf_transmission(V3 wo, V3 wh, V3 wi, float eta) const {
        float pwh = pdf_wh(wo, wh);
        float d = dotp(wi, wh) * eta + dotp(wo, wh);
        float denom = d * d;
        float dwh_dwi = std::fabs(dotp(wi, wh)) / denom;
        return pwh * dwh_dwi;
    }
    float brdf(V3 wo, V3 wh, V3 wi) const { return D(wh) * G(wo, wi) / std::fabs(4.f * wo.z * wi.z); }
    // BTDF_div_ft in radiance transport (microfacet.cpp:160-174)
    float btdf(V3 wo, V3 wh, V3 wi, float eta) const {
        float numerator = D(wh) * G(wo, wi) * std::fabs(dotp(wi, wh) * dotp(wo, wh));
        float d = dotp(wi, wh) * eta + dotp(wo, wh);
        float denom = (d * d) * std::fabs(wi.z * wo.z);
        float ft = numerator / denom;
        float factor = 1.f / (eta * eta);
        ft = denom == 0.f ? 0.f : ft;
        return ft * factor;
    }
};

// ---------------------------------------------------------------- optics (src/math/optics.h)
V3 face_forward(V3 v, V3 n) { return dotp(v, n) < 0.f ? neg(v) : v; }
V3 reflect_dir(V3 wo, V3 n) { return add(neg(wo), mul(n, 2.f * dotp(wo, n))); }
bool refract_dir(V3 wi, V3 n, float eta, V3 &wt) {
    float cos_i = dotp(n, wi);
    float sin2_i = std::max(0.f, 1.f - cos_i * cos_i);
    float sin2_t = sin2_i / (eta * eta);
    bool valid = sin2_t < 1.f;
    float cos_t = std::sqrt(std::max(0.f, 1.f - sin2_t));
    wt = add(v3(-wi.x / eta, -wi.y / eta, -wi.z / eta), mul(n, cos_i / eta - cos_t));
    return valid;
}
float fresnel_dielectric(float abs_cos_i, float eta) {
    float sin2_i = 1.f - abs_cos_i * abs_cos_i;
    float sin2_t = sin2_i / (eta * eta);
    float cos_t = std::sqrt(std::max(0.f, 1.f - sin2_t));
    float r_parl = (eta * abs_cos_i - cos_t) / (eta * abs_cos_i + cos_t);
    float r_perp = (abs_cos_i - eta * cos_t) / (abs_cos_i + eta * cos_t);
    return sin2_t >= 1.f ? 1.f : (r_parl * r_parl + r_perp * r_perp) * 0.5f;
}
// Complex arithmetic of src/math/complex.h:10-62
struct Cx {
    float re, im;
};
Cx cx_add(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
Cx cx_sub(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
Cx cx_mul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
Cx cx_div(Cx a, Cx z) {
    float scale = 1.f / (z.re * z.re + z.im * z.im);
    return {scale * (a.re * z.re + a.im * z.im), scale * (a.im * z.re - a.re * z.im)};
}
float cx_norm2(Cx z) { return z.re * z.re + z.im * z.im; }
Cx cx_sqrt(Cx z) {
    float n = std::sqrt(cx_norm2(z));
    if (n == 0.f) return {0.f, 0.f};
    float t1 = std::sqrt(0.5f * (n + std::fabs(z.re)));
    float t2 = 0.5f * z.im / t1;
    if (z.re >= 0.f) return {t1, t2};
    return {std::fabs(t2), std::copysign(t1, z.im)};
}
// fresnel_complex (optics.h:92-115)
float fresnel_complex(float cos_i, float eta_re, float k) {
    Cx eta{eta_re, k};
    float sin2_i = 1.f - cos_i * cos_i;
    Cx sin2_t = cx_div({sin2_i, 0.f}, cx_mul(eta, eta));
    Cx cos_t = cx_sqrt(cx_sub({1.f, 0.f}, sin2_t));
    Cx ec = cx_mul({cos_i, 0.f}, eta);
    Cx r_parl = cx_div(cx_sub(ec, cos_t), cx_add(ec, cos_t));
    Cx ect = cx_mul(eta, cos_t);
    Cx r_perp = cx_div(cx_sub({cos_i, 0.f}, ect), cx_add({cos_i, 0.f}, ect));
    return (cx_norm2(r_parl) + cx_norm2(r_perp)) * .5f;
}
float schlick_f0_from_ior(float ior) {
    float r = (ior - 1.0f) / (ior + 1.0f);
    return r * r;
}

// Fresnel kinds of src/base/scattering/fresnel.h and metal.cpp:9-21
enum FresnelKind { FR_CONDUCTOR, FR_DIELECTRIC, FR_SCHLICK, FR_F82, FR_ONE };
struct Fresnel {
    int kind{FR_DIELECTRIC};
    Spec eta, k;// conductor eta / k; dielectric and Schlick: eta per wavelength
    Spec F0, B; // Schlick F0; F82-tint F0 and B
};
Spec fresnel_eval(const Fresnel &fr, float c, int n) {
    Spec r;
    switch (fr.kind) {
        case FR_ONE:// FresnelConstant (fresnel.h:132-137)
            return sfill(1.f);
        case FR_CONDUCTOR:
            for (int i = 0; i < n; ++i) r.v[i] = fresnel_complex(c, fr.eta.v[i], fr.k.v[i]);
            return r;
        case FR_DIELECTRIC:
            for (int i = 0; i < n; ++i) r.v[i] = fresnel_dielectric(c, fr.eta.v[i]);
            return r;
        case FR_SCHLICK: {// FresnelSchlick::evaluate (fresnel.h:56-63)
            float F_real = fresnel_dielectric(c, fr.eta.v[0]);
            float F0_real = schlick_f0_from_ior(fr.eta.v[0]);
            float t = (F_real - F0_real) / (1.f - F0_real);
            t = std::min(std::max(t, 0.f), 1.f);
            for (int i = 0; i < n; ++i) r.v[i] = fr.F0.v[i] + t * (1.f - fr.F0.v[i]);
            return r;
        }
        default: {// FresnelF82Tint::evaluate (fresnel.h:121-127)
            float mu = std::min(std::max(1.f - c, 0.f), 1.f);
            float mu2 = mu * mu;
            float mu5 = mu2 * mu2 * mu;
            for (int i = 0; i < n; ++i) {
                float fs = fr.F0.v[i] + mu5 * (1.f - fr.F0.v[i]);
                float v = fs - fr.B.v[i] * c * mu5 * mu;
                r.v[i] = std::min(std::max(v, 0.f), 1.f);
            }
            return r;
        }
    }
}

// ---------------------------------------------------------------- lobes
// One lobe of a material's lobe set (src/base/scattering/lobe.h): Lambert /
// Oren-Nayar (DiffuseLobe), MicrofacetReflection with optional
// pure-reflection compensation (ConductorLobe, MetallicLobe, CoatLobe,
// SpecularLobe), DielectricLobe, SheenLTC.
enum LobeKind { LB_LAMBERT, LB_OREN, LB_REFLECTION, LB_DIELECTRIC, LB_SHEEN };
struct Lobe {
    int kind{LB_LAMBERT};
    float sample_weight{1.f}, weight{1.f};
    Spec kr;// Lambert / Oren-Nayar R, MicrofacetReflection kr, dielectric kt, sheen tint
    float onA{0.f}, onB{0.f};
    Ggx mf;
    Fresnel fr;
    bool compensate{false};
    float ltc_a{1.f}, ltc_b{0.f};
    // the albedo of the principled coat and specular lobes reads a table at
    // (lut_x, cos_theta, lut_z) (CoatLobe / SpecularLobe::albedo)
    int albedo_lut{0};// 0 none, 1 coat, 2 specular
    float lut_x{0.f}, lut_z{0.f};
};

struct LEval {
    Spec f;
    float pdf{0.f};
    bool trans{false};
};

struct Bsdf {
    bool black{false};
    bool set{false};// a LobeSet (principled) rather than a single lobe
    int nlobes{0};
    Lobe lobes[12];// a mix / add of two principled stacks flattens to 12
    Frame fr;
    V3 ng, wo_w;
    bool dispersive{false};
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

// DielectricLobe::sample_lut / refl_compensate (lobe.cpp:263-306): 1 / E(x, |cos|, eta)
float dielectric_compensation(const Spectral &sd, const Ggx &g, float abs_cos_o, float eta) {
    const float *t = eta > 1.f ? sd.lut_diel : sd.lut_diel_inv;
    float x = std::sqrt(std::sqrt(g.ax * g.ay));
    float e = eta > 1.f ? eta : 1.f / eta;
    float z = (e - 1.003f) / (5.f - 1.003f);
    float v[2];
    lut3(t, 32, 2, x, abs_cos_o, z, v);
    return 1.f / v[0];
}

// refl_prob (lobe.cpp:312-316)
float refl_prob(const Spec &F, const Spec &kt, int n) {
    Spec T = 1.f - F;
    Spec total = T * kt + F;
    return savg(F, n) / savg(total, n);
}

// Lobe::evaluate_local_impl of each kind; f without the cosine factor
LEval lobe_eval_local(const Spectral &sd, const Lobe &L, V3 wo, V3 wi, int n, float *eta) {
    LEval e;
    bool same = wo.z * wi.z > 0.f;
    switch (L.kind) {
        case LB_LAMBERT:
        case LB_OREN: {// BxDF::safe_evaluate (bxdf.cpp:27-40), LambertReflection / OrenNayar::f
            Spec f;
            if (L.kind == LB_LAMBERT) {
                f = L.kr * kInvPi;
            } else {
                float sti = sin_theta(wi), sto = sin_theta(wo);
                float d_cos = cos_phi(wi) * cos_phi(wo) + sin_phi(wi) * sin_phi(wo);
                float max_cos = std::max(0.f, d_cos);
                bool cond = std::fabs(wi.z) > std::fabs(wo.z);
                float sin_a = cond ? sto : sti;
                float tan_b = cond ? sti / std::fabs(wi.z) : sto / std::fabs(wo.z);
                f = L.kr * kInvPi * (L.onA + L.onB * max_cos * sin_a * tan_b);
            }
            e.f = same ? f : sfill(0.f);
            e.pdf = same ? std::fabs(wi.z) * kInvPi : 0.f;
            return e;
        }
        case LB_REFLECTION: {// MicrofacetReflection::f / PDF (bxdf.cpp:64-79)
            V3 wh = unit(add(wo, wi));
            V3 whf = face_forward(wh, v3(0.f, 0.f, 1.f));
            Spec F = fresnel_eval(L.fr, std::fabs(dotp(wo, whf)), n);
            Spec fr = L.mf.brdf(wo, whf, wi) * F;
            e.f = same ? fr * L.kr : sfill(0.f);
            e.pdf = same ? L.mf.pdf_reflection(wo, wh) : 0.f;
            if (L.compensate) {// PureReflectionLobe::compensate_factor (lobe.cpp:716-721)
                float v;
                lut2(sd.lut_pure, 32, 1, std::sqrt(L.mf.ax * L.mf.ay), wo.z, &v);
                e.f = e.f * (1.f / v);
            }
            return e;
        }
        case LB_DIELECTRIC: {// DielectricLobe::evaluate_local_impl with eta (lobe.cpp:318-415)
            float eta0 = L.fr.eta.v[0];
            float eta_p = same ? 1.f : eta0;
            if (eta) *eta = eta_p;
            V3 wh = face_forward(unit(add(wo, mul(wi, eta_p))), wo);
            Spec F = fresnel_eval(L.fr, std::fabs(dotp(wh, wo)), n);
            float rp = refl_prob(F, L.kr, n);
            float comp = dielectric_compensation(sd, L.mf, std::fabs(wo.z), eta0);
            if (same) {
                e.f = L.mf.brdf(wo, wh, wi) * F;
                e.pdf = L.mf.pdf_reflection(wo, wh) * rp;
            } else {
                V3 wh2 = unit(add(wo, mul(wi, eta0)));
                Spec tr = L.mf.btdf(wo, wh2, wi, eta0) * (1.f - F);
                e.f = tr * L.kr;
                e.pdf = L.mf.pdf_transmission(wo, face_forward(wh, wo), wi, eta0) * (1.f - rp);
                e.trans = true;
            }
            e.f = e.f * comp;
            return e;
        }
        default: {// SheenLTC::evaluate_local_impl / eval_ltc (principled_bsdf.cpp:59-104)
            V3 w = v3(L.ltc_a * wi.x + L.ltc_b * wi.z, L.ltc_a * wi.y, wi.z);
            float length = len(w);
            w = v3(w.x / length, w.y / length, w.z / length);
            float jac = (L.ltc_a * L.ltc_a) / (length * length * length);
            float ltc = w.z * kInvPi * jac;
            Spec f = (ltc * L.kr) / wi.z;
            e.f = (wi.z < 0.f || wo.z < 0.f) ? sfill(0.f) : f;
            e.pdf = ltc;
            return e;
        }
    }
}

// lobe direction sampling (sample_wi_local_impl of each kind); returns validity
bool lobe_sample_local(const Lobe &L, V3 wo, Sampler &sp, int n, V3 &wi) {
    switch (L.kind) {
        case LB_LAMBERT:
        case LB_OREN: {// BxDF::sample_wi (bxdf.cpp:42-46)
            float u0 = sp.next1d(), u1 = sp.next1d();
            wi = cosine_hemisphere(u0, u1);
            if (wo.z < 0.f) wi.z = -wi.z;
            return true;
        }
        case LB_REFLECTION: {// MicrofacetReflection::sample_wi (bxdf.cpp:80-84)
            float u0 = sp.next1d(), u1 = sp.next1d();
            V3 wh = L.mf.sample_wh(wo, u0, u1);
            wi = reflect_dir(wo, wh);
            return wo.z * wi.z > 0.f;
        }
        case LB_DIELECTRIC: {// DielectricLobe::sample_wi_local_impl (lobe.cpp:452-470)
            float u0 = sp.next1d(), u1 = sp.next1d();
            V3 wh = L.mf.sample_wh(wo, u0, u1);
            float d = dotp(wo, wh);
            Spec F = fresnel_eval(L.fr, std::fabs(d), n);
            float uc = sp.next1d();
            if (uc < refl_prob(F, L.kr, n)) {
                wi = reflect_dir(wo, wh);
                return wo.z * wi.z > 0.f;
            }
            bool ok = refract_dir(wo, wh, L.fr.eta.v[0], wi);
            return ok && !(wo.z * wi.z > 0.f);
        }
        default: {// SheenLTC::sample_wi_local_impl (principled_bsdf.cpp:88-95)
            float u0 = sp.next1d(), u1 = sp.next1d();
            V3 c = cosine_hemisphere(u0, u1);
            V3 m = v3(c.x / L.ltc_a - c.z * L.ltc_b / L.ltc_a, c.y / L.ltc_a, c.z);
            wi = unit(m);
            return true;
        }
    }
}

struct BEval {
    Spec f;// f * |cos wi|
    float pdf{0.f};
};

// Lobe::evaluate (f *= |cos wi|), LobeSet::evaluate_impl (lobe.cpp:668-688)
// and the hemisphere discard of MaterialEvaluator::evaluate (material.cpp:132-148)
BEval bsdf_eval(const Spectral &sd, const Bsdf &b, V3 wi_w, int n, float *eta) {
    BEval r;
    bool trans = false;
    if (b.black) {// BlackBodyLobe::evaluate_local_impl (black_body.cpp:33-40)
        r.f = sfill(0.f);
        r.pdf = 1.f;
    } else {
        V3 wo = b.fr.local(b.wo_w), wi = b.fr.local(wi_w);
        if (!b.set) {
            LEval e = lobe_eval_local(sd, b.lobes[0], wo, wi, n, eta);
            r.f = e.f * std::fabs(wi.z);
            r.pdf = e.pdf;
            trans = e.trans;
        } else {
            r.f = sfill(0.f);
            r.pdf = 0.f;
            for (int k = 0; k < b.nlobes; ++k) {
                const Lobe &L = b.lobes[k];
                LEval e = lobe_eval_local(sd, L, wo, wi, n, eta);
                Spec f = e.f * std::fabs(wi.z);
                float factor = L.kind == LB_DIELECTRIC ? 1.f : (wo.z * wi.z > 0.f ? 1.f : 0.f);
                f = f * (L.weight * factor);
                r.f = r.f + f;
                r.pdf += e.pdf * (L.sample_weight * factor);
                trans = trans || e.trans;
            }
        }
    }
    bool same_geo = dotp(b.wo_w, b.ng) * dotp(wi_w, b.ng) > 0.f;
    if (same_geo == trans) r.pdf = 0.f;
    return r;
}

struct BSample {
    BEval e;
    V3 wi;
    float eta{1.f};
};

// MaterialEvaluator::sample (material.cpp:162-184) through Lobe::sample
// (lobe.cpp:97-104) and LobeSet::sample_wi_impl (lobe.cpp:597-627)
BSample bsdf_sample(const Spectral &sd, const Bsdf &b, Sampler &sp, int n) {
    BSample s;
    if (b.black) {// BlackBodyLobe::sample_local: wi = wo, pdf 1
        s.wi = b.wo_w;
        s.e.f = sfill(0.f);
        s.e.pdf = 1.f;
        if (!(dotp(b.wo_w, b.ng) * dotp(s.wi, b.ng) > 0.f)) s.e.pdf = 0.f;
        return s;
    }
    V3 wo = b.fr.local(b.wo_w);
    V3 wi;
    bool valid;
    if (!b.set) {
        valid = lobe_sample_local(b.lobes[0], wo, sp, n, wi);
    } else {
        float uc = sp.next1d();
        (void)sp.next1d();// LobeSet draws a 2D sample it does not use
        (void)sp.next1d();
        int strategy = 0;
        float sum = 0.f;
        for (int k = 0; k < b.nlobes; ++k) {
            strategy = uc > sum ? k : strategy;
            sum += b.lobes[k].sample_weight;
        }
        valid = lobe_sample_local(b.lobes[strategy], wo, sp, n, wi);
    }
    s.wi = b.fr.world(wi);
    s.e = bsdf_eval(sd, b, s.wi, n, &s.eta);
    if (!valid) s.e.pdf = 0.f;
    return s;
}

// ---------------------------------------------------------------- materials
// record layouts (material records of the scene export)
//  diffuse / oren-nayar: p[0..3] colour slot, p[4] sigma
//  metal: p[0..2] eta rgb, p[3..5] k rgb, p[6] ax, p[7] ay, p[8], p[9] SPD offsets
//         of eta and k (uint bits), p[10] SPD interval
//  glass: p[0..3] colour slot, p[4] ior, p[5] ax, p[6] ay, p[7] SPD offset,
//         p[8] SPD interval, flags bit 0: dispersive ior curve
//  principled: p[0..3] color, p[4..7] spec_tint, p[8..11] sheen_tint,
//         p[12..15] coat_tint, p[16] metallic, p[17] ior, p[18] roughness,
//         p[19] anisotropic, p[20] sheen_weight, p[21] sheen_roughness,
//         p[22] coat_weight, p[23] coat_roughness, p[24] coat_ior,
//         p[25] transmission_weight
//  mix / add: p[0], p[1] indices (uint bits) of the two component
//         records, p[2] the mix fraction
enum MatType {
    M_DIFFUSE = 0, M_OREN = 1, M_BLACK = 2, M_METAL = 3, M_GLASS = 4, M_PRINCIPLED = 5, M_MIRROR = 6, M_MIX = 7,
    M_ADD = 8
};

float layering_div(float a, float w) { return w != 0.f ? a / w : 0.f; }
// layering_weight (principled_bsdf.cpp:174-179)
Spec layering_weight(const Spec &albedo, const Spec &weight, int n) {
    float m = layering_div(albedo.v[0], weight.v[0]);
    for (int i = 1; i < n; ++i) m = std::max(m, layering_div(albedo.v[i], weight.v[i]));
    return weight * std::min(std::max(1.f - m, 0.f), 1.f);
}

Bsdf make_leaf_bsdf(const Spectral &sd, const Scene &s, uint32_t m, const Surf &it, const Swl &w);

// MixMaterial::create_lobe_set (mix.cpp) / AddMaterial::create_lobe_set
// (add.cpp): LobeSet::create_mix(frac, A, B) gives A and B the weights
// 1 - frac and frac (sample and evaluation alike); create_add gives both
// 1 and normalises the sample weights to 1/2 (lobe.cpp:495-521).  Both then
// flatten (lobe.cpp:533-562): a component that is itself a lobe set (the
// principled stack) is replaced by its lobes, each lobe's sample and
// evaluation weight multiplied by the component's sample weight.
Bsdf make_bsdf(const Spectral &sd, const Scene &s, uint32_t m, const Surf &it, const Swl &w) {
    const uint32_t *rec = s.mat + 32 * m;
    if (rec[0] != M_MIX && rec[0] != M_ADD) return make_leaf_bsdf(sd, s, m, it, w);
    float frac;
    std::memcpy(&frac, rec + 2 + 2, 4);
    const bool add = rec[0] == M_ADD;
    Bsdf b;
    b.fr = it.fr;
    b.ng = it.ng;
    b.wo_w = it.wo;
    b.set = true;
    b.nlobes = 0;
    for (int k = 0; k < 2; ++k) {
        Bsdf c = make_leaf_bsdf(sd, s, rec[2 + k], it, w);
        b.dispersive = b.dispersive || c.dispersive;
        const float wk = k == 0 ? 1.f - frac : frac;
        const float parent_sample = add ? 0.5f : wk;     // after normalize_sampled_weight
        const float parent_eval = add ? 1.f : wk;
        for (int i = 0; i < c.nlobes; ++i) {
            Lobe L = c.lobes[i];
            if (c.set) {// flattened: both weights times the parent's sample weight
                L.sample_weight *= parent_sample;
                L.weight *= parent_sample;
            } else {
                L.sample_weight = parent_sample;
                L.weight = parent_eval;
            }
            b.lobes[b.nlobes++] = L;
        }
    }
    return b;
}

// Material::create_lobe_set of diffuse.cpp:16-24, metal.cpp:124-140,
// glass.cpp:215-228, principled_bsdf.cpp:352-461 and black_body.cpp
Bsdf make_leaf_bsdf(const Spectral &sd, const Scene &s, uint32_t m, const Surf &it, const Swl &w) {
    Bsdf b;
    b.fr = it.fr;
    b.ng = it.ng;
    b.wo_w = it.wo;
    const uint32_t *rec = s.mat + 32 * m;
    uint32_t type = rec[0], flags = rec[1];
    float p[30];
    std::memcpy(p, rec + 2, sizeof(p));
    int n = w.n;
    b.nlobes = 1;
    Lobe &L = b.lobes[0];
    switch (type) {
        case M_BLACK: b.black = true; break;
        case M_DIFFUSE:
        case M_OREN: {
            L.kind = type == M_DIFFUSE ? LB_LAMBERT : LB_OREN;
            L.kr = albedo_at(sd, flags, p, 0, it.u, it.v, w);
            if (type == M_OREN) {// OrenNayar::OrenNayar (bxdf.cpp:86-92)
                float sigma = p[4] * (kPi * 0.5f);
                float sigma2 = (sigma * sigma) * (sigma * sigma);
                L.onA = 1.f - (sigma2 / (2.f * (sigma2 + 0.33f)));
                L.onB = 0.45f * sigma2 / (sigma2 + 0.09f);
            }
            break;
        }
        case M_METAL: {
            L.kind = LB_REFLECTION;
            L.kr = sfill(1.f);
            L.mf = {p[6], p[7]};
            L.fr.kind = FR_CONDUCTOR;
            if (sd.rgb) {
                for (int i = 0; i < 3; ++i) {
                    L.fr.eta.v[i] = p[i];
                    L.fr.k.v[i] = p[3 + i];
                }
            } else {
                uint32_t oe = rec[2 + 8], ok = rec[2 + 9];
                for (int i = 0; i < n; ++i) {
                    L.fr.eta.v[i] = spd_eval(sd.spd + oe, p[10], w.lambda[i]);
                    L.fr.k.v[i] = spd_eval(sd.spd + ok, p[10], w.lambda[i]);
                }
            }
            L.compensate = true;
            break;
        }
        case M_MIRROR: {// MirrorMaterial::create_lobe_set (mirror.cpp:60-72)
            L.kind = LB_REFLECTION;
            L.kr = albedo_at(sd, flags, p, 0, it.u, it.v, w);
            L.mf = {p[6], p[7]};
            L.fr.kind = FR_ONE;
            L.compensate = true;
            break;
        }
        case M_GLASS: {
            L.kind = LB_DIELECTRIC;
            L.kr = albedo_at(sd, flags, p, 0, it.u, it.v, w);
            L.mf = {p[5], p[6]};
            L.fr.kind = FR_DIELECTRIC;
            bool front = dotp(it.wo, it.ng) > 0.f;// Interaction::correct_eta (interaction.cpp:85-88)
            if (!sd.rgb && (flags & 1u)) {
                b.dispersive = true;
                uint32_t off = rec[2 + 7];
                for (int i = 0; i < n; ++i) {
                    float ior = spd_eval(sd.spd + off, p[8], w.lambda[i]);
                    L.fr.eta.v[i] = front ? ior : 1.f / ior;
                }
            } else {
                L.fr.eta = sfill(front ? p[4] : 1.f / p[4]);
            }
            break;
        }
        default: {// principled
            b.set = true;
            Spec color = albedo_at(sd, flags, p, 0, it.u, it.v, w);
            Spec spec_tint = albedo_at(sd, flags, p, 4, it.u, it.v, w);
            float ior = p[17];
            float roughness = std::min(std::max(p[18], 0.0001f), 1.f);
            float anisotropic = p[19];
            float aspect = std::sqrt(1.f - anisotropic * 0.9f);
            Ggx mf{std::max(0.001f, (roughness * roughness) / aspect), std::max(0.001f, (roughness * roughness) * aspect)};
            Spec weight = sfill(1.f);
            float cos_theta = dotp(it.wo, it.ng);
            float front = cos_theta > 0.f ? 1.f : 0.f;
            int k = 0;
            {// sheen
                Lobe &S = b.lobes[k++];
                S.kind = LB_SHEEN;
                Spec tint = albedo_at(sd, flags, p, 8, it.u, it.v, w) * (p[20] * front) * weight;
                float ltc[4];
                lut2(sd.lut_sheen, 32, 4, cos_theta, p[21], ltc);
                S.ltc_a = ltc[0];
                S.ltc_b = ltc[1];
                S.kr = tint * ltc[2];
                S.sample_weight = savg(S.kr, n);
                weight = layering_weight(S.kr, weight, n);
            }
            {// coat
                Lobe &C = b.lobes[k++];
                float cc_weight = p[22] * front;
                float cc_rough = std::min(std::max(p[23], 0.0001f), 1.f);
                cc_rough = cc_rough * cc_rough;
                C.kind = LB_REFLECTION;
                C.mf = {cc_rough, cc_rough};
                C.fr.kind = FR_DIELECTRIC;
                C.fr.eta = sfill(p[24]);
                C.kr = cc_weight * weight * albedo_at(sd, flags, p, 12, it.u, it.v, w);
                float x = std::sqrt(std::sqrt(C.mf.ax * C.mf.ay));
                float z = (p[24] - 1.003f) / (4.f - 1.003f);
                C.albedo_lut = 1, C.lut_x = x, C.lut_z = z;
                float sv;
                lut3(sd.lut_coat, 32, 1, x, cos_theta, z, &sv);
                Spec alb = sv * C.kr;
                C.sample_weight = savg(alb, n);
                weight = layering_weight(alb, weight, n);
            }
            {// metallic
                Lobe &M = b.lobes[k++];
                float metallic = p[16] * front;
                M.kind = LB_REFLECTION;
                M.mf = mf;
                M.fr.kind = FR_F82;
                M.fr.F0 = color;
                const float f = 6.f / 7.f;
                const float f5 = f * f * f * f * f;
                Spec one = sfill(1.f);
                Spec f_schlick = color + f5 * (one - color);
                M.fr.B = f_schlick * (7.f / (f5 * f)) * (one - spec_tint);
                M.kr = weight * metallic;
                M.compensate = true;
                M.sample_weight = metallic * savg(weight, n);
                weight = weight * (1.0f - metallic);
            }
            {// transmission
                Lobe &T = b.lobes[k++];
                float tw = p[25];
                float eta = cos_theta > 0.f ? ior : 1.f / ior;
                T.kind = LB_DIELECTRIC;
                T.mf = mf;
                T.fr.kind = FR_SCHLICK;
                T.fr.eta = sfill(eta);
                T.fr.F0 = schlick_f0_from_ior(eta) * spec_tint;
                T.kr = color;
                Spec t_weight = tw * weight;
                T.sample_weight = savg(t_weight, n);
                T.weight = savg(t_weight, n);
                weight = weight * (1.0f - tw);
            }
            {// specular
                Lobe &P = b.lobes[k++];
                P.kind = LB_REFLECTION;
                P.mf = mf;
                P.fr.kind = FR_SCHLICK;
                P.fr.eta = sfill(ior);
                P.fr.F0 = schlick_f0_from_ior(ior) * spec_tint;
                P.kr = weight;
                float x = std::sqrt(std::sqrt(mf.ax * mf.ay));
                float z = std::sqrt(std::fabs((ior - 1.0f) / (ior + 1.0f)));
                P.albedo_lut = 2, P.lut_x = x, P.lut_z = z;
                float sv;
                lut3(sd.lut_spec, 32, 1, x, cos_theta, z, &sv);
                Spec alb = (P.fr.F0 + sv * (1.f - P.fr.F0)) * P.kr;
                P.sample_weight = savg(alb, n);
                weight = layering_weight(alb, weight, n);
            }
            {// diffuse
                Lobe &D = b.lobes[k++];
                D.kind = LB_LAMBERT;
                D.kr = color * weight * front;
                D.sample_weight = savg(D.kr, n);
            }
            b.nlobes = k;
            float sum = 0.f;// LobeSet::normalize_sampled_weight (lobe.cpp:448-457)
            for (int i = 0; i < k; ++i) sum += b.lobes[i].sample_weight;
            for (int i = 0; i < k; ++i) b.lobes[i].sample_weight = b.lobes[i].sample_weight / sum;
            break;
        }
    }
    return b;
}

// ---------------------------------------------------------------- lights
float luminance(V3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }

struct EnvMap {
    bool present{false};
    uint32_t light{kInv};
    uint32_t w{0}, h{0};
    const float *rgb{nullptr};
    std::vector<float> coef;// decode_unbound per texel (hero mode)
    std::vector<Alias> rows;
    Alias marg;
    float w2o[9]{}, o2w[9]{};
    float scale{1.f};
    float diameter{0.f};
};

struct LightTables {
    std::vector<Alias> tri_alias;// per light (area)
    Alias select;
    bool power{false};
    EnvMap env;
};

// light power (area.cpp:87-89, point.cpp:32-34, spot.cpp:44-46, spherical.cpp:57-60)
V3 light_power(const Scene &s, uint32_t l, const LightTables &lt, float world_radius) {
    V3 avg = mul(v3(s.lightf(l, 8), s.lightf(l, 9), s.lightf(l, 10)), s.lightf(l, 11));
    uint32_t type = s.lightu(l, 0);
    if (type == 0) {
        float area = 0.f;
        for (float a : lt.tri_alias[l].func) area += a;
        return mul(avg, (s.lightu(l, 1) ? 2.f : 1.f) * area * kPi);
    }
    if (type == 1) return mul(avg, 4.f * kPi);
    if (type == 2) return mul(avg, 2.f * kPi * (1.f - .5f * (s.lightf(l, 15) * 2.f + s.lightf(l, 19))));
    if (type == 3) {
        double acc[3] = {0, 0, 0};
        size_t n = size_t(lt.env.w) * lt.env.h;
        for (size_t i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) acc[c] += lt.env.rgb[3 * i + c];
        float sc = s.lightf(l, 11);
        V3 mean = v3(float(acc[0] / double(n)) * sc, float(acc[1] / double(n)) * sc, float(acc[2] / double(n)) * sc);
        return mul(mean, kPi * (world_radius * world_radius));
    }
    return v3(0, 0, 0);
}

LightTables build_lights(const Scene &s, const Spectral &sd, uint32_t env_light, uint32_t env_w, uint32_t env_h,
                         const float *env_rgb, const float *w2o, const float *o2w, float env_scale, float world_radius) {
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
    if (env_light != kInv) {
        EnvMap &e = lt.env;
        e.present = true;
        e.light = env_light;
        e.w = env_w;
        e.h = env_h;
        e.rgb = env_rgb;
        e.scale = env_scale;
        e.diameter = 2.f * world_radius;
        std::memcpy(e.w2o, w2o, sizeof(e.w2o));
        std::memcpy(e.o2w, o2w, sizeof(e.o2w));
        size_t n = size_t(env_w) * env_h;
        if (!sd.rgb) {
            e.coef.resize(4 * n);
            for (size_t i = 0; i < n; ++i)
                decode_unbound(sd.rgb2spec, v3(env_rgb[3 * i], env_rgb[3 * i + 1], env_rgb[3 * i + 2]), &e.coef[4 * i]);
        }
        // SphericalMap::calculate_weights (spherical.cpp:170-196); the row
        // index is the integer quotient idx / res.y, the reference's own form
        std::vector<float> weights(n);
        for (size_t i = 0; i < n; ++i) {
            float v = float(uint32_t(i) / env_h) + 0.5f;
            float theta = v / float(env_w);
            float sin_t = std::sin(kPi * theta);
            weights[i] = luminance(v3(env_rgb[3 * i], env_rgb[3 * i + 1], env_rgb[3 * i + 2])) * sin_t;
        }
        // AliasTable2D::build (alias2d.cpp:32-71)
        std::vector<float> marg(env_h);
        e.rows.resize(env_h);
        for (uint32_t y = 0; y < env_h; ++y) {
            e.rows[y] = alias_build(weights.data() + size_t(y) * env_w, env_w);
            marg[y] = e.rows[y].integral;
        }
        e.marg = alias_build(marg.data(), env_h);
    }
    lt.power = s.render[4] != 0.f;
    if (lt.power) {
        std::vector<float> w(s.nl);
        for (uint32_t l = 0; l < s.nl; ++l) w[l] = luminance(light_power(s, l, lt, world_radius));
        lt.select = alias_build(w.data(), s.nl);
    }
    return lt;
}

float select_pmf(const Scene &s, const LightTables &lt, uint32_t l) {
    return lt.power ? alias_pmf(lt.select, l) : 1.f / float(s.nl);
}

V3 m3(const float *m, V3 v) {
    return v3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z, m[6] * v.x + m[7] * v.y + m[8] * v.z);
}

// SphericalMap::calculate_uv (spherical.cpp:62-65) with spherical_theta/phi
void env_uv(V3 local, float &u, float &v) {
    float theta = vs::em::acos(std::min(std::max(local.z, -1.f), 1.f));
    float phi = vs::em::atan2(local.y, local.x);
    if (phi < 0.f) phi += 2.f * kPi;
    u = phi * (0.5f * kInvPi);
    v = theta * kInvPi;
}
// texel lookup of the environment image: nearest texel
Spec env_L(const Spectral &sd, const EnvMap &e, V3 local, const Swl &w) {
    float u, v;
    env_uv(local, u, v);
    uint32_t x = std::min(uint32_t(std::max(u, 0.f) * float(e.w)), e.w - 1u);
    uint32_t y = std::min(uint32_t(std::max(v, 0.f) * float(e.h)), e.h - 1u);
    size_t i = size_t(y) * e.w + x;
    return illumination(sd, e.rgb + 3 * i, e.coef.empty() ? nullptr : &e.coef[4 * i], e.scale, w);
}
// AliasTable2D::PDF (alias2d.cpp:101-105)
float env_map_pdf(const EnvMap &e, float u, float v) {
    uint32_t iu = std::min(uint32_t(std::max(u * float(e.w), 0.f)), e.w - 1u);
    uint32_t iv = std::min(uint32_t(std::max(v * float(e.h), 0.f)), e.h - 1u);
    float integral = e.marg.integral;
    return integral > 0.f ? e.rows[iv].func[iu] / integral : 0.f;
}
// SphericalMap::evaluate_wi (spherical.cpp:76-93) for a world direction
void env_eval(const Spectral &sd, const EnvMap &e, V3 dir, const Swl &w, Spec &L, float &pdf) {
    V3 local = m3(e.w2o, dir);
    float theta = vs::em::acos(std::min(std::max(local.z, -1.f), 1.f));
    float sin_t = vs::em::sin(theta);
    float u, v;
    env_uv(local, u, v);
    L = env_L(sd, e, local, w);
    float p = env_map_pdf(e, u, v) / (2.f * kPi * kPi * sin_t);
    pdf = sin_t == 0.f ? 0.f : p;
}

struct LSample {
    Spec L;
    float pdf;
    V3 p;
};

// a light's radiance; an area light with an emission image (word 7: its
// texture) takes the nearest texel at the light surface's (u, v) as its
// colour, wrapped, v = 0 the first row (AreaLight::L, area.cpp:90-94)
Spec light_color(const Spectral &sd, const Scene &s, uint32_t l, const Swl &w, float u = 0.f, float v = 0.f) {
    const uint32_t t = s.lightu(l, 7);
    if (t != 0xFFFFFFFFu) {
        if (t >= sd.ntex) return sfill(0.f);
        const uint32_t *info = sd.tex_info + 4 * t;
        float uu = u - std::floor(u), vv = v - std::floor(v);
        uint32_t x = std::min(uint32_t(uu * float(info[1])), info[1] - 1u);
        uint32_t y = std::min(uint32_t(vv * float(info[2])), info[2] - 1u);
        const float *q = sd.tex + 4 * (size_t(info[0]) + size_t(y) * info[1] + x);
        return illumination(sd, q, q, s.lightf(l, 11), w);
    }
    float rgb[3] = {s.lightf(l, 8), s.lightf(l, 9), s.lightf(l, 10)};
    float coef[4] = {s.lightf(l, 24), s.lightf(l, 25), s.lightf(l, 26), s.lightf(l, 27)};
    return illumination(sd, rgb, coef, s.lightf(l, 11), w);
}

// LightSampler::sample_wi (src/base/illumination/lightsampler.cpp:199-216),
// AreaLight::sample_wi (src/render_core/lights/area.cpp:120-149),
// SphericalMap::sample_wi / evaluate (spherical.cpp:94-163),
// AliasTable2D::sample_continuous (alias2d.cpp:109-129),
// IPointLight::sample_wi (src/base/illumination/light.cpp:42-51),
// PointLight::Le (point.cpp:39-43), SpotLight::Le/falloff (spot.cpp:51-70)
LSample sample_light(const Spectral &sd, const Scene &s, const LightTables &lt, const Surf &it, Sampler &sp,
                     const Swl &w) {
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
        V3 wv = sub(it.pos, p.pos);
        bool front = dotp(wv, p.ng) > 0.f || s.lightu(l, 1);
        ls.L = front ? light_color(sd, s, l, w, p.u, p.v) : sfill(0.f);
        float pdf = pdf_pos * dotp(wv, wv) / std::fabs(dotp(p.ng, unit(wv)));
        ls.pdf = (std::isinf(pdf) || std::isnan(pdf)) ? 0.f : pdf;
        float f = dotp(p.ng, wv) > 0.f ? 1.f : -1.f;
        ls.p = offset_origin(p.pos, mul(p.ng, f * s.render[5]));
    } else if (type == 3) {
        const EnvMap &e = lt.env;
        float uvr;
        uint32_t iv = alias_sample(e.marg, uy, &uvr);
        float fv = (float(iv) + uvr) / float(e.h);
        float pdf_v = e.marg.integral > 0.f ? e.marg.func[iv] / e.marg.integral : 0.f;
        const Alias &row = e.rows[iv];
        float uur;
        uint32_t iu = alias_sample(row, ux, &uur);
        float fu = (float(iu) + uur) / float(e.w);
        float integral_u = e.marg.func[iv];
        float pdf_u = integral_u > 0.f ? row.func[iu] / integral_u : 0.f;
        float pdf_map = pdf_u * pdf_v;
        float theta = fv * kPi, phi = fu * 2.f * kPi;
        float sin_t, cos_t, sin_p, cos_p;
        vs::em::sincos(theta, &sin_t, &cos_t);
        vs::em::sincos(phi, &sin_p, &cos_p);
        V3 local = v3(sin_t * cos_p, sin_t * sin_p, cos_t);
        V3 world = unit(m3(e.o2w, local));
        float pdf_dir = pdf_map / (2.f * kPi * kPi * sin_t);
        ls.pdf = std::isinf(pdf_dir) ? 0.f : pdf_dir;
        ls.L = env_L(sd, e, local, w);
        ls.p = add(it.pos, mul(world, e.diameter));
    } else {
        V3 pos = v3(s.lightf(l, 12), s.lightf(l, 13), s.lightf(l, 14));
        V3 wu = sub(it.pos, pos);
        Spec val = light_color(sd, s, l, w) / dotp(wu, wu);
        if (type == 2) {
            V3 dir = v3(s.lightf(l, 16), s.lightf(l, 17), s.lightf(l, 18));
            float angle = s.lightf(l, 15), fall = s.lightf(l, 19);
            float start = std::max(0.f, angle - fall);
            float ca = std::cos(angle), cs = std::cos(start);
            float ct = std::min(std::max(dotp(dir, unit(wu)), ca), cs);
            float f = (ct - ca) / (cs - ca);
            float f2 = f * f;
            val = val * (f2 * f2);
        }
        ls.L = val;
        ls.pdf = -1.f;
        ls.p = pos;
    }
    ls.pdf *= pmf_sel;
    return ls;
}

// LightSampler::evaluate_hit_wi (lightsampler.cpp:252-267)
void eval_hit(const Spectral &sd, const Scene &s, const LightTables &lt, const Surf &it, uint32_t l, uint32_t tri,
              V3 ref, const Swl &w, Spec &L, float &pdf) {
    uint32_t prim = tri - s.lightu(l, 3);
    float pdf_pos = (1.f / it.area) * alias_pmf(lt.tri_alias[l], prim);
    V3 wv = sub(ref, it.pos);
    bool front = dotp(wv, it.ng) > 0.f || s.lightu(l, 1);
    L = front ? light_color(sd, s, l, w, it.u, it.v) : sfill(0.f);
    float p = pdf_pos * dotp(wv, wv) / std::fabs(dotp(it.ng, unit(wv)));
    p = (std::isinf(p) || std::isnan(p)) ? 0.f : p;
    pdf = p * select_pmf(s, lt, l);
}

// ---------------------------------------------------------------- integrator
struct Ctx {
    const Scene *s;
    const Bvh2 *bvh;
    const LightTables *lt;
    const Spectral *sd;
    bool brute;
    Hit trace(const Ray &r, bool any) const { return brute ? trace_brute(*s, r, any) : trace_bvh(*s, *bvh, r, any); }
};

float mis_w(float a, float b) { return a / (a + b); }

// The surface HitContext::it receives (integrator.cpp:181-183): the one hit
// at bounce 0 (including the bounce-0 hit of the max_depth < 2 extra ray)
struct FirstHit {
    bool set{false};
    V3 pos{0.f, 0.f, 0.f}, ng{0.f, 0.f, 0.f};
};

// IlluminationIntegrator::Li (src/base/integrator.cpp:134-283), with
// evaluate_miss (integrator.cpp:111-132) and direct_lighting (integrator.cpp:22-38).
// thr0 / pdf0 / depth: the starting throughput, scatter pdf and max_depth
// (1, 1e16 and the scene's for a camera path; ReSTIR GI starts its paths at
// the primary surface, indirect.cpp:65-83); `first` records the bounce-0 hit.
V3 Li(const Ctx &c, V3 ro, V3 rd, Sampler &sp, Swl &w, const Spec *thr0 = nullptr, float pdf0 = 1e16f,
      int depth = -1, FirstHit *first = nullptr) {
    const Scene &s = *c.s;
    const Spectral &sd = *c.sd;
    const LightTables &lt = *c.lt;
    int n = w.n;
    uint32_t max_depth = depth >= 0 ? uint32_t(depth) : uint32_t(s.render[0]), min_depth = uint32_t(s.render[1]);
    float rr_th = s.render[2];
    int mis_mode = int(s.render[3]);
    V3 ret = v3(0, 0, 0);
    Spec thr = thr0 ? *thr0 : sfill(1.f);
    float scatter_pdf = pdf0, eta_scale = 1.f;
    uint32_t skips = 0;
    auto correct = [&](float wgt, uint32_t b) {// correct_bsdf_weight
        if (mis_mode == 2) return 1.f;
        if (mis_mode == 1) return b == 0 ? wgt : 0.f;
        return wgt;
    };
    const V3 primary = rd;
    for (uint32_t b = 0; b < max_depth; ++b) {
        Hit h = c.trace({ro, rd, 0.f, FLT_MAX}, false);
        if (h.tri == kInv) {
            if (lt.env.present) {
                Spec L;
                float lpdf;
                env_eval(sd, lt.env, unit(sub(add(ro, rd), ro)), w, L, lpdf);
                lpdf *= select_pmf(s, lt, lt.env.light);
                float wgt = correct(mis_w(scatter_pdf, lpdf), b);
                Spec d = (L * wgt) * thr;
                ret = add(ret, linear_srgb(sd, d, w));
            }
            break;
        }
        Surf it = surface(s, h.tri, h.b1, h.b2, true);
        it.wo = unit(neg(rd));
        uint32_t mat = s.inst[4 * it.inst], light = s.inst[4 * it.inst + 1];
        if (mat == kInv) {
            V3 nn = dotp(it.ng, rd) > 0.f ? it.ng : neg(it.ng);
            ro = offset_origin(it.pos, nn);
            if (++skips > 64u) break;
            --b;
            continue;
        }
        if (first && b == 0) *first = FirstHit{true, it.pos, it.ng};
        if (light != kInv) {
            Spec Le;
            float lpdf;
            eval_hit(sd, s, lt, it, light, h.tri, ro, w, Le, lpdf);
            float wgt = correct(mis_w(scatter_pdf, lpdf), b);
            Spec d = (Le * thr) * wgt;
            ret = add(ret, linear_srgb(sd, d, w));
        }
        LSample ls = sample_light(sd, s, lt, it, sp, w);
        Bsdf bsdf = make_bsdf(sd, s, mat, it, w);
        if (!sd.rgb && bsdf.dispersive) {// SampledWavelengths::check_dispersive (spectrum.cpp:10-39)
            bool secondary_valid = n > 1;
            for (int i = 1; i < n; ++i) secondary_valid = secondary_valid && w.pdf[i] != 0.f;
            if (secondary_valid)
                for (int i = 1; i < n; ++i) w.pdf[i] = 0.f;
        }
        BEval se = bsdf_eval(sd, bsdf, unit(sub(ls.p, it.pos)), n, nullptr);
        BSample bs = bsdf_sample(sd, bsdf, sp, n);
        bool delta = ls.pdf < 0.f;
        float wgt = delta ? 1.f : mis_w(ls.pdf, se.pdf);
        if (mis_mode == 1) wgt = 1.f;
        float lpdf = delta ? -ls.pdf : ls.pdf;
        if (se.pdf > 0.f && lpdf > 0.f && mis_mode != 2) {
            V3 dir = sub(ls.p, it.pos);
            V3 nn = dotp(it.ng, dir) > 0.f ? it.ng : neg(it.ng);
            Hit sh = c.trace({offset_origin(it.pos, nn), dir, 0.f, 1.f - kShadowEps}, true);
            if (sh.tri == kInv) {
                Spec Ld = ((ls.L * se.f) * wgt) / lpdf;
                ret = add(ret, linear_srgb(sd, thr * Ld, w));
            }
        }
        eta_scale *= bs.eta * bs.eta;
        float lum = smax(thr, n);
        if (!(bs.e.pdf > 0.f) || lum == 0.f) break;
        thr = thr * (bs.e.f / bs.e.pdf);
        if (eta_scale * lum < rr_th && b >= min_depth) {
            float q = std::min(0.95f, lum);
            float rr = sp.next1d();
            if (q < rr) break;
            thr = thr / q;
        }
        scatter_pdf = bs.e.pdf;
        V3 nn = dotp(it.ng, bs.wi) > 0.f ? it.ng : neg(it.ng);
        ro = offset_origin(it.pos, nn);
        rd = bs.wi;
    }
    // only_direct (max_depth < 2) with both MIS strategies: one more
    // BSDF-sampled ray collects emitters and the environment
    // (integrator.cpp:276-281 running mis_bsdf of :150-204 with inner = false):
    // nothing when the ray is still the primary one, material-less surfaces
    // passed through, the bounce index 0 for the MIS correction
    if (max_depth < 2 && mis_mode == 0 && !(rd.x == primary.x && rd.y == primary.y && rd.z == primary.z)) {
        while (true) {
            Hit h = c.trace({ro, rd, 0.f, FLT_MAX}, false);
            if (h.tri == kInv) {
                if (lt.env.present) {
                    Spec L;
                    float lpdf;
                    env_eval(sd, lt.env, unit(sub(add(ro, rd), ro)), w, L, lpdf);
                    lpdf *= select_pmf(s, lt, lt.env.light);
                    Spec d = (L * correct(mis_w(scatter_pdf, lpdf), 0)) * thr;
                    ret = add(ret, linear_srgb(sd, d, w));
                }
                break;
            }
            Surf it = surface(s, h.tri, h.b1, h.b2, true);
            uint32_t mat = s.inst[4 * it.inst], light = s.inst[4 * it.inst + 1];
            if (mat == kInv) {
                V3 nn = dotp(it.ng, rd) > 0.f ? it.ng : neg(it.ng);
                ro = offset_origin(it.pos, nn);
                if (++skips > 64u) break;
                continue;
            }
            if (first) *first = FirstHit{true, it.pos, it.ng};
            if (light != kInv) {
                Spec Le;
                float lpdf;
                eval_hit(sd, s, lt, it, light, h.tri, ro, w, Le, lpdf);
                Spec d = (Le * thr) * correct(mis_w(scatter_pdf, lpdf), 0);
                ret = add(ret, linear_srgb(sd, d, w));
            }
            break;
        }
    }
    return ret;
}

// ---------------------------------------------------------------- pixel filters
// The fitted-curve filters' evaluate: gaussian.cpp:38-42 (a normal density,
// src/math/util.h:37-42, less its value at the radius, clamped at 0, per
// axis), mitchell.cpp:29-45 (Mitchell-Netravali cubic of 2x/r per axis) and
// sinc.cpp:21-25 (Lanczos: sinc(x) sinc(x/tau) inside the radius, per axis,
// times 4).
float normal_density(float x, float sigma) {
    float var = sigma * sigma;
    return 1.f / std::sqrt(2 * kPi * var) * std::exp(-(x * x) / (2 * var));
}
float mitchell(float x, float B, float C) {
    x = std::fabs(x);
    float x2 = x * x, x3 = x2 * x;
    if (x <= 1) return ((12 - 9 * B - 6 * C) * x3 + (-18 + 12 * B + 6 * C) * x2 + (6 - 2 * B)) * (1.f / 6.f);
    if (x <= 2) return ((-B - 6 * C) * x3 + (6 * B + 30 * C) * x2 + (-12 * B - 48 * C) * x + (8 * B + 24 * C)) * (1.f / 6.f);
    return 0.f;
}
float sinc_pi(float x) {
    float t = kPi * x;
    return 1 + t * t == 1 ? 1.f : std::sin(t) / t;
}
float filter_value(const float *cam, float x, float y) {
    const int kind = int(cam[24]);
    const float r = cam[22];
    if (kind == 2) {
        float sigma = cam[25], edge = normal_density(r, sigma);
        return std::max(0.f, normal_density(x, sigma) - edge) * std::max(0.f, normal_density(y, sigma) - edge);
    }
    if (kind == 3) return mitchell(2 * x / r, cam[26], cam[27]) * mitchell(2 * y / r, cam[26], cam[27]);
    if (kind == 4) {
        float tau = cam[28];
        float fx = std::fabs(x) > r ? 0.f : sinc_pi(x) * sinc_pi(x / tau);
        float fy = std::fabs(y) > r ? 0.f : sinc_pi(y) * sinc_pi(y / tau);
        return fx * fy * 4;
    }
    return 1.f;
}

// FilterSampler::build (src/render_core/filters/fitted_curve.h:33-53): the
// filter at the centres of a 20 x 20 grid over [0, r)^2, the table normalised
// by its mean signed value (summed in float, as std::accumulate does), |f|
// sampled by a 2D alias table (AliasTable2D::build, alias2d.cpp:32-71)
constexpr uint32_t kFis = 20;
struct FisTable {
    bool on{false};
    std::vector<float> lut;
    std::vector<Alias> rows;
    Alias marg;
};
FisTable fis_build(const float *cam) {
    FisTable t;
    if (cam[21] != 2.f) return t;
    t.on = true;
    const float r = cam[22];
    std::vector<float> absf(kFis * kFis);
    t.lut.resize(kFis * kFis);
    for (uint32_t y = 0; y < kFis; ++y)
        for (uint32_t x = 0; x < kFis; ++x) {
            float v = filter_value(cam, (float(x) + 0.5f) / float(kFis) * r, (float(y) + 0.5f) / float(kFis) * r);
            t.lut[y * kFis + x] = v;
            absf[y * kFis + x] = std::fabs(v);
        }
    float total = 0.f;
    for (float v : t.lut) total += v;
    const float mean = total / float(kFis * kFis);
    for (float &v : t.lut) v = v / mean;
    std::vector<float> row_mass(kFis);
    for (uint32_t y = 0; y < kFis; ++y) {
        t.rows.push_back(alias_build(absf.data() + y * kFis, kFis));
        row_mass[y] = t.rows.back().integral;
    }
    t.marg = alias_build(row_mass.data(), kFis);
    return t;
}

// FilterSampler::sample (fitted_curve.h:67-73) with
// AliasTable2D::sample_continuous (alias2d.cpp:106-129): u -> 2u - 1, the
// table sampled at |u|, the point mirrored by the sign of u (sign(0) = +1),
// weight = lut / pdf; FittedCurveFilter::sample (:90-94) scales by the radius
void fis_sample(const FisTable &t, float radius, float ux, float uy, float &fx, float &fy, float &weight) {
    float sx = ux * 2.f - 1.f, sy = uy * 2.f - 1.f;
    float ry;
    uint32_t iv = alias_sample(t.marg, std::fabs(sy), &ry);
    float v = (float(iv) + ry) / float(kFis);
    float pdf_v = t.marg.integral > 0.f ? t.marg.func[iv] / t.marg.integral : 0.f;
    const Alias &row = t.rows[iv];
    float rx;
    uint32_t iu = alias_sample(row, std::fabs(sx), &rx);
    float u = (float(iu) + rx) / float(kFis);
    float row_integral = t.marg.func[iv];
    float pdf_u = row_integral > 0.f ? row.func[iu] / row_integral : 0.f;
    weight = t.lut[iv * kFis + iu] / (pdf_u * pdf_v);
    fx = (sx < 0.f ? -u : u) * radius;
    fy = (sy < 0.f ? -v : v) * radius;
}

// Sensor::generate_ray (src/base/sensor/sensor.cpp:44-56), thin lens
// (src/render_core/sensors/thin_lens.cpp:34-42), Sampler::sensor_sample
// (src/base/sampler.h:65-73), box / tent filters (box.cpp:16-20, triangle.cpp:16-18)
void camera_ray(const Scene &s, const FisTable &fis, uint32_t px, uint32_t py, Sampler &sp, V3 &o, V3 &d,
                float &weight, float *p_film = nullptr) {
    const float *c = s.cam;
    float ux = sp.next1d(), uy = sp.next1d();
    float r = c[22];
    float fx, fy;
    weight = 1.f;
    if (c[21] == 0.f) {
        fx = -r + ux * (r + r);
        fy = -r + uy * (r + r);
    } else if (fis.on) {
        fis_sample(fis, r, ux, uy, fx, fy, weight);
    } else {
        fx = sample_tent(ux, r);
        fy = sample_tent(uy, r);
    }
    float lx = sp.next1d(), ly = sp.next1d();
    (void)sp.next1d();
    float ffx = float(px) + 0.5f + fx, ffy = float(py) + 0.5f + fy;
    if (p_film) p_film[0] = ffx, p_film[1] = ffy;
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

// ---------------------------------------------------------------- ReSTIR DI
// The direct half of the real-time integrator
// (src/render_core/integrators/ReSTIR/direct.cpp): shader0 (primary surface,
// RIS over light and BSDF candidates, visibility, temporal reuse) and shader1
// (spatial reuse, shading), DIReservoir (ReSTIR/reservoir.h:15-115) and the
// point-measure light functions of LightSampler (lightsampler.cpp:218-313).
// The state lives in float4 records laid out as the device buffers:
// reservoirs [4][3][P] (weight_sum C W p_hat | light prim bary | pos age),
// surfaces [2][3][P] (tri b1 b2 flag|replaced<<16 | normal depth | pos diff),
// extensions [2][2][P] (throughput t_max | view position).
struct RParams {
    uint32_t W, H, P, frame;
    uint32_t M_light, M_bsdf, max_age, max_recursion;
    float diff;
    bool pairwise;
    bool spatial;
    uint32_t s_num;
    float s_dot, s_depth, s_radius;
    bool temporal;
    uint32_t t_limit, t_N;
    float t_radius, t_dot, t_depth;
    bool t_mis;
    uint32_t extra;
    float w2c[12], prev_w2c[12];
    V3 cam_pos, prev_cam_pos;
};

struct RState {
    float *rsv, *surf, *ext;
    uint32_t P;
    float *at(float *base, uint32_t buf, uint32_t nfield, uint32_t field, uint32_t i) const {
        return base + 4 * ((size_t(buf) * nfield + field) * P + i);
    }
};

constexpr uint32_t kFlagUnset = 1, kFlagRefl = 2, kFlagDiffuse = 8, kFlagGlossy = 16, kFlagNearSpec = 64;

struct DSample {
    uint32_t light{kInv}, prim{0};
    float bx{0.f}, by{0.f};
    uint32_t age{0};
    float p_hat{0.f};
    V3 pos{0.f, 0.f, 0.f};
};
struct Rsv {
    float ws{0.f}, C{0.f}, W{0.f};
    DSample s;
    // DIReservoir::update / truncation / process_occluded / update_W (reservoir.h:80-103)
    void update(float u, const DSample &v, float w, float c) {
        ws += w;
        C += c;
        if (u * ws < w) s = v;
    }
    void truncation(float limit) { C = std::min(limit, C); }
    void occlude(bool o) {
        W = o ? 0.f : W;
        ws = o ? 0.f : ws;
    }
    void update_W(float p) { W = p == 0.f ? 0.f : ws / p; }
};
float weight_of(float mis, float p_hat, float W) { return mis * p_hat * W; }
float safe_weight_of(float mis, float p_hat, float W) {
    float r = weight_of(mis, p_hat, W);
    return std::isnan(r) ? 0.f : r;
}

struct SRec {
    uint32_t tri{kInv};
    float b1{0.f}, b2{0.f};
    uint32_t flag{kFlagUnset}, replaced{0};
    V3 n{0.f, 0.f, 0.f};
    float depth{0.f};
    V3 pos{0.f, 0.f, 0.f};
    float diff{0.f};
    bool hit() const { return tri != kInv; }
};
struct SExt {
    V3 thr{1.f, 1.f, 1.f};
    V3 view{0.f, 0.f, 0.f};
    float t_max{0.f};
};

uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
float bitsf(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

Rsv rsv_load(const RState &st, uint32_t buf, uint32_t i) {
    const float *a = st.at(st.rsv, buf, 3, 0, i), *b = st.at(st.rsv, buf, 3, 1, i), *c = st.at(st.rsv, buf, 3, 2, i);
    Rsv r;
    r.ws = a[0], r.C = a[1], r.W = a[2], r.s.p_hat = a[3];
    r.s.light = fbits(b[0]), r.s.prim = fbits(b[1]), r.s.bx = b[2], r.s.by = b[3];
    r.s.pos = v3(c[0], c[1], c[2]);
    r.s.age = fbits(c[3]);
    return r;
}
void rsv_store(const RState &st, uint32_t buf, uint32_t i, const Rsv &r) {
    float *a = st.at(st.rsv, buf, 3, 0, i), *b = st.at(st.rsv, buf, 3, 1, i), *c = st.at(st.rsv, buf, 3, 2, i);
    a[0] = r.ws, a[1] = r.C, a[2] = r.W, a[3] = r.s.p_hat;
    b[0] = bitsf(r.s.light), b[1] = bitsf(r.s.prim), b[2] = r.s.bx, b[3] = r.s.by;
    c[0] = r.s.pos.x, c[1] = r.s.pos.y, c[2] = r.s.pos.z, c[3] = bitsf(r.s.age);
}
SRec srec_load(const RState &st, uint32_t buf, uint32_t i) {
    const float *a = st.at(st.surf, buf, 3, 0, i), *b = st.at(st.surf, buf, 3, 1, i), *c = st.at(st.surf, buf, 3, 2, i);
    SRec r;
    r.tri = fbits(a[0]), r.b1 = a[1], r.b2 = a[2];
    uint32_t fr = fbits(a[3]);
    r.flag = fr & 0xFFFFu, r.replaced = fr >> 16;
    r.n = v3(b[0], b[1], b[2]), r.depth = b[3];
    r.pos = v3(c[0], c[1], c[2]), r.diff = c[3];
    return r;
}
void srec_store(const RState &st, uint32_t buf, uint32_t i, const SRec &r) {
    float *a = st.at(st.surf, buf, 3, 0, i), *b = st.at(st.surf, buf, 3, 1, i), *c = st.at(st.surf, buf, 3, 2, i);
    a[0] = bitsf(r.tri), a[1] = r.b1, a[2] = r.b2, a[3] = bitsf(r.flag | (r.replaced << 16));
    b[0] = r.n.x, b[1] = r.n.y, b[2] = r.n.z, b[3] = r.depth;
    c[0] = r.pos.x, c[1] = r.pos.y, c[2] = r.pos.z, c[3] = r.diff;
}
SExt sext_load(const RState &st, uint32_t buf, uint32_t i) {
    const float *a = st.at(st.ext, buf, 2, 0, i), *b = st.at(st.ext, buf, 2, 1, i);
    SExt e;
    e.thr = v3(a[0], a[1], a[2]), e.t_max = a[3];
    e.view = v3(b[0], b[1], b[2]);
    return e;
}
void sext_store(const RState &st, uint32_t buf, uint32_t i, const SExt &e) {
    float *a = st.at(st.ext, buf, 2, 0, i), *b = st.at(st.ext, buf, 2, 1, i);
    a[0] = e.thr.x, a[1] = e.thr.y, a[2] = e.thr.z, a[3] = e.t_max;
    b[0] = e.view.x, b[1] = e.view.y, b[2] = e.view.z, b[3] = 0.f;
}

// the DSL's float -> integer casts, saturated
uint32_t cast_u32(float v) { return v > 0.f ? (v < 4.0e9f ? uint32_t(v) : 0xFFFFFFFFu) : 0u; }
int cast_i32(float v) { return v > -2.0e9f ? (v < 2.0e9f ? int(v) : 2000000000) : -2000000000; }
V3 affine(const float *m, V3 p) {
    return v3(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
              m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}

// luminance of f.vec3(): the first three channels
float lum3(const Spec &f, int n) { return luminance(v3(f.v[0], n > 1 ? f.v[1] : 0.f, n > 2 ? f.v[2] : 0.f)); }

// material flag and diffuse factor (MaterialEvaluator::flag / diffuse_factor,
// material.cpp:110-130; near-specular below alpha 0.022, metal.cpp:46,147-148,
// glass.cpp:178,251-252; MicrofacetLobe sqrt(alpha_average), lobe.cpp:212-214;
// DielectricLobe sqrt(alpha_x), lobe.cpp:427-429); the principled lobe set
// holds a diffuse lobe, never near-specular
uint32_t mat_flag(const Scene &s, uint32_t m) {
    if (m == kInv) return kFlagDiffuse | kFlagRefl;
    uint32_t t = s.mat_type(m);
    if (t == M_METAL || t == M_MIRROR) return std::min(s.matp(m, 6), s.matp(m, 7)) < 0.022f ? kFlagNearSpec : kFlagGlossy;
    if (t == M_GLASS) return std::min(s.matp(m, 5), s.matp(m, 6)) < 0.022f ? kFlagNearSpec : kFlagGlossy;
    if (t == M_PRINCIPLED) return kFlagDiffuse | kFlagRefl | kFlagGlossy;
    return kFlagDiffuse | kFlagRefl;
}
float mat_diffuse_factor(const Scene &s, uint32_t m) {
    if (m == kInv) return 1.f;
    uint32_t t = s.mat_type(m);
    if (t == M_METAL || t == M_MIRROR) return std::sqrt(std::sqrt(s.matp(m, 6) * s.matp(m, 7)));
    if (t == M_GLASS) return std::sqrt(s.matp(m, 5));
    return 1.f;
}

// create_evaluator at a surface, with check_dispersive (spectrum.cpp:21-39)
Bsdf bsdf_at(const Spectral &sd, const Scene &s, const Surf &it, Swl &w, bool check) {
    uint32_t m = s.inst[4 * it.inst];
    if (m == kInv) {
        Bsdf b;
        b.black = true;
        b.fr = it.fr;
        b.ng = it.ng;
        b.wo_w = it.wo;
        return b;
    }
    Bsdf b = make_bsdf(sd, s, m, it, w);
    if (check && !sd.rgb && b.dispersive) {
        bool secondary_valid = w.n > 1;
        for (int i = 1; i < w.n; ++i) secondary_valid = secondary_valid && w.pdf[i] != 0.f;
        if (secondary_valid)
            for (int i = 1; i < w.n; ++i) w.pdf[i] = 0.f;
    }
    return b;
}

// env map coordinate -> direction (SphericalMap::evaluate, spherical.cpp:104-124)
void env_dir(const EnvMap &e, float u, float v, V3 &local, V3 &world, float &sin_t) {
    float theta = v * kPi, phi = u * 2.f * kPi;
    float cos_t, sin_p, cos_p;
    vs::em::sincos(theta, &sin_t, &cos_t);
    vs::em::sincos(phi, &sin_p, &cos_p);
    local = v3(sin_t * cos_p, sin_t * sin_p, cos_t);
    world = unit(m3(e.o2w, local));
}
// 2D alias sample of the env map (alias2d.cpp:106-129)
void env_uv_sample(const EnvMap &e, float ux, float uy, float &u, float &v, float &pdf_map) {
    float uvr;
    uint32_t iv = alias_sample(e.marg, uy, &uvr);
    v = (float(iv) + uvr) / float(e.h);
    float pdf_v = e.marg.integral > 0.f ? e.marg.func[iv] / e.marg.integral : 0.f;
    const Alias &row = e.rows[iv];
    float uur;
    uint32_t iu = alias_sample(row, ux, &uur);
    u = (float(iu) + uur) / float(e.w);
    float integral_u = e.marg.func[iv];
    float pdf_u = integral_u > 0.f ? row.func[iu] / integral_u : 0.f;
    pdf_map = pdf_u * pdf_v;
}
V3 robust_pos(V3 pos, V3 ng, V3 dir, float factor) {
    float f = dotp(ng, dir) > 0.f ? 1.f : -1.f;
    return offset_origin(pos, mul(mul(ng, f), factor));
}

// LightSampler::sample_only (lightsampler.cpp:218-228): select 1, surface 2
DSample light_sample_only(const Scene &s, const LightTables &lt, Sampler &sp) {
    DSample d;
    float ul = sp.next1d();
    uint32_t l = !lt.power ? uint32_t(std::min(ul * float(s.nl), float(s.nl) - 1.f)) : alias_sample(lt.select, ul, nullptr);
    float ux = sp.next1d(), uy = sp.next1d();
    uint32_t type = s.lightu(l, 0);
    if (type == 0) {// AreaLight::sample_only (area.cpp:139-145)
        float urem;
        d.prim = alias_sample(lt.tri_alias[l], ux, &urem);
        float su = std::sqrt(urem);
        d.bx = 1.f - su, d.by = uy * su;
    } else if (type == 3) {// SphericalMap::sample_only (spherical.cpp:126-131)
        float pdf_map;
        env_uv_sample(lt.env, ux, uy, d.bx, d.by, pdf_map);
    } else {// Light::sample_only (light.h:177-182)
        d.bx = ux, d.by = uy;
    }
    d.light = l;
    return d;
}

struct PEval {
    Spec L;
    float pdf{0.f};
    V3 p{0.f, 0.f, 0.f};
};
// LightSampler::evaluate_point (lightsampler.cpp:230-240): AreaLight Le * G
// and area-measure pdf (area.cpp:96-118,157-168), point / spot Li and -1
// (light.h:237-243), environment in solid angle (spherical.cpp:104-141)
PEval light_eval_point(const Spectral &sd, const Scene &s, const LightTables &lt, V3 ref, const DSample &d,
                       const Swl &w) {
    PEval r;
    uint32_t l = d.light, type = s.lightu(l, 0);
    if (type == 0) {
        Surf p = surface(s, s.lightu(l, 3) + d.prim, d.bx, d.by, false);
        V3 wv = sub(ref, p.pos);
        Spec Le = light_color(sd, s, l, w, p.u, p.v) * ((dotp(wv, p.ng) > 0.f || s.lightu(l, 1)) ? 1.f : 0.f);
        float G = dotp(unit(wv), p.ng) / dotp(wv, wv);
        r.L = Le * G;
        r.pdf = (1.f / p.area) * alias_pmf(lt.tri_alias[l], d.prim);
        r.p = robust_pos(p.pos, p.ng, wv, s.render[5]);
    } else if (type == 3) {
        const EnvMap &e = lt.env;
        float pdf_map = env_map_pdf(e, d.bx, d.by);
        V3 local, world;
        float sin_t;
        env_dir(e, d.bx, d.by, local, world, sin_t);
        float pdf_dir = pdf_map / (2.f * kPi * kPi * sin_t);
        r.pdf = std::isinf(pdf_dir) ? 0.f : pdf_dir;
        r.L = env_L(sd, e, local, w);
        r.p = add(ref, mul(world, e.diameter));
    } else {
        V3 pos = v3(s.lightf(l, 12), s.lightf(l, 13), s.lightf(l, 14));
        V3 wu = sub(ref, pos);
        Spec val = light_color(sd, s, l, w) / dotp(wu, wu);
        if (type == 2) {
            V3 dir = v3(s.lightf(l, 16), s.lightf(l, 17), s.lightf(l, 18));
            float angle = s.lightf(l, 15), fall = s.lightf(l, 19);
            float start = std::max(0.f, angle - fall);
            float ca = std::cos(angle), cs = std::cos(start);
            float ct = std::min(std::max(dotp(dir, unit(wu)), ca), cs);
            float f = (ct - ca) / (cs - ca);
            float f2 = f * f;
            val = val * (f2 * f2);
        }
        r.L = val;
        r.pdf = -1.f;
        r.p = pos;
    }
    r.pdf *= select_pmf(s, lt, l);
    return r;
}
// LightSampler::PDF_point (lightsampler.cpp:242-250; Light::PDF_point,
// light.h:141-151; warp.h:101-105; spherical.cpp:143-146)
float light_pdf_point(const Scene &s, V3 ref, const DSample &d, float pdf_wi) {
    uint32_t l = d.light, type = s.lightu(l, 0);
    if (type == 3) return pdf_wi;
    V3 lp, ln;
    if (type == 0) {
        Surf p = surface(s, s.lightu(l, 3) + d.prim, d.bx, d.by, false);
        lp = p.pos, ln = p.ng;
    } else {
        lp = v3(s.lightf(l, 12), s.lightf(l, 13), s.lightf(l, 14));
        ln = unit(sub(ref, lp));
    }
    V3 wv = sub(ref, lp);
    float r = pdf_wi * std::fabs(dotp(ln, unit(wv))) / dotp(wv, wv);
    return std::isinf(r) ? 0.f : r;
}

// ReSTIRDI::Li from the light (direct.cpp:123-159): f * L for a sample
Spec light_contrib(const Spectral &sd, const Scene &s, const LightTables &lt, const Surf &it, const Bsdf &b,
                   const DSample &d, const Swl &w) {
    PEval le;
    if (d.light != kInv) le = light_eval_point(sd, s, lt, it.pos, d, w);
    else le.L = sfill(0.f);
    V3 wi = unit(sub(le.p, it.pos));
    BEval e = bsdf_eval(sd, b, wi, w.n, nullptr);
    return e.f * le.L;
}
float p_hat_of(const Spectral &sd, const Scene &s, const LightTables &lt, const Surf &it, const Bsdf &b,
               const DSample &d, const Swl &w) {
    return lum3(light_contrib(sd, s, lt, it, b, d, w), w.n);
}

bool valid_neighbor(const SRec &c, const SRec &o, float dt, float depth_t, float diff_t) {
    bool c0 = std::fabs(dotp(c.n, o.n)) > dt;
    bool c1 = (std::fabs(c.depth - o.depth) / c.depth) < depth_t;
    bool c2 = std::fabs(c.diff - o.diff) / c.diff < diff_t;
    return c0 && c1 && c2 && c.hit() && o.hit();
}

bool occluded_to(const Ctx &c, const Surf &it, V3 p) {
    V3 dir = sub(p, it.pos);
    V3 nn = dotp(it.ng, dir) > 0.f ? it.ng : neg(it.ng);
    return c.trace({offset_origin(it.pos, nn), dir, 0.f, 1.f - kShadowEps}, true).tri != kInv;
}

// delta sampling of a near-specular metal or glass (MaterialEvaluator::
// sample_delta, material.cpp:154-160; lobe.cpp:223-230, 404-425): the mirror
// or refracted direction and f |cos| / pdf there (1 draw for glass)
// (the lobes evaluated without energy compensation, as BxDF::evaluate is)
bool delta_sample(const Spectral &sd, const Bsdf &b, Sampler &sp, int n, V3 &wi_w, V3 &thr) {
    if (b.black || b.set || (b.lobes[0].kind != LB_REFLECTION && b.lobes[0].kind != LB_DIELECTRIC)) {
        wi_w = thr = v3(0.f, 0.f, 0.f);
        return false;
    }
    Lobe L = b.lobes[0];
    V3 wo = b.fr.local(b.wo_w);
    V3 wi = v3(-wo.x, -wo.y, wo.z);
    LEval e;
    if (L.kind == LB_DIELECTRIC) {
        Spec F = fresnel_eval(L.fr, std::fabs(wo.z), n);
        float rp = refl_prob(F, L.kr, n);
        float uc = sp.next1d();
        float eta0 = L.fr.eta.v[0];
        if (!(uc < rp)) {
            V3 nz = v3(0.f, 0.f, 1.f);
            refract_dir(wo, dotp(nz, wo) < 0.f ? neg(nz) : nz, eta0, wi);
        }
        bool same = wo.z * wi.z > 0.f;
        float eta_p = same ? 1.f : eta0;
        V3 wh = face_forward(unit(add(wo, mul(wi, eta_p))), wo);
        Spec F2 = fresnel_eval(L.fr, std::fabs(dotp(wh, wo)), n);
        float rp2 = refl_prob(F2, L.kr, n);
        if (same) {
            e.f = L.mf.brdf(wo, wh, wi) * F2;
            e.pdf = L.mf.pdf_reflection(wo, wh) * rp2;
        } else {
            V3 wh2 = unit(add(wo, mul(wi, eta0)));
            e.f = (L.mf.btdf(wo, wh2, wi, eta0) * (1.f - F2)) * L.kr;
            e.pdf = L.mf.pdf_transmission(wo, face_forward(wh, wo), wi, eta0) * (1.f - rp2);
        }
    } else {
        L.compensate = false;
        e = lobe_eval_local(sd, L, wo, wi, n, nullptr);
    }
    Spec t = (e.f * std::fabs(wi.z)) / e.pdf;
    thr = v3(t.v[0], t.v[n > 1 ? 1 : 0], t.v[n > 2 ? 2 : 0]);
    wi_w = b.fr.world(wi);
    return true;
}

// ReSTIRDI::Li from the BSDF (direct.cpp:57-121)
Spec bsdf_candidate(const Ctx &c, const Surf &it, const Bsdf &b, Sampler &sp, const Swl &w, DSample &d, float &pdf,
                    float &light_pdf, float *hit_bsdf = nullptr) {
    const Scene &s = *c.s;
    const Spectral &sd = *c.sd;
    const LightTables &lt = *c.lt;
    BSample bs = bsdf_sample(sd, b, sp, w.n);
    if (hit_bsdf) {// HitBSDF {wi, f.vec3(), pdf} (direct.cpp:82-85)
        hit_bsdf[0] = bs.wi.x, hit_bsdf[1] = bs.wi.y, hit_bsdf[2] = bs.wi.z, hit_bsdf[3] = bs.e.pdf;
        hit_bsdf[4] = bs.e.f.v[0], hit_bsdf[5] = bs.e.f.v[1], hit_bsdf[6] = bs.e.f.v[2];
    }
    V3 nn = dotp(it.ng, bs.wi) > 0.f ? it.ng : neg(it.ng);
    Hit h = c.trace({offset_origin(it.pos, nn), bs.wi, 0.f, FLT_MAX}, false);
    pdf = bs.e.pdf;
    light_pdf = 0.f;
    Spec f = sfill(0.f);
    if (h.tri != kInv) {
        Surf nx = surface(s, h.tri, h.b1, h.b2, false);
        uint32_t l = s.inst[4 * nx.inst + 1];
        if (l != kInv) {// LightSampler::evaluate_hit_point (lightsampler.cpp:269-288)
            V3 wv = sub(it.pos, nx.pos);
            Spec Le = light_color(sd, s, l, w, nx.u, nx.v) * ((dotp(wv, nx.ng) > 0.f || s.lightu(l, 1)) ? 1.f : 0.f);
            float G = dotp(unit(wv), nx.ng) / dotp(wv, wv);
            float pp = pdf * std::fabs(dotp(nx.ng, unit(wv))) / dotp(wv, wv);
            uint32_t prim = h.tri - s.lightu(l, 3);
            light_pdf = (1.f / nx.area) * alias_pmf(lt.tri_alias[l], prim) * select_pmf(s, lt, l);
            d.light = l, d.prim = prim, d.bx = h.b1, d.by = h.b2;
            d.pos = robust_pos(nx.pos, nx.ng, wv, s.render[5]);
            pdf = std::isinf(pp) ? 0.f : pp;
            f = bs.e.f * (Le * G);
        }
    } else if (lt.env.present) {// LightSampler::evaluate_miss_point (lightsampler.cpp:302-313)
        V3 wi = mul(bs.wi, lt.env.diameter);
        V3 pl = add(it.pos, wi);
        Spec Lv;
        float epdf;
        env_eval(sd, lt.env, unit(sub(pl, it.pos)), w, Lv, epdf);
        light_pdf = epdf * select_pmf(s, lt, lt.env.light);
        d.light = lt.env.light;
        env_uv(unit(m3(lt.env.w2o, bs.wi)), d.bx, d.by);
        d.pos = pl;
        f = bs.e.f * Lv;
    }
    return f;
}

// ReSTIRDI::RIS (direct.cpp:161-233)
Rsv ris(const Ctx &c, const RParams &rp, const Surf &it, Sampler &sp, Swl &w, V3 thr, float *hit_bsdf = nullptr) {
    const Scene &s = *c.s;
    const Spectral &sd = *c.sd;
    const LightTables &lt = *c.lt;
    Rsv r;
    if (rp.M_light + rp.M_bsdf == 0) return r;
    Bsdf b = bsdf_at(sd, s, it, w, true);
    const float lum = luminance(thr);
    const float Ml = float(rp.M_light), Mb = float(rp.M_bsdf);
    for (uint32_t i = 0; i < rp.M_light; ++i) {
        DSample d = light_sample_only(s, lt, sp);
        PEval le = light_eval_point(sd, s, lt, it.pos, d, w);
        V3 wi = unit(sub(le.p, it.pos));
        BEval e = bsdf_eval(sd, b, wi, w.n, nullptr);
        Spec f = e.f * le.L;
        float bp = light_pdf_point(s, it.pos, d, e.pdf);
        d.p_hat = lum3(f, w.n);
        d.pos = le.p;
        bool delta = le.pdf < 0.f;
        float lpdf = delta ? -le.pdf : le.pdf;
        float mis = delta ? 1.f / Ml : lpdf / (Ml * lpdf + Mb * bp);
        r.update(sp.next1d(), d, safe_weight_of(mis, d.p_hat, 1.f / lpdf) * lum, 1.f);
    }
    for (uint32_t i = 0; i < rp.M_bsdf; ++i) {
        DSample d;
        float pdf, lpdf;
        Spec f = bsdf_candidate(c, it, b, sp, w, d, pdf, lpdf, hit_bsdf);
        d.p_hat = lum3(f, w.n);
        r.update(sp.next1d(), d, safe_weight_of(pdf / (Ml * lpdf + Mb * pdf), d.p_hat, 1.f / pdf) * lum, 1.f);
    }
    return r;
}

// ReSTIRDI::combine_temporal (direct.cpp:327-380)
Rsv combine_temporal(const Ctx &c, const RParams &rp, const Rsv &cur, const SRec &cs, Rsv other, V3 view,
                     V3 prev_view, Sampler &sp, Swl &w) {
    const Scene &s = *c.s;
    const Spectral &sd = *c.sd;
    const LightTables &lt = *c.lt;
    other.s.age += 1;
    Surf it = surface(s, cs.tri, cs.b1, cs.b2, true);
    float mis_cur, mis_prev, p_nn;
    if (rp.t_mis) {
        it.wo = unit(sub(prev_view, it.pos));
        Bsdf bp = bsdf_at(sd, s, it, w, true);
        float p_cn = p_hat_of(sd, s, lt, it, bp, cur.s, w);
        p_nn = p_hat_of(sd, s, lt, it, bp, other.s, w);
        it.wo = unit(sub(view, it.pos));
        Bsdf bc = bsdf_at(sd, s, it, w, true);
        float p_cc = p_hat_of(sd, s, lt, it, bc, cur.s, w);
        float p_nc = p_hat_of(sd, s, lt, it, bc, other.s, w);
        float nc = float(int(cur.C)), no = float(int(other.C));
        mis_cur = (nc * p_cc) / (nc * p_cc + no * p_cn);
        mis_prev = (no * p_nn) / (no * p_nn + nc * p_nc);
    } else {
        it.wo = unit(sub(prev_view, it.pos));
        Bsdf bp = bsdf_at(sd, s, it, w, true);
        p_nn = p_hat_of(sd, s, lt, it, bp, other.s, w);
        mis_cur = cur.C / (cur.C + other.C);
        mis_prev = other.C / (other.C + cur.C);
    }
    Rsv ret;
    ret.update(0.5f, cur.s, safe_weight_of(mis_cur, cur.s.p_hat, cur.W), cur.C);
    DSample os = other.s;
    os.p_hat = p_nn;
    ret.update(sp.next1d(), os, safe_weight_of(mis_prev, os.p_hat, other.W), other.C);
    ret.update_W(ret.s.p_hat);
    return ret;
}

// ReSTIRDI::temporal_reuse (direct.cpp:382-432)
Rsv temporal_reuse(const Ctx &c, const RParams &rp, const RState &st, Rsv rsv, const SRec &cs, uint32_t pix,
                   float mx, float my, const float *p_film, Sampler &sp, Swl &w) {
    float prev_x = p_film[0] - mx, prev_y = p_film[1] - my;
    uint32_t pslot = rp.frame & 1u, cslot = (rp.frame + 1u) & 1u;
    float limit = rsv.C * float(rp.t_limit);
    V3 view = cs.replaced ? sext_load(st, cslot, pix).view : rp.cam_pos;
    V3 prev_view = rp.prev_cam_pos;
    int ix = cast_i32(prev_x), iy = cast_i32(prev_y);
    if (!(ix >= 0 && iy >= 0 && ix < int(rp.W) && iy < int(rp.H) && rp.temporal)) return rsv;
    if (!cs.hit()) return rsv;
    auto visit = [&](uint32_t j, Rsv &pr, SRec &ps) {
        pr = rsv_load(st, pslot, j);
        pr.truncation(limit);
        ps = srec_load(st, pslot, j);
        if (ps.replaced) prev_view = sext_load(st, pslot, j).view;
    };
    Rsv pr;
    SRec ps;
    visit(std::min(cast_u32(prev_y), rp.H - 1) * rp.W + std::min(cast_u32(prev_x), rp.W - 1), pr, ps);
    if (valid_neighbor(cs, ps, rp.t_dot, rp.t_depth, rp.diff) && pr.s.age < rp.max_age)
        return combine_temporal(c, rp, rsv, cs, pr, view, prev_view, sp, w);
    for (uint32_t i = 0; i < rp.t_N; ++i) {
        float ux = sp.next1d(), uy = sp.next1d();
        float dx, dy;
        square_to_disk(ux, uy, dx, dy);
        float qx = dx * rp.t_radius + prev_x, qy = dy * rp.t_radius + prev_y;
        visit(std::min(cast_u32(qy), rp.H - 1) * rp.W + std::min(cast_u32(qx), rp.W - 1), pr, ps);
        if (valid_neighbor(cs, ps, rp.t_dot, rp.t_depth, rp.diff) && pr.s.age < rp.max_age)
            return combine_temporal(c, rp, rsv, cs, pr, view, prev_view, sp, w);
    }
    return rsv;
}

// Sensor::prev_raster_coord (sensor.cpp:105-110): the point in the previous
// camera's space divided by its depth, mapped to raster coordinates
void prev_raster(const float *cam, const float *prev_w2c, V3 p, float &rx, float &ry) {
    V3 q = affine(prev_w2c, p);
    float qx = q.x / q.z, qy = q.y / q.z;
    float sx = qx / cam[16], sy = qy / cam[16];
    rx = (sx - cam[12]) / cam[14];
    ry = (cam[13] - sy) / cam[15];
}

// shader0 for one pixel (direct.cpp:519-555; compute_hit :434-517)
void restir_initial_px(const Ctx &c, const FisTable &fis, const RParams &rp, const RState &st, uint32_t x,
                       uint32_t y, float *aux = nullptr) {
    const Scene &s = *c.s;
    const Spectral &sd = *c.sd;
    const uint32_t pix = y * rp.W + x, cslot = (rp.frame + 1u) & 1u;
    Sampler sp;
    sp.start(x, y, rp.frame, 0u);
    Swl w = sample_wavelengths(sd, x, y, rp.frame);
    V3 o, d;
    float fw, pf[2];
    camera_ray(s, fis, x, y, sp, o, d, fw, pf);
    SRec cs;
    SExt ext;
    ext.view = o;
    V3 ro = o, rd = d;
    Hit h = c.trace({ro, rd, 0.f, FLT_MAX}, false);
    Surf it{};
    uint32_t counter = 0;
    while (true) {
        cs.tri = h.tri, cs.b1 = h.b1, cs.b2 = h.b2;
        if (h.tri == kInv) break;
        it = surface(s, h.tri, h.b1, h.b2, true);
        it.wo = unit(neg(rd));
        ext.t_max += len(sub(it.pos, ro)) / len(rd);
        V3 vpos = add(o, mul(d, ext.t_max));
        uint32_t m = s.inst[4 * it.inst];
        cs.flag = mat_flag(s, m);
        cs.diff = mat_diffuse_factor(s, m);
        V3 wnext = v3(0.f, 0.f, 0.f);
        if (cs.flag == kFlagNearSpec) {
            Bsdf b = bsdf_at(sd, s, it, w, false);
            V3 t;
            delta_sample(sd, b, sp, w.n, wnext, t);
            ext.thr = v3(ext.thr.x * t.x, ext.thr.y * t.y, ext.thr.z * t.z);
        }
        if (counter == 0) {
            cs.depth = affine(rp.w2c, vpos).z;
            cs.n = it.fr.z;
            cs.pos = it.pos;
        }
        ++counter;
        if (counter >= rp.max_recursion) break;
        if (cs.flag != kFlagNearSpec) break;
        ext.view = it.pos;
        V3 nn = dotp(it.ng, wnext) > 0.f ? it.ng : neg(it.ng);
        ro = offset_origin(it.pos, nn);
        rd = wnext;
        h = c.trace({ro, rd, 0.f, FLT_MAX}, false);
        cs.replaced = 1;
    }
    bool hit = h.tri != kInv;
    srec_store(st, cslot, pix, cs);
    if (cs.replaced) sext_store(st, cslot, pix, ext);
    Rsv r;
    // the HitBSDF of the last BSDF candidate (for ReSTIR GI; its default
    // when there is none: wi 0, f 0, pdf -1, interaction.h:60-66)
    float hb[7] = {0.f, 0.f, 0.f, -1.f, 0.f, 0.f, 0.f};
    if (hit) r = ris(c, rp, it, sp, w, ext.thr, hb);
    r.update_W(r.s.p_hat);
    r.truncation(1.f);
    float mx = 0.f, my = 0.f;
    if (hit) {// FrameBuffer::compute_motion_vec (frame_buffer.cpp:325-333)
        float rx, ry;
        prev_raster(s.cam, rp.prev_w2c, add(o, mul(d, ext.t_max)), rx, ry);
        mx = pf[0] - rx, my = pf[1] - ry;
    }
    if (aux) {// [3][P]: wi pdf | f 0 | motion 0 0
        float *a0 = aux + 4 * size_t(pix), *a1 = aux + 4 * (size_t(rp.P) + pix), *a2 = aux + 4 * (2 * size_t(rp.P) + pix);
        a0[0] = hb[0], a0[1] = hb[1], a0[2] = hb[2], a0[3] = hb[3];
        a1[0] = hb[4], a1[1] = hb[5], a1[2] = hb[6], a1[3] = 0.f;
        a2[0] = mx, a2[1] = my, a2[2] = 0.f, a2[3] = 0.f;
    }
    if (hit) r.occlude(occluded_to(c, it, r.s.pos));
    r = temporal_reuse(c, rp, st, r, cs, pix, mx, my, pf, sp, w);
    rsv_store(st, 2, pix, r);
}

uint32_t spatial_pick(const RParams &rp, uint32_t x, uint32_t y, Sampler &sp) {
    float ux = sp.next1d(), uy = sp.next1d();
    float dx, dy;
    square_to_disk(ux, uy, dx, dy);
    float ox = dx * rp.s_radius, oy = dy * rp.s_radius;
    int ax = std::min(std::max(int(x) + int(std::round(ox)), 0), int(rp.W) - 1);
    int ay = std::min(std::max(int(y) + int(std::round(oy)), 0), int(rp.H) - 1);
    return uint32_t(ay) * rp.W + uint32_t(ax);
}

// ReSTIRDI::spatial_reuse with constant_combine / pairwise_combine
// (direct.cpp:559-581, 287-309, 264-285, 235-262)
Rsv spatial_reuse(const Ctx &c, const RParams &rp, const RState &st, Rsv r, const SRec &cs, uint32_t x, uint32_t y,
                  uint32_t buf, Sampler &sp, Swl &w) {
    if (!rp.spatial) return r;
    const Scene &s = *c.s;
    const Spectral &sd = *c.sd;
    const LightTables &lt = *c.lt;
    const uint32_t cslot = (rp.frame + 1u) & 1u, pix = y * rp.W + x;
    std::vector<uint32_t> idx;
    for (uint32_t i = 0; i < rp.s_num; ++i) {
        uint32_t j = spatial_pick(rp, x, y, sp);
        if (valid_neighbor(cs, srec_load(st, cslot, j), rp.s_dot, rp.s_depth, rp.diff)) idx.push_back(j);
    }
    if (!cs.hit()) return r;
    V3 view = cs.replaced ? sext_load(st, cslot, pix).view : rp.cam_pos;
    Surf cit = surface(s, cs.tri, cs.b1, cs.b2, true);
    cit.wo = unit(sub(view, cit.pos));
    Bsdf cb;
    if (!idx.empty()) cb = bsdf_at(sd, s, cit, w, true);
    Rsv ret;
    if (!rp.pairwise) {
        float inv = 1.f / float(uint32_t(idx.size()) + 1u);
        ret.update(0.5f, r.s, weight_of(inv, r.s.p_hat, r.W), r.C);
        for (uint32_t j : idx) {
            Rsv nr = rsv_load(st, buf, j);
            nr.s.p_hat = p_hat_of(sd, s, lt, cit, cb, nr.s, w);
            ret.update(sp.next1d(), nr.s, weight_of(inv, nr.s.p_hat, nr.W), nr.C);
        }
    } else {
        float num = float(idx.size()), num_i = float(int(num));
        float canonical = 0.f;
        for (uint32_t j : idx) {
            Rsv nr = rsv_load(st, buf, j);
            SRec ns = srec_load(st, cslot, j);
            Surf nit = surface(s, ns.tri, ns.b1, ns.b2, true);
            nit.wo = unit(sub(view, nit.pos));
            Bsdf nb = bsdf_at(sd, s, nit, w, true);
            float p_cc = p_hat_of(sd, s, lt, cit, cb, r.s, w);
            float p_cn = p_hat_of(sd, s, lt, nit, nb, r.s, w);
            float p_nn = p_hat_of(sd, s, lt, nit, nb, nr.s, w);
            float p_nc = p_hat_of(sd, s, lt, cit, cb, nr.s, w);
            float mi = (1.f * p_nn) / (1.f * p_nn + num_i * p_nc);
            ret.update(sp.next1d(), nr.s, safe_weight_of(mi, nr.s.p_hat, nr.W), nr.C);
            float cw = ((1.f * p_cc) / (1.f * p_cc + num_i * p_cn)) / num;
            canonical += std::isnan(cw) ? 0.f : cw;
        }
        canonical = canonical == 0.f ? 1.f : canonical;
        ret.update(sp.next1d(), r.s, safe_weight_of(canonical, r.s.p_hat, r.W), r.C);
    }
    ret.update_W(ret.s.p_hat);
    return ret;
}

// ReSTIRDI::shading (direct.cpp:583-636); the dead BSDF-sampled emitter term
// draws after every other use of the sampler and is left out
V3 restir_shading(const Ctx &c, const RParams &rp, const RState &st, Rsv r, const SRec &cs, uint32_t pix, Swl &w) {
    const Scene &s = *c.s;
    const Spectral &sd = *c.sd;
    const LightTables &lt = *c.lt;
    V3 view = rp.cam_pos, thr = v3(1.f, 1.f, 1.f);
    if (cs.replaced) {
        SExt e = sext_load(st, (rp.frame + 1u) & 1u, pix);
        thr = e.thr, view = e.view;
    }
    Surf it = surface(s, cs.tri, cs.b1, cs.b2, true);
    it.wo = unit(sub(view, it.pos));
    Spec value = sfill(0.f), Le = sfill(0.f);
    uint32_t l = s.inst[4 * it.inst + 1];
    if (l != kInv) {
        if (s.lightu(l, 0) == 0) {
            V3 wv = sub(rp.cam_pos, it.pos);
            Le = light_color(sd, s, l, w, it.u, it.v) * ((dotp(wv, it.ng) > 0.f || s.lightu(l, 1)) ? 1.f : 0.f);
        }
    } else {
        Bsdf b = bsdf_at(sd, s, it, w, true);
        Spec t;
        for (int i = 0; i < w.n; ++i) t.v[i] = i == 0 ? thr.x : (i == 1 ? thr.y : thr.z);
        value = (light_contrib(sd, s, lt, it, b, r.s, w) * t) / luminance(thr);
        r.occlude(occluded_to(c, it, r.s.pos));
        value = value * r.W;
    }
    return linear_srgb(sd, value + Le, w);
}

// shader1 for one pixel (direct.cpp:638-671); returns the direct radiance
V3 restir_final_px(const Ctx &c, const FisTable &fis, const RParams &rp, const RState &st, uint32_t x, uint32_t y,
                   uint32_t buf) {
    const Scene &s = *c.s;
    const Spectral &sd = *c.sd;
    const LightTables &lt = *c.lt;
    const uint32_t pix = y * rp.W + x, cslot = (rp.frame + 1u) & 1u;
    Sampler sp;
    sp.start(x, y, rp.frame, 0u);
    Swl w = sample_wavelengths(sd, x, y, rp.frame);
    V3 o, d;
    float fw;
    camera_ray(s, fis, x, y, sp, o, d, fw);
    sp.start(x, y, rp.frame, 1u);
    SRec cs = srec_load(st, cslot, pix);
    Rsv r = rsv_load(st, buf, pix);
    r = spatial_reuse(c, rp, st, r, cs, x, y, buf, sp, w);
    V3 L = v3(0.f, 0.f, 0.f);
    if (cs.hit()) {
        L = restir_shading(c, rp, st, r, cs, pix, w);
    } else if (lt.env.present) {
        Spec Lv;
        float pdf;
        env_eval(sd, lt.env, unit(sub(add(o, d), o)), w, Lv, pdf);
        L = linear_srgb(sd, Lv, w);
    }
    rsv_store(st, cslot, pix, r);
    return L;
}

void restir_spatial_px(const Ctx &c, const RParams &rp, const RState &st, uint32_t x, uint32_t y, uint32_t in,
                       uint32_t out, uint32_t dim) {
    const uint32_t pix = y * rp.W + x;
    Sampler sp;
    sp.start(x, y, rp.frame, dim);
    Swl w = sample_wavelengths(*c.sd, x, y, rp.frame);
    SRec cs = srec_load(st, (rp.frame + 1u) & 1u, pix);
    Rsv r = rsv_load(st, in, pix);
    r = spatial_reuse(c, rp, st, r, cs, x, y, in, sp, w);
    rsv_store(st, out, pix, r);
}

// ---------------------------------------------------------------- ReSTIR GI
// The indirect half of the real-time integrator
// (src/render_core/integrators/ReSTIR/indirect.cpp): an initial sample per
// pixel (a path from the primary surface along the direction of ReSTIR DI's
// BSDF candidate), temporal reuse, spatial reuse with visibility, shading;
// GIReservoir / GISample (ReSTIR/reservoir.h:157-222).  State as the device
// buffers: GI reservoirs [3][4][P] (weight_sum C W age | sample position |
// sample normal | Lo), samples [4][P] (the same four fields), and the
// per-pixel record ReSTIR DI leaves for it [3][P]: HitBSDF (wi, pdf | f) and
// the motion vector.
struct GParams {
    bool open;
    uint32_t max_age;
    float diff;
    bool spatial;
    uint32_t s_num;
    float s_dot, s_depth, s_radius;
    bool temporal;
    uint32_t t_limit, t_N;
    float t_dot, t_depth, t_radius;
    int depth;// the real-time integrator's max_depth - 1 (rt.cpp:30)
};
struct GState {
    float *rsv, *smp, *aux, *direct, *indirect;
    uint32_t P;
    float *at(float *base, uint32_t buf, uint32_t nfield, uint32_t field, uint32_t i) const {
        return base + 4 * ((size_t(buf) * nfield + field) * P + i);
    }
};
struct GSample {
    V3 pos{0.f, 0.f, 0.f}, n{0.f, 0.f, 0.f}, Lo{0.f, 0.f, 0.f};
    float age{0.f};
};
struct GRsv {
    float ws{0.f}, C{0.f}, W{0.f};
    GSample s;
    // GIReservoir::update / truncation / cal_W (reservoir.h:196-219)
    void update(float u, const GSample &v, float w, float c = 1.f) {
        ws += w;
        C += c;
        if (u * ws < w || ws == 0.f) s = v;
    }
    void truncation(float limit) {
        float factor = limit / C;
        C = std::min(limit, C);
        ws = factor < 1.f ? ws * factor : ws;
    }
    void update_W(float p) { W = p == 0.f ? 0.f : ws / (p * C); }
};
GSample gsample_load(const GState &g, float *base, uint32_t buf, uint32_t i, float *age_out = nullptr) {
    const float *a = g.at(base, buf, 4, 0, i), *b = g.at(base, buf, 4, 1, i), *c = g.at(base, buf, 4, 2, i),
                *d = g.at(base, buf, 4, 3, i);
    GSample s;
    s.age = a[3];
    s.pos = v3(b[0], b[1], b[2]);
    s.n = v3(c[0], c[1], c[2]);
    s.Lo = v3(d[0], d[1], d[2]);
    if (age_out) *age_out = a[3];
    return s;
}
GRsv grsv_load(const GState &g, uint32_t buf, uint32_t i) {
    GRsv r;
    r.s = gsample_load(g, g.rsv, buf, i);
    const float *a = g.at(g.rsv, buf, 4, 0, i);
    r.ws = a[0], r.C = a[1], r.W = a[2];
    return r;
}
void gfields_store(const GState &g, float *base, uint32_t buf, uint32_t i, float f0, float f1, float f2,
                   const GSample &s) {
    float *a = g.at(base, buf, 4, 0, i), *b = g.at(base, buf, 4, 1, i), *c = g.at(base, buf, 4, 2, i),
          *d = g.at(base, buf, 4, 3, i);
    a[0] = f0, a[1] = f1, a[2] = f2, a[3] = s.age;
    b[0] = s.pos.x, b[1] = s.pos.y, b[2] = s.pos.z, b[3] = 0.f;
    c[0] = s.n.x, c[1] = s.n.y, c[2] = s.n.z, c[3] = 0.f;
    d[0] = s.Lo.x, d[1] = s.Lo.y, d[2] = s.Lo.z, d[3] = 0.f;
}
void grsv_store(const GState &g, uint32_t buf, uint32_t i, const GRsv &r) { gfields_store(g, g.rsv, buf, i, r.ws, r.C, r.W, r.s); }

// Pipeline::compute_surface_interaction(hit, view_pos): the surface seen
// from view_pos
Surf surface_from(const Scene &s, const SRec &cs, V3 view) {
    Surf it = surface(s, cs.tri, cs.b1, cs.b2, true);
    it.wo = unit(sub(view, it.pos));
    return it;
}

// ReSTIRGI::eval_bsdf / compute_p_hat (indirect.cpp:112-130): f |cos| of the
// surface's material towards the sample point, luminance of Lo f (sRGB)
V3 gi_bsdf(const Ctx &c, const Surf &it, const GSample &g, Swl &w) {
    const Scene &s = *c.s;
    if (s.inst[4 * it.inst] == kInv) return v3(0.f, 0.f, 0.f);
    Bsdf b = bsdf_at(*c.sd, s, it, w, true);
    BEval e = bsdf_eval(*c.sd, b, unit(sub(g.pos, it.pos)), w.n, nullptr);
    return v3(e.f.v[0], e.f.v[1], e.f.v[2]);
}
float gi_p_hat(const Ctx &c, const Surf &it, const GSample &g, Swl &w) {
    V3 f = gi_bsdf(c, it, g, w);
    return luminance(v3(g.Lo.x * f.x, g.Lo.y * f.y, g.Lo.z * f.z));
}

// ReSTIRGI::Jacobian_det (indirect.cpp:15-31)
float gi_jacobian(V3 cur, V3 nb, const GSample &g) {
    V3 cv = sub(cur, g.pos), nv = sub(nb, g.pos);
    float cpc = std::fabs(dotp(unit(cv), g.n)), cpn = std::fabs(dotp(unit(nv), g.n));
    float r = (cpc * dotp(nv, nv)) / (cpn * dotp(cv, cv));
    r = zero_nan_inf(r);
    const float lower = 0.6f;
    return std::min(std::max(r, lower), 1.f / lower);
}

// the initial sample (ReSTIRGI::init_sample / initial_samples,
// indirect.cpp:65-108): the path from the primary surface along the DI BSDF
// candidate's direction, its radiance divided by that candidate's
// throughput, the bounce-0 surface as the sample point
void gi_initial_px(const Ctx &c, const RParams &rp, const GParams &gp, const RState &st, const GState &g,
                   uint32_t x, uint32_t y) {
    const Scene &s = *c.s;
    const uint32_t pix = y * rp.W + x, cslot = (rp.frame + 1u) & 1u;
    float *rad = g.indirect + 4 * size_t(pix);
    SRec cs = srec_load(st, cslot, pix);
    if (!cs.hit()) {
        rad[0] = rad[1] = rad[2] = rad[3] = 0.f;
        return;
    }
    Swl w = sample_wavelengths(*c.sd, x, y, rp.frame);
    SExt ext;
    if (cs.replaced) ext = sext_load(st, cslot, pix);
    V3 view = cs.replaced ? ext.view : rp.cam_pos;
    Surf it = surface_from(s, cs, view);
    const float *h0 = g.at(g.aux, 0, 3, 0, pix), *h1 = g.at(g.aux, 0, 3, 1, pix);
    V3 wi = v3(h0[0], h0[1], h0[2]);
    const float pdf = h0[3];
    V3 t = v3(zero_nan_inf(h1[0] / pdf), zero_nan_inf(h1[1] / pdf), zero_nan_inf(h1[2] / pdf));
    Sampler sp;
    sp.start(x, y, rp.frame, 3u);
    V3 nn = dotp(it.ng, wi) > 0.f ? it.ng : neg(it.ng);
    Spec thr = sfill(0.f);
    thr.v[0] = t.x, thr.v[1] = t.y, thr.v[2] = t.z;
    FirstHit fh;
    V3 L = Li(c, offset_origin(it.pos, nn), wi, sp, w, &thr, pdf, gp.depth, &fh);
    GSample smp;
    smp.Lo = v3(zero_nan_inf(L.x / t.x), zero_nan_inf(L.y / t.y), zero_nan_inf(L.z / t.z));
    if (fh.set) smp.pos = fh.pos, smp.n = fh.ng;
    if (cs.replaced) smp.Lo = v3(smp.Lo.x * ext.thr.x, smp.Lo.y * ext.thr.y, smp.Lo.z * ext.thr.z);
    gfields_store(g, g.smp, 0, pix, 0.f, 0.f, 0.f, smp);
}

// ReSTIRGI::combine_temporal (indirect.cpp:132-152)
GRsv gi_combine_temporal(const Ctx &c, const GRsv &cur, const SRec &cs, GRsv other, const SRec *nb, V3 view,
                         V3 prev_view, Sampler &sp, Swl &w) {
    const Scene &s = *c.s;
    other.s.age += 1.f;
    Surf it = surface_from(s, cs, view);
    GRsv ret;
    float cur_p = gi_p_hat(c, it, cur.s, w);
    ret.update(sp.next1d(), cur.s, safe_weight_of(cur.C, cur_p, cur.W));
    float other_p = gi_p_hat(c, it, other.s, w);
    ret.update(sp.next1d(), other.s, safe_weight_of(other.C, other_p, other.W), other.C);
    float p = gi_p_hat(c, it, ret.s, w);
    if (nb) p = p * gi_jacobian(it.pos, surface_from(s, *nb, prev_view).pos, other.s);
    ret.update_W(p);
    return ret;
}

// the temporal pass (compile_temporal_reuse / temporal_reuse,
// indirect.cpp:154-238): the initial sample as a reservoir, combined with the
// reprojected pixel's previous reservoir, or failing that with one of up to
// temporal.N random pixels within temporal.radius of it (2 draws each)
void gi_temporal_px(const Ctx &c, const FisTable &fis, const RParams &rp, const GParams &gp, const RState &st,
                    const GState &g, uint32_t x, uint32_t y) {
    const Scene &s = *c.s;
    const uint32_t pix = y * rp.W + x, cslot = (rp.frame + 1u) & 1u, pslot = rp.frame & 1u;
    SRec cs = srec_load(st, cslot, pix);
    if (!cs.hit()) return;
    Swl w = sample_wavelengths(*c.sd, x, y, rp.frame);
    float pf[2];
    {
        Sampler tmp;
        tmp.start(x, y, rp.frame, 0u);
        V3 o, d;
        float fw;
        camera_ray(s, fis, x, y, tmp, o, d, fw, pf);
    }
    Sampler sp;
    sp.start(x, y, rp.frame, 4u);
    GSample smp = gsample_load(g, g.smp, 0, pix);
    const float *h0 = g.at(g.aux, 0, 3, 0, pix), *h1 = g.at(g.aux, 0, 3, 1, pix), *mv = g.at(g.aux, 0, 3, 2, pix);
    GRsv rsv;
    float p_hat = luminance(v3(smp.Lo.x * h1[0], smp.Lo.y * h1[1], smp.Lo.z * h1[2]));
    rsv.update(0.5f, smp, safe_weight_of(1.f, p_hat, 1.f / h0[3]));
    rsv.update_W(p_hat);
    const float prev_x = pf[0] - mv[0], prev_y = pf[1] - mv[1];
    const float limit = rsv.C * float(gp.t_limit);
    V3 view = cs.replaced ? sext_load(st, cslot, pix).view : rp.cam_pos;
    V3 prev_view = rp.prev_cam_pos;
    int ix = cast_i32(prev_x), iy = cast_i32(prev_y);
    if (ix >= 0 && iy >= 0 && ix < int(rp.W) && iy < int(rp.H) && gp.temporal) {
        auto visit = [&](uint32_t j, GRsv &pr, SRec &ps) {
            pr = grsv_load(g, pslot, j);
            pr.truncation(limit);
            ps = srec_load(st, pslot, j);
            if (ps.replaced) prev_view = sext_load(st, pslot, j).view;
        };
        auto valid = [&](const SRec &ps, const GRsv &pr) {
            return pr.s.age < float(gp.max_age) && valid_neighbor(cs, ps, gp.t_dot, gp.t_depth, gp.diff);
        };
        GRsv pr;
        SRec ps;
        visit(std::min(cast_u32(prev_y), rp.H - 1) * rp.W + std::min(cast_u32(prev_x), rp.W - 1), pr, ps);
        if (valid(ps, pr)) {
            rsv = gi_combine_temporal(c, rsv, cs, pr, nullptr, view, prev_view, sp, w);
        } else {
            for (uint32_t i = 0; i < gp.t_N; ++i) {
                float ux = sp.next1d(), uy = sp.next1d();
                float dx, dy;
                square_to_disk(ux, uy, dx, dy);
                float qx = dx * gp.t_radius + prev_x, qy = dy * gp.t_radius + prev_y;
                GRsv ar;
                SRec as;
                visit(std::min(cast_u32(qy), rp.H - 1) * rp.W + std::min(cast_u32(qx), rp.W - 1), ar, as);
                if (valid(as, ar)) {
                    rsv = gi_combine_temporal(c, rsv, cs, ar, &as, view, prev_view, sp, w);
                    break;
                }
            }
        }
    }
    grsv_store(g, 2, pix, rsv);
}

// spatial reuse with visibility, then shading (compile_spatial_shading,
// spatial_reuse, constant_combine, shading; indirect.cpp:240-332)
void gi_spatial_px(const Ctx &c, const RParams &rp, const GParams &gp, const RState &st, const GState &g, uint32_t x,
                   uint32_t y) {
    const Scene &s = *c.s;
    const uint32_t pix = y * rp.W + x, cslot = (rp.frame + 1u) & 1u;
    SRec cs = srec_load(st, cslot, pix);
    if (!cs.hit()) return;
    Swl w = sample_wavelengths(*c.sd, x, y, rp.frame);
    Sampler sp;
    sp.start(x, y, rp.frame, 5u);
    GRsv rsv = grsv_load(g, 2, pix);
    V3 view = rp.cam_pos;
    if (cs.replaced) view = sext_load(st, cslot, pix).view;
    if (gp.spatial) {
        uint32_t idx[64];
        uint32_t cnt = 0;
        for (uint32_t i = 0; i < gp.s_num; ++i) {
            float ux = sp.next1d(), uy = sp.next1d();
            float dx, dy;
            square_to_disk(ux, uy, dx, dy);
            float ox = dx * gp.s_radius, oy = dy * gp.s_radius;
            int ax = std::min(std::max(int(x) + int(std::round(ox)), 0), int(rp.W) - 1);
            int ay = std::min(std::max(int(y) + int(std::round(oy)), 0), int(rp.H) - 1);
            uint32_t j = uint32_t(ay) * rp.W + uint32_t(ax);
            if (valid_neighbor(cs, srec_load(st, cslot, j), gp.s_dot, gp.s_depth, gp.diff)) idx[cnt++] = j;
        }
        Surf cit = surface_from(s, cs, view);
        GRsv ret = rsv;
        for (uint32_t k = 0; k < cnt; ++k) {
            GRsv nr = grsv_load(g, 2, idx[k]);
            if (!(luminance(nr.s.Lo) > 0.f)) continue;
            SRec ns = srec_load(st, cslot, idx[k]);
            Surf nit = surface_from(s, ns, view);
            float p = gi_p_hat(c, cit, nr.s, w) * gi_jacobian(cit.pos, nit.pos, nr.s);
            float v = occluded_to(c, cit, nr.s.pos) ? 0.f : 1.f;
            float wt = safe_weight_of(nr.C, p, nr.W);
            ret.update(sp.next1d(), nr.s, wt * v, nr.C * v);
        }
        ret.update_W(gi_p_hat(c, cit, ret.s, w));
        rsv = ret;
    }
    // ReSTIRGI::shading: Lo f W (the near-specular chain's throughput is read
    // but unused there, indirect.cpp:305-320)
    Surf it = surface_from(s, cs, view);
    V3 f = gi_bsdf(c, it, rsv.s, w);
    float *rad = g.indirect + 4 * size_t(pix);
    rad[0] = rsv.s.Lo.x * f.x * rsv.W;
    rad[1] = rsv.s.Lo.y * f.y * rsv.W;
    rad[2] = rsv.s.Lo.z * f.z * rsv.W;
    rad[3] = 1.f;
    grsv_store(g, (rp.frame + 1u) & 1u, pix, rsv);
}

// Lobe::albedo of each lobe kind: MicrofacetLobe (lobe.cpp:208-210: kr x
// F(cos)), CoatLobe / SpecularLobe (principled_bsdf.cpp:154-160, :198-205:
// the table value s, then s kr and (F0 + s (1 - F0)) kr), DielectricLobe
// (lobe.cpp:308-313: kt (1 - F(|cos|)) + F), DiffuseLobe / SheenLTC (their
// R / tint)
Spec lobe_albedo(const Spectral &sd, const Lobe &L, float cos_theta, int n) {
    switch (L.kind) {
        case LB_REFLECTION: {
            if (L.albedo_lut != 0) {
                float sv;
                lut3(L.albedo_lut == 1 ? sd.lut_coat : sd.lut_spec, 32, 1, L.lut_x, cos_theta, L.lut_z, &sv);
                if (L.albedo_lut == 1) return sv * L.kr;
                return (L.fr.F0 + sv * (1.f - L.fr.F0)) * L.kr;
            }
            return L.kr * fresnel_eval(L.fr, cos_theta, n);
        }
        case LB_DIELECTRIC: {
            Spec F = fresnel_eval(L.fr, std::fabs(cos_theta), n);
            return L.kr * (1.f - F) + F;
        }
        default: return L.kr;
    }
}

// MaterialEvaluator::albedo (material.cpp:91-98): cos_theta against the
// shading normal; a lobe set sums its lobes' albedos times their evaluation
// weights (LobeSet::albedo, lobe.cpp:564-570); the black body's is zero
Spec bsdf_albedo(const Spectral &sd, const Bsdf &b, V3 wo_w, int n) {
    if (b.black) return sfill(0.f);
    const float c = dotp(b.fr.z, wo_w);
    if (!b.set) return lobe_albedo(sd, b.lobes[0], c, n);
    Spec r = sfill(0.f);
    for (int i = 0; i < b.nlobes; ++i) r = r + lobe_albedo(sd, b.lobes[i], c, n) * b.lobes[i].weight;
    return r;
}

// FrameBuffer::compile_compute_geom's kernel (frame_buffer.cpp:72-122) for
// one pixel: the camera ray of the pixel's first sampler dimension, its
// closest hit's shading normal and camera-space depth (Sensor::linear_depth,
// sensor.cpp:198-201), the material albedo and the emitted radiance toward
// the camera in linear sRGB.  geo: the PixelGeometry record of 8 floats
// (fwidth and the depth gradient are filled by or_gbuffer_gradient).
void gbuffer_px(const Ctx &c, const FisTable &fis, const RParams &rp, uint32_t x, uint32_t y, float *geo, float *alb,
                float *emi) {
    const Scene &s = *c.s;
    const Spectral &sd = *c.sd;
    const size_t pix = size_t(y) * rp.W + x;
    Swl w = sample_wavelengths(sd, x, y, rp.frame);
    Sampler sp;
    sp.start(x, y, rp.frame, 0u);
    V3 o, d;
    float fw, pf[2];
    camera_ray(s, fis, x, y, sp, o, d, fw, pf);
    Hit h = c.trace({o, d, 0.f, FLT_MAX}, false);
    float *g = geo + 8 * pix;
    g[0] = g[1] = g[2] = g[3] = 0.f;
    g[4] = pf[0], g[5] = pf[1], g[6] = 0.f, g[7] = -1.f;
    V3 a = v3(0.f, 0.f, 0.f), e = v3(0.f, 0.f, 0.f);
    if (h.tri != kInv) {
        Surf it = surface(s, h.tri, h.b1, h.b2, true);
        it.wo = unit(neg(d));
        g[0] = it.fr.z.x, g[1] = it.fr.z.y, g[2] = it.fr.z.z;
        g[7] = affine(rp.w2c, it.pos).z;
        const uint32_t m = s.inst[4 * it.inst], light = s.inst[4 * it.inst + 1];
        if (m != kInv) {
            Bsdf b = make_bsdf(sd, s, m, it, w);
            a = linear_srgb(sd, bsdf_albedo(sd, b, it.wo, w.n), w);
        }
        if (light != kInv) {
            Spec Le;
            float lpdf;
            eval_hit(sd, s, *c.lt, it, light, h.tri, o, w, Le, lpdf);
            e = linear_srgb(sd, Le, w);
        }
    }
    alb[4 * pix] = a.x, alb[4 * pix + 1] = a.y, alb[4 * pix + 2] = a.z, alb[4 * pix + 3] = 1.f;
    emi[4 * pix] = e.x, emi[4 * pix + 1] = e.y, emi[4 * pix + 2] = e.z, emi[4 * pix + 3] = 1.f;
}

}// namespace

extern "C" {

static float tone_map(float v, uint32_t tm);
void or_gbuffer_gradient(float *gbuf, uint32_t W, uint32_t H);
void or_variance_stats_add(float *stats, const float *x, uint32_t n);

// Spectral inputs of or_render / or_bsdf_*; field order mirrored by
// oracle/oracle.py (OrSpectral).
typedef struct or_spectral {
    uint32_t rgb;
    uint32_t ns;
    const float *spd;
    float cie_y_integral;
    const float *lut_diel, *lut_diel_inv, *lut_spec, *lut_coat, *lut_pure, *lut_sheen;
    const float *rgb2spec;
    uint32_t env_light, env_w, env_h;
    const float *env_rgb;
    float env_w2o[9], env_o2w[9];
    float env_scale, world_radius;
    const float *tex;
    const uint32_t *tex_info;
    uint32_t ntex;
} or_spectral;

static Spectral to_spectral(const or_spectral *x) {
    Spectral sd;
    if (!x) return sd;
    sd.rgb = x->rgb != 0;
    sd.ns = sd.rgb ? 3 : int(x->ns);
    sd.spd = x->spd;
    sd.cie_y_integral = x->cie_y_integral;
    sd.lut_diel = x->lut_diel;
    sd.lut_diel_inv = x->lut_diel_inv;
    sd.lut_spec = x->lut_spec;
    sd.lut_coat = x->lut_coat;
    sd.lut_pure = x->lut_pure;
    sd.lut_sheen = x->lut_sheen;
    sd.rgb2spec = x->rgb2spec;
    sd.tex = x->tex;
    sd.tex_info = x->tex_info;
    sd.ntex = x->ntex;
    return sd;
}

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

// pixel filter of a 32-float camera record (fields 21, 22, 24-28: sampling
// mode, radius, filter kind, sigma, b, c, tau)
void or_filter_eval(const float *cam, const float *xy, uint32_t n, float *out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = filter_value(cam, xy[2 * i], xy[2 * i + 1]);
}
void or_filter_table(const float *cam, float *lut) {
    FisTable t = fis_build(cam);
    for (uint32_t i = 0; i < kFis * kFis; ++i) lut[i] = t.on ? t.lut[i] : 0.f;
}
void or_filter_sample(const float *cam, const float *u, uint32_t n, float *p, float *weight) {
    FisTable t = fis_build(cam);
    for (uint32_t i = 0; i < n; ++i) {
        if (!t.on) {
            p[2 * i] = p[2 * i + 1] = weight[i] = 0.f;
            continue;
        }
        fis_sample(t, cam[22], u[2 * i], u[2 * i + 1], p[2 * i], p[2 * i + 1], weight[i]);
    }
}

// reservoir streaming (DIReservoir::update, reservoir.h:80-86): n weights
// with their uniforms; returns the kept index (-1: none), weight_sum, C
int32_t or_reservoir_stream(const float *w, const float *u, uint32_t n, float *weight_sum, float *C) {
    Rsv r;
    int32_t kept = -1;
    for (uint32_t i = 0; i < n; ++i) {
        DSample d;
        d.light = i;
        r.update(u[i], d, w[i], 1.f);
    }
    if (r.s.light != kInv) kept = int32_t(r.s.light);
    *weight_sum = r.ws;
    *C = r.C;
    return kept;
}

// raster coordinates of world points in a camera given as the 32-float
// camera record (screen window, tan(fov/2)) and a world-to-camera 3x4
void or_prev_raster(const float *cam, const float *w2c, const float *p, uint32_t n, float *out) {
    for (uint32_t i = 0; i < n; ++i) prev_raster(cam, w2c, v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]), out[2 * i], out[2 * i + 1]);
}

void or_offset_ray_origin(const float *p, const float *n, uint32_t count, float *out) {
    for (uint32_t i = 0; i < count; ++i) {
        V3 r = offset_origin(v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]), v3(n[3 * i], n[3 * i + 1], n[3 * i + 2]));
        out[3 * i] = r.x, out[3 * i + 1] = r.y, out[3 * i + 2] = r.z;
    }
}

// CIE tables from the 1 nm data (4 x 471: X, Y, Z, D65): the 94-sample
// downsampling of SPD::create_cie_* (spd.cpp:18-28, 96-110) and
// SPD::cie_y_integral (spd.cpp:9-17, 111-114)
void or_cie_tables(const float *raw, float *spd_out, float *y_integral) {
    for (int c = 0; c < 4; ++c)
        for (int x = 0; x < 94; ++x) spd_out[94 * c + x] = raw[471 * c + 5 * x];
    float sum = 0.f;
    const float *Y = raw + 471;
    for (uint32_t i = 0; i < 94u; ++i) sum += 0.5f * (Y[i * 5] + Y[(i + 1u) * 5]);
    *y_integral = sum * 5.f;
}

void or_spd_eval(const float *tab, float interval, const float *lambda, uint32_t n, float *out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = spd_eval(tab, interval, lambda[i]);
}

void or_wavelengths(const float *u, uint32_t n, float *lambda, float *pdf) {
    for (uint32_t i = 0; i < n; ++i) {
        lambda[i] = sample_visible_wavelength(u[i]);
        pdf[i] = visible_wavelength_pdf(lambda[i]);
    }
}
void or_wavelength_pdf(const float *lambda, uint32_t n, float *pdf) {
    for (uint32_t i = 0; i < n; ++i) pdf[i] = visible_wavelength_pdf(lambda[i]);
}

// sigmoid spectra of coefficient rows (c0, c1, c2, scale) at wavelengths
void or_sigmoid(const float *coef, uint32_t ncoef, const float *lambda, uint32_t nl, float *out) {
    for (uint32_t c = 0; c < ncoef; ++c)
        for (uint32_t i = 0; i < nl; ++i) out[c * nl + i] = sigmoid_poly(coef + 4 * c, lambda[i]);
}

void or_rgb_to_spectrum(const float *table, const float *rgb, uint32_t n, int unbound, float *out) {
    for (uint32_t i = 0; i < n; ++i) {
        V3 c = v3(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
        if (unbound) decode_unbound(table, c, out + 4 * i);
        else decode_albedo(table, c, out + 4 * i);
    }
}

// kind 0: conductor (eta, k); 1: dielectric (eta)
void or_fresnel(int kind, const float *cos_t, uint32_t n, float eta, float k, float *out) {
    for (uint32_t i = 0; i < n; ++i)
        out[i] = kind == 0 ? fresnel_complex(cos_t[i], eta, k) : fresnel_dielectric(cos_t[i], eta);
}

// GGX terms for direction batches: D(wh), G1(wo), visible-normal pdf of wh
void or_ggx(float ax, float ay, const float *wo, const float *wh, uint32_t n, float *D, float *G1, float *pdf) {
    Ggx g{ax, ay};
    for (uint32_t i = 0; i < n; ++i) {
        V3 o = v3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), h = v3(wh[3 * i], wh[3 * i + 1], wh[3 * i + 2]);
        D[i] = g.D(h);
        G1[i] = g.G1(o);
        pdf[i] = g.pdf_wh(o, h);
    }
}
void or_ggx_sample(float ax, float ay, const float *wo, const float *u, uint32_t n, float *wh) {
    Ggx g{ax, ay};
    for (uint32_t i = 0; i < n; ++i) {
        V3 h = g.sample_wh(v3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), u[2 * i], u[2 * i + 1]);
        wh[3 * i] = h.x, wh[3 * i + 1] = h.y, wh[3 * i + 2] = h.z;
    }
}

// Material record `mat` (32 words) at a surface whose shading frame is the
// world axes and whose geometric normal is +z: evaluate f * |cos| and pdf
// for direction pairs (wo, wi), or sample wi for each wo from the sampler
// stream (0, i, frame, 0).  Hero mode reads 4 wavelengths per pair.
static void bsdf_setup(const uint32_t *mat, const float *wo, uint32_t i, Surf &it, Scene &s, uint32_t *inst) {
    inst[0] = 0, inst[1] = kInv, inst[2] = 0, inst[3] = 0;
    s = Scene{};
    s.mat = mat;
    s.inst = inst;
    it = Surf{};
    it.fr = {v3(1, 0, 0), v3(0, 1, 0), v3(0, 0, 1)};
    it.ng = v3(0, 0, 1);
    it.wo = v3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
}
static Swl bsdf_swl(const Spectral &sd, const float *lambda, uint32_t i) {
    Swl w;
    w.n = sd.ns;
    for (int k = 0; k < w.n; ++k) {
        w.lambda[k] = sd.rgb ? (k == 0 ? 602.785f : (k == 1 ? 539.285f : 445.772f)) : lambda[4 * i + k];
        w.pdf[k] = sd.rgb ? 1.f : visible_wavelength_pdf(w.lambda[k]);
    }
    return w;
}
void or_bsdf_eval(const or_spectral *spec, const uint32_t *mat, const float *wo, const float *wi, const float *lambda,
                  uint32_t n, float *f, float *pdf) {
    Spectral sd = to_spectral(spec);
    for (uint32_t i = 0; i < n; ++i) {
        Surf it;
        Scene s;
        uint32_t inst[4];
        bsdf_setup(mat, wo, i, it, s, inst);
        Swl w = bsdf_swl(sd, lambda, i);
        Bsdf b = make_bsdf(sd, s, 0, it, w);
        BEval e = bsdf_eval(sd, b, v3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]), w.n, nullptr);
        for (int k = 0; k < 4; ++k) f[4 * i + k] = e.f.v[k];
        pdf[i] = e.pdf;
    }
}
void or_bsdf_sample(const or_spectral *spec, const uint32_t *mat, const float *wo, const float *lambda, uint32_t n,
                    uint32_t frame, float *wi, float *f, float *pdf) {
    Spectral sd = to_spectral(spec);
    for (uint32_t i = 0; i < n; ++i) {
        Surf it;
        Scene s;
        uint32_t inst[4];
        bsdf_setup(mat, wo, i, it, s, inst);
        Swl w = bsdf_swl(sd, lambda, i);
        Bsdf b = make_bsdf(sd, s, 0, it, w);
        Sampler sp;
        sp.start(0u, i, frame, 0u);
        BSample bs = bsdf_sample(sd, b, sp, w.n);
        wi[3 * i] = bs.wi.x, wi[3 * i + 1] = bs.wi.y, wi[3 * i + 2] = bs.wi.z;
        for (int k = 0; k < 4; ++k) f[4 * i + k] = bs.e.f.v[k];
        pdf[i] = bs.e.pdf;
    }
}

// Oren-Nayar / Lambert evaluation in sRGB mode (closed-form checks)
void or_eval_diffuse(uint32_t type, const float *color, float sigma, const float *wo, const float *wi, uint32_t n,
                     float *f, float *pdf) {
    uint32_t mat[32] = {};
    mat[0] = type;
    std::memcpy(&mat[2], color, 12);
    std::memcpy(&mat[2 + 4], &sigma, 4);
    or_spectral sp{};
    sp.rgb = 1;
    sp.ns = 3;
    std::vector<float> f4(4 * size_t(n));
    or_bsdf_eval(&sp, mat, wo, wi, nullptr, n, f4.data(), pdf);
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) f[3 * i + k] = f4[4 * i + k];
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
// force, 1 BVH2.  threads = 0 uses all hardware threads.  spec = null
// renders sRGB mode without an environment light.
// One ReSTIR DI frame: shader0 over all pixels, the extra spatial passes,
// shader1 with the film's running mean.  restir: the exported settings
// (32 floats: flag, M_light, ..., extra passes, world-to-camera 3x4);
// prev_w2c / prev_pos: the camera of the previous frame; state buffers as
// described above (rsv [4][3][P][4], surf [2][3][P][4], ext [2][2][P][4]);
// rt [P][4] the film.
uint64_t or_restir_frame(const float *P, const float *N, const float *UV, const uint32_t *tri,
                         const uint32_t *tri_inst, const uint32_t *inst, const uint32_t *mat, const uint32_t *light,
                         const float *cam, const float *render, uint32_t nv, uint32_t nt, uint32_t ni, uint32_t nm,
                         uint32_t nl, const or_spectral *spec, const float *restir, const float *prev_w2c,
                         const float *prev_pos, uint32_t frame, uint32_t threads, float *rsv, float *surf, float *ext,
                         float *rt, float *gi_rsv, float *gi_smp, float *gi_aux, float *gi_rad, float *gbuf) {
    Scene s{P, N, UV, tri, tri_inst, inst, mat, light, cam, render, nv, nt, ni, nm, nl};
    Spectral sd = to_spectral(spec);
    Bvh2 b;
    bvh2_build(s, b);
    LightTables lt = spec ? build_lights(s, sd, spec->env_light, spec->env_w, spec->env_h, spec->env_rgb, spec->env_w2o,
                                         spec->env_o2w, spec->env_scale, spec->world_radius)
                          : build_lights(s, sd, kInv, 0, 0, nullptr, nullptr, nullptr, 1.f, 1.f);
    Ctx c{&s, &b, &lt, &sd, false};
    const FisTable fis = fis_build(cam);
    RParams rp{};
    rp.W = uint32_t(cam[19]);
    rp.H = uint32_t(cam[20]);
    rp.P = rp.W * rp.H;
    rp.frame = frame;
    rp.M_light = uint32_t(restir[1]);
    rp.M_bsdf = uint32_t(restir[2]);
    rp.max_age = uint32_t(restir[3]);
    rp.max_recursion = uint32_t(restir[4]);
    rp.diff = restir[5];
    rp.pairwise = restir[6] != 0.f;
    rp.spatial = restir[7] != 0.f;
    rp.s_num = uint32_t(restir[8]);
    rp.s_dot = restir[9];
    rp.s_depth = restir[10];
    rp.s_radius = restir[11];
    rp.temporal = restir[12] != 0.f;
    rp.t_limit = uint32_t(restir[13]);
    rp.t_N = uint32_t(restir[14]);
    rp.t_radius = restir[15];
    rp.t_dot = restir[16];
    rp.t_depth = restir[17];
    rp.t_mis = restir[18] != 0.f;
    rp.extra = uint32_t(restir[19]);
    std::memcpy(rp.w2c, restir + 20, sizeof(rp.w2c));
    std::memcpy(rp.prev_w2c, prev_w2c, sizeof(rp.prev_w2c));
    rp.cam_pos = v3(cam[3], cam[7], cam[11]);
    rp.prev_cam_pos = v3(prev_pos[0], prev_pos[1], prev_pos[2]);
    RState st{rsv, surf, ext, rp.P};
    const bool accumulate = render[6] != 0.f;
    uint32_t nth = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
    auto rows = [&](auto &&per_pixel) {
        std::atomic<uint32_t> next{0};
        auto work = [&]() {
            while (true) {
                uint32_t y = next.fetch_add(1);
                if (y >= rp.H) break;
                for (uint32_t x = 0; x < rp.W; ++x) per_pixel(x, y);
            }
        };
        std::vector<std::thread> th;
        for (uint32_t t = 0; t < nth; ++t) th.emplace_back(work);
        for (auto &t : th) t.join();
    };
    // ReSTIR GI settings (restir[32..47]) and state
    GParams gp{};
    gp.open = restir[32] != 0.f && gi_rsv != nullptr;
    gp.max_age = uint32_t(restir[33]);
    gp.diff = restir[34];
    gp.spatial = restir[35] != 0.f;
    gp.s_num = uint32_t(restir[36]);
    gp.s_dot = restir[37];
    gp.s_depth = restir[38];
    gp.s_radius = restir[39];
    gp.temporal = restir[40] != 0.f;
    gp.t_limit = uint32_t(restir[41]);
    gp.t_N = uint32_t(restir[42]);
    gp.t_dot = restir[43];
    gp.t_depth = restir[44];
    gp.t_radius = restir[45];
    gp.depth = int(restir[46]);
    GState g{gi_rsv, gi_smp, gi_aux, nullptr, gi_rad, rp.P};
    std::vector<float> direct;
    if (gp.open) direct.assign(4 * size_t(rp.P), 0.f);
    if (gbuf) {
        // RealTimeIntegrator::render (rt.cpp:144-149): the G-buffer of this
        // frame (geometry [P][8], albedo [P][4], emission [P][4]) before
        // ReSTIR DI; DI's shader0 then writes the motion vectors (gi_aux[2])
        float *geo = gbuf, *alb = gbuf + 8 * size_t(rp.P), *emi = gbuf + 12 * size_t(rp.P);
        rows([&](uint32_t x, uint32_t y) { gbuffer_px(c, fis, rp, x, y, geo, alb, emi); });
        or_gbuffer_gradient(geo, rp.W, rp.H);
    }
    rows([&](uint32_t x, uint32_t y) {
        restir_initial_px(c, fis, rp, st, x, y, gp.open || gbuf ? gi_aux : nullptr);
    });
    uint32_t in = 2;
    for (uint32_t k = 0; k < rp.extra; ++k) {
        rows([&](uint32_t x, uint32_t y) { restir_spatial_px(c, rp, st, x, y, in, in ^ 1u, 2u + k); });
        in ^= 1u;
    }
    auto film_add = [&](uint32_t x, uint32_t y, V3 L) {
        float v[4] = {zero_nan_inf(L.x), zero_nan_inf(L.y), zero_nan_inf(L.z), 1.f};
        float *acc = rt + 4 * (size_t(y) * rp.W + x);
        if (accumulate) {
            float a = 1.f / float(frame + 1u);
            for (int k = 0; k < 4; ++k) acc[k] = acc[k] + a * (v[k] - acc[k]);
        } else {
            for (int k = 0; k < 4; ++k) acc[k] = v[k];
        }
    };
    rows([&](uint32_t x, uint32_t y) {
        V3 L = restir_final_px(c, fis, rp, st, x, y, in);
        if (!gp.open) {
            film_add(x, y, L);
            return;
        }
        float *dl = direct.data() + 4 * (size_t(y) * rp.W + x);
        dl[0] = L.x, dl[1] = L.y, dl[2] = L.z, dl[3] = 1.f;
    });
    if (!gp.open) return 0;
    // ReSTIRGI::dispatch (indirect.cpp:357-365), then the combine of
    // RealTimeIntegrator::compile (rt.cpp:119-126): direct x 1 + indirect x 1
    rows([&](uint32_t x, uint32_t y) { gi_initial_px(c, rp, gp, st, g, x, y); });
    rows([&](uint32_t x, uint32_t y) { gi_temporal_px(c, fis, rp, gp, st, g, x, y); });
    rows([&](uint32_t x, uint32_t y) { gi_spatial_px(c, rp, gp, st, g, x, y); });
    rows([&](uint32_t x, uint32_t y) {
        const float *dl = direct.data() + 4 * (size_t(y) * rp.W + x), *il = gi_rad + 4 * (size_t(y) * rp.W + x);
        film_add(x, y, v3(dl[0] * 1.f + il[0] * 1.f, dl[1] * 1.f + il[1] * 1.f, dl[2] * 1.f + il[2] * 1.f));
    });
    return 0;
}

uint64_t or_render(const float *P, const float *N, const float *UV, const uint32_t *tri, const uint32_t *tri_inst,
                   const uint32_t *inst, const uint32_t *mat, const uint32_t *light, const float *cam,
                   const float *render, uint32_t nv, uint32_t nt, uint32_t ni, uint32_t nm, uint32_t nl,
                   uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1, uint32_t frame_start, uint32_t spp, int mode,
                   uint32_t threads, const or_spectral *spec, float *rt, float *vstats) {
    Scene s{P, N, UV, tri, tri_inst, inst, mat, light, cam, render, nv, nt, ni, nm, nl};
    Spectral sd = to_spectral(spec);
    Bvh2 b;
    if (mode == 1) bvh2_build(s, b);
    LightTables lt = spec ? build_lights(s, sd, spec->env_light, spec->env_w, spec->env_h, spec->env_rgb, spec->env_w2o,
                                         spec->env_o2w, spec->env_scale, spec->world_radius)
                          : build_lights(s, sd, kInv, 0, 0, nullptr, nullptr, nullptr, 1.f, 1.f);
    Ctx c{&s, &b, &lt, &sd, mode == 0};
    const FisTable fis = fis_build(cam);
    uint32_t W = uint32_t(cam[19]);
    bool accumulate = render[6] != 0.f;
    // ConvergenceInspector (pt.cpp:43-48, :59-64; inspector.cpp:26-38):
    // render[8] tone mapper, [9] on, [10] threshold, [11] min sample number;
    // vstats (avg, var, N bits, 0) per pixel
    const bool adaptive = vstats != nullptr && render[9] != 0.f;
    const float exposure = render[7], threshold = render[10];
    const uint32_t tm = uint32_t(render[8]), min_n = uint32_t(render[11]);
    std::atomic<uint32_t> next_row{y0};
    auto work = [&]() {
        while (true) {
            uint32_t y = next_row.fetch_add(1);
            if (y >= y1) break;
            for (uint32_t x = x0; x < x1; ++x) {
                float *acc = rt + 4 * (size_t(y) * W + x);
                float *vst = adaptive ? vstats + 4 * (size_t(y) * W + x) : nullptr;
                for (uint32_t f = frame_start; f < frame_start + spp; ++f) {
                    if (vst) {// is_convergence: the pixel adds no more samples
                        uint32_t n;
                        std::memcpy(&n, vst + 2, 4);
                        if (n > min_n && vst[1] / vst[0] < threshold) continue;
                    }
                    Swl w = sample_wavelengths(sd, x, y, f);
                    Sampler sp;
                    sp.start(x, y, f, 0u);
                    V3 o, d;
                    float fw;
                    camera_ray(s, fis, x, y, sp, o, d, fw);
                    V3 L = mul(Li(c, o, d, sp, w), fw);
                    float v[4] = {zero_nan_inf(L.x), zero_nan_inf(L.y), zero_nan_inf(L.z), 1.f};
                    if (accumulate) {
                        float a = 1.f / float(f + 1u);
                        for (int k = 0; k < 4; ++k) acc[k] = acc[k] + a * (v[k] - acc[k]);
                    } else {
                        for (int k = 0; k < 4; ++k) acc[k] = v[k];
                    }
                    if (vst) {// the output buffer's value (normal.cpp:96-99) feeds the statistics
                        float o[3];
                        for (int k = 0; k < 3; ++k) o[k] = tone_map(1.f - vs::em::exp(-acc[k] * exposure), tm);
                        float sum = o[0] + o[1] + o[2];
                        or_variance_stats_add(vst, &sum, 1);
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


// ---- film output and the saved picture ----------------------------------
// ToneMapper::apply of the linear / aces / reinhard plugins
// (src/render_core/tonemappers/impl.cpp:10-46); tm 0, 1, 2
static float tone_map(float v, uint32_t tm) {
    if (tm == 1) {
        float r = (v * (2.51f * v + 0.03f)) / (v * (2.43f * v + 0.59f) + 0.14f);
        return std::min(std::max(r, 0.f), 1.f);
    }
    if (tm == 2) return v / (v + 1.f);
    return v;
}
// sRGB encoding (IEC 61966-2-1), the linear_to_srgb of Postprocessor::compile_tone_mapping
static float encode_srgb(float x) {
    return x <= 0.0031308f ? 12.92f * x : 1.055f * vs::em::pow(x, 1.f / 2.4f) - 0.055f;
}
// the collector's output buffer from its running mean:
// tone_map(1 - exp(-v * exposure)), w = 1 (radiance_collector.cpp:24-26, normal.cpp:96-99)
void or_film_output(const float *rt, uint32_t n, float exposure, uint32_t tm, float *out) {
    for (uint32_t i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) out[4 * i + c] = tone_map(1.f - vs::em::exp(-rt[4 * i + c] * exposure), tm);
        out[4 * i + 3] = 1.f;
    }
}
// Pipeline::final_picture (src/base/mgr/pipeline.cpp:313-330) without a denoiser:
// the output buffer tone-mapped again, then sRGB-encoded when gamma, w = 1
// (src/base/mgr/postprocessor.cpp:13-21)
void or_final_picture(const float *out, uint32_t n, uint32_t tm, int32_t gamma, float *pic) {
    for (uint32_t i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) {
            float v = tone_map(out[4 * i + c], tm);
            pic[4 * i + c] = gamma ? encode_srgb(v) : v;
        }
        pic[4 * i + 3] = 1.f;
    }
}


// Camera rays of pixel (px, py) for frames frame0 .. frame0 + n - 1 (the
// sensor sample drawn first on each frame's stream, as the path tracer does):
// origins o[3n], directions d[3n] and film positions pf[2n].
void or_camera_rays(const float *cam, uint32_t px, uint32_t py, uint32_t frame0, uint32_t n, float *o, float *d,
                    float *pf) {
    Scene s{};
    s.cam = cam;
    FisTable fis = fis_build(cam);
    for (uint32_t i = 0; i < n; ++i) {
        Sampler sp;
        sp.start(px, py, frame0 + i, 0);
        V3 ro, rd;
        float w;
        camera_ray(s, fis, px, py, sp, ro, rd, w, pf + 2 * i);
        o[3 * i] = ro.x, o[3 * i + 1] = ro.y, o[3 * i + 2] = ro.z;
        d[3 * i] = rd.x, d[3 * i + 1] = rd.y, d[3 * i + 2] = rd.z;
    }
}

// The shared elementary functions over arrays, for their accuracy tests:
// fn 0 sin, 1 cos, 2 exp, 3 log, 4 log1p, 5 atanh, 6 cosh, 7 acos,
// 8 atan2(x, y), 9 pow(x, y).
void or_elementary(int fn, const float *x, const float *y, uint32_t n, float *out) {
    namespace em = vs::em;
    for (uint32_t i = 0; i < n; ++i) {
        const float a = x[i];
        switch (fn) {
        case 0: out[i] = em::sin(a); break;
        case 1: out[i] = em::cos(a); break;
        case 2: out[i] = em::exp(a); break;
        case 3: out[i] = em::log(a); break;
        case 4: out[i] = em::log1p(a); break;
        case 5: out[i] = em::atanh(a); break;
        case 6: out[i] = em::cosh(a); break;
        case 7: out[i] = em::acos(a); break;
        case 8: out[i] = em::atan2(a, y[i]); break;
        case 9: out[i] = em::pow(a, y[i]); break;
        default: out[i] = 0.f;
        }
    }
}

}// extern "C"
