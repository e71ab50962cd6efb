// ReSTIR DI: per-pixel reservoirs of light samples, filled from candidates,
// from the previous frame and from screen-space neighbours.
//
// Restates the direct-lighting half of the reference's real-time integrator
// (src/render_core/integrators/ReSTIR/direct.cpp): shader0 = primary hit,
// RIS over light and BSDF candidates, visibility, temporal reuse; shader1 =
// spatial reuse and shading.  The reservoir is DIReservoir
// (src/render_core/integrators/ReSTIR/reservoir.h:15-115), the surface record
// SurfaceData (src/base/scattering/interaction.h:15-55), and the light side
// uses the point-measure entry points of LightSampler
// (src/base/illumination/lightsampler.cpp:218-313).
//
// HBM layout (all float4, field-major so a wave's loads coalesce):
//   reservoirs: 4 buffers (previous / current ping-pong by frame parity,
//               pass-through, second pass-through) x 3 fields x pixels
//     r0 = weight_sum, C, W, p_hat
//     r1 = light index bits, prim bits, bary.x, bary.y
//     r2 = pos.xyz, age bits
//   surfaces:   2 buffers x 3 fields x pixels
//     s0 = triangle bits (0xFFFFFFFF: miss), b1, b2, flag | replaced << 16
//     s1 = shading normal xyz, linear depth
//     s2 = position xyz, diffuse factor
//   extensions: 2 buffers x 2 fields x pixels (written when replaced)
//     e0 = throughput xyz, t_max;  e1 = view position xyz, 0
#pragma once

#include "dev_pt.h"
#include "restir.h"

namespace vs {
namespace dev {

// BxDFFlag values (src/base/sample.h:23-39)
constexpr uint32_t BXDF_UNSET = 1u;
constexpr uint32_t BXDF_REFLECTION = 2u;
constexpr uint32_t BXDF_DIFFUSE = 8u;
constexpr uint32_t BXDF_GLOSSY = 16u;
constexpr uint32_t BXDF_NEAR_SPEC = 64u;
// alpha below which metal and glass count as near-specular (metal.cpp:46,
// glass.cpp:178)
constexpr float kNearSpecAlpha = 0.022f;

// buffers by frame parity (frame_buffer.h:44-49, direct.h:108-116)
__device__ __forceinline__ uint32_t prev_slot(uint32_t frame) { return frame & 1u; }
__device__ __forceinline__ uint32_t cur_slot(uint32_t frame) { return (frame + 1u) & 1u; }
constexpr uint32_t kPassThrough = 2u;

struct DISample {
    uint32_t light;// kInvalid: no sample
    uint32_t prim;
    f2 bary;
    uint32_t age;
    float p_hat;
    f3 pos;
};

__device__ __forceinline__ DISample empty_sample() {
    DISample s;
    s.light = kInvalid;
    s.prim = 0u;
    s.bary = mk2(0.f, 0.f);
    s.age = 0u;
    s.p_hat = 0.f;
    s.pos = mk3(0.f, 0.f, 0.f);
    return s;
}

struct Reservoir {
    float weight_sum, C, W;
    DISample s;
    // DIReservoir::update / truncation / process_occluded / update_W
    // (reservoir.h:80-103)
    __device__ __forceinline__ bool update(float u, const DISample &v, float weight, float new_C) {
        weight_sum += weight;
        C += new_C;
        bool take = u * weight_sum < weight;
        if (take) s = v;
        return take;
    }
    __device__ __forceinline__ void truncation(float limit) { C = fminf(limit, C); }
    __device__ __forceinline__ void process_occluded(bool occluded) {
        W = occluded ? 0.f : W;
        weight_sum = occluded ? 0.f : weight_sum;
    }
    __device__ __forceinline__ void update_W(float p_hat) { W = p_hat == 0.f ? 0.f : weight_sum / p_hat; }
};

__device__ __forceinline__ Reservoir empty_reservoir() {
    Reservoir r;
    r.weight_sum = 0.f;
    r.C = 0.f;
    r.W = 0.f;
    r.s = empty_sample();
    return r;
}

// DIReservoir::cal_weight / safe_weight (reservoir.h:54-67)
__device__ __forceinline__ float cal_weight(float mis, float p_hat, float W) { return mis * p_hat * W; }
__device__ __forceinline__ float safe_weight(float mis, float p_hat, float W) {
    float r = cal_weight(mis, p_hat, W);
    return isnan(r) ? 0.f : r;
}

struct SurfaceRec {
    uint32_t tri;// kInvalid on a miss
    float b1, b2;
    uint32_t flag;
    uint32_t replaced;
    f3 normal;
    float depth;
    f3 pos;
    float diff;
    __device__ __forceinline__ bool hit() const { return tri != kInvalid; }
    __device__ __forceinline__ bool near_specular() const { return flag == BXDF_NEAR_SPEC; }
};

__device__ __forceinline__ SurfaceRec empty_surface() {
    SurfaceRec s;
    s.tri = kInvalid;
    s.b1 = s.b2 = 0.f;
    s.flag = BXDF_UNSET;
    s.replaced = 0u;
    s.normal = mk3(0.f, 0.f, 0.f);
    s.depth = 0.f;
    s.pos = mk3(0.f, 0.f, 0.f);
    s.diff = 0.f;
    return s;
}

struct SurfaceExt {
    f3 throughput;
    f3 view_pos;
    float t_max;
};

// ---- HBM records
__device__ __forceinline__ Reservoir load_reservoir(const RestirView &v, uint32_t buf, uint32_t i) {
    const float4 *b = v.rsv + size_t(buf) * 3u * v.pixels;
    float4 a = b[i], c = b[v.pixels + i], d = b[2u * size_t(v.pixels) + i];
    Reservoir r;
    r.weight_sum = a.x;
    r.C = a.y;
    r.W = a.z;
    r.s.p_hat = a.w;
    r.s.light = __float_as_uint(c.x);
    r.s.prim = __float_as_uint(c.y);
    r.s.bary = mk2(c.z, c.w);
    r.s.pos = mk3(d.x, d.y, d.z);
    r.s.age = __float_as_uint(d.w);
    return r;
}
__device__ __forceinline__ void store_reservoir(const RestirView &v, uint32_t buf, uint32_t i, const Reservoir &r) {
    float4 *b = v.rsv + size_t(buf) * 3u * v.pixels;
    b[i] = make_float4(r.weight_sum, r.C, r.W, r.s.p_hat);
    b[v.pixels + i] = make_float4(__uint_as_float(r.s.light), __uint_as_float(r.s.prim), r.s.bary.x, r.s.bary.y);
    b[2u * size_t(v.pixels) + i] = make_float4(r.s.pos.x, r.s.pos.y, r.s.pos.z, __uint_as_float(r.s.age));
}
__device__ __forceinline__ SurfaceRec load_surface(const RestirView &v, uint32_t buf, uint32_t i) {
    const float4 *b = v.surf + size_t(buf) * 3u * v.pixels;
    float4 a = b[i], c = b[v.pixels + i], d = b[2u * size_t(v.pixels) + i];
    SurfaceRec s;
    s.tri = __float_as_uint(a.x);
    s.b1 = a.y;
    s.b2 = a.z;
    uint32_t fr = __float_as_uint(a.w);
    s.flag = fr & 0xFFFFu;
    s.replaced = fr >> 16;
    s.normal = mk3(c.x, c.y, c.z);
    s.depth = c.w;
    s.pos = mk3(d.x, d.y, d.z);
    s.diff = d.w;
    return s;
}
__device__ __forceinline__ void store_surface(const RestirView &v, uint32_t buf, uint32_t i, const SurfaceRec &s) {
    float4 *b = v.surf + size_t(buf) * 3u * v.pixels;
    b[i] = make_float4(__uint_as_float(s.tri), s.b1, s.b2, __uint_as_float(s.flag | (s.replaced << 16)));
    b[v.pixels + i] = make_float4(s.normal.x, s.normal.y, s.normal.z, s.depth);
    b[2u * size_t(v.pixels) + i] = make_float4(s.pos.x, s.pos.y, s.pos.z, s.diff);
}
__device__ __forceinline__ SurfaceExt load_ext(const RestirView &v, uint32_t buf, uint32_t i) {
    const float4 *b = v.ext + size_t(buf) * 2u * v.pixels;
    float4 a = b[i], c = b[v.pixels + i];
    SurfaceExt e;
    e.throughput = mk3(a.x, a.y, a.z);
    e.t_max = a.w;
    e.view_pos = mk3(c.x, c.y, c.z);
    return e;
}
__device__ __forceinline__ void store_ext(const RestirView &v, uint32_t buf, uint32_t i, const SurfaceExt &e) {
    float4 *b = v.ext + size_t(buf) * 2u * v.pixels;
    b[i] = make_float4(e.throughput.x, e.throughput.y, e.throughput.z, e.t_max);
    b[v.pixels + i] = make_float4(e.view_pos.x, e.view_pos.y, e.view_pos.z, 0.f);
}

// ---- conversions with the conventions of the DSL's casts
__device__ __forceinline__ uint32_t to_u32(float v) {
    return v > 0.f ? (v < 4.0e9f ? uint32_t(v) : 0xFFFFFFFFu) : 0u;
}
__device__ __forceinline__ int to_i32(float v) {
    return v > -2.0e9f ? (v < 2.0e9f ? int(v) : 2000000000) : -2000000000;
}
__device__ __forceinline__ f3 xform3x4(const float *m, f3 p) {
    return mk3(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
               m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}

// f.vec3(): the first three channels
template <int NS>
__device__ __forceinline__ float luminance_of(const Spec<NS> &f) {
    return luminance(mk3(f.v[0], NS > 1 ? f.v[1 < NS ? 1 : 0] : 0.f, NS > 2 ? f.v[2 < NS ? 2 : 0] : 0.f));
}

// ---- material properties read by ReSTIR (a surface without material
// evaluates as the black body)

// MaterialEvaluator::flag (material.cpp:110-116): the lobe flags; metal and
// glass are near-specular below the alpha threshold (metal.cpp:147-148,
// glass.cpp:251-252); diffuse lobes carry DiffRefl (bxdf.h:88); the
// principled lobe set always holds its diffuse lobe, so it is never
// near-specular
__device__ __forceinline__ uint32_t material_flag(const SceneView &sc, uint32_t mat) {
    if (mat == kInvalid) return BXDF_DIFFUSE | BXDF_REFLECTION;
    const MaterialRecord &m = sc.mat[mat];
    if (m.type == MAT_METAL || m.type == MAT_MIRROR)
        return fminf(m.p[6], m.p[7]) < kNearSpecAlpha ? BXDF_NEAR_SPEC : BXDF_GLOSSY;
    if (m.type == MAT_GLASS) return fminf(m.p[5], m.p[6]) < kNearSpecAlpha ? BXDF_NEAR_SPEC : BXDF_GLOSSY;
    if (m.type == MAT_PRINCIPLED) return BXDF_DIFFUSE | BXDF_REFLECTION | BXDF_GLOSSY;
    return BXDF_DIFFUSE | BXDF_REFLECTION;
}

// MaterialEvaluator::diffuse_factor (material.cpp:124-130):
// MicrofacetLobe sqrt(sqrt(ax ay)) (lobe.cpp:212-214), DielectricLobe
// sqrt(ax) (lobe.cpp:427-429), 1 otherwise (lobe.h:84)
__device__ __forceinline__ float material_diffuse_factor(const SceneView &sc, uint32_t mat) {
    if (mat == kInvalid) return 1.f;
    const MaterialRecord &m = sc.mat[mat];
    if (m.type == MAT_METAL || m.type == MAT_MIRROR) return sqrtf(sqrtf(m.p[6] * m.p[7]));
    if (m.type == MAT_GLASS) return sqrtf(m.p[5]);
    return 1.f;
}

// SampledWavelengths::check_dispersive (spectrum.cpp:21-39)
template <int NS, bool RGB>
__device__ __forceinline__ void collapse_if_dispersive(Swl<NS> &swl, bool dispersive) {
    if (RGB || !dispersive) return;
    bool secondary_valid = NS > 1;
#pragma unroll
    for (int i = 1; i < NS; ++i) secondary_valid = secondary_valid && swl.pdf[i] != 0.f;
    if (secondary_valid) {
#pragma unroll
        for (int i = 1; i < NS; ++i) swl.pdf[i] = 0.f;
    }
}

template <int NS, bool RGB, int MS>
__device__ __forceinline__ MatCtx<NS> material_at(const SceneView &sc, const Surface &it, Swl<NS> &swl) {
    uint32_t m = sc.inst[it.inst].mat;
    if (m == kInvalid) {
        MatCtx<NS> c;
        c.type = MAT_BLACK;
        c.fr = it.frame;
        c.ng = it.ng;
        c.wo_w = it.wo;
        c.wo = it.frame.to_local(it.wo);
        return c;
    }
    bool dispersive;
    MatCtx<NS> mc = make_material<NS, RGB, MS>(sc, sc.mat[m], it.frame, it.ng, it.wo, swl, dispersive, it.uv);
    collapse_if_dispersive<NS, RGB>(swl, dispersive);
    return mc;
}

// ---- point-measure light entry points

// LightSampler::sample_only (lightsampler.cpp:218-228) with AreaLight::
// sample_only (area.cpp:139-145), SphericalMap::sample_only
// (spherical.cpp:126-131) and Light::sample_only (light.h:177-182).
// Draws: select 1, surface 2.
__device__ __forceinline__ DISample light_sample_only(const SceneView &sc, Sampler &sp) {
    DISample s = empty_sample();
    uint32_t li = select_light(sc, sp.next1d());
    f2 u = sp.next2d();
    const LightRecord &L = sc.lights[li];
    if (L.type == LIGHT_AREA) {
        float ux;
        s.prim = alias_sample(sc, L.alias_offset, L.tri_count, u.x, &ux);
        s.bary = square_to_triangle(mk2(ux, u.y));
    } else if (L.type == LIGHT_ENV) {
        float pdf_map;
        s.bary = env_sample_uv(sc, u, pdf_map);
    } else {
        s.bary = u;
    }
    s.light = li;
    return s;
}

template <int NS>
struct PointEval {
    Spec<NS> L;// area lights: Le * G (radiance towards ref over the area measure)
    float pdf; // area measure (solid angle for the environment); < 0 for delta lights
    f3 p_light;
};

__device__ __forceinline__ f3 robust_point(f3 pos, f3 ng, f3 dir, float factor) {
    // LightEvalContext::robust_pos / Interaction::robust_position
    // (interaction.h:314-317, interaction.cpp:140-142)
    float s = dot(ng, dir) > 0.f ? 1.f : -1.f;
    return offset_ray_origin(pos, ng * s * factor);
}

// spherical direction of a map coordinate (SphericalMap::evaluate,
// spherical.cpp:104-124)
__device__ __forceinline__ void env_direction(const SceneView &sc, f2 uv, f3 &local, f3 &world, float &sin_t) {
    float theta = uv.y * kPi, phi = uv.x * 2.f * kPi;
    float cos_t;
    em::sincos(theta, &sin_t, &cos_t);
    float sp, cp;
    em::sincos(phi, &sp, &cp);
    local = mk3(sin_t * cp, sin_t * sp, cos_t);
    world = normalize(mat3_apply(sc.env_o2w, local));
}

// LightSampler::evaluate_point (lightsampler.cpp:230-240) with
// AreaLight::evaluate_point (area.cpp:157-168), IPointLight::evaluate_point
// (light.h:237-243) and SphericalMap::evaluate_point (spherical.cpp:133-141)
template <int NS, bool RGB>
__device__ __forceinline__ PointEval<NS> light_eval_point(const SceneView &sc, f3 ref, const DISample &s,
                                                          const Swl<NS> &swl) {
    PointEval<NS> r;
    const LightRecord &L = sc.lights[s.light];
    if (L.type == LIGHT_AREA) {
        Surface p = surface_at(sc, L.tri_base + s.prim, s.bary.x, s.bary.y, false);
        f3 w = ref - p.pos;
        Spec<NS> Le = light_color<NS, RGB>(sc, L, swl, p.uv) * ((dot(w, p.ng) > 0.f || L.two_sided) ? 1.f : 0.f);
        float G = dot(normalize(w), p.ng) / length_squared(w);
        r.L = Le * G;
        r.pdf = (1.f / p.prim_area) * alias_pmf(sc, L.alias_offset, L.tri_count, L.alias_integral, s.prim);
        r.p_light = robust_point(p.pos, p.ng, w, sc.ray_offset_factor);
    } else if (L.type == LIGHT_ENV) {
        float pdf_map = env_map_pdf(sc, s.bary);
        f3 local, world;
        float sin_t;
        env_direction(sc, s.bary, local, world, sin_t);
        float pdf_dir = pdf_map / (2.f * kPi * kPi * sin_t);
        r.pdf = isinf(pdf_dir) ? 0.f : pdf_dir;
        r.L = env_texel<NS, RGB>(sc, spherical_uv(local), swl);
        r.p_light = ref + world * sc.world_diameter;
    } else {
        f3 pos = mk3(L.position[0], L.position[1], L.position[2]);
        f3 w_un = ref - pos;
        Spec<NS> val = light_color<NS, RGB>(sc, L, swl, mk2(0.f, 0.f)) / length_squared(w_un);
        if (L.type == LIGHT_SPOT) {
            f3 dir = mk3(L.direction[0], L.direction[1], L.direction[2]);
            float ct = fminf(fmaxf(dot(dir, normalize(w_un)), L.cos_angle), L.cos_falloff_start);
            float f = (ct - L.cos_angle) / (L.cos_falloff_start - L.cos_angle);
            float f2v = f * f;
            val = val * (f2v * f2v);
        }
        r.L = val;
        r.pdf = -1.f;
        r.p_light = pos;
    }
    r.pdf *= light_select_pmf(sc, s.light);
    return r;
}

// LightSampler::PDF_point (lightsampler.cpp:242-250): a solid-angle pdf at
// ref converted to the light's point measure (Light::PDF_point,
// light.h:141-151, warp.h:101-105); the environment keeps solid angle
// (spherical.cpp:143-146)
__device__ __forceinline__ float light_pdf_point(const SceneView &sc, f3 ref, const DISample &s, float pdf_wi) {
    const LightRecord &L = sc.lights[s.light];
    if (L.type == LIGHT_ENV) return pdf_wi;
    f3 lp, ln;
    if (L.type == LIGHT_AREA) {
        Surface p = surface_at(sc, L.tri_base + s.prim, s.bary.x, s.bary.y, false);
        lp = p.pos;
        ln = p.ng;
    } else {
        lp = mk3(L.position[0], L.position[1], L.position[2]);
        ln = normalize(ref - lp);
    }
    f3 w = ref - lp;
    float r = pdf_wi * fabsf(dot(ln, normalize(w))) / length_squared(w);
    return isinf(r) ? 0.f : r;
}

// p_hat = luminance of f * L for a sample seen from `it` (ReSTIRDI::Li from
// the light, direct.cpp:123-159, and compute_p_hat, direct.h:128-134).  An
// empty sample evaluates with zero radiance from the origin.
template <int NS, bool RGB, int MS>
__device__ __forceinline__ Spec<NS> light_contribution(const SceneView &sc, const Surface &it, const MatCtx<NS> &mc,
                                                       const DISample &s, const Swl<NS> &swl, float *eval_pdf) {
    PointEval<NS> le;
    if (s.light != kInvalid) {
        le = light_eval_point<NS, RGB>(sc, it.pos, s, swl);
    } else {
        le.L = Spec<NS>::fill(0.f);
        le.pdf = 0.f;
        le.p_light = mk3(0.f, 0.f, 0.f);
    }
    f3 wi = normalize(le.p_light - it.pos);
    EvalN<NS> e = mat_eval<NS, MS>(mc, wi, nullptr);
    if (eval_pdf) *eval_pdf = e.pdf;
    return e.f * le.L;
}
template <int NS, bool RGB, int MS>
__device__ __forceinline__ float p_hat_of(const SceneView &sc, const Surface &it, const MatCtx<NS> &mc,
                                          const DISample &s, const Swl<NS> &swl) {
    return luminance_of(light_contribution<NS, RGB, MS>(sc, it, mc, s, swl, nullptr));
}

// ---- validity of a neighbour (is_valid_neighbor, ReSTIR/common.h:72-80)
__device__ __forceinline__ bool valid_neighbor(const SurfaceRec &cur, const SurfaceRec &other, float dot_threshold,
                                               float depth_threshold, float diff_threshold) {
    bool c0 = fabsf(dot(cur.normal, other.normal)) > dot_threshold;
    bool c1 = (fabsf(cur.depth - other.depth) / cur.depth) < depth_threshold;
    bool c2 = fabsf(cur.diff - other.diff) / cur.diff < diff_threshold;
    return c0 && c1 && c2 && cur.hit() && other.hit();
}

// ---- sample_delta of near-specular lobes (MaterialEvaluator::sample_delta,
// material.cpp:154-160; MicrofacetLobe::sample_delta_local, lobe.cpp:223-230;
// DielectricLobe::sample_delta_local, lobe.cpp:404-425).  The mirror or
// refracted direction and the throughput f |cos| / pdf of the lobe evaluated
// there; dielectric draws 1.
template <int NS>
__device__ __forceinline__ bool sample_delta(const MatCtx<NS> &c, Sampler &sp, f3 &wi_w, f3 &throughput) {
    f3 wo = c.wo;
    EvalN<NS> e;
    f3 wi = mk3(-wo.x, -wo.y, wo.z);
    if (c.type == MAT_METAL) {
        eval_conductor_like(c.ggx, wo, wi, false, c.a0, c.a1, Spec<NS>::fill(1.f), 1.f, e);
    } else if (c.type == MAT_MIRROR) {
        eval_mirror(c.ggx, wo, wi, c.a0, 1.f, e);
    } else if (c.type == MAT_GLASS) {
        Spec<NS> F;
#pragma unroll
        for (int i = 0; i < NS; ++i) F.v[i] = fresnel_dielectric(fabsf(wo.z), c.a1.v[i]);
        float rp = F.average() / ((1.f - F) * c.a0 + F).average();
        float uc = sp.next1d();
        if (!(uc < rp)) refract_dir(wo, face_forward(mk3(0.f, 0.f, 1.f), wo), c.eta0, wi);
        eval_dielectric(c.ggx, wo, wi, c.a1, c.eta0, false, c.a1, c.a0, 1.f, e, nullptr);
    } else {
        wi_w = mk3(0.f, 0.f, 0.f);
        throughput = mk3(0.f, 0.f, 0.f);
        return false;
    }
    Spec<NS> t = (e.f * fabsf(wi.z)) / e.pdf;
    throughput = mk3(t.v[0], t.v[NS > 1 ? 1 : 0], t.v[NS > 2 ? 2 : 0]);
    wi_w = c.fr.to_world(wi);
    return true;
}


// ---- camera conventions (Sensor::linear_depth, sensor.cpp:198-201;
// Sensor::prev_raster_coord, sensor.cpp:105-110: the point in the previous
// camera's space, divided by its depth, mapped back to raster coordinates)
__device__ __forceinline__ float linear_depth(const RestirView &rv, f3 p) { return xform3x4(rv.w2c, p).z; }
__device__ __forceinline__ f2 prev_raster(const CameraView &cam, const RestirView &rv, f3 p) {
    f3 c = xform3x4(rv.prev_w2c, p);
    float x = c.x / c.z, y = c.y / c.z;
    float sx = x / cam.tan_half_fov, sy = y / cam.tan_half_fov;
    return mk2((sx - cam.lower_x) / cam.raster_sx, (cam.upper_y - sy) / cam.raster_sy);
}

// shadow ray towards a point (Geometry::occluded, geometry.cpp:175-184)
__device__ __forceinline__ bool occluded(const SceneView &sc, const Surface &it, f3 p, uint2 *stack, uint32_t stride,
                                         PathCounters &pc) {
    f3 dir = p - it.pos;
    f3 n = dot(it.ng, dir) > 0.f ? it.ng : -it.ng;
    uint32_t dn = 0, dt = 0;
    TraceHit h = traverse<true>(sc, offset_ray_origin(it.pos, n), dir, 0.f, 1.f - kShadowEpsilon, stack, stride, dn, dt);
    pc.nodes += dn;
    pc.tris += dt;
    ++pc.rays;
    return h.tri != kInvalid;
}

// ReSTIRDI::Li from the BSDF (direct.cpp:57-121): one BSDF-sampled
// candidate; `s` becomes the emitter point it reaches (area emitter or
// environment), `pdf` the BSDF pdf in the light's point measure and
// `light_pdf` the light's own point-measure pdf
// The evaluator is rebuilt here (the same arithmetic as the caller's), so
// that no evaluator stays live across the traversal below.
// HitBSDF of a BSDF candidate {wi, f, pdf} (ReSTIRDI::Li, direct.cpp:82-85),
// kept for ReSTIR GI; its default without a candidate is wi 0, f 0, pdf -1
// (interaction.h:60-66)
struct HitBsdfRec {
    f3 wi, f;
    float pdf;
};

template <int NS, bool RGB, int MS>
__device__ __forceinline__ Spec<NS> bsdf_candidate(const SceneView &sc, const Surface &it, Sampler &sp,
                                                   Swl<NS> &swl, DISample &s, float &pdf, float &light_pdf,
                                                   uint2 *stack, uint32_t stride, PathCounters &pc,
                                                   HitBsdfRec *hb) {
    SampleN<NS> bs;
    {
        MatCtx<NS> mc = material_at<NS, RGB, MS>(sc, it, swl);
        bs = mat_sample<NS, MS>(mc, sp);
    }
    if (hb) {
        hb->wi = bs.wi;
        hb->pdf = bs.e.pdf;
        hb->f = mk3(bs.e.f.v[0], bs.e.f.v[NS > 1 ? 1 : 0], bs.e.f.v[NS > 2 ? 2 : 0]);
    }
    f3 n = dot(it.ng, bs.wi) > 0.f ? it.ng : -it.ng;
    TraceHit h = traverse<false>(sc, offset_ray_origin(it.pos, n), bs.wi, 0.f, 3.402823466e38f, stack, stride,
                                 pc.nodes, pc.tris);
    ++pc.rays;
    pdf = bs.e.pdf;
    light_pdf = 0.f;
    Spec<NS> f = Spec<NS>::fill(0.f);
    if (h.tri != kInvalid) {
        Surface nx = surface_at(sc, h.tri, h.b1, h.b2, false);
        const uint32_t li = sc.inst[nx.inst].light;
        if (li != kInvalid) {
            // LightSampler::evaluate_hit_point (lightsampler.cpp:269-288)
            const LightRecord &L = sc.lights[li];
            f3 w = it.pos - nx.pos;
            Spec<NS> Le = light_color<NS, RGB>(sc, L, swl, nx.uv) * ((dot(w, nx.ng) > 0.f || L.two_sided) ? 1.f : 0.f);
            float G = dot(normalize(w), nx.ng) / length_squared(w);
            float pp = pdf * fabsf(dot(nx.ng, normalize(w))) / length_squared(w);
            uint32_t prim = h.tri - L.tri_base;
            light_pdf = (1.f / nx.prim_area) * alias_pmf(sc, L.alias_offset, L.tri_count, L.alias_integral, prim) *
                        light_select_pmf(sc, li);
            s.light = li;
            s.prim = prim;
            s.bary = mk2(h.b1, h.b2);
            s.pos = robust_point(nx.pos, nx.ng, w, sc.ray_offset_factor);
            pdf = isinf(pp) ? 0.f : pp;
            f = bs.e.f * (Le * G);
        }
    } else if (sc.env_light != kInvalid) {
        // LightSampler::evaluate_miss_point (lightsampler.cpp:302-313)
        f3 wi = bs.wi * sc.world_diameter;
        f3 pl = it.pos + wi;
        Spec<NS> Lv;
        float epdf;
        env_eval<NS, RGB>(sc, normalize(pl - it.pos), swl, Lv, epdf);
        light_pdf = epdf * light_select_pmf(sc, sc.env_light);
        s.light = sc.env_light;
        s.bary = spherical_uv(normalize(mat3_apply(sc.env_w2o, bs.wi)));
        s.pos = pl;
        f = bs.e.f * Lv;
    }
    return f;
}

// ReSTIRDI::RIS (direct.cpp:161-233): M_light light candidates (sample_only
// + evaluate_point; MIS against the BSDF in the light's point measure, delta
// lights 1 / M_light), then M_bsdf BSDF candidates, each weighted by the
// luminance of the path throughput.  Draws per light candidate: 3 + 1, per
// BSDF candidate: the BSDF's + 1.
template <int NS, bool RGB, int MS>
__device__ __forceinline__ Reservoir ris(const SceneView &sc, const RestirView &rv, const Surface &it, Sampler &sp,
                                         Swl<NS> &swl, f3 throughput, uint2 *stack, uint32_t stride,
                                         PathCounters &pc, HitBsdfRec *hb) {
    Reservoir r = empty_reservoir();
    if (rv.M_light + rv.M_bsdf == 0u) return r;
    const float lum = luminance(throughput);
    const float Ml = float(rv.M_light), Mb = float(rv.M_bsdf);
    if (rv.M_light > 0u) {
    MatCtx<NS> mc = material_at<NS, RGB, MS>(sc, it, swl);
    for (uint32_t i = 0; i < rv.M_light; ++i) {
        DISample s = light_sample_only(sc, sp);
        PointEval<NS> le = light_eval_point<NS, RGB>(sc, it.pos, s, swl);
        f3 wi = normalize(le.p_light - it.pos);
        EvalN<NS> e = mat_eval<NS, MS>(mc, wi, nullptr);
        Spec<NS> f = e.f * le.L;
        float bsdf_point = light_pdf_point(sc, it.pos, s, e.pdf);
        s.p_hat = luminance_of(f);
        s.pos = le.p_light;
        bool delta = le.pdf < 0.f;
        float lpdf = delta ? -le.pdf : le.pdf;
        float mis = delta ? 1.f / Ml : lpdf / (Ml * lpdf + Mb * bsdf_point);
        float w = safe_weight(mis, s.p_hat, 1.f / lpdf) * lum;
        r.update(sp.next1d(), s, w, 1.f);
    }
    }
    for (uint32_t i = 0; i < rv.M_bsdf; ++i) {
        DISample s = empty_sample();
        float pdf, lpdf;
        Spec<NS> f = bsdf_candidate<NS, RGB, MS>(sc, it, sp, swl, s, pdf, lpdf, stack, stride, pc, hb);
        s.p_hat = luminance_of(f);
        float w = safe_weight(pdf / (Ml * lpdf + Mb * pdf), s.p_hat, 1.f / pdf) * lum;
        r.update(sp.next1d(), s, w, 1.f);
    }
    return r;
}

// ReSTIRDI::combine_temporal (direct.cpp:327-380)
template <int NS, bool RGB, int MS>
__device__ __forceinline__ Reservoir combine_temporal(const SceneView &sc, const RestirView &rv, const Reservoir &cur,
                                                      const SurfaceRec &cs, Reservoir other, f3 view_pos,
                                                      f3 prev_view_pos, Sampler &sp, Swl<NS> &swl) {
    other.s.age += 1u;
    Surface it = surface_at(sc, cs.tri, cs.b1, cs.b2, true);
    float mis_cur, mis_prev, p_nn;
    if (rv.temporal_mis) {
        it.wo = normalize(prev_view_pos - it.pos);
        MatCtx<NS> mp = material_at<NS, RGB, MS>(sc, it, swl);
        float p_cn = p_hat_of<NS, RGB, MS>(sc, it, mp, cur.s, swl);
        p_nn = p_hat_of<NS, RGB, MS>(sc, it, mp, other.s, swl);
        it.wo = normalize(view_pos - it.pos);
        MatCtx<NS> mc = material_at<NS, RGB, MS>(sc, it, swl);
        float p_cc = p_hat_of<NS, RGB, MS>(sc, it, mc, cur.s, swl);
        float p_nc = p_hat_of<NS, RGB, MS>(sc, it, mc, other.s, swl);
        // MIS_weight_n with the confidences as integers (warp.h:186-192)
        float nc = float(int(cur.C)), no = float(int(other.C));
        mis_cur = (nc * p_cc) / (nc * p_cc + no * p_cn);
        mis_prev = (no * p_nn) / (no * p_nn + nc * p_nc);
    } else {
        it.wo = normalize(prev_view_pos - it.pos);
        MatCtx<NS> mp = material_at<NS, RGB, MS>(sc, it, swl);
        p_nn = p_hat_of<NS, RGB, MS>(sc, it, mp, other.s, swl);
        mis_cur = cur.C / (cur.C + other.C);
        mis_prev = other.C / (other.C + cur.C);
    }
    Reservoir ret = empty_reservoir();
    ret.update(0.5f, cur.s, safe_weight(mis_cur, cur.s.p_hat, cur.W), cur.C);
    DISample os = other.s;
    os.p_hat = p_nn;
    ret.update(sp.next1d(), os, safe_weight(mis_prev, os.p_hat, other.W), other.C);
    ret.update_W(ret.s.p_hat);
    return ret;
}

// ReSTIRDI::temporal_reuse (direct.cpp:382-432): the reprojected pixel of
// the previous frame, or failing that up to temporal.N random pixels within
// temporal.radius of it (2 draws each)
template <int NS, bool RGB, int MS>
__device__ __forceinline__ Reservoir temporal_reuse(const SceneView &sc, const RestirView &rv, Reservoir rsv,
                                                    const SurfaceRec &cs, uint32_t pix, f2 motion, f2 p_film,
                                                    Sampler &sp, Swl<NS> &swl) {
    const f2 prev = mk2(p_film.x - motion.x, p_film.y - motion.y);
    const uint32_t W = rv.width, H = rv.height;
    const uint32_t pslot = prev_slot(rv.frame);
    const float limit = rsv.C * float(rv.history_limit);
    f3 view_pos = cs.replaced ? load_ext(rv, cur_slot(rv.frame), pix).view_pos
                              : mk3(rv.cam_pos[0], rv.cam_pos[1], rv.cam_pos[2]);
    f3 prev_view_pos = mk3(rv.prev_cam_pos[0], rv.prev_cam_pos[1], rv.prev_cam_pos[2]);
    const int ix = to_i32(prev.x), iy = to_i32(prev.y);
    if (!(ix >= 0 && iy >= 0 && ix < int(W) && iy < int(H) && rv.temporal)) return rsv;
    // every candidate below is invalid for a pixel that sees no surface, and
    // its draws are the last of the kernel
    if (!cs.hit()) return rsv;
    uint32_t idx = min(to_u32(prev.y), H - 1u) * W + min(to_u32(prev.x), W - 1u);
    Reservoir pr = load_reservoir(rv, pslot, idx);
    pr.truncation(limit);
    SurfaceRec ps = load_surface(rv, pslot, idx);
    if (ps.replaced) prev_view_pos = load_ext(rv, pslot, idx).view_pos;
    if (valid_neighbor(cs, ps, rv.t_dot, rv.t_depth, rv.diff_factor) && pr.s.age < rv.max_age)
        return combine_temporal<NS, RGB, MS>(sc, rv, rsv, cs, pr, view_pos, prev_view_pos, sp, swl);
    for (uint32_t i = 0; i < rv.temporal_N; ++i) {
        f2 dk = square_to_disk(sp.next2d());
        f2 q = mk2(dk.x * rv.t_radius + prev.x, dk.y * rv.t_radius + prev.y);
        uint32_t j = min(to_u32(q.y), H - 1u) * W + min(to_u32(q.x), W - 1u);
        Reservoir ar = load_reservoir(rv, pslot, j);
        ar.truncation(limit);
        SurfaceRec as = load_surface(rv, pslot, j);
        if (as.replaced) prev_view_pos = load_ext(rv, pslot, j).view_pos;
        if (valid_neighbor(cs, as, rv.t_dot, rv.t_depth, rv.diff_factor) && ar.s.age < rv.max_age)
            return combine_temporal<NS, RGB, MS>(sc, rv, rsv, cs, ar, view_pos, prev_view_pos, sp, swl);
    }
    return rsv;
}

// shader0 for one pixel (direct.cpp:519-555 with compute_hit, :434-517)
template <int NS, bool RGB, int MS>
__device__ __forceinline__ void restir_initial(const SceneView &sc, const CameraView &cam, const RestirView &rv, uint32_t px,
                               uint32_t py, uint2 *stack, uint32_t stride, PathCounters &pc) {
    const uint32_t pix = py * rv.width + px;
    const uint32_t cslot = cur_slot(rv.frame);
    Sampler sp;
    sp.start(px, py, rv.frame, 0u);
    Swl<NS> swl = sample_wavelengths<NS, RGB>(px, py, rv.frame);
    f3 o, d;
    float fw;
    f2 p_film;
    camera_ray(cam, px, py, sp, o, d, fw, &p_film);
    // compute_hit: the first surface, followed through up to max_recursion
    // near-specular surfaces (each one draws what its delta lobe draws)
    SurfaceRec cs = empty_surface();
    SurfaceExt ext;
    ext.throughput = mk3(1.f, 1.f, 1.f);
    ext.view_pos = o;
    ext.t_max = 0.f;
    f3 ro = o, rd = d;
    TraceHit h = traverse<false>(sc, ro, rd, 0.f, 3.402823466e38f, stack, stride, pc.nodes, pc.tris);
    ++pc.rays;
    Surface it;
    uint32_t counter = 0;
    while (true) {
        cs.tri = h.tri;
        cs.b1 = h.b1;
        cs.b2 = h.b2;
        if (h.tri == kInvalid) break;
        it = surface_at(sc, h.tri, h.b1, h.b2, true);
        it.wo = normalize(-rd);
        ext.t_max += length(it.pos - ro) / length(rd);
        f3 v_pos = o + d * ext.t_max;
        const uint32_t mat = sc.inst[it.inst].mat;
        cs.flag = material_flag(sc, mat);
        cs.diff = material_diffuse_factor(sc, mat);
        f3 w_next = mk3(0.f, 0.f, 0.f);
        if (MS == MS_ALL && cs.near_specular()) {
            // metal and glass only: the materials with enable_delta whose
            // flag can be near-specular (material.h:264)
            bool dispersive;
            MatCtx<NS> mc = make_material<NS, RGB, MS>(sc, sc.mat[mat], it.frame, it.ng, it.wo, swl, dispersive,
                                                           it.uv);
            f3 t;
            sample_delta(mc, sp, w_next, t);
            ext.throughput = ext.throughput * t;
        }
        if (counter == 0u) {
            cs.depth = linear_depth(rv, v_pos);
            cs.normal = it.frame.z;
            cs.pos = it.pos;
        }
        ++counter;
        if (counter >= rv.max_recursion) break;
        if (!cs.near_specular()) break;
        ext.view_pos = it.pos;
        f3 n = dot(it.ng, w_next) > 0.f ? it.ng : -it.ng;
        ro = offset_ray_origin(it.pos, n);
        rd = w_next;
        h = traverse<false>(sc, ro, rd, 0.f, 3.402823466e38f, stack, stride, pc.nodes, pc.tris);
        ++pc.rays;
        cs.replaced = 1u;
    }
    const bool hit = h.tri != kInvalid;
    store_surface(rv, cslot, pix, cs);
    if (cs.replaced) store_ext(rv, cslot, pix, ext);
    Reservoir r = empty_reservoir();
    HitBsdfRec hb;
    hb.wi = mk3(0.f, 0.f, 0.f);
    hb.f = mk3(0.f, 0.f, 0.f);
    hb.pdf = -1.f;
    if (hit) r = ris<NS, RGB, MS>(sc, rv, it, sp, swl, ext.throughput, stack, stride, pc, rv.gi_aux ? &hb : nullptr);
    r.update_W(r.s.p_hat);
    r.truncation(1.f);
    // FrameBuffer::compute_motion_vec (frame_buffer.cpp:325-333)
    f2 motion = mk2(0.f, 0.f);
    if (hit) {
        f2 pr = prev_raster(cam, rv, o + d * ext.t_max);
        motion = mk2(p_film.x - pr.x, p_film.y - pr.y);
    }
    if (rv.gi_aux) {
        // the HitBSDF and motion vector ReSTIR GI reads
        // (frame_buffer.h hit_bsdfs / motion_vectors)
        rv.gi_aux[pix] = make_float4(hb.wi.x, hb.wi.y, hb.wi.z, hb.pdf);
        rv.gi_aux[size_t(rv.pixels) + pix] = make_float4(hb.f.x, hb.f.y, hb.f.z, 0.f);
        rv.gi_aux[2u * size_t(rv.pixels) + pix] = make_float4(motion.x, motion.y, 0.f, 0.f);
    }
    if (hit) r.process_occluded(occluded(sc, it, r.s.pos, stack, stride, pc));
    r = temporal_reuse<NS, RGB, MS>(sc, rv, r, cs, pix, motion, p_film, sp, swl);
    store_reservoir(rv, kPassThrough, pix, r);
}

// the neighbour pixel of a spatial draw (spatial_reuse, direct.cpp:563-575)
__device__ __forceinline__ uint32_t spatial_neighbor(const RestirView &rv, uint32_t px, uint32_t py, Sampler &sp) {
    f2 dk = square_to_disk(sp.next2d());
    f2 off = mk2(dk.x * rv.s_radius, dk.y * rv.s_radius);
    int ax = min(max(int(px) + int(roundf(off.x)), 0), int(rv.width) - 1);
    int ay = min(max(int(py) + int(roundf(off.y)), 0), int(rv.height) - 1);
    return uint32_t(ay) * rv.width + uint32_t(ax);
}

// ReSTIRDI::spatial_reuse with constant_combine or pairwise_combine
// (direct.cpp:559-581, :287-309, :264-285): spatial.sample_num neighbour
// draws (2 each) first, then the combination over the valid ones (1 draw
// per neighbour, pairwise one more for the canonical sample).  Neighbour
// reservoirs come from buffer `buf`.
template <int NS, bool RGB, int MS>
__device__ __forceinline__ Reservoir spatial_reuse(const SceneView &sc, const RestirView &rv, Reservoir r,
                                                   const SurfaceRec &cs, uint32_t px, uint32_t py, uint32_t buf,
                                                   Sampler &sp, Swl<NS> &swl) {
    if (!rv.spatial) return r;
    const uint32_t cslot = cur_slot(rv.frame);
    const uint32_t pix = py * rv.width + px;
    // first pass over the draws counts the valid neighbours; the second
    // replays them from the saved state while the combination draws from sp
    Sampler replay = sp;
    uint32_t count = 0;
    for (uint32_t i = 0; i < rv.spatial_num; ++i) {
        uint32_t j = spatial_neighbor(rv, px, py, sp);
        if (valid_neighbor(cs, load_surface(rv, cslot, j), rv.s_dot, rv.s_depth, rv.diff_factor)) ++count;
    }
    if (!cs.hit()) return r;
    f3 view_pos = cs.replaced ? load_ext(rv, cslot, pix).view_pos : mk3(rv.cam_pos[0], rv.cam_pos[1], rv.cam_pos[2]);
    Surface cit = surface_at(sc, cs.tri, cs.b1, cs.b2, true);
    cit.wo = normalize(view_pos - cit.pos);
    MatCtx<NS> mc;
    if (count > 0u) mc = material_at<NS, RGB, MS>(sc, cit, swl);
    Reservoir ret = empty_reservoir();
    if (!rv.pairwise) {
        const float inv = 1.f / float(count + 1u);
        ret.update(0.5f, r.s, cal_weight(inv, r.s.p_hat, r.W), r.C);
        for (uint32_t i = 0; i < rv.spatial_num; ++i) {
            uint32_t j = spatial_neighbor(rv, px, py, replay);
            if (!valid_neighbor(cs, load_surface(rv, cslot, j), rv.s_dot, rv.s_depth, rv.diff_factor)) continue;
            Reservoir nr = load_reservoir(rv, buf, j);
            nr.s.p_hat = p_hat_of<NS, RGB, MS>(sc, cit, mc, nr.s, swl);
            ret.update(sp.next1d(), nr.s, cal_weight(inv, nr.s.p_hat, nr.W), nr.C);
        }
    } else {
        const float num = float(count);// M - 1
        const float num_i = float(int(num));
        float canonical_weight = 0.f;
        for (uint32_t i = 0; i < rv.spatial_num; ++i) {
            uint32_t j = spatial_neighbor(rv, px, py, replay);
            SurfaceRec ns = load_surface(rv, cslot, j);
            if (!valid_neighbor(cs, ns, rv.s_dot, rv.s_depth, rv.diff_factor)) continue;
            Reservoir nr = load_reservoir(rv, buf, j);
            Surface nit = surface_at(sc, ns.tri, ns.b1, ns.b2, true);
            nit.wo = normalize(view_pos - nit.pos);
            MatCtx<NS> nm = material_at<NS, RGB, MS>(sc, nit, swl);
            // ReSTIRDI::neighbor_pairwise_MIS (direct.cpp:235-256)
            float p_cc = p_hat_of<NS, RGB, MS>(sc, cit, mc, r.s, swl);
            float p_cn = p_hat_of<NS, RGB, MS>(sc, nit, nm, r.s, swl);
            float p_nn = p_hat_of<NS, RGB, MS>(sc, nit, nm, nr.s, swl);
            float p_nc = p_hat_of<NS, RGB, MS>(sc, cit, mc, nr.s, swl);
            float mi = (1.f * p_nn) / (1.f * p_nn + num_i * p_nc);
            ret.update(sp.next1d(), nr.s, safe_weight(mi, nr.s.p_hat, nr.W), nr.C);
            float cw = ((1.f * p_cc) / (1.f * p_cc + num_i * p_cn)) / num;
            canonical_weight += isnan(cw) ? 0.f : cw;
        }
        canonical_weight = canonical_weight == 0.f ? 1.f : canonical_weight;
        // ReSTIRDI::canonical_pairwise_MIS (direct.cpp:258-262)
        ret.update(sp.next1d(), r.s, safe_weight(canonical_weight, r.s.p_hat, r.W), r.C);
    }
    ret.update_W(ret.s.p_hat);
    return ret;
}

// ReSTIRDI::shading (direct.cpp:583-636).  Its BSDF-sampled emitter term is
// overwritten before use and its draws come after every other use of the
// sampler, so it is left out.
template <int NS, bool RGB, int MS>
__device__ __forceinline__ f3 restir_shading(const SceneView &sc, const RestirView &rv, Reservoir r,
                                             const SurfaceRec &cs, uint32_t pix, Swl<NS> &swl, uint2 *stack,
                                             uint32_t stride, PathCounters &pc) {
    const f3 cam_pos = mk3(rv.cam_pos[0], rv.cam_pos[1], rv.cam_pos[2]);
    f3 view_pos = cam_pos;
    f3 thr = mk3(1.f, 1.f, 1.f);
    if (cs.replaced) {
        SurfaceExt e = load_ext(rv, cur_slot(rv.frame), pix);
        thr = e.throughput;
        view_pos = e.view_pos;
    }
    Surface it = surface_at(sc, cs.tri, cs.b1, cs.b2, true);
    it.wo = normalize(view_pos - it.pos);
    Spec<NS> value = Spec<NS>::fill(0.f), Le = Spec<NS>::fill(0.f);
    const uint32_t li = sc.inst[it.inst].light;
    if (li != kInvalid) {
        const LightRecord &L = sc.lights[li];
        if (L.type == LIGHT_AREA) {
            f3 w = cam_pos - it.pos;
            Le = light_color<NS, RGB>(sc, L, swl, it.uv) * ((dot(w, it.ng) > 0.f || L.two_sided) ? 1.f : 0.f);
        }
    } else {
        MatCtx<NS> mc = material_at<NS, RGB, MS>(sc, it, swl);
        Spec<NS> t;
#pragma unroll
        for (int i = 0; i < NS; ++i) t.v[i] = i == 0 ? thr.x : (i == 1 ? thr.y : thr.z);
        value = (light_contribution<NS, RGB, MS>(sc, it, mc, r.s, swl, nullptr) * t) / luminance(thr);
        r.process_occluded(occluded(sc, it, r.s.pos, stack, stride, pc));
        value = value * r.W;
    }
    return to_linear_srgb<NS, RGB>(sc, value + Le, swl);
}

// shader1 for one pixel (direct.cpp:638-671): spatial reuse over buffer
// `buf`, shading; the reservoir is kept for the next frame.  Returns the
// pixel's direct radiance for the film.
template <int NS, bool RGB, int MS>
__device__ __forceinline__ f3 restir_final(const SceneView &sc, const CameraView &cam, const RestirView &rv, uint32_t px,
                           uint32_t py, uint32_t buf, uint2 *stack, uint32_t stride, PathCounters &pc) {
    const uint32_t pix = py * rv.width + px;
    Sampler sp;
    sp.start(px, py, rv.frame, 0u);
    Swl<NS> swl = sample_wavelengths<NS, RGB>(px, py, rv.frame);
    f3 o, d;
    float fw;
    camera_ray(cam, px, py, sp, o, d, fw);
    sp.start(px, py, rv.frame, 1u);
    SurfaceRec cs = load_surface(rv, cur_slot(rv.frame), pix);
    Reservoir r = load_reservoir(rv, buf, pix);
    r = spatial_reuse<NS, RGB, MS>(sc, rv, r, cs, px, py, buf, sp, swl);
    f3 L = mk3(0.f, 0.f, 0.f);
    if (cs.hit()) {
        L = restir_shading<NS, RGB, MS>(sc, rv, r, cs, pix, swl, stack, stride, pc);
    } else if (sc.env_light != kInvalid) {
        Spec<NS> Lv;
        float pdf;
        env_eval<NS, RGB>(sc, normalize((o + d) - o), swl, Lv, pdf);
        L = to_linear_srgb<NS, RGB>(sc, Lv, swl);
    }
    store_reservoir(rv, cur_slot(rv.frame), pix, r);
    return L;
}

// an extra spatial pass (beyond the reference's one): buffer `in` to `out`,
// its own sampler dimension
template <int NS, bool RGB, int MS>
__device__ __forceinline__ void restir_spatial_pass(const SceneView &sc, const RestirView &rv, uint32_t px, uint32_t py,
                                    uint32_t in, uint32_t out, uint32_t dim) {
    const uint32_t pix = py * rv.width + px;
    Sampler sp;
    sp.start(px, py, rv.frame, dim);
    Swl<NS> swl = sample_wavelengths<NS, RGB>(px, py, rv.frame);
    SurfaceRec cs = load_surface(rv, cur_slot(rv.frame), pix);
    Reservoir r = load_reservoir(rv, in, pix);
    r = spatial_reuse<NS, RGB, MS>(sc, rv, r, cs, px, py, in, sp, swl);
    store_reservoir(rv, out, pix, r);
}

}// namespace dev
}// namespace vs
