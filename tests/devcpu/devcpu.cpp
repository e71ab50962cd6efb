// Host build of the device path-tracing code (tests only): the same headers
// the HIP kernels use, compiled by g++ against tests/devcpu/hip/hip_runtime.h,
// driven like the kernels drive them, for bit-level comparison with the oracle.
#include "dev_pt.h"
#include "dev_restir_gi.h"
#include "data_tables.h"
#include <vector>
using namespace vs;
using namespace vs::dev;

template <int NS, bool RGB>
static void run(const uint32_t *mat, const float *spd, const float *wo, const float *wi, const float *lambda,
                uint32_t n, int sample, uint32_t frame, float *wi_out, float *f, float *pdf) {
    SceneView sc{};
    sc.spd = spd;
    sc.lut = lut_tables().data();
    // a mix / add brings its two components as records 1 and 2
    MaterialRecord recs[3];
    std::memcpy(&recs[0], mat, sizeof(MaterialRecord));
    if (recs[0].type == MAT_MIX || recs[0].type == MAT_ADD) std::memcpy(&recs[1], mat + 32, 2 * sizeof(MaterialRecord));
    sc.mat = recs;
    const MaterialRecord &rec = recs[0];
    for (uint32_t i = 0; i < n; ++i) {
        Swl<NS> swl;
        for (int k = 0; k < NS; ++k) {
            swl.lambda[k] = RGB ? (k == 0 ? 602.785f : (k == 1 ? 539.285f : 445.772f)) : lambda[4 * i + k];
            swl.pdf[k] = RGB ? 1.f : visible_wavelength_pdf(swl.lambda[k]);
        }
        Frame fr{mk3(1, 0, 0), mk3(0, 1, 0), mk3(0, 0, 1)};
        f3 wo3 = mk3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
        bool disp;
        MatCtx<NS> c = make_material<NS, RGB>(sc, rec, fr, mk3(0, 0, 1), wo3, swl, disp, mk2(0, 0));
        if (!sample) {
            EvalN<NS> e = mat_eval_any<NS, RGB>(sc, c, mk3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]), nullptr, swl);
            for (int k = 0; k < 4; ++k) f[4 * i + k] = k < NS ? e.f.v[k] : 0.f;
            pdf[i] = e.pdf;
        } else {
            Sampler sp;
            sp.start(0u, i, frame, 0u);
            SampleN<NS> s = mat_sample_any<NS, RGB>(sc, c, sp, swl);
            wi_out[3 * i] = s.wi.x, wi_out[3 * i + 1] = s.wi.y, wi_out[3 * i + 2] = s.wi.z;
            for (int k = 0; k < 4; ++k) f[4 * i + k] = k < NS ? s.e.f.v[k] : 0.f;
            pdf[i] = s.e.pdf;
        }
    }
}

extern "C" void dc_bsdf(int rgb, int ns, const uint32_t *mat, const float *spd, const float *wo, const float *wi,
                        const float *lambda, uint32_t n, int sample, uint32_t frame, float *wi_out, float *f, float *pdf) {
    if (rgb) run<3, true>(mat, spd, wo, wi, lambda, n, sample, frame, wi_out, f, pdf);
    else if (ns == 4) run<4, false>(mat, spd, wo, wi, lambda, n, sample, frame, wi_out, f, pdf);
    else run<3, false>(mat, spd, wo, wi, lambda, n, sample, frame, wi_out, f, pdf);
}

// ---- full-path CPU render of the device code
#include "bvh.h"
#include "scene.h"
#include "tables.h"
#include <string>

struct DcScene {
    SceneHost h;
    PackedScene pk;
    Bvh8 bvh;
    SceneView v{};
    CameraView cam{};
    FilterTable fis;
};

extern "C" void *dc_load(const char *text, const char *base) {
    auto *s = new DcScene();
    s->h = load_scene(text, base);
    s->pk = pack_scene(s->h);
    s->bvh = build_bvh8(s->h.wP, s->h.wTri);
    SceneView &v = s->v;
    const SceneHost &h = s->h;
    const PackedScene &pk = s->pk;
    v.nodes = reinterpret_cast<const uint32_t *>(s->bvh.nodes.data());
    v.leaf_tris = reinterpret_cast<const float *>(s->bvh.tris.data());
    v.P = h.wP.data();
    v.N = h.wN.data();
    v.UV = h.wUV.data();
    v.tri = h.wTri.data();
    v.tri_inst = h.triInst.data();
    v.inst = h.instRecords.data();
    v.mat = pk.mats.data();
    v.lights = pk.lights.data();
    v.alias = pk.alias.data();
    v.alias_func = pk.func.data();
    v.num_lights = uint32_t(h.lights.size());
    v.light_sampler = h.light_sampler_type == "power" ? 1u : 0u;
    v.select_offset = pk.select_offset;
    v.select_integral = pk.select_integral;
    v.spd = pk.spd.data();
    v.lut = lut_tables().data();
    v.cie_interval = pk.cie_interval;
    v.cie_y_integral = pk.cie_y_integral;
    v.spectrum_rgb = pk.rgb ? 1u : 0u;
    v.env_light = pk.env_light;
    v.env_tex = pk.env_tex.data();
    v.env_cond = pk.env_cond.data();
    v.env_cond_func = pk.env_cond_func.data();
    v.env_marg = pk.env_marg.data();
    v.env_marg_func = pk.env_marg_func.data();
    v.env_w = pk.env_w;
    v.env_h = pk.env_h;
    v.env_marg_integral = pk.env_marg_integral;
    v.world_diameter = 2.f * pk.world_radius;
    for (int i = 0; i < 9; ++i) v.env_w2o[i] = pk.env_w2o[i], v.env_o2w[i] = pk.env_o2w[i];
    v.env_scale = pk.env_scale;
    v.tex = reinterpret_cast<const float4 *>(pk.tex.data());
    v.tex_info = reinterpret_cast<const uint4 *>(pk.tex_info.data());
    v.stack_size = s->bvh.max_depth;
    v.ray_offset_factor = h.ray_offset_factor;
    v.num_tris = uint32_t(h.wTri.size());
    return s;
}

extern "C" void dc_set_camera(void *p, const float *c) {
    auto *s = static_cast<DcScene *>(p);
    CameraView &v = s->cam;
    for (int i = 0; i < 12; ++i) v.c2w[i] = c[i];
    v.lower_x = c[12]; v.upper_y = c[13]; v.raster_sx = c[14]; v.raster_sy = c[15]; v.tan_half_fov = c[16];
    v.lens_radius = c[17]; v.focal_distance = c[18]; v.width = uint32_t(c[19]); v.height = uint32_t(c[20]);
    v.filter = uint32_t(c[21]); v.filter_radius = c[22]; v.thin_lens = uint32_t(c[23]);
    if (v.filter == 2u) {
        s->fis = build_filter_table(s->h.camera);
        v.fis_lut = s->fis.lut.data();
        v.fis_cond = s->fis.cond.data();
        v.fis_cond_func = s->fis.cond_func.data();
        v.fis_marg = s->fis.marg.data();
        v.fis_marg_func = s->fis.marg_func.data();
        v.fis_marg_integral = s->fis.marg_integral;
    }
}

// every sample of one pixel through render_column, the loop the PT and
// serving megakernels run
template <int NS, bool RGB>
static void render_px(DcScene *s, const RenderView &rv, uint32_t px, uint32_t py, float4 *rt) {
    std::vector<uint2> stack(64);
    PathCounters pc{0, 0, 0};
    render_column<NS, RGB>(s->v, s->cam, rv, rt, nullptr, px, py, 1u, 1u, rv.frame_start, rv.spp, stack.data(), 1u,
                           pc);
}

extern "C" void dc_render(void *p, const float *render, uint32_t spp, float *film) {
    auto *s = static_cast<DcScene *>(p);
    RenderView rv{};
    rv.max_depth = uint32_t(render[0]);
    rv.min_depth = uint32_t(render[1]);
    rv.rr_threshold = render[2];
    rv.mis_mode = int(render[3]);
    rv.frame_start = 0;
    rv.spp = spp;
    rv.accumulation = 1;
    rv.spectrum_dim = s->pk.rgb ? 3u : s->h.spectrum_dim;
    std::vector<float4> rt(size_t(s->cam.width) * s->cam.height, make_float4(0.f, 0.f, 0.f, 0.f));
    for (uint32_t y = 0; y < s->cam.height; ++y)
        for (uint32_t x = 0; x < s->cam.width; ++x) {
            if (s->pk.rgb) render_px<3, true>(s, rv, x, y, rt.data());
            else if (rv.spectrum_dim == 4) render_px<4, false>(s, rv, x, y, rt.data());
            else render_px<3, false>(s, rv, x, y, rt.data());
        }
    for (size_t i = 0; i < rt.size(); ++i) {
        film[4 * i + 0] = rt[i].x;
        film[4 * i + 1] = rt[i].y;
        film[4 * i + 2] = rt[i].z;
        film[4 * i + 3] = rt[i].w;
    }
}

// ---- ReSTIR DI frame of the device code (restir.hip's three kernels, in order)
template <int NS, bool RGB, int MS>
static void restir_frame(DcScene *s, const float *e, const float *prev_w2c, const float *prev_pos, uint32_t frame,
                         float *rsv, float *surf, float *ext, float *film, const float *render, float *gi_rsv,
                         float *gi_aux, float *gi_rad) {
    RestirView v{};
    v.rsv = reinterpret_cast<float4 *>(rsv);
    v.surf = reinterpret_cast<float4 *>(surf);
    v.ext = reinterpret_cast<float4 *>(ext);
    v.width = s->cam.width;
    v.height = s->cam.height;
    v.pixels = v.width * v.height;
    v.frame = frame;
    v.M_light = uint32_t(e[1]);
    v.M_bsdf = uint32_t(e[2]);
    v.max_age = uint32_t(e[3]);
    v.max_recursion = uint32_t(e[4]);
    v.diff_factor = e[5];
    v.pairwise = e[6] != 0.f;
    v.spatial = e[7] != 0.f;
    v.spatial_num = uint32_t(e[8]);
    v.s_dot = e[9];
    v.s_depth = e[10];
    v.s_radius = e[11];
    v.temporal = e[12] != 0.f;
    v.history_limit = uint32_t(e[13]);
    v.temporal_N = uint32_t(e[14]);
    v.t_radius = e[15];
    v.t_dot = e[16];
    v.t_depth = e[17];
    v.temporal_mis = e[18] != 0.f;
    for (int i = 0; i < 12; ++i) v.w2c[i] = e[20 + i], v.prev_w2c[i] = prev_w2c[i];
    for (int i = 0; i < 3; ++i) v.cam_pos[i] = s->cam.c2w[4 * i + 3], v.prev_cam_pos[i] = prev_pos[i];
    // ReSTIR GI open (restir[32]): shader0's record and the direct radiance
    const bool gi = RGB && e[32] != 0.f && gi_rsv;
    std::vector<float4> direct;
    if (gi) {
        direct.resize(v.pixels);
        v.gi_aux = reinterpret_cast<float4 *>(gi_aux);
        v.gi_direct = direct.data();
    }
    std::vector<uint2> stack(64);
    PathCounters pc{0, 0, 0};
    for (uint32_t y = 0; y < v.height; ++y)
        for (uint32_t x = 0; x < v.width; ++x) restir_initial<NS, RGB, MS>(s->v, s->cam, v, x, y, stack.data(), 1, pc);
    uint32_t in = 2, extra = uint32_t(e[19]);
    for (uint32_t k = 0; k < extra; ++k) {
        for (uint32_t y = 0; y < v.height; ++y)
            for (uint32_t x = 0; x < v.width; ++x) restir_spatial_pass<NS, RGB, MS>(s->v, v, x, y, in, in ^ 1u, 2u + k);
        in ^= 1u;
    }
    RenderView rv{};
    rv.max_depth = uint32_t(render[0]);
    rv.min_depth = uint32_t(render[1]);
    rv.rr_threshold = render[2];
    rv.mis_mode = int(render[3]);
    rv.accumulation = render[6] != 0.f;
    float4 *f4 = reinterpret_cast<float4 *>(film);
    for (uint32_t y = 0; y < v.height; ++y)
        for (uint32_t x = 0; x < v.width; ++x) {
            f3 L = restir_final<NS, RGB, MS>(s->v, s->cam, v, x, y, in, stack.data(), 1, pc);
            if (gi) direct[y * v.width + x] = make_float4(L.x, L.y, L.z, 1.f);
            else f4[y * v.width + x] = film_add(rv, f4[y * v.width + x], L, frame);
        }
    if (!gi) return;
    // ReSTIR GI: restir.hip's two GI kernels, the second with the combine
    RestirGiView g{};
    g.rsv = reinterpret_cast<float4 *>(gi_rsv);
    g.aux = reinterpret_cast<const float4 *>(gi_aux);
    g.direct = direct.data();
    g.rad = reinterpret_cast<float4 *>(gi_rad);
    g.max_age = uint32_t(e[33]);
    g.diff_factor = e[34];
    g.spatial = e[35] != 0.f;
    g.s_num = uint32_t(e[36]);
    g.s_dot = e[37];
    g.s_depth = e[38];
    g.s_radius = e[39];
    g.temporal = e[40] != 0.f;
    g.t_limit = uint32_t(e[41]);
    g.t_N = uint32_t(e[42]);
    g.t_dot = e[43];
    g.t_depth = e[44];
    g.t_radius = e[45];
    g.depth = uint32_t(e[46]);
    for (uint32_t y = 0; y < v.height; ++y)
        for (uint32_t x = 0; x < v.width; ++x)
            restir_gi_initial_temporal<MS>(s->v, s->cam, rv, v, g, x, y, stack.data(), 1, pc);
    for (uint32_t y = 0; y < v.height; ++y)
        for (uint32_t x = 0; x < v.width; ++x) {
            const uint32_t idx = y * v.width + x;
            const SurfaceRec cs = load_surface(v, cur_slot(v.frame), idx);
            f3 il = mk3(0.f, 0.f, 0.f);
            if (cs.hit()) {
                il = restir_gi_spatial_shade<MS>(s->v, v, g, cs, x, y, stack.data(), 1, pc);
                g.rad[idx] = make_float4(il.x, il.y, il.z, 1.f);
            }
            const float4 dl = direct[idx];
            f3 L = mk3(dl.x * 1.f + il.x * 1.f, dl.y * 1.f + il.y * 1.f, dl.z * 1.f + il.z * 1.f);
            f4[idx] = film_add(rv, f4[idx], L, frame);
        }
}

extern "C" void dc_restir_frame(void *p, const float *restir, const float *prev_w2c, const float *prev_pos,
                                uint32_t frame, float *rsv, float *surf, float *ext, float *film, const float *render,
                                float *gi_rsv, float *gi_aux, float *gi_rad) {
    auto *s = static_cast<DcScene *>(p);
    // the material set the library picks (diffuse family only, or all)
    bool diffuse_only = true;
    for (const MaterialRecord &m : s->pk.mats)
        if (m.type != MAT_DIFFUSE && m.type != MAT_OREN_NAYAR && m.type != MAT_BLACK) diffuse_only = false;
    if (diffuse_only) {
        if (s->pk.rgb) restir_frame<3, true, MS_DIFFUSE>(s, restir, prev_w2c, prev_pos, frame, rsv, surf, ext, film, render, gi_rsv, gi_aux, gi_rad);
        else if (s->h.spectrum_dim == 4) restir_frame<4, false, MS_DIFFUSE>(s, restir, prev_w2c, prev_pos, frame, rsv, surf, ext, film, render, gi_rsv, gi_aux, gi_rad);
        else restir_frame<3, false, MS_DIFFUSE>(s, restir, prev_w2c, prev_pos, frame, rsv, surf, ext, film, render, gi_rsv, gi_aux, gi_rad);
    } else {
        if (s->pk.rgb) restir_frame<3, true, MS_ALL>(s, restir, prev_w2c, prev_pos, frame, rsv, surf, ext, film, render, gi_rsv, gi_aux, gi_rad);
        else if (s->h.spectrum_dim == 4) restir_frame<4, false, MS_ALL>(s, restir, prev_w2c, prev_pos, frame, rsv, surf, ext, film, render, gi_rsv, gi_aux, gi_rad);
        else restir_frame<3, false, MS_ALL>(s, restir, prev_w2c, prev_pos, frame, rsv, surf, ext, film, render, gi_rsv, gi_aux, gi_rad);
    }
}
