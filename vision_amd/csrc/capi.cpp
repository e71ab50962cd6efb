// extern "C" boundary of libvision_amd.so (declared in include/vision_amd.h).
#include "../../include/vision_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "bvh.h"
#include "context.h"
#include "dev_scene.h"
#include "launch.h"
#include "scene.h"
#include "data_tables.h"
#include "image_io.h"
#include "restir.h"
#include "svgf.h"
#include "tables.h"
#include "wavefront.h"

namespace {

using namespace vs_internal;

}// namespace

namespace vs_internal {

void camera_set_resolution(vs::CameraView &v, uint32_t width, uint32_t height) {
    // RadianceCollector::update_screen_window + Sensor::update_resolution_
    // (src/base/sensor/radiance_collector.cpp:33-42, src/base/sensor/sensor.cpp:63-76)
    float ratio = float(width) / float(height);
    float lower_x = -1.f, upper_x = 1.f, lower_y = -1.f, upper_y = 1.f;
    if (ratio > 1.f) {
        lower_x = -ratio;
        upper_x = ratio;
    } else {
        lower_y = -1.f / ratio;
        upper_y = 1.f / ratio;
    }
    v.lower_x = lower_x;
    v.upper_y = upper_y;
    v.raster_sx = (upper_x - lower_x) / float(width);
    v.raster_sy = (upper_y - lower_y) / float(height);
    v.width = width;
    v.height = height;
}

void upload_mode_tables(const vs::PackedScene &pk, ModeTables &t, hipStream_t st) {
    t.mat.upload(pk.mats, st);
    t.light.upload(pk.lights, st);
    t.alias.upload(pk.alias, st);
    t.func.upload(pk.func, st);
    t.spd.upload(pk.spd, st);
    t.env_tex.upload(pk.env_tex, st);
    t.env_cond.upload(pk.env_cond, st);
    t.env_cond_func.upload(pk.env_cond_func, st);
    t.env_marg.upload(pk.env_marg, st);
    t.env_marg_func.upload(pk.env_marg_func, st);
    t.tex.upload(pk.tex, st);
    t.tex_info.upload(pk.tex_info, st);
}

void set_mode_view(vs::SceneView &v, const vs::PackedScene &pk, const ModeTables &t) {
    v.mat = t.mat.as<vs::MaterialRecord>();
    v.lights = t.light.as<vs::LightRecord>();
    v.alias = t.alias.as<vs::AliasEntry>();
    v.alias_func = t.func.as<float>();
    v.select_offset = pk.select_offset;
    v.select_integral = pk.select_integral;
    v.spd = t.spd.as<float>();
    v.cie_interval = pk.cie_interval;
    v.cie_y_integral = pk.cie_y_integral;
    v.spectrum_rgb = pk.rgb ? 1u : 0u;
    v.env_light = pk.env_light;
    v.env_tex = t.env_tex.as<float>();
    v.env_cond = t.env_cond.as<vs::AliasEntry>();
    v.env_cond_func = t.env_cond_func.as<float>();
    v.env_marg = t.env_marg.as<vs::AliasEntry>();
    v.env_marg_func = t.env_marg_func.as<float>();
    v.env_w = pk.env_w;
    v.env_h = pk.env_h;
    v.env_marg_integral = pk.env_marg_integral;
    v.world_diameter = 2.f * pk.world_radius;
    for (int i = 0; i < 9; ++i) {
        v.env_w2o[i] = pk.env_w2o[i];
        v.env_o2w[i] = pk.env_o2w[i];
    }
    v.env_scale = pk.env_scale;
    v.tex = t.tex.as<float4>();
    v.tex_info = t.tex_info.as<uint4>();
}

vs::RenderView render_view(const vs_scene *s) {
    const vs::SceneHost &h = s->host;
    vs::RenderView rv{};
    rv.max_depth = h.integrator.max_depth;
    rv.min_depth = h.integrator.min_depth;
    rv.rr_threshold = h.integrator.rr_threshold;
    rv.mis_mode = h.integrator.mis_mode;
    rv.accumulation = h.camera.accumulation;
    rv.exposure = h.camera.exposure;
    rv.tone_mapper = h.camera.tone_mapper == "aces" ? 1u : (h.camera.tone_mapper == "reinhard" ? 2u : 0u);
    rv.write_output = 1;
    rv.spectrum_dim = s->packed.rgb ? 3u : h.spectrum_dim;
    return rv;
}

void ensure_alt_mode(vs_scene *s) {
    if (s->have_alt) return;
    vs::SceneHost other = s->host;
    const bool rgb = s->packed.rgb;
    other.spectrum_type = rgb ? "hero" : "srgb";
    other.spectrum_dim = rgb ? 4u : 3u;
    s->packed_alt = vs::pack_scene(other);
    upload_mode_tables(s->packed_alt, s->mt_alt, s->ctx->stream);
    hip_check(hipStreamSynchronize(s->ctx->stream), "mode tables");
    s->have_alt = true;
}

vs::SceneView view_in_mode(vs_scene *s, bool rgb) {
    vs::SceneView v = s->view;
    if (rgb != s->packed.rgb) {
        ensure_alt_mode(s);
        set_mode_view(v, s->packed_alt, s->mt_alt);
    }
    return v;
}

}// namespace vs_internal

namespace {

void fill_camera(vs_scene *s) {
    const vs::CameraHost &c = s->host.camera;
    vs::CameraView &v = s->cam;
    // c2w rows [R | t]
    for (int r = 0; r < 3; ++r)
        for (int col = 0; col < 4; ++col) v.c2w[r * 4 + col] = vs::m4_at(c.c2w, r, col);
    camera_set_resolution(v, c.width, c.height);
    v.tan_half_fov = std::tan(c.fov_y * 0.5f * 3.14159265358979323846f / 180.f);
    v.lens_radius = c.lens_radius;
    v.focal_distance = c.focal_distance;
    const int kind = vs::filter_kind(c);
    v.filter = kind <= 1 ? uint32_t(kind) : 2u;
    v.filter_radius = c.filter_radius;
    v.thin_lens = c.type == "thin_lens" ? 1u : 0u;
    s->exp_camera.assign(v.c2w, v.c2w + 12);
    float rest[20] = {v.lower_x, v.upper_y, v.raster_sx, v.raster_sy, v.tan_half_fov, v.lens_radius,
                      v.focal_distance, float(v.width), float(v.height), float(v.filter), v.filter_radius,
                      float(v.thin_lens), float(kind), c.filter_sigma, c.filter_b, c.filter_c, c.filter_tau,
                      0.f, 0.f, 0.f};
    s->exp_camera.insert(s->exp_camera.end(), rest, rest + 20);
}

// inverse of an affine 3x4 camera-to-world (rows), in double precision
void affine_inverse(const float *m, float *out) {
    double a[3][3], t[3];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) a[r][c] = m[r * 4 + c];
        t[r] = m[r * 4 + 3];
    }
    double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    if (det == 0.0) throw std::invalid_argument("camera matrix is singular");
    double inv[3][3];
    inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
    inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
    inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
    inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
    inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
    inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
    inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
    inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
    inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) out[r * 4 + c] = float(inv[r][c]);
        out[r * 4 + 3] = float(-(inv[r][0] * t[0] + inv[r][1] * t[1] + inv[r][2] * t[2]));
    }
}

float radians_to_cos(float deg) { return std::cos(deg * (3.14159265358979323846f / 180.f)); }

void fill_restir_export(vs_scene *s) {
    s->exp_restir.assign(64, 0.f);
    const vs::IntegratorHost &in = s->host.integrator;
    if (in.type != "rt") return;
    const vs::RestirHost &r = in.restir;
    float *e = s->exp_restir.data();
    e[0] = 1.f;
    e[1] = float(r.M_light);
    e[2] = float(r.M_bsdf);
    e[3] = float(r.max_age);
    e[4] = float(r.max_recursion);
    e[5] = r.diff_factor;
    e[6] = r.pairwise ? 1.f : 0.f;
    e[7] = r.s_open ? 1.f : 0.f;
    e[8] = float(r.s_num);
    e[9] = radians_to_cos(r.s_theta);
    e[10] = r.s_depth;
    e[11] = r.s_radius;
    e[12] = r.t_open ? 1.f : 0.f;
    e[13] = float(r.t_limit);
    e[14] = float(r.t_N);
    e[15] = r.t_radius;
    e[16] = radians_to_cos(r.t_theta);
    e[17] = r.t_depth;
    e[18] = r.t_mis ? 1.f : 0.f;
    e[19] = float(r.extra_spatial_passes);
    affine_inverse(s->cam.c2w, e + 20);
    // ReSTIR GI: construct_param (indirect.cpp:334-352) and the path depth
    // (the real-time integrator's max_depth - 1, rt.cpp:30)
    const vs::RestirGiHost &g = in.gi;
    e[32] = g.open ? 1.f : 0.f;
    e[33] = float(g.max_age);
    e[34] = g.diff_factor;
    e[35] = g.s_open ? 1.f : 0.f;
    e[36] = float(g.s_num);
    e[37] = radians_to_cos(g.s_theta);
    e[38] = g.s_depth;
    e[39] = g.s_radius;
    e[40] = g.t_open ? 1.f : 0.f;
    e[41] = float(g.t_limit);
    e[42] = float(g.t_N);
    e[43] = radians_to_cos(g.t_theta);
    e[44] = g.t_depth;
    e[45] = g.t_radius;
    e[46] = float(std::max(int(in.max_depth) - 1, 0));
    // SVGF (restir[48..57]): on, N, alpha, moments_alpha, history_limit,
    // moments_filter_radius, sigma_rt, sigma_normal, reproject, filter moments
    const vs::SvgfHost &d = in.svgf;
    e[48] = d.svgf && d.on ? 1.f : 0.f;
    e[49] = float(d.N);
    e[50] = d.alpha;
    e[51] = d.moments_alpha;
    e[52] = float(d.history_limit);
    e[53] = float(d.moments_radius);
    e[54] = d.sigma_rt;
    e[55] = d.sigma_normal;
    e[56] = d.reproject ? 1.f : 0.f;
    e[57] = d.filter_moments ? 1.f : 0.f;
}

void fill_render_export(vs_scene *s) {
    const vs::SceneHost &h = s->host;
    const vs::AdaptiveHost &a = h.integrator.adaptive;
    s->exp_render = {float(h.integrator.max_depth), float(h.integrator.min_depth), h.integrator.rr_threshold,
                     float(h.integrator.mis_mode), h.light_sampler_type == "power" ? 1.f : 0.f, h.ray_offset_factor,
                     float(h.camera.accumulation), h.camera.exposure,
                     float(h.camera.tone_mapper == "aces" ? 1 : (h.camera.tone_mapper == "reinhard" ? 2 : 0)),
                     a.on && h.integrator.type == "pt" ? 1.f : 0.f, a.threshold, float(a.min_samples)};
}

void alloc_film(vs_scene *s) {
    size_t px = size_t(s->host.camera.width) * s->host.camera.height;
    s->film_rt.alloc(px * 16);
    s->film_out.alloc(px * 16);
    hip_check(hipMemsetAsync(s->film_rt.ptr, 0, px * 16, s->ctx->stream), "film clear");
    hip_check(hipMemsetAsync(s->film_out.ptr, 0, px * 16, s->ctx->stream), "film clear");
    s->frame_index = 0;
}


}// namespace

extern "C" {

int32_t vs_version(void) { return 100; }

vs_status vs_rgb_to_spectrum(const float *rgb, uint32_t n, int32_t unbound, float *coeffs) {
    if ((n && (!rgb || !coeffs))) return VS_ERROR_INVALID_ARGUMENT;
    try {
        for (uint32_t i = 0; i < n; ++i) {
            vs::f3 c = vs::mk3(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
            vs::f4 r = unbound ? vs::rgb_to_unbound_coeffs(c) : vs::rgb_to_albedo_coeffs(c);
            coeffs[4 * i] = r.x;
            coeffs[4 * i + 1] = r.y;
            coeffs[4 * i + 2] = r.z;
            coeffs[4 * i + 3] = r.w;
        }
    } catch (const std::exception &) {
        return VS_ERROR_STATE;
    }
    return VS_OK;
}

vs_status vs_create(int32_t device, vs_context **out) {
    if (!out) return VS_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (device == -1) {
        auto *host_ctx = new (std::nothrow) vs_context();
        if (!host_ctx) return VS_ERROR_OUT_OF_MEMORY;
        host_ctx->device = -1;
        *out = host_ctx;
        return VS_OK;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return VS_ERROR_DEVICE;
    if (device < 0 || device >= count) return VS_ERROR_INVALID_ARGUMENT;
    auto *ctx = new (std::nothrow) vs_context();
    if (!ctx) return VS_ERROR_OUT_OF_MEMORY;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
        hipEventCreate(&ctx->tr0) != hipSuccess || hipEventCreate(&ctx->tr1) != hipSuccess ||
        hipEventCreate(&ctx->tr2) != hipSuccess) {
        vs_destroy(ctx);
        return VS_ERROR_DEVICE;
    }
    try {
        vs_internal::noalloc_setup(device);
    } catch (const std::exception &) {
        vs_destroy(ctx);
        return VS_ERROR_DEVICE;
    }
    *out = ctx;
    return VS_OK;
}

void vs_destroy(vs_context *ctx) {
    if (!ctx) return;
    if (ctx->device < 0) {
        delete ctx;
        return;
    }
    (void)hipSetDevice(ctx->device);
    for (hipEvent_t e : {ctx->ev0, ctx->ev1, ctx->tr0, ctx->tr1, ctx->tr2})
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *vs_last_error(const vs_context *ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

vs_status vs_scene_load(vs_context *ctx, const char *json_text, const char *base_dir, vs_scene **out) {
    if (!ctx || !json_text || !out) return fail(ctx, VS_ERROR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    return guarded(ctx, [&] {
        auto s = std::make_unique<vs_scene>();
        s->ctx = ctx;
        s->host = vs::load_scene(json_text, base_dir ? base_dir : "");
        s->packed = vs::pack_scene(s->host);
        fill_camera(s.get());
        std::memcpy(s->prev_c2w, s->cam.c2w, sizeof(s->prev_c2w));
        fill_render_export(s.get());
        fill_restir_export(s.get());
        for (const auto &ir : s->host.instRecords) {
            s->exp_inst.push_back(ir.mat);
            s->exp_inst.push_back(ir.light);
            s->exp_inst.push_back(ir.tri_base);
            s->exp_inst.push_back(0);
        }
        *out = s.release();
    });
}

vs_status vs_scene_load_file(vs_context *ctx, const char *path, vs_scene **out) {
    if (!ctx || !path || !out) return fail(ctx, VS_ERROR_INVALID_ARGUMENT, "null argument");
    std::string text;
    vs_status st = guarded(ctx, [&] { text = vs::read_text_file(path); });
    if (st != VS_OK) return st;
    std::string p(path);
    size_t slash = p.find_last_of('/');
    std::string dir = slash == std::string::npos ? "." : p.substr(0, slash);
    return vs_scene_load(ctx, text.c_str(), dir.c_str(), out);
}

void vs_scene_destroy(vs_scene *scene) {
    if (!scene) return;
    if (scene->ctx && scene->ctx->device >= 0) {
        (void)hipSetDevice(scene->ctx->device);
        (void)hipStreamSynchronize(scene->ctx->stream);
    }
    delete scene;
}

vs_status vs_scene_set_kernel(vs_scene *s, int32_t mode, uint32_t chunk_frames, uint32_t trace_block) {
    if (!s) return VS_ERROR_INVALID_ARGUMENT;
    if (mode != VS_KERNEL_MEGAKERNEL && mode != VS_KERNEL_WAVEFRONT)
        return fail(s->ctx, VS_ERROR_INVALID_ARGUMENT, "unknown kernel mode");
    if (chunk_frames == 0 || chunk_frames > 1024)
        return fail(s->ctx, VS_ERROR_INVALID_ARGUMENT, "chunk_frames must be in 1..1024");
    if (trace_block != 64 && trace_block != 128 && trace_block != 256)
        return fail(s->ctx, VS_ERROR_INVALID_ARGUMENT, "trace_block must be 64, 128 or 256");
    s->kernel_mode = uint32_t(mode);
    s->wf_chunk_frames = chunk_frames;
    s->wf_trace_block = trace_block;
    return VS_OK;
}

vs_status vs_scene_build_accel(vs_scene *s) {
    if (!s) return VS_ERROR_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] {
        auto t0 = std::chrono::steady_clock::now();
        s->bvh = vs::build_bvh8(s->host.wP, s->host.wTri);
        auto t1 = std::chrono::steady_clock::now();
        s->build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        if (s->bvh.max_depth > 64) throw std::runtime_error("bvh deeper than the 64-entry traversal stack");
        s->host_bvh = true;
        if (s->ctx->device < 0) return;// host-only context: no upload
        hip_check(hipSetDevice(s->ctx->device), "hipSetDevice");
        hipStream_t st = s->ctx->stream;
        const vs::SceneHost &h = s->host;
        s->nodes.upload(s->bvh.nodes, st);
        s->ltris.upload(s->bvh.tris, st);
        s->P.upload(h.wP, st);
        s->N.upload(h.wN, st);
        s->UV.upload(h.wUV, st);
        s->tri.upload(h.wTri, st);
        s->tri_inst.upload(h.triInst, st);
        s->inst.upload(h.instRecords, st);
        s->lut.upload(vs::lut_tables(), st);
        upload_mode_tables(s->packed, s->mt, st);
        vs::FilterTable ft = vs::build_filter_table(h.camera);
        s->fis_lut.upload(ft.lut, st);
        s->fis_cond.upload(ft.cond, st);
        s->fis_cond_func.upload(ft.cond_func, st);
        s->fis_marg.upload(ft.marg, st);
        s->fis_marg_func.upload(ft.marg_func, st);
        s->cam.fis_lut = s->fis_lut.as<float>();
        s->cam.fis_cond = s->fis_cond.as<vs::AliasEntry>();
        s->cam.fis_cond_func = s->fis_cond_func.as<float>();
        s->cam.fis_marg = s->fis_marg.as<vs::AliasEntry>();
        s->cam.fis_marg_func = s->fis_marg_func.as<float>();
        s->cam.fis_marg_integral = ft.marg_integral;
        s->counters.alloc(64);
        alloc_film(s);
        hip_check(hipStreamSynchronize(st), "upload sync");
        vs::SceneView &v = s->view;
        v.nodes = s->nodes.as<uint32_t>();
        v.leaf_tris = s->ltris.as<float>();
        v.P = s->P.as<vs::f3>();
        v.N = s->N.as<vs::f3>();
        v.UV = s->UV.as<vs::f2>();
        v.tri = s->tri.as<vs::u3>();
        v.tri_inst = s->tri_inst.as<uint32_t>();
        v.inst = s->inst.as<vs::InstanceRecord>();
        v.num_lights = uint32_t(h.lights.size());
        v.light_sampler = h.light_sampler_type == "power" ? 1u : 0u;
        v.lut = s->lut.as<float>();
        set_mode_view(v, s->packed, s->mt);
        s->diffuse_only = true;
        for (const vs::MaterialRecord &m : s->packed.mats)
            if (m.type != vs::MAT_DIFFUSE && m.type != vs::MAT_OREN_NAYAR && m.type != vs::MAT_BLACK)
                s->diffuse_only = false;
        v.stack_size = s->bvh.max_depth;
        v.ray_offset_factor = h.ray_offset_factor;
        v.num_tris = uint32_t(h.wTri.size());
        s->built = true;
    });
}

vs_status vs_scene_info_get(const vs_scene *s, vs_scene_info *out) {
    if (!s || !out) return VS_ERROR_INVALID_ARGUMENT;
    std::memset(out, 0, sizeof(*out));
    const vs::SceneHost &h = s->host;
    out->num_triangles = uint32_t(h.wTri.size());
    out->num_vertices = uint32_t(h.wP.size());
    out->num_instances = uint32_t(h.instances.size());
    out->num_materials = uint32_t(h.materials.size());
    out->num_lights = uint32_t(h.lights.size());
    out->width = h.camera.width;
    out->height = h.camera.height;
    out->output_spp = h.output_spp;
    out->max_depth = h.integrator.max_depth;
    out->spectrum_dim = h.spectrum_dim;
    out->bvh_nodes = uint32_t(s->bvh.nodes.size());
    out->bvh_depth = s->bvh.max_depth;
    out->frame_index = s->frame_index;
    out->accel_built = s->built ? 1u : 0u;
    out->bvh_build_ms = s->build_ms;
    out->bvh_sah_cost = s->bvh.sah_cost;
    return VS_OK;
}

vs_status vs_scene_export(const vs_scene *s, vs_scene_arrays *out) {
    if (!s || !out) return VS_ERROR_INVALID_ARGUMENT;
    const vs::SceneHost &h = s->host;
    out->positions = reinterpret_cast<const float *>(h.wP.data());
    out->normals = reinterpret_cast<const float *>(h.wN.data());
    out->uvs = reinterpret_cast<const float *>(h.wUV.data());
    out->triangles = reinterpret_cast<const uint32_t *>(h.wTri.data());
    out->tri_instance = h.triInst.data();
    out->instances = s->exp_inst.data();
    out->materials = reinterpret_cast<const uint32_t *>(s->packed.mats.data());
    out->lights = reinterpret_cast<const uint32_t *>(s->packed.lights.data());
    static_assert(sizeof(vs::LightRecord) == 28 * 4, "exported light record is 28 words");
    out->camera = s->exp_camera.data();
    out->render = s->exp_render.data();
    out->num_vertices = uint32_t(h.wP.size());
    out->num_triangles = uint32_t(h.wTri.size());
    out->num_instances = uint32_t(h.instances.size());
    out->num_materials = uint32_t(s->packed.mats.size());
    out->num_lights = uint32_t(s->packed.lights.size());
    const vs::PackedScene &pk = s->packed;
    out->spd = pk.spd.data();
    out->num_spd = uint32_t(pk.spd.size());
    out->spectrum_rgb = pk.rgb ? 1u : 0u;
    out->spectrum_dim = s->host.spectrum_dim;
    out->env_light = pk.env_light;
    out->env_w = pk.env_w;
    out->env_h = pk.env_h;
    const vs::LightHost *env = pk.env_light != vs::kInvalid ? &s->host.lights[pk.env_light] : nullptr;
    out->env_rgb = env ? env->env_rgb.data() : nullptr;
    for (int i = 0; i < 9; ++i) {
        out->env_w2o[i] = pk.env_w2o[i];
        out->env_o2w[i] = pk.env_o2w[i];
    }
    out->env_scale = pk.env_scale;
    out->world_radius = pk.world_radius;
    out->restir = s->exp_restir.data();
    out->num_restir = uint32_t(s->exp_restir.size());
    out->textures = s->packed.tex.data();
    out->num_texels = uint32_t(s->packed.tex.size() / 4);
    out->texture_info = s->packed.tex_info.data();
    out->num_textures = uint32_t(s->packed.tex_info.size() / 4);
    return VS_OK;
}

vs_status vs_scene_bvh_export(const vs_scene *s, const uint32_t **nodes, uint32_t *num_nodes,
                              const float **leaf_tris, uint32_t *num_leaf_tris) {
    if (!s || !nodes || !num_nodes || !leaf_tris || !num_leaf_tris) return VS_ERROR_INVALID_ARGUMENT;
    if (!s->host_bvh) return fail(s->ctx, VS_ERROR_STATE, "bvh not built");
    *nodes = reinterpret_cast<const uint32_t *>(s->bvh.nodes.data());
    *num_nodes = uint32_t(s->bvh.nodes.size());
    *leaf_tris = reinterpret_cast<const float *>(s->bvh.tris.data());
    *num_leaf_tris = uint32_t(s->bvh.tris.size());
    return VS_OK;
}

vs_status vs_scene_set_resolution(vs_scene *s, uint32_t width, uint32_t height) {
    if (!s || width == 0 || height == 0 || width > 16384 || height > 16384)
        return fail(s ? s->ctx : nullptr, VS_ERROR_INVALID_ARGUMENT, "bad resolution");
    return guarded(s->ctx, [&] {
        s->host.camera.width = width;
        s->host.camera.height = height;
        fill_camera(s);
        if (s->built && s->ctx->device >= 0) {
            hip_check(hipSetDevice(s->ctx->device), "hipSetDevice");
            alloc_film(s);
            hip_check(hipStreamSynchronize(s->ctx->stream), "film sync");
        }
    });
}

vs_status vs_scene_update_camera(vs_scene *s, const float *c2w, float fov_y) {
    if (!s || !c2w || !(fov_y > 0.f && fov_y < 180.f))
        return fail(s ? s->ctx : nullptr, VS_ERROR_INVALID_ARGUMENT, "bad camera");
    return guarded(s->ctx, [&] {
        vs::CameraHost &c = s->host.camera;
        // Sensor::set_mat (sensor.cpp:86-89): the matrix as given
        for (int col = 0; col < 4; ++col) c.c2w.c[col] = vs::mk4(c2w[4 * col], c2w[4 * col + 1], c2w[4 * col + 2], c2w[4 * col + 3]);
        c.position = vs::mk3(c2w[12], c2w[13], c2w[14]);
        c.fov_y = fov_y;
        float probe[12];
        for (int r = 0; r < 3; ++r)
            for (int col = 0; col < 4; ++col) probe[r * 4 + col] = c2w[4 * col + r];
        affine_inverse(probe, probe);// rejects a singular matrix before anything changes
        fill_camera(s);
        fill_restir_export(s);
    });
}

vs_status vs_scene_set_spectrum(vs_scene *s, int32_t spectral, uint32_t dimension) {
    if (!s) return VS_ERROR_INVALID_ARGUMENT;
    if (spectral && dimension != 3 && dimension != 4)
        return fail(s->ctx, VS_ERROR_INVALID_ARGUMENT, "hero dimension must be 3 or 4");
    return guarded(s->ctx, [&] {
        require_built(s);
        hip_check(hipSetDevice(s->ctx->device), "hipSetDevice");
        const bool rgb = !spectral;
        if (rgb != s->packed.rgb) {
            ensure_alt_mode(s);
            s->mt.swap(s->mt_alt);
            std::swap(s->packed, s->packed_alt);
            set_mode_view(s->view, s->packed, s->mt);
        }
        s->host.spectrum_type = rgb ? "srgb" : "hero";
        s->host.spectrum_dim = rgb ? 3u : dimension;
    });
}

vs_status vs_scene_set_integrator(vs_scene *s, uint32_t max_depth, uint32_t min_depth, float rr_threshold) {
    if (!s || max_depth > 1024) return fail(s ? s->ctx : nullptr, VS_ERROR_INVALID_ARGUMENT, "bad integrator settings");
    s->host.integrator.max_depth = max_depth;
    s->host.integrator.min_depth = min_depth;
    s->host.integrator.rr_threshold = rr_threshold;
    fill_render_export(s);
    return VS_OK;
}

static vs_status trace_host(vs_scene *s, const vs_ray *rays, uint32_t n, vs_hit *hits, uint8_t *occ, bool any) {
    if (!s || (n && !rays) || (n && !any && !hits) || (n && any && !occ))
        return fail(s ? s->ctx : nullptr, VS_ERROR_INVALID_ARGUMENT, "null argument");
    return guarded(s->ctx, [&] {
        require_built(s);
        if (n == 0) return;
        hip_check(hipSetDevice(s->ctx->device), "hipSetDevice");
        hipStream_t st = s->ctx->stream;
        DevBuf dr, dh;
        dr.alloc(size_t(n) * sizeof(vs::RayRec));
        dh.alloc(size_t(n) * (any ? 1 : sizeof(vs::HitRec)));
        hip_check(hipMemcpyAsync(dr.ptr, rays, size_t(n) * sizeof(vs::RayRec), hipMemcpyHostToDevice, st), "rays H2D");
        hip_check(vs::launch_trace(s->view, dr.as<vs::RayRec>(), any ? nullptr : dh.as<vs::HitRec>(),
                                   any ? dh.as<uint8_t>() : nullptr, n, any, nullptr, st),
                  "trace launch");
        if (any) hip_check(hipMemcpyAsync(occ, dh.ptr, n, hipMemcpyDeviceToHost, st), "occ D2H");
        else hip_check(hipMemcpyAsync(hits, dh.ptr, size_t(n) * sizeof(vs::HitRec), hipMemcpyDeviceToHost, st), "hits D2H");
        hip_check(hipStreamSynchronize(st), "trace sync");
    });
}

vs_status vs_trace_closest(vs_scene *s, const vs_ray *rays, uint32_t n, vs_hit *hits) {
    static_assert(sizeof(vs_ray) == sizeof(vs::RayRec), "ray layout");
    static_assert(sizeof(vs_hit) == sizeof(vs::HitRec), "hit layout");
    return trace_host(s, rays, n, hits, nullptr, false);
}

vs_status vs_trace_any(vs_scene *s, const vs_ray *rays, uint32_t n, uint8_t *occluded) {
    return trace_host(s, rays, n, nullptr, occluded, true);
}

static vs_status trace_device(vs_scene *s, const vs_ray *rays, uint32_t n, vs_hit *hits, uint8_t *occ, bool any,
                              void *stream) {
    if (!s || (n && !rays) || (n && !any && !hits) || (n && any && !occ))
        return fail(s ? s->ctx : nullptr, VS_ERROR_INVALID_ARGUMENT, "null argument");
    return guarded(s->ctx, [&] {
        require_built(s);
        hip_check(hipSetDevice(s->ctx->device), "hipSetDevice");
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s->ctx->stream;
        unsigned long long *ctr = nullptr;
        if (s->counting) {
            ctr = s->counters.as<unsigned long long>();
            hip_check(hipMemsetAsync(ctr, 0, 64, st), "counter clear");
        }
        hip_check(vs::launch_trace(s->view, reinterpret_cast<const vs::RayRec *>(rays),
                                   reinterpret_cast<vs::HitRec *>(hits), occ, n, any, ctr, st),
                  "trace launch");
    });
}

vs_status vs_trace_closest_device(vs_scene *s, const vs_ray *d_rays, uint32_t n, vs_hit *d_hits, void *stream) {
    return trace_device(s, d_rays, n, d_hits, nullptr, false, stream);
}

vs_status vs_trace_any_device(vs_scene *s, const vs_ray *d_rays, uint32_t n, uint8_t *d_occ, void *stream) {
    return trace_device(s, d_rays, n, nullptr, d_occ, true, stream);
}

vs_status vs_scene_bsdf(vs_scene *s, uint32_t mat, uint32_t n, const float *wo, const float *wi, const float *lambda,
                        int32_t sample, uint32_t frame, float *wi_out, float *f, float *pdf) {
    if (!s || (n && (!wo || !f || !pdf || (sample ? !wi_out : !wi))))
        return fail(s ? s->ctx : nullptr, VS_ERROR_INVALID_ARGUMENT, "null argument");
    return guarded(s->ctx, [&] {
        require_built(s);
        if (mat >= s->packed.mats.size()) throw std::invalid_argument("material index out of range");
        if (!s->packed.rgb && n && !lambda) throw std::invalid_argument("hero mode needs 4 wavelengths per direction");
        if (n == 0) return;
        hip_check(hipSetDevice(s->ctx->device), "hipSetDevice");
        hipStream_t st = s->ctx->stream;
        DevBuf d_wo, d_wi, d_lam, d_wio, d_f, d_pdf;
        d_wo.alloc(size_t(n) * 12);
        d_wi.alloc(size_t(n) * 12);
        d_lam.alloc(size_t(n) * 16);
        d_wio.alloc(size_t(n) * 12);
        d_f.alloc(size_t(n) * 16);
        d_pdf.alloc(size_t(n) * 4);
        hip_check(hipMemcpyAsync(d_wo.ptr, wo, size_t(n) * 12, hipMemcpyHostToDevice, st), "copy wo");
        if (!sample) hip_check(hipMemcpyAsync(d_wi.ptr, wi, size_t(n) * 12, hipMemcpyHostToDevice, st), "copy wi");
        if (!s->packed.rgb)
            hip_check(hipMemcpyAsync(d_lam.ptr, lambda, size_t(n) * 16, hipMemcpyHostToDevice, st), "copy lambda");
        uint32_t dim = s->packed.rgb ? 3u : s->host.spectrum_dim;
        hip_check(vs::launch_bsdf(s->view, dim, mat, d_wo.as<float>(), d_wi.as<float>(), d_lam.as<float>(), n,
                                  sample ? 1 : 0, frame, d_wio.as<float>(), d_f.as<float>(), d_pdf.as<float>(), st),
                  "bsdf launch");
        if (sample) hip_check(hipMemcpyAsync(wi_out, d_wio.ptr, size_t(n) * 12, hipMemcpyDeviceToHost, st), "copy wi");
        hip_check(hipMemcpyAsync(f, d_f.ptr, size_t(n) * 16, hipMemcpyDeviceToHost, st), "copy f");
        hip_check(hipMemcpyAsync(pdf, d_pdf.ptr, size_t(n) * 4, hipMemcpyDeviceToHost, st), "copy pdf");
        hip_check(hipStreamSynchronize(st), "bsdf sync");
    });
}

namespace {

}// namespace

}// extern "C"

uint32_t vs_internal::wavefront_render(const vs::SceneView &view, const vs::CameraView &cam,
                                       const vs::RenderView &rv0, bool rgb, bool diffuse_only, float4 *rt,
                                       float4 *out,
                                       uint32_t frame_index, uint32_t spp, uint32_t chunk_frames,
                                       uint32_t trace_block, WavefrontBuffers &wb, hipStream_t st,
                                       unsigned long long *ctr, vs_render_stats *ts, hipEvent_t *ev) {
    const uint32_t P = cam.width * cam.height;
    const uint32_t chunk = std::min<uint32_t>(chunk_frames, spp);
    const size_t cap = size_t(P) * chunk;
    if (cap > 0xFFFFFFFFull) throw std::invalid_argument("wavefront chunk exceeds 2^32 paths");
    if (wb.capacity < cap) {
        for (DevBuf *b : {&wb.o, &wb.d, &wb.thr, &wb.L, &wb.meta, &wb.lam, &wb.lpdf, &wb.hit, &wb.sh_o, &wb.sh_d,
                          &wb.sh_c})
            b->alloc(cap * 16);
        wb.q0.alloc(cap * 4);
        wb.q1.alloc(cap * 4);
        wb.counts.alloc(64);
        if (!wb.host_counts) {
            void *hp = nullptr;
            hip_check(hipHostMalloc(&hp, 64, hipHostMallocDefault), "pinned counts");
            wb.host_counts = static_cast<uint32_t *>(hp);
        }
        wb.capacity = cap;
    }
    vs::WavefrontView w{};
    w.o = wb.o.as<float4>();
    w.d = wb.d.as<float4>();
    w.thr = wb.thr.as<float4>();
    w.L = wb.L.as<float4>();
    w.meta = wb.meta.as<uint4>();
    w.lam = wb.lam.as<float4>();
    w.lpdf = wb.lpdf.as<float4>();
    w.hit = wb.hit.as<float4>();
    w.sh_o = wb.sh_o.as<float4>();
    w.sh_d = wb.sh_d.as<float4>();
    w.sh_c = wb.sh_c.as<float4>();
    w.pixels = P;
    w.width = cam.width;
    const uint32_t dim = rv0.spectrum_dim;
    uint32_t *q[2] = {wb.q0.as<uint32_t>(), wb.q1.as<uint32_t>()};
    uint32_t *counts = wb.counts.as<uint32_t>();
    const bool timed = ts && ev;
    uint32_t launches = 0;
    for (uint32_t f0 = 0; f0 < spp; f0 += chunk) {
        const uint32_t fc = std::min(chunk, spp - f0);
        const uint32_t n = P * fc;
        const uint32_t frame0 = frame_index + f0;
        hip_check(vs::wf_begin(cam, rv0, w, dim, rgb, frame0, n, q[0], st), "wf begin");
        ++launches;
        uint32_t count = rv0.max_depth > 0 ? n : 0u;
        int cur = 0;
        while (count) {
            hip_check(vs::wf_zero(counts, st), "wf zero");
            if (timed) hip_check(hipEventRecord(ev[0], st), "event");
            hip_check(vs::wf_trace(view, w, q[cur], count, ctr, trace_block, st), "wf trace");
            if (timed) hip_check(hipEventRecord(ev[1], st), "event");
            hip_check(vs::wf_shade(view, rv0, w, dim, rgb, diffuse_only ? 1 : 0, q[cur], count, q[cur ^ 1], counts, st), "wf shade");
            if (timed) hip_check(hipEventRecord(ev[2], st), "event");
            hip_check(vs::wf_shadow(view, w, counts, count, ctr ? ctr + 4 : nullptr, trace_block, st), "wf shadow");
            if (timed) hip_check(hipEventRecord(ev[3], st), "event");
            launches += 4;
            hip_check(hipMemcpyAsync(wb.host_counts, counts, 8, hipMemcpyDeviceToHost, st), "wf counts");
            hip_check(hipStreamSynchronize(st), "wf sync");
            if (timed) {
                float t_trace = 0.f, t_shade = 0.f, t_shadow = 0.f;
                hip_check(hipEventElapsedTime(&t_trace, ev[0], ev[1]), "event time");
                hip_check(hipEventElapsedTime(&t_shade, ev[1], ev[2]), "event time");
                hip_check(hipEventElapsedTime(&t_shadow, ev[2], ev[3]), "event time");
                ts->stage_ms[0] += t_trace;
                ts->stage_ms[1] += t_shade;
                ts->stage_ms[2] += t_shadow;
                for (int k = 0; k < 3; ++k) ts->stage_launches[k] += 1;
            }
            count = wb.host_counts[vs::WF_LIVE];
            if (count > n) throw std::logic_error("wavefront queue overflow");
            cur ^= 1;
        }
        hip_check(vs::wf_accumulate(rv0, w, frame0, fc, rt, out, st), "wf accumulate");
        ++launches;
    }
    return launches;
}

extern "C" {

namespace {

// Wavefront render of `spp` frames of a scene's film from its frame index
uint32_t render_wavefront(vs_scene *s, const vs::RenderView &rv0, uint32_t spp, unsigned long long *ctr,
                          vs_render_stats &ts) {
    hipEvent_t ev[4] = {s->ctx->tr0, s->ctx->tr1, s->ctx->tr2, s->ctx->ev1};
    return wavefront_render(s->view, s->cam, rv0, s->packed.rgb, s->diffuse_only, s->film_rt.as<float4>(),
                            s->film_out.as<float4>(),
                            s->frame_index, spp, s->wf_chunk_frames, s->wf_trace_block, s->wf, s->ctx->stream, ctr,
                            &ts, ev);
}

vs::RestirView restir_view(vs_scene *s, uint32_t frame) {
    const vs::RestirHost &r = s->host.integrator.restir;
    const float *e = s->exp_restir.data();
    vs::RestirView v{};
    v.rsv = s->rs_rsv.as<float4>();
    v.surf = s->rs_surf.as<float4>();
    v.ext = s->rs_ext.as<float4>();
    v.width = s->cam.width;
    v.height = s->cam.height;
    v.pixels = v.width * v.height;
    v.frame = frame;
    v.M_light = r.M_light;
    v.M_bsdf = r.M_bsdf;
    v.max_age = r.max_age;
    v.diff_factor = r.diff_factor;
    v.spatial = r.s_open ? 1u : 0u;
    v.spatial_num = r.s_num;
    v.s_dot = e[9];
    v.s_depth = r.s_depth;
    v.s_radius = r.s_radius;
    v.temporal = r.t_open ? 1u : 0u;
    v.history_limit = r.t_limit;
    v.t_dot = e[16];
    v.t_depth = r.t_depth;
    v.t_radius = r.t_radius;
    v.temporal_N = r.t_N;
    v.temporal_mis = r.t_mis ? 1u : 0u;
    v.pairwise = r.pairwise ? 1u : 0u;
    v.max_recursion = r.max_recursion;
    for (int i = 0; i < 3; ++i) {
        v.cam_pos[i] = s->cam.c2w[i * 4 + 3];
        v.prev_cam_pos[i] = s->prev_c2w[i * 4 + 3];
    }
    for (int i = 0; i < 12; ++i) v.w2c[i] = e[20 + i];
    affine_inverse(s->prev_c2w, v.prev_w2c);
    if (s->aux_pixels == v.pixels) v.gi_aux = s->gi_aux.as<float4>();
    if (s->host.integrator.gi.open && s->gi_pixels == v.pixels) v.gi_direct = s->gi_direct.as<float4>();
    return v;
}

bool svgf_active(const vs_scene *s) {
    const vs::SvgfHost &d = s->host.integrator.svgf;
    return s->host.integrator.type == "rt" && d.svgf && d.on;
}

vs::SvgfView svgf_view(vs_scene *s, uint32_t frame) {
    vs::SvgfView v{};
    v.geo = s->sv_geo.as<float4>();
    v.alb = s->sv_alb.as<float4>();
    v.emi = s->sv_emi.as<float4>();
    v.rec = s->sv_rec.as<float4>();
    v.width = s->cam.width;
    v.height = s->cam.height;
    v.pixels = v.width * v.height;
    v.frame = frame;
    return v;
}

vs::SvgfParams svgf_params(const vs_scene *s) {
    const vs::SvgfHost &d = s->host.integrator.svgf;
    vs::SvgfParams p{};
    p.N = d.N;
    p.alpha = d.alpha;
    p.moments_alpha = d.moments_alpha;
    p.history_limit = d.history_limit;
    p.moments_radius = d.moments_radius;
    p.sigma_rt = d.sigma_rt;
    p.sigma_normal = d.sigma_normal;
    p.reproject = d.reproject ? 1u : 0u;
    p.filter_moments = d.filter_moments ? 1u : 0u;
    return p;
}

// ReSTIR GI settings and state (construct_param, indirect.cpp:335-352, as
// exported in restir[32..46])
vs::RestirGiView restir_gi_view(vs_scene *s) {
    const float *e = s->exp_restir.data();
    vs::RestirGiView g{};
    g.rsv = s->gi_rsv.as<float4>();
    g.aux = s->gi_aux.as<float4>();
    g.direct = s->gi_direct.as<float4>();
    g.rad = s->gi_rad.as<float4>();
    g.max_age = uint32_t(e[33]);
    g.diff_factor = e[34];
    g.spatial = e[35] != 0.f ? 1u : 0u;
    g.s_num = uint32_t(e[36]);
    g.s_dot = e[37];
    g.s_depth = e[38];
    g.s_radius = e[39];
    g.temporal = e[40] != 0.f ? 1u : 0u;
    g.t_limit = uint32_t(e[41]);
    g.t_N = uint32_t(e[42]);
    g.t_dot = e[43];
    g.t_depth = e[44];
    g.t_radius = e[45];
    g.depth = uint32_t(e[46]);
    return g;
}

// RealTimeIntegrator::render (rt.cpp:144-160), direct half: per frame
// shader0, extra spatial passes, shader1 with the film update; the camera
// becomes the previous camera after each frame (Sensor::after_render)
uint32_t render_restir(vs_scene *s, const vs::RenderView &rv, uint32_t spp, unsigned long long *ctr,
                       vs_render_stats &ts) {
    hipStream_t st = s->ctx->stream;
    const uint32_t P = s->cam.width * s->cam.height;
    if (s->rs_pixels != P) {
        s->rs_rsv.alloc(size_t(P) * vs::kRestirReservoirBuffers * 3u * 16u);
        s->rs_surf.alloc(size_t(P) * 2u * 3u * 16u);
        s->rs_ext.alloc(size_t(P) * 2u * 2u * 16u);
        s->rs_pixels = P;
        hip_check(vs::restir_clear(restir_view(s, 0), st), "restir clear");
    }
    const bool gi = s->host.integrator.gi.open;
    const bool svgf = svgf_active(s);
    if ((gi || svgf) && s->aux_pixels != P) {
        s->gi_aux.alloc(size_t(P) * 3u * 16u);
        hip_check(hipMemsetAsync(s->gi_aux.ptr, 0, s->gi_aux.bytes, st), "aux clear");
        s->aux_pixels = P;
    }
    if (gi && s->gi_pixels != P) {
        s->gi_rsv.alloc(size_t(P) * vs::kGiReservoirBuffers * 4u * 16u);
        s->gi_direct.alloc(size_t(P) * 16u);
        s->gi_rad.alloc(size_t(P) * 16u);
        s->gi_pixels = P;
        hip_check(vs::restir_gi_clear(restir_gi_view(s), P, st), "restir gi clear");
    }
    if (svgf && s->sv_pixels != P) {
        s->sv_geo.alloc(size_t(P) * 2u * 2u * 16u);
        s->sv_alb.alloc(size_t(P) * 16u);
        s->sv_emi.alloc(size_t(P) * 16u);
        s->sv_rec.alloc(size_t(P) * 3u * 2u * 16u);
        s->sv_pixels = P;
        hip_check(vs::svgf_clear(svgf_view(s, 0), st), "svgf clear");
    }
    const bool rgb = s->packed.rgb;
    const uint32_t dim = rv.spectrum_dim;
    const uint32_t extra = s->host.integrator.restir.extra_spatial_passes;
    // the diffuse-family kernels when no material needs more
    int ms = 1;
    for (const vs::MaterialRecord &m : s->packed.mats)
        if (m.type != vs::MAT_DIFFUSE && m.type != vs::MAT_OREN_NAYAR && m.type != vs::MAT_BLACK) ms = 0;
    uint32_t launches = 0;
    for (uint32_t f = 0; f < spp; ++f) {
        vs::RestirView v = restir_view(s, s->frame_index + f);
        hip_check(hipEventRecord(s->ctx->tr0, st), "event");
        // the G-buffer pass (rt.cpp:147-148), then shader0
        if (svgf)
            hip_check(vs::gbuffer_compute(s->view, s->cam, v, svgf_view(s, v.frame), dim, rgb, ms, ctr, st),
                      "g-buffer");
        hip_check(vs::restir_initial(s->view, s->cam, v, dim, rgb, ms, ctr, st), "restir initial");
        hip_check(hipEventRecord(s->ctx->tr1, st), "event");
        uint32_t in = 2u;
        for (uint32_t k = 0; k < extra; ++k) {
            hip_check(vs::restir_spatial(s->view, v, dim, rgb, ms, in, in ^ 1u, 2u + k, st), "restir spatial");
            in ^= 1u;
        }
        hip_check(hipEventRecord(s->ctx->tr2, st), "event");
        hip_check(vs::restir_final(s->view, s->cam, v, rv, dim, rgb, ms, in, s->film_rt.as<float4>(),
                                   s->film_out.as<float4>(), ctr ? ctr + 4 : nullptr, st),
                  "restir final");
        if (gi) {
            // ReSTIR GI and the combine (rt.cpp:153-158), timed with shader1
            vs::RestirGiView g = restir_gi_view(s);
            hip_check(vs::restir_gi_initial_temporal(s->view, s->cam, v, g, rv, ms, ctr ? ctr + 4 : nullptr, st),
                      "restir gi initial / temporal");
            hip_check(vs::restir_gi_spatial_shade(s->view, v, g, rv, ms, s->film_rt.as<float4>(),
                                                  s->film_out.as<float4>(), ctr ? ctr + 4 : nullptr, st),
                      "restir gi spatial / shading");
            launches += 2u;
            ts.stage_launches[2] += 2u;
        }
        if (svgf) {
            // the denoiser after the combine (rt.cpp:157-158): the film's
            // running mean in, the modulated picture into the output buffer
            hip_check(vs::svgf_denoise(svgf_view(s, v.frame), svgf_params(s), s->gi_aux.as<float4>() + 2u * size_t(P),
                                       s->film_rt.as<float4>(), s->film_out.as<float4>(), st),
                      "svgf");
            const vs::SvgfParams sp = svgf_params(s);
            const uint32_t n = 1u + sp.filter_moments + sp.N + (sp.N == 0 ? 1u : 0u);
            launches += 2u + n;
            ts.stage_launches[2] += 2u + n;
        }
        hip_check(hipEventRecord(s->ctx->ev1, st), "event");
        hip_check(hipEventSynchronize(s->ctx->ev1), "restir sync");
        float t0 = 0.f, t1 = 0.f, t2 = 0.f;
        hip_check(hipEventElapsedTime(&t0, s->ctx->tr0, s->ctx->tr1), "event time");
        hip_check(hipEventElapsedTime(&t1, s->ctx->tr1, s->ctx->tr2), "event time");
        hip_check(hipEventElapsedTime(&t2, s->ctx->tr2, s->ctx->ev1), "event time");
        ts.stage_ms[0] += t0;
        ts.stage_ms[1] += t1;
        ts.stage_ms[2] += t2;
        ts.stage_launches[0] += 1;
        ts.stage_launches[1] += extra;
        ts.stage_launches[2] += 1;
        launches += 2u + extra;
        std::memcpy(s->prev_c2w, s->cam.c2w, sizeof(s->prev_c2w));
    }
    return launches;
}

}// namespace

vs_status vs_scene_restir_state(vs_scene *s, float *reservoirs, float *surfaces) {
    if (!s || !reservoirs || !surfaces) return fail(s ? s->ctx : nullptr, VS_ERROR_INVALID_ARGUMENT, "null argument");
    return guarded(s->ctx, [&] {
        require_built(s);
        if (s->rs_pixels == 0) throw std::logic_error("no ReSTIR frame rendered");
        hip_check(hipSetDevice(s->ctx->device), "hipSetDevice");
        hip_check(hipStreamSynchronize(s->ctx->stream), "restir sync");
        hip_check(hipMemcpy(reservoirs, s->rs_rsv.ptr, s->rs_rsv.bytes, hipMemcpyDeviceToHost), "restir D2H");
        hip_check(hipMemcpy(surfaces, s->rs_surf.ptr, s->rs_surf.bytes, hipMemcpyDeviceToHost), "restir D2H");
    });
}

vs_status vs_scene_restir_gi_state(vs_scene *s, float *reservoirs, float *radiance) {
    if (!s || !reservoirs || !radiance)
        return fail(s ? s->ctx : nullptr, VS_ERROR_INVALID_ARGUMENT, "null argument");
    return guarded(s->ctx, [&] {
        require_built(s);
        if (s->gi_pixels == 0) throw std::logic_error("no ReSTIR GI frame rendered");
        hip_check(hipSetDevice(s->ctx->device), "hipSetDevice");
        hip_check(hipStreamSynchronize(s->ctx->stream), "restir gi sync");
        hip_check(hipMemcpy(reservoirs, s->gi_rsv.ptr, s->gi_rsv.bytes, hipMemcpyDeviceToHost), "restir gi D2H");
        hip_check(hipMemcpy(radiance, s->gi_rad.ptr, s->gi_rad.bytes, hipMemcpyDeviceToHost), "restir gi D2H");
    });
}

vs_status vs_scene_set_adaptive(vs_scene *s, int32_t on, float threshold, uint32_t min_samples) {
    if (!s) return VS_ERROR_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] {
        if (on && s->host.integrator.type != "pt") throw std::invalid_argument("adaptive sampling needs the pt integrator");
        if (!(threshold >= 0.f)) throw std::invalid_argument("threshold must be >= 0");
        vs::AdaptiveHost &a = s->host.integrator.adaptive;
        a.on = on != 0;
        a.threshold = threshold;
        a.min_samples = min_samples;
        fill_render_export(s);
    });
}

vs_status vs_scene_variance_stats(vs_scene *s, float *stats) {
    if (!s || !stats) return fail(s ? s->ctx : nullptr, VS_ERROR_INVALID_ARGUMENT, "null argument");
    return guarded(s->ctx, [&] {
        if (s->vstats_pixels == 0) throw std::logic_error("no adaptive frame rendered");
        hip_check(hipSetDevice(s->ctx->device), "hipSetDevice");
        hip_check(hipStreamSynchronize(s->ctx->stream), "stats sync");
        hip_check(hipMemcpy(stats, s->vstats.ptr, s->vstats.bytes, hipMemcpyDeviceToHost), "stats D2H");
    });
}

vs_status vs_scene_set_denoiser(vs_scene *s, int32_t on, int32_t reproject, int32_t filter_moments) {
    if (!s) return VS_ERROR_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] {
        vs::SvgfHost &d = s->host.integrator.svgf;
        if (on && !d.svgf) throw std::invalid_argument("the scene's denoiser is not svgf");
        d.on = on != 0;
        d.reproject = reproject != 0;
        d.filter_moments = filter_moments != 0;
        fill_restir_export(s);
    });
}

vs_status vs_scene_svgf_state(vs_scene *s, float *gbuffer, float *albedo, float *emission, float *records,
                              float *motion) {
    if (!s) return VS_ERROR_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] {
        require_built(s);
        if (s->sv_pixels == 0) throw std::logic_error("no denoised frame rendered");
        hip_check(hipSetDevice(s->ctx->device), "hipSetDevice");
        hip_check(hipStreamSynchronize(s->ctx->stream), "svgf sync");
        const size_t P = s->sv_pixels;
        if (gbuffer) hip_check(hipMemcpy(gbuffer, s->sv_geo.ptr, s->sv_geo.bytes, hipMemcpyDeviceToHost), "svgf D2H");
        if (albedo) hip_check(hipMemcpy(albedo, s->sv_alb.ptr, s->sv_alb.bytes, hipMemcpyDeviceToHost), "svgf D2H");
        if (emission) hip_check(hipMemcpy(emission, s->sv_emi.ptr, s->sv_emi.bytes, hipMemcpyDeviceToHost), "svgf D2H");
        if (records) hip_check(hipMemcpy(records, s->sv_rec.ptr, s->sv_rec.bytes, hipMemcpyDeviceToHost), "svgf D2H");
        if (motion)
            hip_check(hipMemcpy(motion, s->gi_aux.as<float4>() + 2u * P, P * 16u, hipMemcpyDeviceToHost), "svgf D2H");
    });
}

vs_status vs_render(vs_scene *s, uint32_t spp, int32_t fused) {
    if (!s) return VS_ERROR_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] {
        require_built(s);
        if (spp == 0) return;
        hip_check(hipSetDevice(s->ctx->device), "hipSetDevice");
        hipStream_t st = s->ctx->stream;
        const vs::SceneHost &h = s->host;
        vs::RenderView rv = render_view(s);
        unsigned long long *ctr = nullptr;
        if (s->counting) {
            ctr = s->counters.as<unsigned long long>();
            hip_check(hipMemsetAsync(ctr, 0, 64, st), "counter clear");
        }
        uint32_t launches = 0;
        vs_render_stats ts{};
        const bool restir = h.integrator.type == "rt";
        // adaptive sampling runs in the megakernel form (its statistics live
        // with the pixel's running mean)
        const bool adaptive = !restir && h.integrator.adaptive.on;
        const bool wavefront = !restir && !adaptive && s->kernel_mode == VS_KERNEL_WAVEFRONT;
        const bool staged = restir || wavefront;// two counter sets, [0..2] and [4..6]
        if (!restir) {
            // ConvergenceInspector::reset at frame 0 (pt.cpp:93-95)
            const uint32_t P = s->cam.width * s->cam.height;
            if (adaptive && s->vstats_pixels != P) {
                s->vstats.alloc(size_t(P) * 16u);
                s->vstats_pixels = P;
                hip_check(hipMemsetAsync(s->vstats.ptr, 0, s->vstats.bytes, st), "stats clear");
            }
            if (s->frame_index == 0 && s->vstats_pixels == P)
                hip_check(hipMemsetAsync(s->vstats.ptr, 0, s->vstats.bytes, st), "stats clear");
            if (adaptive) {
                rv.vstats = s->vstats.as<float4>();
                rv.adapt_threshold = h.integrator.adaptive.threshold;
                rv.adapt_min_samples = h.integrator.adaptive.min_samples;
            }
        }
        hip_check(hipEventRecord(s->ctx->ev0, st), "event");
        if (restir) {
            launches = render_restir(s, rv, spp, ctr, ts);
        } else if (wavefront) {
            launches = render_wavefront(s, rv, spp, ctr, ts);
        } else if (fused) {
            rv.frame_start = s->frame_index;
            rv.spp = spp;
            hip_check(vs::launch_pt(s->view, s->cam, rv, s->film_rt.as<float4>(), s->film_out.as<float4>(), ctr,
                                      s->diffuse_only, st),
                      "pt launch");
            ++launches;
        } else {
            for (uint32_t f = 0; f < spp; ++f) {
                rv.frame_start = s->frame_index + f;
                rv.spp = 1;
                hip_check(vs::launch_pt(s->view, s->cam, rv, s->film_rt.as<float4>(), s->film_out.as<float4>(), ctr,
                                      s->diffuse_only, st),
                          "pt launch");
                ++launches;
            }
        }
        hip_check(hipEventRecord(s->ctx->ev1, st), "event");
        hip_check(hipEventSynchronize(s->ctx->ev1), "render sync");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, s->ctx->ev0, s->ctx->ev1), "event time");
        s->frame_index += spp;
        s->stats = ts;
        s->stats.kernel_ms = ms;
        s->stats.launches = launches;
        s->stats.samples_per_pixel = spp;
        if (ctr) {
            // [0..2] rays / node / triangle fetches; staged renders keep their
            // second kernel's counts apart in [4..6]
            unsigned long long c[8];
            hip_check(hipMemcpy(c, ctr, sizeof(c), hipMemcpyDeviceToHost), "counters D2H");
            s->stats.rays = c[0] + (staged ? c[4] : 0ull);
            s->stats.node_fetches = c[1] + (staged ? c[5] : 0ull);
            s->stats.tri_fetches = c[2] + (staged ? c[6] : 0ull);
            if (staged) {
                for (int k = 0; k < 2; ++k) {// the first and the last stage trace rays
                    s->stats.stage_rays[2 * k] = c[4 * k];
                    s->stats.stage_node_fetches[2 * k] = c[4 * k + 1];
                    s->stats.stage_tri_fetches[2 * k] = c[4 * k + 2];
                }
            }
        }
    });
}

vs_status vs_set_counting(vs_scene *s, int32_t enable) {
    if (!s) return VS_ERROR_INVALID_ARGUMENT;
    s->counting = enable != 0;
    return VS_OK;
}

vs_status vs_last_render_stats(const vs_scene *s, vs_render_stats *out) {
    if (!s || !out) return VS_ERROR_INVALID_ARGUMENT;
    *out = s->stats;
    return VS_OK;
}

vs_status vs_reset_film(vs_scene *s) {
    if (!s) return VS_ERROR_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] {
        require_built(s);
        hip_check(hipSetDevice(s->ctx->device), "hipSetDevice");
        alloc_film(s);
        hip_check(hipStreamSynchronize(s->ctx->stream), "film sync");
    });
}

vs_status vs_film_download(vs_scene *s, int32_t which, float *rgba) {
    if (!s || !rgba || (which != 0 && which != 1)) return fail(s ? s->ctx : nullptr, VS_ERROR_INVALID_ARGUMENT, "bad argument");
    return guarded(s->ctx, [&] {
        require_built(s);
        hip_check(hipSetDevice(s->ctx->device), "hipSetDevice");
        size_t bytes = size_t(s->host.camera.width) * s->host.camera.height * 16;
        const DevBuf &b = which == 0 ? s->film_rt : s->film_out;
        hip_check(hipMemcpyAsync(rgba, b.ptr, bytes, hipMemcpyDeviceToHost, s->ctx->stream), "film D2H");
        hip_check(hipStreamSynchronize(s->ctx->stream), "film sync");
    });
}

}// extern "C"

namespace {

// Pipeline::final_picture (src/base/mgr/pipeline.cpp:313-330): the output
// buffer tone-mapped again, sRGB-encoded when gamma, downloaded to rgba
void final_picture(vs_scene *s, bool gamma, float *rgba) {
    require_built(s);
    hip_check(hipSetDevice(s->ctx->device), "hipSetDevice");
    uint32_t n = s->host.camera.width * s->host.camera.height;
    if (s->film_pic.bytes < size_t(n) * 16) s->film_pic.alloc(size_t(n) * 16);
    uint32_t tm = s->host.camera.tone_mapper == "aces" ? 1u : (s->host.camera.tone_mapper == "reinhard" ? 2u : 0u);
    hip_check(vs::launch_final_picture(s->film_out.as<float4>(), s->film_pic.as<float4>(), n, tm, gamma,
                                       s->ctx->stream),
              "final picture");
    hip_check(hipMemcpyAsync(rgba, s->film_pic.ptr, size_t(n) * 16, hipMemcpyDeviceToHost, s->ctx->stream),
              "picture D2H");
    hip_check(hipStreamSynchronize(s->ctx->stream), "picture sync");
}

}// namespace

extern "C" {

vs_status vs_film_final_picture(vs_scene *s, int32_t gamma, float *rgba) {
    if (!s || !rgba) return fail(s ? s->ctx : nullptr, VS_ERROR_INVALID_ARGUMENT, "bad argument");
    return guarded(s->ctx, [&] { final_picture(s, gamma != 0, rgba); });
}

vs_status vs_film_save(vs_scene *s, const char *path) {
    if (!s) return VS_ERROR_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] {
        // Pipeline::save_result (pipeline.cpp:174-182): output.fn under the scene's directory
        std::string fn = (path && *path) ? std::string(path) : vs::join_path(s->host.base_dir, s->host.output_fn);
        std::string e = vs::image_extension(fn);
        if (e != "exr" && e != "hdr" && e != "png" && e != "pfm")
            throw std::runtime_error("unsupported image format ." + e + " (exr, hdr, png, pfm): " + fn);
        bool gamma = !(e == "exr" || e == "hdr" || e == "pfm");
        std::vector<float> pic(size_t(s->host.camera.width) * s->host.camera.height * 4);
        final_picture(s, gamma, pic.data());
        vs::write_image(fn, s->host.camera.width, s->host.camera.height, pic.data());
    });
}

vs_status vs_image_write(vs_context *ctx, const char *path, uint32_t width, uint32_t height, const float *rgba) {
    if (!path || !rgba || width == 0 || height == 0) return fail(ctx, VS_ERROR_INVALID_ARGUMENT, "bad argument");
    return guarded(ctx, [&] { vs::write_image(path, width, height, rgba); });
}

vs_status vs_image_read(vs_context *ctx, const char *path, uint32_t *width, uint32_t *height, float *rgba) {
    if (!path || !width || !height) return fail(ctx, VS_ERROR_INVALID_ARGUMENT, "bad argument");
    return guarded(ctx, [&] {
        vs::ImageRGBA img = vs::read_image(path);
        *width = img.width;
        *height = img.height;
        if (rgba) std::memcpy(rgba, img.rgba.data(), img.rgba.size() * sizeof(float));
    });
}

}// extern "C"
