// Internal state behind the C ABI, shared by the translation units that
// implement it (capi.cpp, serve.cpp).  Not installed.
#pragma once

#include "../../include/vision_amd.h"

#include <hip/hip_runtime.h>

#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "bvh.h"
#include "dev_scene.h"
#include "json.h"
#include "scene.h"
#include "tables.h"

struct vs_context {
    int device{0};// -1: host-only context (scene parsing, BVH build, export)
    hipStream_t stream{nullptr};
    hipEvent_t ev0{nullptr}, ev1{nullptr};
    hipEvent_t tr0{nullptr}, tr1{nullptr}, tr2{nullptr};// around wavefront traversal kernels
    std::string last_error;
};

namespace vs_internal {

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

struct DevBuf {
    void *ptr{nullptr};
    size_t bytes{0};
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    void alloc(size_t n) {
        release();
        if (n == 0) n = 16;
        hipError_t e = hipMalloc(&ptr, n);
        if (e != hipSuccess) {
            ptr = nullptr;
            throw DeviceError(std::string("hipMalloc failed: ") + hipGetErrorString(e));
        }
        bytes = n;
    }
    template <typename T>
    void upload(const std::vector<T> &v, hipStream_t s) {
        alloc(v.size() * sizeof(T));
        if (!v.empty()) hip_check(hipMemcpyAsync(ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
    }
    template <typename T>
    T *as() const { return static_cast<T *>(ptr); }
    void swap(DevBuf &o) noexcept {
        std::swap(ptr, o.ptr);
        std::swap(bytes, o.bytes);
    }
};

// the device tables that depend on the spectrum mode (sRGB or hero
// wavelengths): materials, lights and their alias tables, spectral curves,
// the environment map
struct ModeTables {
    DevBuf mat, light, alias, func, spd, env_tex, env_cond, env_cond_func, env_marg, env_marg_func, tex, tex_info;
    void swap(ModeTables &o) noexcept {
        tex.swap(o.tex);
        tex_info.swap(o.tex_info);
        mat.swap(o.mat);
        light.swap(o.light);
        alias.swap(o.alias);
        func.swap(o.func);
        spd.swap(o.spd);
        env_tex.swap(o.env_tex);
        env_cond.swap(o.env_cond);
        env_cond_func.swap(o.env_cond_func);
        env_marg.swap(o.env_marg);
        env_marg_func.swap(o.env_marg_func);
    }
};

// per-path state and queues of the wavefront form, sized on first use
struct WavefrontBuffers {
    size_t capacity{0};
    DevBuf o, d, thr, L, meta, lam, lpdf, hit, sh_o, sh_d, sh_c, q0, q1, counts;
    uint32_t *host_counts{nullptr};// pinned
    WavefrontBuffers() = default;
    WavefrontBuffers(const WavefrontBuffers &) = delete;
    WavefrontBuffers &operator=(const WavefrontBuffers &) = delete;
    ~WavefrontBuffers() {
        if (host_counts) (void)hipHostFree(host_counts);
    }
};

}// namespace vs_internal

using vs_internal::DevBuf;

struct vs_scene {
    vs_context *ctx{nullptr};
    vs::SceneHost host;
    vs::Bvh8 bvh;
    bool host_bvh{false};// BVH built on the host
    bool built{false};   // BVH and scene resident in HBM
    double build_ms{0.0};
    vs::PackedScene packed;
    // device buffers
    DevBuf nodes, ltris, P, N, UV, tri, tri_inst, inst, lut;
    vs_internal::ModeTables mt;// materials, lights and environment in the scene's spectrum mode
    // the other spectrum mode, packed and uploaded on first use; kept until
    // the scene is destroyed, so views that point at it stay valid
    vs_internal::ModeTables mt_alt;
    vs::PackedScene packed_alt;
    bool have_alt{false};
    bool diffuse_only{false};// only diffuse / Oren-Nayar / black materials
    DevBuf fis_lut, fis_cond, fis_cond_func, fis_marg, fis_marg_func;
    DevBuf film_rt, film_out, film_pic, counters;
    DevBuf scratch_rays, scratch_hits;
    // wavefront form: per-path state, queues, append counters
    uint32_t kernel_mode{VS_KERNEL_MEGAKERNEL};
    uint32_t wf_chunk_frames{8};
    uint32_t wf_trace_block{128};
    vs_internal::WavefrontBuffers wf;
    // ReSTIR DI ("rt" integrator): reservoirs, surfaces, surface extensions,
    // and the camera of the previous frame (3x4 rows of camera-to-world)
    DevBuf rs_rsv, rs_surf, rs_ext;
    uint32_t rs_pixels{0};
    // ReSTIR GI (the "indirect" block open): reservoirs, shader0's record,
    // direct and indirect radiance
    DevBuf gi_rsv, gi_aux, gi_direct, gi_rad;
    uint32_t gi_pixels{0};
    // shader0's per-pixel record (HitBSDF, motion vector) when GI or SVGF reads it
    uint32_t aux_pixels{0};
    // G-buffer and SVGF records (svgf.h), allocated on the first denoised frame
    DevBuf sv_geo, sv_alb, sv_emi, sv_rec;
    uint32_t sv_pixels{0};
    // ConvergenceInspector's VarianceStats (avg, var, N, 0) per pixel
    DevBuf vstats;
    uint32_t vstats_pixels{0};
    float prev_c2w[12]{};
    vs::SceneView view{};
    vs::CameraView cam{};
    uint32_t frame_index{0};
    bool counting{false};
    vs_render_stats stats{};
    // exported flat arrays
    std::vector<uint32_t> exp_inst;
    std::vector<float> exp_camera, exp_render, exp_restir;
};

namespace vs_internal {

inline vs_status fail(vs_context *ctx, vs_status code, const std::string &msg) {
    if (ctx) ctx->last_error = msg;
    return code;
}

template <typename F>
vs_status guarded(vs_context *ctx, F &&f) {
    try {
        f();
        if (ctx) ctx->last_error.clear();
        return VS_OK;
    } catch (const vs::JsonError &e) {
        return fail(ctx, VS_ERROR_PARSE, e.what());
    } catch (const DeviceError &e) {
        return fail(ctx, VS_ERROR_DEVICE, e.what());
    } catch (const std::bad_alloc &) {
        return fail(ctx, VS_ERROR_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::invalid_argument &e) {
        return fail(ctx, VS_ERROR_INVALID_ARGUMENT, e.what());
    } catch (const std::logic_error &e) {
        return fail(ctx, VS_ERROR_STATE, e.what());
    } catch (const std::exception &e) {
        std::string m = e.what();
        vs_status code = (m.find("unsupported") != std::string::npos || m.find("not supported") != std::string::npos)
                             ? VS_ERROR_UNSUPPORTED
                             : VS_ERROR_INVALID_ARGUMENT;
        return fail(ctx, code, m);
    }
}


}// namespace vs_internal

namespace vs_internal {

// screen window and raster scale of a resolution (camera fields only)
void camera_set_resolution(vs::CameraView &v, uint32_t width, uint32_t height);
// upload the mode-dependent tables of a packed scene / point a view at them
void upload_mode_tables(const vs::PackedScene &pk, ModeTables &t, hipStream_t st);
void set_mode_view(vs::SceneView &v, const vs::PackedScene &pk, const ModeTables &t);
// the render settings of a scene's integrator and film
vs::RenderView render_view(const vs_scene *s);
// pack and upload the tables of the spectrum mode the scene is not in
void ensure_alt_mode(vs_scene *s);
// a view of the scene in sRGB (rgb) or hero-wavelength mode
vs::SceneView view_in_mode(vs_scene *s, bool rgb);

// device work needs a GPU context and a built scene
// resolves this device's address of the staging array behind
// vs_trace_closest_noalloc (noalloc.hip); `device` must be current
void noalloc_setup(int device);

inline void require_built(vs_scene *s) {
    if (s->ctx->device < 0) throw std::logic_error("host-only context: no device work possible");
    if (!s->built) throw std::logic_error("accel not built: call vs_scene_build_accel first");
}

// The wavefront form over one film: `spp` frames from frame_index in
// chunks of `chunk` frames, each a loop of trace / shade / shadow over the
// live-path queue until it empties, then the chunk folds into the film.
// Synchronises `st` once per bounce (the live count).  Returns launches;
// adds per-stage times to *ts when it and the three events are given.
uint32_t wavefront_render(const vs::SceneView &view, const vs::CameraView &cam, const vs::RenderView &rv0,
                          bool rgb, bool diffuse_only, float4 *rt, float4 *out, uint32_t frame_index, uint32_t spp, uint32_t chunk,
                          uint32_t trace_block, WavefrontBuffers &wb, hipStream_t st, unsigned long long *ctr,
                          vs_render_stats *ts, hipEvent_t *ev4);

}// namespace vs_internal
