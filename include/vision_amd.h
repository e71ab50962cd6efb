/*
 * vision_amd — MI355X-native megakernel path tracer behind Vision's scene
 * format.  Plain C ABI: opaque handles, caller-owned arrays, integer status
 * codes and a per-context error string; nothing throws across this boundary.
 *
 * The reference's only C ABI is the SDK vtable returned by create()/destroy()
 * (src/sdk/vision.h:66-99, src/sdk/vision.cpp:222-227).  Each entry point below
 * names the reference interface it stands in for.
 */
#ifndef VISION_AMD_H
#define VISION_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t vs_status;
#define VS_OK 0
#define VS_ERROR_INVALID_ARGUMENT 1
#define VS_ERROR_PARSE 2
#define VS_ERROR_UNSUPPORTED 3
#define VS_ERROR_DEVICE 4
#define VS_ERROR_OUT_OF_MEMORY 5
#define VS_ERROR_STATE 6

typedef struct vs_context vs_context;
typedef struct vs_scene vs_scene;

/* 32-byte ray: world origin, t_min, direction (need not be unit), t_max.
 * Valid hits satisfy t_min < t < t_max.  Mirrors ocarina's Ray as used by
 * Geometry::trace_closest (src/base/mgr/geometry.cpp:167-173). */
typedef struct vs_ray {
    float org[3];
    float t_min;
    float dir[3];
    float t_max;
} vs_ray;

/* 24-byte hit: TriangleHit{inst_id, prim_id, bary} of the reference
 * (used at src/base/mgr/geometry.cpp:78-96) plus t and the scene-global
 * triangle id.  inst == 0xFFFFFFFF marks a miss.  bary weights (v1, v2):
 * p = (1 - b0 - b1) * v0 + b0 * v1 + b1 * v2. */
typedef struct vs_hit {
    uint32_t inst;
    uint32_t prim;
    float bary[2];
    float t;
    uint32_t tri;
} vs_hit;

typedef struct vs_scene_info {
    uint32_t num_triangles;
    uint32_t num_vertices;
    uint32_t num_instances;
    uint32_t num_materials;
    uint32_t num_lights;
    uint32_t width;
    uint32_t height;
    uint32_t output_spp;
    uint32_t max_depth;
    uint32_t spectrum_dim;
    uint32_t bvh_nodes;
    uint32_t bvh_depth;
    uint32_t frame_index;
    uint32_t accel_built;
    double bvh_build_ms;
    double bvh_sah_cost;
} vs_scene_info;

typedef struct vs_render_stats {
    double kernel_ms;      /* HIP-event time of the last vs_render call's kernels */
    uint64_t rays;         /* rays cast (only when counting was enabled) */
    uint64_t node_fetches; /* BVH8 node fetches (counting runs) */
    uint64_t tri_fetches;  /* leaf triangle fetches (counting runs) */
    uint32_t launches;     /* kernel launches of the last call */
    uint32_t samples_per_pixel;
    /* three kernels timed on their own (HIP events), with their launch
     * counts and (counting runs) their rays / node / triangle fetches: in
     * the wavefront form [0] closest-hit traversal, [1] shading, [2] any-hit
     * traversal; with the "rt" integrator [0] shader0 (candidates,
     * visibility, temporal reuse), [1] the extra spatial passes, [2] shader1
     * (spatial reuse, shading) and, with ReSTIR GI open, the two GI kernels
     * and the combine; zeros otherwise */
    double stage_ms[3];
    uint32_t stage_launches[3];
    uint64_t stage_rays[3], stage_node_fetches[3], stage_tri_fetches[3];
} vs_render_stats;

/* Flattened world-space scene handed to tests and the CPU oracle.
 * All pointers are owned by the scene and stay valid until vs_scene_destroy. */
typedef struct vs_scene_arrays {
    const float *positions;     /* [num_vertices][3] */
    const float *normals;       /* [num_vertices][3], unnormalized, may be 0 */
    const float *uvs;           /* [num_vertices][2] */
    const uint32_t *triangles;  /* [num_triangles][3] */
    const uint32_t *tri_instance;/* [num_triangles] */
    const uint32_t *instances;  /* [num_instances][4]: material, light, tri_base, 0 */
    const uint32_t *materials;  /* [num_materials][32]: type, flags, 30 float params */
    const uint32_t *lights;     /* [num_lights][28]: light records (112 B) */
    const float *camera;        /* [32]: c2w rows [R|t] (12), screen lower_x, upper_y,
                                   raster step x, y, tan(fov_y/2), lens_radius,
                                   focal_distance, width, height, filter (0 box,
                                   1 triangle, 2 fitted curve), filter_radius,
                                   thin_lens flag, filter kind (0 box, 1 triangle,
                                   2 gaussian, 3 mitchell, 4 sinc), gaussian sigma,
                                   mitchell b, c, sinc tau, 3 unused */
    const float *render;        /* [12]: max_depth, min_depth, rr_threshold, mis_mode,
                                   light_sampler(0 uniform,1 power), ray_offset_factor,
                                   accumulation, exposure, tone mapper (0 linear,
                                   1 aces, 2 reinhard), adaptive sampling on,
                                   its threshold, its minimum sample number */
    uint32_t num_vertices, num_triangles, num_instances, num_materials, num_lights;
    const float *spd;           /* [num_spd]: CIE X, Y, Z, D65 (4 x 94) then material curves */
    uint32_t num_spd;
    uint32_t spectrum_rgb;      /* 1: sRGB mode, 0: hero wavelengths */
    uint32_t spectrum_dim;
    uint32_t env_light;         /* index of the environment light, 0xFFFFFFFF if none */
    uint32_t env_w, env_h;
    const float *env_rgb;       /* [env_h][env_w][3] linear RGB, top row first */
    float env_w2o[9], env_o2w[9];/* row-major 3x3 */
    float env_scale;
    float world_radius;
    /* ReSTIR DI settings of an "rt" integrator ([0] = 1) or zeros: M_light,
     * M_bsdf, max_age, max_recursion, diffuse_factor, pairwise, spatial open,
     * sample_num, cos(spatial theta), spatial depth, spatial radius, temporal
     * open, history_limit, N, temporal radius, cos(temporal theta), temporal
     * depth, temporal mis, extra spatial passes ([1..19]); world-to-camera
     * 3x4 rows of the current camera ([20..31]) */
    const float *restir;
    uint32_t num_restir;
    /* images of textured colour slots (a material record's flag bit 16 + k
     * marks slot p[4k..4k+3] as textured, p[4k] holding the texture index):
     * float4 texels in the scene's spectrum mode (linear sRGB, or albedo
     * sigmoid coefficients) and per texture {first texel, width, height, 0} */
    const float *textures;
    uint32_t num_texels;
    const uint32_t *texture_info;
    uint32_t num_textures;
} vs_scene_arrays;

/* library version (major*10000 + minor*100 + patch) */
int32_t vs_version(void);

/* RGB -> sigmoid-polynomial spectrum coefficients of hero-wavelength mode
 * (RGBToSpectrumTable::decode_albedo / decode_unbound,
 * src/render_core/spectra/hero.cpp:91-139).  unbound = 0: albedo (rgb clamped
 * to [0, 1]; out[3] = luminance); unbound = 1: illumination (out[3] = scale). */
vs_status vs_rgb_to_spectrum(const float *rgb, uint32_t n, int32_t unbound, float *coeffs);

/* Device context on HIP device `device` with its own stream; device = -1
 * makes a host-only context that can parse, build the BVH on the host and
 * export, but not render.  Stands in for the SDK's create()
 * (src/sdk/vision.cpp:222-224) and device selection (src/core/cli_parser.cpp:21-22). */
vs_status vs_create(int32_t device, vs_context **out);
void vs_destroy(vs_context *ctx);
/* last error message of the context ("" if none) */
const char *vs_last_error(const vs_context *ctx);

/* Parse a Vision scene (JSON text, comments allowed) and resolve it to a
 * flattened world-space scene.  base_dir resolves relative mesh paths.
 * Replaces SceneDesc::from_json + Scene::init
 * (src/core/scene_desc.cpp:37-61, src/base/mgr/scene.cpp:16-34). */
vs_status vs_scene_load(vs_context *ctx, const char *json_text, const char *base_dir, vs_scene **out);
vs_status vs_scene_load_file(vs_context *ctx, const char *path, vs_scene **out);
void vs_scene_destroy(vs_scene *scene);

/* Build the BVH8 (on the host) and, on a device context, upload geometry,
 * materials, lights and film to HBM.
 * Replaces Geometry::build_accel + Pipeline upload
 * (src/base/mgr/geometry.cpp:36-70) and the SDK's build_accel. */
vs_status vs_scene_build_accel(vs_scene *scene);
vs_status vs_scene_info_get(const vs_scene *scene, vs_scene_info *out);
vs_status vs_scene_export(const vs_scene *scene, vs_scene_arrays *out);
/* The host copy of the BVH8: 80-byte nodes (20 uint32 each) and 48-byte leaf
 * triangles (12 floats: v0, id bits, v1, 0, v2, 0), valid until destroy. */
vs_status vs_scene_bvh_export(const vs_scene *scene, const uint32_t **nodes, uint32_t *num_nodes,
                              const float **leaf_tris, uint32_t *num_leaf_tris);

/* Override film resolution (SDK update_resolution, src/sdk/vision.h:78). */
vs_status vs_scene_set_resolution(vs_scene *scene, uint32_t width, uint32_t height);
/* Move the camera: camera-to-world as 16 floats, four columns of four, and
 * the vertical field of view in degrees.  Replaces the SDK's update_camera
 * (src/sdk/vision.h:77, vision.cpp:204-210, which ignores fov_y and sets 45);
 * the camera of the previous rendered frame stays the reference for ReSTIR's
 * reprojection (Sensor::store_prev_data, src/base/sensor/sensor.cpp:95-103). */
vs_status vs_scene_update_camera(vs_scene *scene, const float *c2w, float fov_y);
/* Override max_depth / min_depth / rr_threshold of the integrator. */
vs_status vs_scene_set_integrator(vs_scene *scene, uint32_t max_depth, uint32_t min_depth, float rr_threshold);
/* Switch the spectrum mode of a built scene, as its JSON "spectrum" block
 * would (SRGBSpectrum / HeroWavelengthSpectrum, src/render_core/spectra/srgb.cpp,
 * hero.cpp:227-348): spectral = 0 sRGB, 1 hero wavelengths with dimension 3 or 4.
 * The other mode's material, light and environment tables are packed on first
 * use and kept, so switching back and forth costs nothing after that. */
vs_status vs_scene_set_spectrum(vs_scene *scene, int32_t spectral, uint32_t dimension);

/* Closest-hit / any-hit for a batch of rays in host memory (the parity hook
 * for Geometry::trace_closest / trace_occlusion, geometry.cpp:167-184). */
vs_status vs_trace_closest(vs_scene *scene, const vs_ray *rays, uint32_t n, vs_hit *hits);
vs_status vs_trace_any(vs_scene *scene, const vs_ray *rays, uint32_t n, uint8_t *occluded);
/* Same on device-resident buffers, enqueued on the context stream (or on
 * `stream` when non-null, a hipStream_t); returns without synchronizing. */
vs_status vs_trace_closest_device(vs_scene *scene, const vs_ray *d_rays, uint32_t n, vs_hit *d_hits, void *stream);
vs_status vs_trace_any_device(vs_scene *scene, const vs_ray *d_rays, uint32_t n, uint8_t *d_occluded, void *stream);
/* vs_trace_closest without device allocations.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * around hipGraph capture, serving loops, harnesses that forbid hipMalloc).
 * Rays and hits stay in host memory; they are staged through a 14 MiB array
 * that the library holds on each GPU until the process ends, so at most
 * 262144 rays per call (more is VS_ERROR_INVALID_ARGUMENT).  Results equal
 * vs_trace_closest's.  Calls that overlap on one GPU (several threads, or
 * several contexts of the device) race on that array: serialise them. */
vs_status vs_trace_closest_noalloc(vs_scene *scene, const vs_ray *rays, uint32_t n, vs_hit *hits);

/* Material-level parity hook: MaterialEvaluator::evaluate (sample = 0) or
 * ::sample (sample = 1) of scene material `mat` at a surface whose shading
 * frame is the world axes and whose geometric normal is +z
 * (src/base/scattering/material.cpp:132-184).  Host arrays: wo, wi, wi_out
 * [n][3]; lambda [n][4] (hero mode only: the sampled wavelengths); f [n][4]
 * (f * |cos wi| per wavelength, unused channels 0); pdf [n].  Sampling draws
 * from the sampler stream (pixel (0, i), index `frame`, dim 0). */
vs_status vs_scene_bsdf(vs_scene *scene, uint32_t mat, uint32_t n, const float *wo, const float *wi,
                        const float *lambda, int32_t sample, uint32_t frame, float *wi_out, float *f, float *pdf);

/* Kernel form of vs_render.  MEGAKERNEL: one thread per pixel runs whole
 * paths (fused or per-frame launches).  WAVEFRONT: the bounce loop split into
 * closest-hit / shade / any-hit kernels over queues of live paths held in
 * HBM, `chunk_frames` frames of all pixels per chunk, traversal blocks of
 * `trace_block` (64, 128 or 256) threads.  Both give the same film.  No
 * reference counterpart: the reference runs one kernel per frame
 * (pt.cpp:90-99); this is the choice the autotuner makes per scene. */
#define VS_KERNEL_MEGAKERNEL 0
#define VS_KERNEL_WAVEFRONT 1
vs_status vs_scene_set_kernel(vs_scene *scene, int32_t mode, uint32_t chunk_frames, uint32_t trace_block);

/* Add `spp` progressive frames to the film (frame indices continue from the
 * current one).  fused = 0: one launch per frame, as
 * PathTracingIntegrator::render does (src/render_core/integrators/pt.cpp:90-99);
 * fused = 1: one launch adding all frames with the running mean in registers.
 * Both produce the same film.  With the "rt" integrator each frame runs ReSTIR
 * DI (RealTimeIntegrator::render, src/render_core/integrators/rt.cpp:144-160,
 * direct half): initial candidates + temporal reuse, any extra spatial
 * passes, spatial reuse + shading; with the "indirect" block open (sRGB
 * mode) then ReSTIR GI (ReSTIRGI::dispatch,
 * src/render_core/integrators/ReSTIR/indirect.cpp:354-362: initial samples
 * fused with temporal reuse, spatial reuse + shading) and the combine of
 * both halves into the film (rt.cpp:119-126).  Synchronizes before
 * returning. */
vs_status vs_render(vs_scene *scene, uint32_t spp, int32_t fused);
/* ReSTIR DI state after the last rendered frame (the "rt" integrator): the
 * reservoir buffers [4][3][pixels][4] (previous / current by frame parity,
 * pass-through, second pass-through; fields weight_sum C W p_hat | light
 * prim bary.xy | pos.xyz age) and the surface buffers [2][3][pixels][4]
 * (triangle b1 b2 flag|replaced<<16 | normal depth | position
 * diffuse_factor).  Integer fields are stored as their bit patterns.
 * Test and inspection hook; no reference counterpart. */
vs_status vs_scene_restir_state(vs_scene *scene, float *reservoirs, float *surfaces);
/* ReSTIR GI state after the last rendered frame: the GI reservoir buffers
 * [3][4][pixels][4] (previous / current by frame parity, pass-through;
 * fields weight_sum C W age | sample position | sample normal | Lo, the
 * GIReservoir of src/render_core/integrators/ReSTIR/reservoir.h:157-222)
 * and the indirect radiance [pixels][4].  Test and inspection hook; no
 * reference counterpart. */
vs_status vs_scene_restir_gi_state(vs_scene *scene, float *reservoirs, float *radiance);
/* Adaptive sampling of the "pt" integrator (ConvergenceInspector,
 * src/render_core/integrators/adaptive/inspector.cpp:9-38, used by
 * PathTracingIntegrator, src/render_core/integrators/pt.cpp:43-64): the GUI's
 * switch, threshold and minimum sample number (the reference's file sets only
 * the latter two, as the integrator's "adaptive" block {threshold,
 * start_index}, which the loader reads).  When on, a pixel whose statistics
 * (the running mean, variance and count of the output buffer's x + y + z)
 * show more than min_samples samples and variance / mean below the threshold
 * adds no more samples; the statistics reset at frame 0.  Renders with it
 * on use the megakernel form. */
vs_status vs_scene_set_adaptive(vs_scene *scene, int32_t on, float threshold, uint32_t min_samples);
/* The per-pixel VarianceStats (avg, var, N as uint32 bits, 0), float[W*H*4].
 * Test and inspection hook; no reference counterpart. */
vs_status vs_scene_variance_stats(vs_scene *scene, float *stats);
/* Switch the SVGF denoiser of the "rt" integrator on or off, with its
 * reprojection and moment-filter passes: the check boxes of SVGF::render_sub_UI
 * (src/render_core/denoisers/SVGF/svgf.cpp:29-47; the reference has no file
 * setting for them, its GUI sets them).  The other SVGF parameters come from
 * the integrator's "denoiser" block (SVGF::SVGF, SVGF/svgf.h:44-53).  When
 * on, each rt frame first computes the G-buffer (FrameBuffer::compute_GBuffer,
 * src/base/frame_buffer.cpp:221-229) and after the combine runs
 * SVGF::dispatch (svgf.cpp:89-106), writing the modulated picture to the
 * output buffer (film which = 1). */
vs_status vs_scene_set_denoiser(vs_scene *scene, int32_t on, int32_t reproject, int32_t filter_moments);
/* G-buffer and SVGF state after the last denoised frame; any pointer may be
 * null: gbuffer [2][2][pixels][4] (slots by frame parity, current =
 * (frame + 1) & 1; normal.xyz fwidth | p_film.xy depth_gradient
 * linear_depth, PixelGeometry of frame_buffer.h:15-20), albedo and emission
 * [pixels][4], records [3][2][pixels][4] (SVGFData of SVGF/utils.h:22-26:
 * illumination.rgb variance | history moments.xy 0; buffers 0 / 1 by frame
 * parity, 2 the ping-pong spare) and the motion vectors [pixels][4] (.xy).
 * Test and inspection hook; no reference counterpart. */
vs_status vs_scene_svgf_state(vs_scene *scene, float *gbuffer, float *albedo, float *emission, float *records,
                              float *motion);
/* Enable per-ray counters (rays, node and triangle fetches) for later renders. */
vs_status vs_set_counting(vs_scene *scene, int32_t enable);
vs_status vs_last_render_stats(const vs_scene *scene, vs_render_stats *out);
/* Reset frame index and film (Integrator::invalidation, src/base/integrator.cpp:13-19). */
vs_status vs_reset_film(vs_scene *scene);
/* Download the film: which = 0 running-mean radiance buffer (rt_buffer),
 * 1 exposure/tone-mapped output buffer (normal.cpp:87-101); rgba float[W*H*4]. */
vs_status vs_film_download(vs_scene *scene, int32_t which, float *rgba);
/* The picture a save writes (Pipeline::final_picture, src/base/mgr/pipeline.cpp:313-330,
 * with Postprocessor::tone_mapping, src/base/mgr/postprocessor.cpp:13-30): the output
 * buffer passed through the collector's tone mapper again, then linear_to_srgb when
 * gamma != 0, w = 1; rgba float[W*H*4] on the host. */
vs_status vs_film_final_picture(vs_scene *scene, int32_t gamma, float *rgba);
/* Write the final picture to path (Pipeline::save_result, pipeline.cpp:174-182); NULL or
 * "" writes the scene's output.fn (default "output.png") under its base directory.
 * gamma is off for .exr, .hdr and .pfm and on for .png, as final_picture decides by the
 * file name.  Formats: .exr (float RGBA, uncompressed), .hdr (RGBE), .png (8-bit RGBA),
 * .pfm; others are VS_ERROR_UNSUPPORTED. */
vs_status vs_film_save(vs_scene *scene, const char *path);
/* Host-only image I/O (ocarina's Image::save_image / Image::load, used by
 * pipeline.cpp:176 and image_pool.cpp:28).  Rows top to bottom, float RGBA.
 * vs_image_write: .exr .hdr .png .pfm as above.  vs_image_read: .exr (NONE/ZIPS/ZIP,
 * HALF/FLOAT/UINT), .hdr (flat or RLE), .png (8/16-bit, any colour type, no interlace),
 * .pfm; call with rgba = NULL to get the size, then with a W*H*4 buffer.  Values of
 * 8/16-bit files are code / max with no transfer function.  ctx may be a CPU context
 * (device -1) and only receives the error message. */
vs_status vs_image_write(vs_context *ctx, const char *path, uint32_t width, uint32_t height, const float *rgba);
vs_status vs_image_read(vs_context *ctx, const char *path, uint32_t *width, uint32_t *height, float *rgba);

/* ---- Batched serving -----------------------------------------------------
 *
 * No reference counterpart: the reference renders one scene per process, one
 * frame per launch with a host sync per frame (Pipeline::display,
 * src/base/mgr/pipeline.cpp:296-311).  A server keeps any number of built
 * scenes resident and renders path-tracing jobs on them: a worker thread
 * admits queued jobs while job slots and its HBM budget allow (continuous
 * batching: jobs that arrive while others run join at the next step), and
 * each step runs every admitted job forward by some frames inside persistent
 * kernels whose waves claim 64-pixel tiles of any job from a device counter.
 * A step is a captured hipGraph replayed with a new step table.  A job's film
 * is bit-identical to vs_render(scene at the job's resolution, spp) of the
 * same scene, alone or in any mix.  Submit, wait, result and stats calls are
 * thread-safe. */
typedef struct vs_server vs_server;

typedef struct vs_server_config {
    uint64_t hbm_budget;   /* device bytes for job films (32 B per pixel); 0: a quarter of free memory */
    uint32_t max_jobs;     /* jobs admitted at once, 1..256; 0: 64 */
    float step_ms;         /* target duration of a step; 0: 16 */
    uint32_t tile_w;       /* tile width in pixels (tiles hold 64 pixels): 8, 16, 32 or 64; 0: 8 */
    uint32_t waves_per_cu; /* persistent waves per CU for each kernel variant; 0: the occupancy limit */
    int32_t use_graphs;    /* 1: replay captured hipGraphs; 0: launch each step's kernels directly */
    int32_t count_rays;    /* 1: count rays and BVH fetches (slower) */
    uint32_t min_waves_per_eu; /* register budget of the kernels: 0 the compiler's choice, 6 the measured
                                  budget: at least six resident waves per SIMD for the diffuse-family
                                  kernels, four for the every-material ones (spilling what does not fit) */
    uint64_t lane_min_samples; /* jobs of at least this many samples (width * height * spp) render on the
                                  wavefront lane whatever their route; 0: by route only */
} vs_server_config;

typedef struct vs_job_desc {
    uint32_t scene;        /* index from vs_server_add_scene */
    uint32_t width, height;/* film resolution (the scene's camera, re-aspected) */
    uint32_t spp;          /* frames, 1..65536 */
    uint32_t max_depth;    /* 0: the scene integrator's */
    uint32_t spectral;     /* 0: sRGB mode, 1: hero wavelengths */
    uint32_t spectrum_dim; /* hero wavelengths per path, 3 or 4 (0: 4); ignored in sRGB mode */
    uint32_t reserved;
} vs_job_desc;

#define VS_JOB_QUEUED 0
#define VS_JOB_RUNNING 1
#define VS_JOB_DONE 2
#define VS_JOB_FAILED 3

typedef struct vs_job_info {
    int32_t state;
    uint32_t frames_done;
    uint32_t route;        /* 0 persistent megakernel, 1 wavefront lane */
    uint32_t pad;
    uint64_t samples;      /* width * height * spp */
    double queue_ms;       /* submit -> admitted */
    double latency_ms;     /* submit -> film on the host (done jobs) */
} vs_job_info;

typedef struct vs_server_stats {
    uint64_t jobs_submitted, jobs_done, jobs_failed;
    uint64_t samples_done;
    uint64_t rays, node_fetches, tri_fetches; /* when count_rays */
    uint64_t steps, graph_launches;
    double busy_ms;        /* device time of the steps */
    double wall_ms;        /* first submit -> last completion */
    double msamples_per_s; /* samples_done / wall */
    double latency_p50_ms, latency_p99_ms, latency_mean_ms, latency_max_ms;
    uint32_t max_active;   /* most jobs admitted at once */
    uint32_t active;       /* admitted now */
    uint32_t queued;       /* waiting for admission now */
    uint32_t pad;
    uint64_t hbm_budget, hbm_peak;
    /* the wavefront lanes: samples rendered there, device time and launches
     * of the closest-hit [0], shading [1] and any-hit [2] kernels, and (when
     * count_rays) rays / node / triangle fetches of the closest-hit [0] and
     * any-hit [1] traversals */
    uint64_t lane_samples;
    double lane_stage_ms[3];
    uint64_t lane_stage_launches[3];
    uint64_t lane_rays[2], lane_node_fetches[2], lane_tri_fetches[2];
} vs_server_stats;

/* Create a server on ctx's device (cfg NULL: defaults).  The worker thread
 * starts at once and waits for jobs. */
vs_status vs_server_create(vs_context *ctx, const vs_server_config *cfg, vs_server **out);
/* Make a built scene resident; it must outlive the server.  Both spectrum
 * modes become available (the other mode's material, light and environment
 * tables are packed and uploaded here). */
vs_status vs_server_add_scene(vs_server *srv, vs_scene *scene, uint32_t *index);
/* Queue a job; VS_ERROR_OUT_OF_MEMORY if its film can never fit the budget. */
vs_status vs_server_submit(vs_server *srv, const vs_job_desc *job, uint64_t *job_id);
/* Wait up to timeout_ms (< 0: forever) for a job; *done = 1 when finished (or failed). */
vs_status vs_server_wait(vs_server *srv, uint64_t job_id, double timeout_ms, int32_t *done);
vs_status vs_server_job_info(vs_server *srv, uint64_t job_id, vs_job_info *info);
/* A finished job's film: which = 0 running mean, 1 output buffer; rgba float[W*H*4]. */
vs_status vs_server_result(vs_server *srv, uint64_t job_id, int32_t which, float *rgba);
/* Forget a finished job and free its host copy. */
vs_status vs_server_release(vs_server *srv, uint64_t job_id);
/* Wait until no job is queued or running (timeout_ms < 0: forever). */
vs_status vs_server_drain(vs_server *srv, double timeout_ms);
vs_status vs_server_stats_get(vs_server *srv, vs_server_stats *out);
/* Clear latency samples and counters (between benchmark phases). */
vs_status vs_server_reset_stats(vs_server *srv);
/* Change tile_w / waves_per_cu / step_ms / use_graphs while idle (the
 * budget and slot count are fixed at creation). */
vs_status vs_server_set_config(vs_server *srv, const vs_server_config *cfg);
vs_status vs_server_get_config(vs_server *srv, vs_server_config *cfg);
/* Routing: jobs of a scene in one spectrum mode go to the persistent
 * megakernel steps (route 0, the default) or to the wavefront lane (route 1),
 * a second worker that renders one job at a time in the wavefront form
 * (trace / shade / shadow kernels over path queues, as vs_scene_set_kernel's
 * VS_KERNEL_WAVEFRONT) beside the steps.  Set while idle.  The lane's path
 * state (up to 16 M paths, about 3 GB) lives outside the HBM budget. */
vs_status vs_server_set_route(vs_server *srv, uint32_t scene, int32_t spectral, uint32_t route);
vs_status vs_server_get_route(vs_server *srv, uint32_t scene, int32_t spectral, uint32_t *route);
/* Autotuner: render the probe jobs under each candidate launch shape
 * (register budget 0 / 6 x tile width 8 / 16 x persistent waves per CU at
 * the occupancy limit and half of it) and keep the one with the highest
 * Msamples/s; then, for each (scene, spectrum mode) of the probe list, time
 * its largest probe job alone on each route and route it to the wavefront
 * lane only when that is more than 2x faster (the lane renders one job at a
 * time, so under load a smaller edge alone does not carry over).  The
 * server must be idle; the probe jobs' results are discarded. */
#define VS_AUTOTUNE_MAX_TRIALS 16
typedef struct vs_autotune_result {
    uint32_t tile_w, waves_per_cu, min_waves_per_eu; /* chosen (now the server's config) */
    double msamples_per_s;         /* of the chosen shape */
    uint32_t trials;
    uint32_t trial_tile_w[VS_AUTOTUNE_MAX_TRIALS];
    uint32_t trial_waves_per_cu[VS_AUTOTUNE_MAX_TRIALS];
    uint32_t trial_min_waves_per_eu[VS_AUTOTUNE_MAX_TRIALS];
    double trial_msamples_per_s[VS_AUTOTUNE_MAX_TRIALS];
    uint32_t routes;               /* (scene, mode) pairs timed */
    uint32_t route_scene[VS_AUTOTUNE_MAX_TRIALS], route_spectral[VS_AUTOTUNE_MAX_TRIALS];
    uint32_t route_choice[VS_AUTOTUNE_MAX_TRIALS];
    double route_megakernel_ms[VS_AUTOTUNE_MAX_TRIALS], route_wavefront_ms[VS_AUTOTUNE_MAX_TRIALS];
} vs_autotune_result;
vs_status vs_server_autotune(vs_server *srv, const vs_job_desc *probe, uint32_t n, vs_autotune_result *out);
/* The message of the last failed server call on this thread. */
const char *vs_server_last_error(void);
void vs_server_destroy(vs_server *srv);

/* ---- Radiance cache ---------------------------------------------------------
 *
 * No reference counterpart (the reference's integrators/sharc is a stub).  An
 * online-trained cache of scattered radiance for sRGB path-tracing scenes: a
 * multiresolution hash-grid encoding of position (16 levels x 2 FP16
 * features, 2^19 entries per level) and an MLP 32 -> 64 -> 64 -> 64 -> 3
 * (ReLU, no biases) whose hidden layers run on BF16 MFMA tiles with FP32
 * accumulation.  Paths stop at bounce term_bounce and take the cache's
 * answer there; training paths of every stride-th pixel run to full length
 * and teach the cache the radiance scattered at each of their vertices
 * (relative L2 loss, Adam).  Without VS_NRC_TERMINATE the same call renders
 * the plain path tracer (the fallback), bit for bit as vs_render. */
typedef struct vs_nrc vs_nrc;
typedef struct vs_nrc_config {
    uint32_t term_bounce;  /* paths stop into the cache at this bounce (2) */
    uint32_t train_paths;  /* training paths per frame (65536) */
    uint32_t train_steps;  /* Adam steps per frame (4) */
    uint32_t train_batch;  /* samples per step: consecutive slices of the frame's samples, wrapping; 0: all (65536) */
    float learning_rate;   /* Adam learning rate (1e-2) */
    uint32_t min_res;      /* coarsest level: cells per axis of the scene box (16) */
    uint32_t max_res;      /* finest level (2048) */
    uint32_t seed;         /* initial weights (1) */
} vs_nrc_config;

typedef struct vs_nrc_stats {
    uint32_t frames, adam_steps;
    uint32_t queries_last_frame, train_samples_last_frame, train_paths_per_frame;
    float last_loss;
    /* device time of the last vs_nrc_render call, per stage, ms */
    float render_ms, train_paths_ms, infer_ms, train_ms, accumulate_ms;
} vs_nrc_stats;

#define VS_NRC_TERMINATE 1u /* stop paths into the cache */
#define VS_NRC_TRAIN 2u     /* trace training paths and take Adam steps */

void vs_nrc_default_config(vs_nrc_config *cfg);
/* A cache for a built sRGB scene (cfg NULL: defaults); the scene outlives it. */
vs_status vs_nrc_create(vs_scene *scene, const vs_nrc_config *cfg, vs_nrc **out);
/* Render `frames` frames of `spp` samples into the scene's film (its running
 * mean continues from the scene's frame index, as vs_render's does). */
vs_status vs_nrc_render(vs_nrc *nrc, uint32_t frames, uint32_t spp, uint32_t flags);
vs_status vs_nrc_stats_get(vs_nrc *nrc, vs_nrc_stats *out);
/* Parity and inspection hooks.  Levels: cells per axis of the 16 levels, the
 * mask of directly indexed levels, and the map position -> unit cube
 * (u = (p - lo) * inv_ext, clamped).  Parameters: the FP32 masters, MLP as
 * W1[32][64], W2[64][64], W3[64][64], W4[64][3] (10432 floats), grid as
 * [16][2^19][2] (grid may be NULL); set_params also clears the Adam state.
 * forward: n positions -> n x 3 outputs.  gradients: n samples of (p[3],
 * t[3]) -> mean loss and its gradients (grid_grad may be NULL).  train:
 * `steps` Adam steps on the given samples. */
vs_status vs_nrc_levels(vs_nrc *nrc, uint32_t *res16, uint32_t *dense_mask, float *lo3, float *inv_ext3);
vs_status vs_nrc_get_params(vs_nrc *nrc, float *mlp, float *grid);
vs_status vs_nrc_set_params(vs_nrc *nrc, const float *mlp, const float *grid);
vs_status vs_nrc_forward(vs_nrc *nrc, const float *pos, uint32_t n, float *out);
vs_status vs_nrc_gradients(vs_nrc *nrc, const float *samples, uint32_t n, float *loss, float *mlp_grad,
                           float *grid_grad);
vs_status vs_nrc_train(vs_nrc *nrc, const float *samples, uint32_t n, uint32_t steps, float *last_loss);
/* The message of the last failed radiance-cache call on this thread. */
const char *vs_nrc_last_error(void);
void vs_nrc_destroy(vs_nrc *nrc);

#ifdef __cplusplus
}
#endif

#endif /* VISION_AMD_H */
