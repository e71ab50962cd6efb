// Batched serving of path-tracing jobs: the device-side records shared by
// the host scheduler (serve.cpp) and the persistent kernels (serve.hip).
//
// A step runs every admitted job forward by a number of frames.  The host
// writes one StepEntry per job into a small table; the kernel's waves claim
// work items (one 8x8, 16x4, 32x2 or 64x1 pixel tile of one job) from a
// device counter until the step's items are exhausted, so jobs of different
// scenes, resolutions and sample counts share one launch.  Every pixel's
// frames run in order inside one thread, so a job's film is the one a
// single-job render with the same frame indices produces.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "dev_scene.h"

namespace vs {

// one admitted job on the device (slot table entry)
struct ServeJob {
    CameraView cam;
    RenderView rv;
    float4 *rt;      // running-mean film, width * height
    uint32_t scene;  // index into the step's SceneView table
    uint32_t tiles_x;// items per row at the step's tile width
    uint32_t pad[2];
};

// one job's share of a step: frames [frame_start, frame_start + frames) over
// all its items [first_item, first_item + items); an item is a tile of
// tile_w x (64 / tile_w) pixels
struct StepEntry {
    uint32_t job;
    uint32_t frame_start;
    uint32_t frames;
    uint32_t first_item;
};

constexpr uint32_t kServeMaxEntries = 256;

// the table a step reads; variants (sRGB / hero 4 / hero 3, all materials /
// diffuse only) each have their own
struct StepTable {
    uint32_t num_entries;
    uint32_t num_items;
    uint32_t pad[2];
    StepEntry e[kServeMaxEntries];
};

// kernel variants: spectrum form x material set
enum ServeVariant : uint32_t {
    SV_RGB = 0,       // sRGB, all materials
    SV_RGB_DIFFUSE,   // sRGB, diffuse / Oren-Nayar / black only
    SV_HERO4,         // 4 hero wavelengths
    SV_HERO4_DIFFUSE,
    SV_HERO3,         // 3 hero wavelengths
    SV_HERO3_DIFFUSE,
    SV_COUNT
};

// register budgets the kernels are compiled for: 0 = the compiler's choice,
// 6 = the measured budget: at least 6 resident waves per SIMD for the
// diffuse-family kernels, 4 for the every-material ones
inline bool serve_occ_supported(uint32_t occ) { return occ == 0 || occ == 6; }

// the persistent kernel of one variant: `blocks` one-wave workgroups claim
// items through *next (zeroed by serve_reset before the step)
hipError_t serve_launch(uint32_t variant, uint32_t occ, const SceneView *scenes, const ServeJob *jobs,
                        const StepTable *table, uint32_t *next, uint32_t blocks, uint32_t tile_w,
                        uint32_t stack_entries, unsigned long long *counters, hipStream_t stream);
// zero n counters (a kernel, so the step can be captured in a graph)
hipError_t serve_reset(uint32_t *next, uint32_t n, hipStream_t stream);
// clear a film (n float4)
hipError_t serve_clear(float4 *film, uint32_t n, hipStream_t stream);
// the collector's output buffer from a finished job's film
hipError_t serve_output(const RenderView &rv, const float4 *rt, float4 *out, uint32_t n, hipStream_t stream);
// resident one-wave workgroups per CU of a variant's kernel with its LDS stack
int serve_max_blocks_per_cu(uint32_t variant, uint32_t occ, uint32_t stack_entries);
// registers per lane of a variant's kernel
int serve_kernel_vgprs(uint32_t variant, uint32_t occ);

}// namespace vs
