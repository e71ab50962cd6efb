// ReSTIR DI kernels for gfx950: one thread per pixel in 8x8 tiles (one wave
// per tile keeps neighbouring rays, reservoirs and surfaces together),
// traversal stacks in LDS.
#include <hip/hip_runtime.h>

#include "dev_restir_gi.h"
#include "xcd.h"

namespace vs {

using namespace dev;

namespace {

__global__ void __launch_bounds__(256) clear_kernel(RestirView v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.pixels) return;
    Reservoir r = empty_reservoir();
    for (uint32_t b = 0; b < kRestirReservoirBuffers; ++b) store_reservoir(v, b, i, r);
    SurfaceRec s = empty_surface();
    SurfaceExt e;
    e.throughput = mk3(1.f, 1.f, 1.f);
    e.view_pos = mk3(0.f, 0.f, 0.f);
    e.t_max = 0.f;
    for (uint32_t b = 0; b < 2u; ++b) {
        store_surface(v, b, i, s);
        store_ext(v, b, i, e);
    }
}

__device__ __forceinline__ void add_counters(unsigned long long *counters, const PathCounters &pc) {
    if (!counters) return;
    atomicAdd(counters + 0, (unsigned long long)pc.rays);
    atomicAdd(counters + 1, (unsigned long long)pc.nodes);
    atomicAdd(counters + 2, (unsigned long long)pc.tris);
}

// Register budget of shader0 and shader1: four resident waves per SIMD for
// the diffuse family (128 VGPRs; the compiler alone takes 161-187 and
// three or two waves).  Measured on the 1080p many-lights job (64 frames):
// shader0 150.7 -> 139.6 ms, shader1 59.9 -> 54.2 ms; five and six waves
// were slower, and the spatial pass is fastest left free.  Kernels for
// every material keep the compiler's choice (not measured).
constexpr int restir_waves(int ms) { return ms == MS_ALL ? 1 : 4; }

template <int NS, bool RGB, int MS>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(restir_waves(MS)))) initial_kernel(SceneView sc, CameraView cam, RestirView v,
                                                     unsigned long long *__restrict__ counters) {
    extern __shared__ uint2 lds_stack[];
    const uint2 tile = xcd_tile_2d();
    uint32_t px = tile.x * 8u + (threadIdx.x & 7u);
    uint32_t py = tile.y * 8u + (threadIdx.x >> 3);
    if (px >= v.width || py >= v.height) return;
    PathCounters pc{0, 0, 0};
    restir_initial<NS, RGB, MS>(sc, cam, v, px, py, lds_stack + threadIdx.x, blockDim.x, pc);
    add_counters(counters, pc);
}

template <int NS, bool RGB, int MS>
__global__ void __launch_bounds__(64) spatial_kernel(SceneView sc, RestirView v, uint32_t in, uint32_t out,
                                                     uint32_t dim) {
    const uint2 tile = xcd_tile_2d();
    uint32_t px = tile.x * 8u + (threadIdx.x & 7u);
    uint32_t py = tile.y * 8u + (threadIdx.x >> 3);
    if (px >= v.width || py >= v.height) return;
    restir_spatial_pass<NS, RGB, MS>(sc, v, px, py, in, out, dim);
}

// shader1 plus the film's running mean (RealTimeIntegrator's combine,
// rt.cpp:121-130, with the indirect half absent)
template <int NS, bool RGB, int MS>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(restir_waves(MS)))) final_kernel(SceneView sc, CameraView cam, RestirView v, RenderView rv,
                                                   uint32_t in, float4 *__restrict__ rt, float4 *__restrict__ out,
                                                   unsigned long long *__restrict__ counters) {
    extern __shared__ uint2 lds_stack[];
    const uint2 tile = xcd_tile_2d();
    uint32_t px = tile.x * 8u + (threadIdx.x & 7u);
    uint32_t py = tile.y * 8u + (threadIdx.x >> 3);
    if (px >= v.width || py >= v.height) return;
    PathCounters pc{0, 0, 0};
    f3 L = restir_final<NS, RGB, MS>(sc, cam, v, px, py, in, lds_stack + threadIdx.x, blockDim.x, pc);
    uint32_t idx = py * v.width + px;
    add_counters(counters, pc);
    if (v.gi_direct) {
        // ReSTIR GI open: the combine adds this after the indirect half
        v.gi_direct[idx] = make_float4(L.x, L.y, L.z, 1.f);
        return;
    }
    float4 acc = film_add(rv, rt[idx], L, v.frame);
    rt[idx] = acc;
    if (rv.write_output) out[idx] = film_output(rv, acc);
}

// ---- ReSTIR GI

__global__ void __launch_bounds__(256) gi_clear_kernel(RestirGiView g, uint32_t pixels) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pixels) return;
    GiReservoir r = empty_gi_reservoir();
    for (uint32_t b = 0; b < kGiReservoirBuffers; ++b) store_gi_reservoir(g, pixels, b, i, r);
    g.rad[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

template <int MS>
__global__ void __launch_bounds__(64) gi_initial_temporal_kernel(SceneView sc, CameraView cam, RestirView v,
                                                                 RestirGiView g, RenderView rv,
                                                                 unsigned long long *__restrict__ counters) {
    extern __shared__ uint2 lds_stack[];
    // launch order: a path's cost varies across the image, and XCD order
    // would leave the XCDs with the cheap bands idle
    uint32_t px = blockIdx.x * 8u + (threadIdx.x & 7u);
    uint32_t py = blockIdx.y * 8u + (threadIdx.x >> 3);
    if (px >= v.width || py >= v.height) return;
    PathCounters pc{0, 0, 0};
    restir_gi_initial_temporal<MS>(sc, cam, rv, v, g, px, py, lds_stack + threadIdx.x, blockDim.x, pc);
    add_counters(counters, pc);
}

// spatial reuse and shading, then RealTimeIntegrator's combine (rt.cpp:119-126):
// direct x 1 + indirect x 1 into the film's running mean
template <int MS>
__global__ void __launch_bounds__(64) gi_spatial_shade_kernel(SceneView sc, RestirView v, RestirGiView g,
                                                              RenderView rv, float4 *__restrict__ rt,
                                                              float4 *__restrict__ out,
                                                              unsigned long long *__restrict__ counters) {
    extern __shared__ uint2 lds_stack[];
    const uint2 tile = xcd_tile_2d();
    uint32_t px = tile.x * 8u + (threadIdx.x & 7u);
    uint32_t py = tile.y * 8u + (threadIdx.x >> 3);
    if (px >= v.width || py >= v.height) return;
    const uint32_t idx = py * v.width + px;
    PathCounters pc{0, 0, 0};
    const SurfaceRec cs = load_surface(v, cur_slot(v.frame), idx);
    f3 il = mk3(0.f, 0.f, 0.f);
    if (cs.hit()) {
        il = restir_gi_spatial_shade<MS>(sc, v, g, cs, px, py, lds_stack + threadIdx.x, blockDim.x, pc);
        g.rad[idx] = make_float4(il.x, il.y, il.z, 1.f);
    }
    const float4 dl = g.direct[idx];
    f3 L = mk3(dl.x * 1.f + il.x * 1.f, dl.y * 1.f + il.y * 1.f, dl.z * 1.f + il.z * 1.f);
    float4 acc = film_add(rv, rt[idx], L, v.frame);
    rt[idx] = acc;
    if (rv.write_output) out[idx] = film_output(rv, acc);
    add_counters(counters, pc);
}

dim3 tiles(const RestirView &v) { return dim3((v.width + 7u) / 8u, (v.height + 7u) / 8u); }

// kernel variant of a launch: spectral layout x material set
template <template <int, bool, int> class K, typename... Args>
hipError_t dispatch(uint32_t spectrum_dim, bool rgb, int material_set, Args &&...args) {
    if (material_set != MS_ALL && material_set != MS_DIFFUSE) return hipErrorInvalidValue;
    const bool all = material_set == MS_ALL;
    if (rgb) return all ? K<3, true, MS_ALL>::launch(args...) : K<3, true, MS_DIFFUSE>::launch(args...);
    if (spectrum_dim == 4) return all ? K<4, false, MS_ALL>::launch(args...) : K<4, false, MS_DIFFUSE>::launch(args...);
    if (spectrum_dim == 3) return all ? K<3, false, MS_ALL>::launch(args...) : K<3, false, MS_DIFFUSE>::launch(args...);
    return hipErrorInvalidValue;
}

template <int NS, bool RGB, int MS>
struct Initial {
    static hipError_t launch(const SceneView &sc, const CameraView &cam, const RestirView &v,
                             unsigned long long *counters, hipStream_t stream) {
        size_t lds = size_t(sc.stack_size) * 64 * sizeof(uint2);
        hipLaunchKernelGGL((initial_kernel<NS, RGB, MS>), tiles(v), dim3(64), lds, stream, sc, cam, v, counters);
        return hipGetLastError();
    }
};
template <int NS, bool RGB, int MS>
struct Spatial {
    static hipError_t launch(const SceneView &sc, const RestirView &v, uint32_t in, uint32_t out, uint32_t dim,
                             hipStream_t stream) {
        hipLaunchKernelGGL((spatial_kernel<NS, RGB, MS>), tiles(v), dim3(64), 0, stream, sc, v, in, out, dim);
        return hipGetLastError();
    }
};
template <int NS, bool RGB, int MS>
struct Final {
    static hipError_t launch(const SceneView &sc, const CameraView &cam, const RestirView &v, const RenderView &rv,
                             uint32_t in, float4 *rt, float4 *out, unsigned long long *counters, hipStream_t stream) {
        size_t lds = size_t(sc.stack_size) * 64 * sizeof(uint2);
        hipLaunchKernelGGL((final_kernel<NS, RGB, MS>), tiles(v), dim3(64), lds, stream, sc, cam, v, rv, in, rt, out,
                           counters);
        return hipGetLastError();
    }
};

}// namespace

hipError_t restir_gi_clear(const RestirGiView &g, uint32_t pixels, hipStream_t stream) {
    if (pixels == 0) return hipSuccess;
    hipLaunchKernelGGL(gi_clear_kernel, dim3((pixels + 255u) / 256u), dim3(256), 0, stream, g, pixels);
    return hipGetLastError();
}

hipError_t restir_gi_initial_temporal(const SceneView &sc, const CameraView &cam, const RestirView &v,
                                      const RestirGiView &g, const RenderView &rv, int material_set,
                                      unsigned long long *counters, hipStream_t stream) {
    if (material_set != MS_ALL && material_set != MS_DIFFUSE) return hipErrorInvalidValue;
    size_t lds = size_t(sc.stack_size) * 64 * sizeof(uint2);
    if (material_set == MS_ALL)
        hipLaunchKernelGGL((gi_initial_temporal_kernel<MS_ALL>), tiles(v), dim3(64), lds, stream, sc, cam, v, g, rv,
                           counters);
    else
        hipLaunchKernelGGL((gi_initial_temporal_kernel<MS_DIFFUSE>), tiles(v), dim3(64), lds, stream, sc, cam, v, g,
                           rv, counters);
    return hipGetLastError();
}

hipError_t restir_gi_spatial_shade(const SceneView &sc, const RestirView &v, const RestirGiView &g,
                                   const RenderView &rv, int material_set, float4 *rt, float4 *out,
                                   unsigned long long *counters, hipStream_t stream) {
    if (material_set != MS_ALL && material_set != MS_DIFFUSE) return hipErrorInvalidValue;
    size_t lds = size_t(sc.stack_size) * 64 * sizeof(uint2);
    if (material_set == MS_ALL)
        hipLaunchKernelGGL((gi_spatial_shade_kernel<MS_ALL>), tiles(v), dim3(64), lds, stream, sc, v, g, rv, rt, out,
                           counters);
    else
        hipLaunchKernelGGL((gi_spatial_shade_kernel<MS_DIFFUSE>), tiles(v), dim3(64), lds, stream, sc, v, g, rv, rt,
                           out, counters);
    return hipGetLastError();
}

hipError_t restir_clear(const RestirView &v, hipStream_t stream) {
    if (v.pixels == 0) return hipSuccess;
    hipLaunchKernelGGL(clear_kernel, dim3((v.pixels + 255u) / 256u), dim3(256), 0, stream, v);
    return hipGetLastError();
}

hipError_t restir_initial(const SceneView &sc, const CameraView &cam, const RestirView &v, uint32_t spectrum_dim,
                          bool rgb, int material_set, unsigned long long *counters, hipStream_t stream) {
    return dispatch<Initial>(spectrum_dim, rgb, material_set, sc, cam, v, counters, stream);
}

hipError_t restir_spatial(const SceneView &sc, const RestirView &v, uint32_t spectrum_dim, bool rgb,
                          int material_set, uint32_t in, uint32_t out, uint32_t dim, hipStream_t stream) {
    if (in >= kRestirReservoirBuffers || out >= kRestirReservoirBuffers || in < 2u || out < 2u || in == out)
        return hipErrorInvalidValue;
    return dispatch<Spatial>(spectrum_dim, rgb, material_set, sc, v, in, out, dim, stream);
}

hipError_t restir_final(const SceneView &sc, const CameraView &cam, const RestirView &v, const RenderView &rv,
                        uint32_t spectrum_dim, bool rgb, int material_set, uint32_t in, float4 *rt, float4 *out,
                        unsigned long long *counters, hipStream_t stream) {
    if (in < 2u || in >= kRestirReservoirBuffers) return hipErrorInvalidValue;
    return dispatch<Final>(spectrum_dim, rgb, material_set, sc, cam, v, rv, in, rt, out, counters, stream);
}

}// namespace vs
