// Workgroup -> tile order for the 8 XCDs of an MI355X.
//
// The dispatcher hands workgroups to the XCDs round-robin by linear
// workgroup id (id mod 8), and each XCD has its own L2.  Read in launch
// order, neighbouring tiles land on different XCDs, so a tile's neighbour
// records (spatial reuse, a-trous taps) and the BVH nodes its primary rays
// share are fetched into eight L2s.  xcd_order renumbers the linear id so
// that XCD k takes the k-th contiguous range of tiles, in order.
//
// The split is static: an XCD whose band is cheap idles once it is done.
// It is used where a pixel's cost is about even and pixels read their
// neighbours (ReSTIR spatial reuse and shading, the SVGF passes); the
// path-tracing kernels keep launch order, which mixes cheap and costly
// tiles on every XCD (measured on one MI355X: ReSTIR spatial pass 43.0 ->
// 28.3 ms and shading 54.2 -> 46.9 ms per 64-frame 1080p job; the PT
// megakernel 428 -> 718 ms and the wavefront kernels 3-11 % slower on the
// procedural scene, whose sky fills the top bands).
#pragma once

#include <hip/hip_runtime.h>

namespace vs {

constexpr uint32_t kXcds = 8;

// Bijection on [0, n): workgroup b (on XCD b mod 8, the (b / 8)-th there)
// -> tile index.  XCD x owns q + (x < r) tiles starting at x*q + min(x, r),
// where q = n / 8, r = n mod 8.
__host__ __device__ __forceinline__ uint32_t xcd_order(uint32_t b, uint32_t n) {
    const uint32_t q = n / kXcds, r = n % kXcds;
    const uint32_t x = b % kXcds, i = b / kXcds;
    return x * q + (x < r ? x : r) + i;
}

#if defined(__HIPCC__)
// 1-D grid: the block's index in XCD order
__device__ __forceinline__ uint32_t xcd_block_1d() { return xcd_order(blockIdx.x, gridDim.x); }

// 2-D grid of tiles: (tile x, tile y) of this block in XCD order, tiles
// numbered row-major
__device__ __forceinline__ uint2 xcd_tile_2d() {
    const uint32_t t = xcd_order(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    return make_uint2(t % gridDim.x, t / gridDim.x);
}
#endif

}// namespace vs
