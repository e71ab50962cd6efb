// Host builder for the 8-wide BVH the HIP traversal kernels walk.
//
// Replaces the reference's OptiX accel build (src/base/mgr/geometry.cpp:36-52):
// a binned-SAH binary BVH is collapsed greedily into 8-wide nodes whose child
// boxes are quantized to 8 bits per plane relative to the node box.
#pragma once

#include <cstdint>
#include <vector>

#include "vmath.h"

namespace vs {

// 80-byte node (5 x 16 B):
//  w[0..2]  origin p (float bits), w[3] = ex | ey<<8 | ez<<16 | imask<<24
//           (ex = biased exponent of the per-axis step 2^(ex-127))
//  w[4]     child_base: index of the first internal child node
//  w[5]     tri_base:   index of the first leaf triangle of this node
//  w[6..7]  meta[8]:    per slot; 0 = empty, (count<<5)|offset = leaf of
//                       `count` (1..3) triangles at tri_base+offset,
//                       1 = internal child (its imask bit is set)
//  w[8..9]  qlo_x[8], w[10..11] qlo_y[8], w[12..13] qlo_z[8]
//  w[14..15] qhi_x[8], w[16..17] qhi_y[8], w[18..19] qhi_z[8]
// Slot s holds the child whose centre lies most toward the octant direction
// ((s&1)?+x:-x, (s&2)?+y:-y, (s&4)?+z:-z); a ray whose direction-sign bits are
// r (bit k set when d_k < 0) visits slots in the order i ^ r, near to far.
struct Bvh8Node {
    uint32_t w[20];
};
static_assert(sizeof(Bvh8Node) == 80, "Bvh8Node layout");

// leaf triangle: three float4 (v0.xyz + original triangle index bits, v1, v2)
struct alignas(16) LeafTri {
    float v0[3];
    uint32_t id;
    float v1[3];
    uint32_t pad1;
    float v2[3];
    uint32_t pad2;
};
static_assert(sizeof(LeafTri) == 48, "LeafTri layout");

struct Bvh8 {
    std::vector<Bvh8Node> nodes;
    std::vector<LeafTri> tris;
    uint32_t max_depth{0};// deepest node level (root = 1)
    uint32_t bvh2_nodes{0};
    double sah_cost{0.0};
};

Bvh8 build_bvh8(const std::vector<f3> &P, const std::vector<u3> &tris);

}// namespace vs
