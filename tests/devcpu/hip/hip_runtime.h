// Host stand-in for <hip/hip_runtime.h> used only by tests/devcpu: it lets
// the device headers of vision_amd compile as plain C++ so the kernels'
// source can be checked against the oracle on the CPU.  Not part of the
// library build.
#pragma once
#include <math.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#define __device__
#define __host__
typedef int hipError_t;
typedef void *hipStream_t;
#define __global__
#define __forceinline__ inline

struct uint2 {
    uint32_t x, y;
};
struct uint4 {
    uint32_t x, y, z, w;
};
struct float2 {
    float x, y;
};
struct float4 {
    float x, y, z, w;
};
inline uint2 make_uint2(uint32_t x, uint32_t y) { return {x, y}; }
inline uint4 make_uint4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) { return {x, y, z, w}; }
inline float4 make_float4(float x, float y, float z, float w) { return {x, y, z, w}; }
inline uint32_t __float_as_uint(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float __uint_as_float(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
inline int __float_as_int(float f) {
    int u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float __int_as_float(int u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
inline int __popc(uint32_t x) { return __builtin_popcount(x); }
using std::isinf;
using std::isnan;
using std::max;
using std::min;
