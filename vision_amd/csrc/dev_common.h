// Device-side helpers for the HIP kernels (gfx950): sampler, warps, ray
// spawning, scene views.  Each helper cites the reference function it
// restates.
#pragma once

#include <hip/hip_runtime.h>

#include "records.h"
#include "elementary.h"
#include "vmath.h"

namespace vs {
namespace dev {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kShadowEpsilon = 1e-4f;
constexpr float kOneMinusEpsilon = 0x1.fffffep-1f;

// ---- sampler: TEA-seeded LCG (src/math/util.h:13-34,
//      src/render_core/samplers/independent.cpp:24-42)
__device__ __forceinline__ uint32_t tea4(uint32_t v0, uint32_t v1) {
    uint32_t s0 = 0;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}

struct Sampler {
    uint32_t state;
    __device__ __forceinline__ void start(uint32_t px, uint32_t py, uint32_t index, uint32_t dim) {
        state = tea4(tea4(px, py), tea4(index, dim));
    }
    __device__ __forceinline__ float next1d() {
        state = 1664525u * state + 1013904223u;
        return float(state & 0x00ffffffu) * (1.f / 16777216.f);
    }
    __device__ __forceinline__ f2 next2d() {
        float x = next1d();
        float y = next1d();
        return mk2(x, y);
    }
};

// ---- warps (src/math/warp.h)
__device__ __forceinline__ f2 square_to_disk(f2 u) {
    float r = sqrtf(u.x);
    float theta = 2.f * kPi * u.y;
    float s, c;
    em::sincos(theta, &s, &c);
    return mk2(r * c, r * s);
}
__device__ __forceinline__ f3 square_to_cosine_hemisphere(f2 u) {
    f2 d = square_to_disk(u);
    float z = sqrtf(fmaxf(0.f, 1.f - d.x * d.x - d.y * d.y));
    return mk3(d.x, d.y, z);
}
__device__ __forceinline__ f2 square_to_triangle(f2 u) {
    float su0 = sqrtf(u.x);
    return mk2(1.f - su0, u.y * su0);
}
__device__ __forceinline__ float sample_linear(float u, float a, float b) {
    float x = u * (a + b) / (a + sqrtf(a * a + u * (b * b - a * a)));
    float r = fminf(x, kOneMinusEpsilon);
    return (u == 0.f && a == 0.f) ? 0.f : r;
}
__device__ __forceinline__ float sample_tent(float u, float r) {
    return u < 0.5f ? -r * sample_linear((0.5f - u) * 2.f, 1.f, 0.f)
                    : r * sample_linear((u - 0.5f) * 2.f, 1.f, 0.f);
}
// balance heuristic (warp.h:149-155,186-199)
__device__ __forceinline__ float mis_weight(float f, float g) { return f / (f + g); }

// ---- ray origin offset (Waechter & Binder, Ray Tracing Gems ch. 6), the
// method behind the reference's offset_ray_origin
__device__ __forceinline__ float offset_comp(float p, float n) {
    const float origin = 1.f / 32.f, float_scale = 1.f / 65536.f, int_scale = 256.f;
    int of_i = int(int_scale * n);
    float p_i = __int_as_float(__float_as_int(p) + (p < 0.f ? -of_i : of_i));
    return fabsf(p) < origin ? p + float_scale * n : p_i;
}
__device__ __forceinline__ f3 offset_ray_origin(f3 p, f3 n) {
    return mk3(offset_comp(p.x, n.x), offset_comp(p.y, n.y), offset_comp(p.z, n.z));
}

// ---- orthonormal frame
struct Frame {
    f3 x, y, z;
    __device__ __forceinline__ f3 to_local(f3 v) const { return mk3(dot(v, x), dot(v, y), dot(v, z)); }
    __device__ __forceinline__ f3 to_world(f3 v) const { return x * v.x + y * v.y + z * v.z; }
};

__device__ __forceinline__ void coordinate_system(f3 v1, f3 &v2, f3 &v3) {
    if (fabsf(v1.x) > fabsf(v1.y)) {
        float inv = 1.f / sqrtf(v1.x * v1.x + v1.z * v1.z);
        v2 = mk3(-v1.z * inv, 0.f, v1.x * inv);
    } else {
        float inv = 1.f / sqrtf(v1.y * v1.y + v1.z * v1.z);
        v2 = mk3(0.f, v1.z * inv, -v1.y * inv);
    }
    v3 = cross(v1, v2);
}

__device__ __forceinline__ float luminance(f3 c) {
    return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z;
}

}// namespace dev
}// namespace vs
