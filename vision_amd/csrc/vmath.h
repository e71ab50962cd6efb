// Small float vector types shared by the host library and the HIP kernels.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIP__)
#define VS_HD __host__ __device__ __forceinline__
#else
#define VS_HD inline
#endif

namespace vs {

struct f2 {
    float x, y;
};
struct f3 {
    float x, y, z;
};
struct alignas(16) f4 {
    float x, y, z, w;
};
struct u3 {
    uint32_t x, y, z;
};

VS_HD f2 mk2(float x, float y) { return {x, y}; }
VS_HD f3 mk3(float x, float y, float z) { return {x, y, z}; }
VS_HD f3 mk3(float s) { return {s, s, s}; }
VS_HD f4 mk4(float x, float y, float z, float w) { return {x, y, z, w}; }

VS_HD f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
VS_HD f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
VS_HD f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
VS_HD f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
VS_HD f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
VS_HD f3 operator*(float s, f3 a) { return {a.x * s, a.y * s, a.z * s}; }
VS_HD f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
VS_HD f3 operator/(f3 a, f3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
VS_HD f3 &operator+=(f3 &a, f3 b) { a = a + b; return a; }
VS_HD f3 &operator*=(f3 &a, f3 b) { a = a * b; return a; }
VS_HD f3 &operator*=(f3 &a, float s) { a = a * s; return a; }
VS_HD f2 operator+(f2 a, f2 b) { return {a.x + b.x, a.y + b.y}; }
VS_HD f2 operator-(f2 a, f2 b) { return {a.x - b.x, a.y - b.y}; }
VS_HD f2 operator*(f2 a, float s) { return {a.x * s, a.y * s}; }

VS_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
VS_HD f3 cross(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
VS_HD float length_squared(f3 a) { return dot(a, a); }
VS_HD float length(f3 a) { return sqrtf(dot(a, a)); }
VS_HD f3 normalize(f3 a) { return a * (1.0f / length(a)); }
VS_HD f3 vmin(f3 a, f3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
VS_HD f3 vmax(f3 a, f3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
VS_HD float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
VS_HD float max_comp(f3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }

// 4x4 matrix, column-major: c[col] is a column.  A scene-file "matrix4x4"
// lists its four columns (the translation is the last inner array).
struct m4 {
    f4 c[4];
};

VS_HD m4 m4_identity() {
    return {{mk4(1, 0, 0, 0), mk4(0, 1, 0, 0), mk4(0, 0, 1, 0), mk4(0, 0, 0, 1)}};
}
VS_HD float m4_at(const m4 &m, int row, int col) {
    const f4 &c = m.c[col];
    return row == 0 ? c.x : (row == 1 ? c.y : (row == 2 ? c.z : c.w));
}
VS_HD f3 xform_point(const m4 &m, f3 p) {
    return {m.c[0].x * p.x + m.c[1].x * p.y + m.c[2].x * p.z + m.c[3].x,
            m.c[0].y * p.x + m.c[1].y * p.y + m.c[2].y * p.z + m.c[3].y,
            m.c[0].z * p.x + m.c[1].z * p.y + m.c[2].z * p.z + m.c[3].z};
}
VS_HD f3 xform_vector(const m4 &m, f3 v) {
    return {m.c[0].x * v.x + m.c[1].x * v.y + m.c[2].x * v.z,
            m.c[0].y * v.x + m.c[1].y * v.y + m.c[2].y * v.z,
            m.c[0].z * v.x + m.c[1].z * v.y + m.c[2].z * v.z};
}

}// namespace vs
