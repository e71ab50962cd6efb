// Plain-data records laid out identically on host and device.
#pragma once

#include <cstdint>

namespace vs {

enum MaterialType : uint32_t {
    MAT_DIFFUSE = 0,   // Lambert (diffuse.cpp:21-30, bxdf.h:82-96)
    MAT_OREN_NAYAR = 1,// diffuse with "sigma" (bxdf.cpp:94-121)
    MAT_BLACK = 2,     // black_body on light quads (zero BSDF)
    MAT_METAL = 3,     // GGX conductor (metal.cpp:39-157)
    MAT_GLASS = 4,     // rough dielectric (glass.cpp:171-258)
    MAT_PRINCIPLED = 5,// principled_bsdf lobe stack
    MAT_MIRROR = 6,    // GGX reflection, Fresnel 1, tinted (mirror.cpp:11-74)
    MAT_MIX = 7,       // two component records blended by a fraction (mix.cpp)
    MAT_ADD = 8,       // two component records summed (add.cpp)
};

enum LightType : uint32_t {
    LIGHT_AREA = 0,
    LIGHT_POINT = 1,
    LIGHT_SPOT = 2,
    LIGHT_ENV = 3,
};

// 128 B material record; meaning of p[] depends on type (see scene.cpp pack).
// A colour slot of four floats at p[4 k] (k = 0..3) that names an image
// holds the texture index (as bits) in p[4 k] and sets flag bit 16 + k.
constexpr uint32_t MATF_TEXTURE_SLOT0 = 1u << 16;
struct MaterialRecord {
    uint32_t type;
    uint32_t flags;
    float p[30];
};
static_assert(sizeof(MaterialRecord) == 128, "MaterialRecord layout");

struct LightRecord {
    uint32_t type;
    uint32_t two_sided;
    uint32_t inst;        // emitting instance (area)
    uint32_t tri_base;    // first global triangle of the emitter (area)
    uint32_t tri_count;   // triangles of the emitter (area)
    uint32_t alias_offset;// offset of its per-triangle alias table
    float alias_integral; // sum(area)/n as float (alias.h:86-122)
    uint32_t tex;         // area: emission image (index into the textures), or 0xFFFFFFFF
    float color[3];
    float scale;
    float position[3];
    float angle;// spot: radians
    float direction[3];
    float falloff;// spot: radians
    float cos_angle;        // spot: cos(angle), computed once on the host
    float cos_falloff_start;// spot: cos(max(0, angle - falloff))
    float pad1, pad2;
    // hero mode: RGBIlluminationSpectrum of the colour (sigmoid coefficients,
    // unbound scale), evaluated times D65 and `scale` (hero.cpp:134-152)
    float illum[4];
};
static_assert(sizeof(LightRecord) == 112, "LightRecord layout");

struct InstanceRecord {
    uint32_t mat;  // material index, 0xFFFFFFFF = none
    uint32_t light;// light index, 0xFFFFFFFF = not emissive
    uint32_t tri_base;
    uint32_t pad;
};

struct AliasEntry {
    float prob;
    uint32_t alias;
};

constexpr uint32_t kInvalid = 0xFFFFFFFFu;

}// namespace vs
