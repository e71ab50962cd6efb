// Host construction of sampling tables and device records.
#pragma once

#include <vector>

#include "records.h"
#include "scene.h"

namespace vs {

struct AliasTableHost {
    std::vector<AliasEntry> table;
    std::vector<float> func;
    float integral{0.f};
};

// Vose alias table with the reference's construction order
// (AliasTable::build, src/render_core/warpers/alias.h:86-122)
AliasTableHost build_alias(const std::vector<float> &weights);

// light power per LightType (area.cpp:87-89, point.cpp:32-34, spot.cpp:44-46)
f3 light_power(const LightHost &l);

// Pack materials into 128-byte device records.
std::vector<MaterialRecord> pack_materials(const SceneHost &s);

struct LightTables {
    std::vector<LightRecord> records;
    std::vector<AliasEntry> alias;
    std::vector<float> func;
    uint32_t select_offset{0};
    float select_integral{0.f};
};
LightTables pack_lights(const SceneHost &s);

}// namespace vs
