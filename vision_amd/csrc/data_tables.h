// Host access to the data tables shipped in vision_amd/data: energy
// compensation LUTs, CIE 1931 / D65 curves, measured metal ior and the
// sRGB -> sigmoid-spectrum coefficient table.
#pragma once

#include <string>
#include <vector>

#include "vmath.h"

namespace vs {

// directory holding the data files (next to libvision_amd.so)
const std::string &data_dir();
// whole file as float32; throws if missing or of the wrong size
std::vector<float> load_f32(const std::string &name, size_t expect_floats);

// energy-compensation tables in the layout of dev_bsdf.h (LUT_*)
const std::vector<float> &lut_tables();

// SPD tables (src/base/color/spd.cpp): 94-sample CIE X, Y, Z, D65 taken every
// 5 nm from the 1 nm curves, their nominal interval (471 / 94) and the
// trapezoid integral of Y used by HeroWavelengthSpectrum::cie_xyz
struct CieSpd {
    std::vector<float> x, y, z, d65;
    float interval;
    float y_integral;
};
const CieSpd &cie_spd();

// piecewise-linear SPD evaluation of SPD::eval (spd.cpp:64-72)
float spd_eval_host(const std::vector<float> &tab, float lambda);

// metal ior (eta, k) tables by name, 95 samples over [360, 830] nm;
// unknown names resolve to the first metal in name order (ComplexIorTable)
void metal_ior(const std::string &name, std::vector<float> &eta, std::vector<float> &k, std::string *resolved);

// RGBToSpectrumTable::decode_albedo / decode_unbound (hero.cpp:91-139)
f4 rgb_to_albedo_coeffs(f3 rgb);
f4 rgb_to_unbound_coeffs(f3 rgb);
bool rgb2spec_available();

}// namespace vs
