// Image files in and out: the formats a Vision scene names for its output
// (`output.fn`, src/core/node_desc.cpp:360) and for its textures and
// environment maps (src/base/mgr/image_pool.cpp:15-30).
//
// Writers take float RGBA rows top to bottom:
//   .exr  OpenEXR scanline file, channels A, B, G, R as 32-bit float, no compression
//   .hdr  Radiance RGBE, flat scanlines
//   .png  8-bit RGBA, value = round(clamp(x, 0, 1) * 255), zlib-compressed rows
//   .pfm  little-endian colour PFM (bottom-to-top rows, as the format stores them)
// Readers return float RGBA rows top to bottom:
//   .exr  scanline files with NONE, ZIPS or ZIP compression; HALF, FLOAT or UINT
//         channels R, G, B, A or Y (missing colour channels copy Y or stay 0, A = 1)
//   .hdr  Radiance RGBE, flat or new-style run-length scanlines
//   .png  8- or 16-bit grey, grey+alpha, RGB, RGBA or palette, no interlace
//   .pfm  colour or grey PFM, either byte order
//   .jpg  baseline, extended sequential or progressive Huffman JPEG, 8-bit
//         grey or YCbCr (jpeg.cpp)
// Values read from 8/16-bit files are code / max; no transfer function is
// applied here (the caller decides the colour space).
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace vs {

struct ImageRGBA {
    uint32_t width{0}, height{0};
    std::vector<float> rgba;// width * height * 4, row 0 = top
};

// sRGB transfer functions (IEC 61966-2-1)
inline float srgb_to_linear(float x) {
    return x <= 0.04045f ? x / 12.92f : std::pow((x + 0.055f) / 1.055f, 2.4f);
}
inline float linear_to_srgb(float x) {
    return x <= 0.0031308f ? 12.92f * x : 1.055f * std::pow(x, 1.f / 2.4f) - 0.055f;
}

// the lower-case extension without the dot ("exr", "png", ...)
std::string image_extension(const std::string &path);

// throws std::runtime_error on an unsupported extension or an I/O error
void write_image(const std::string &path, uint32_t width, uint32_t height, const float *rgba);
ImageRGBA read_image(const std::string &path);

// the JPEG decoder behind read_image (throws on a corrupt or unsupported file)
ImageRGBA decode_jpeg(const std::vector<uint8_t> &data, const std::string &path);

}// namespace vs
