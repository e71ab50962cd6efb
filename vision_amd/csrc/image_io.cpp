// Image readers and writers (EXR, Radiance HDR, PNG, PFM).  The layouts
// follow the formats' published specifications: OpenEXR 2 file layout
// (magic, attribute header, offset table, scanline chunks; ZIP chunks are
// zlib streams of a byte-delta-predicted, even/odd-split buffer), Radiance
// RGBE (Ward, Graphics Gems II), PNG (ISO/IEC 15948) and PFM.
#include "image_io.h"

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace vs {

namespace {

std::vector<uint8_t> read_file(const std::string &path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open image " + path);
    return std::vector<uint8_t>((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
}

void write_file(const std::string &path, const std::vector<uint8_t> &bytes) {
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out) throw std::runtime_error("cannot write image " + path);
    out.write(reinterpret_cast<const char *>(bytes.data()), std::streamsize(bytes.size()));
    if (!out) throw std::runtime_error("write failed: " + path);
}

void put_u32le(std::vector<uint8_t> &b, uint32_t v) {
    for (int i = 0; i < 4; ++i) b.push_back(uint8_t(v >> (8 * i)));
}
void put_u64le(std::vector<uint8_t> &b, uint64_t v) {
    for (int i = 0; i < 8; ++i) b.push_back(uint8_t(v >> (8 * i)));
}
void put_f32le(std::vector<uint8_t> &b, float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    put_u32le(b, u);
}
void put_u32be(std::vector<uint8_t> &b, uint32_t v) {
    for (int i = 3; i >= 0; --i) b.push_back(uint8_t(v >> (8 * i)));
}
void put_str0(std::vector<uint8_t> &b, const char *s) {
    while (*s) b.push_back(uint8_t(*s++));
    b.push_back(0);
}

struct Reader {
    const std::vector<uint8_t> &d;
    size_t pos{0};
    const std::string &path;
    void need(size_t n) const {
        if (pos > d.size() || n > d.size() - pos) throw std::runtime_error("truncated image " + path);
    }
    uint32_t u32le() {
        need(4);
        uint32_t v = uint32_t(d[pos]) | uint32_t(d[pos + 1]) << 8 | uint32_t(d[pos + 2]) << 16 |
                     uint32_t(d[pos + 3]) << 24;
        pos += 4;
        return v;
    }
    uint64_t u64le() {
        uint64_t lo = u32le();
        uint64_t hi = u32le();
        return lo | hi << 32;
    }
    int32_t i32le() { return int32_t(u32le()); }
    uint32_t u32be() {
        need(4);
        uint32_t v = uint32_t(d[pos]) << 24 | uint32_t(d[pos + 1]) << 16 | uint32_t(d[pos + 2]) << 8 |
                     uint32_t(d[pos + 3]);
        pos += 4;
        return v;
    }
    std::string str0() {
        std::string s;
        while (true) {
            need(1);
            char c = char(d[pos++]);
            if (!c) return s;
            s.push_back(c);
            if (s.size() > 255) throw std::runtime_error("bad string in " + path);
        }
    }
};

void check_size(uint64_t w, uint64_t h, const std::string &path) {
    if (w == 0 || h == 0 || w > 65536 || h > 65536) throw std::runtime_error("bad image size in " + path);
}

// ---- half floats ----------------------------------------------------------

float half_to_float(uint16_t h) {
    uint32_t s = uint32_t(h >> 15) << 31, e = (h >> 10) & 0x1F, m = h & 0x3FF;
    uint32_t u;
    if (e == 0) {
        if (m == 0) {
            u = s;
        } else {// subnormal: renormalise
            int sh = 0;
            while (!(m & 0x400)) {
                m <<= 1;
                ++sh;
            }
            m &= 0x3FF;
            u = s | uint32_t(127 - 15 - sh + 1) << 23 | m << 13;
        }
    } else if (e == 31) {
        u = s | 0x7F800000u | m << 13;
    } else {
        u = s | (e - 15 + 127) << 23 | m << 13;
    }
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// ---- OpenEXR ---------------------------------------------------------------

constexpr uint32_t kExrMagic = 20000630u;

void exr_attr(std::vector<uint8_t> &b, const char *name, const char *type, const std::vector<uint8_t> &value) {
    put_str0(b, name);
    put_str0(b, type);
    put_u32le(b, uint32_t(value.size()));
    b.insert(b.end(), value.begin(), value.end());
}

std::vector<uint8_t> encode_exr(uint32_t w, uint32_t h, const float *rgba) {
    std::vector<uint8_t> b;
    put_u32le(b, kExrMagic);
    put_u32le(b, 2u);// version 2, single-part scanline
    std::vector<uint8_t> v;
    for (const char *c : {"A", "B", "G", "R"}) {// channel list in alphabetical order
        put_str0(v, c);
        put_u32le(v, 2u);// FLOAT
        put_u32le(v, 0u);// pLinear + reserved
        put_u32le(v, 1u);
        put_u32le(v, 1u);
    }
    v.push_back(0);
    exr_attr(b, "channels", "chlist", v);
    exr_attr(b, "compression", "compression", {0});
    v.clear();
    for (uint32_t x : {0u, 0u, w - 1, h - 1}) put_u32le(v, x);
    exr_attr(b, "dataWindow", "box2i", v);
    exr_attr(b, "displayWindow", "box2i", v);
    exr_attr(b, "lineOrder", "lineOrder", {0});
    v.clear();
    put_f32le(v, 1.f);
    exr_attr(b, "pixelAspectRatio", "float", v);
    v.clear();
    put_f32le(v, 0.f);
    put_f32le(v, 0.f);
    exr_attr(b, "screenWindowCenter", "v2f", v);
    v.clear();
    put_f32le(v, 1.f);
    exr_attr(b, "screenWindowWidth", "float", v);
    b.push_back(0);// end of header
    size_t table = b.size();
    b.resize(b.size() + size_t(h) * 8);
    const int order[4] = {3, 2, 1, 0};// A, B, G, R from RGBA
    for (uint32_t y = 0; y < h; ++y) {
        uint64_t off = b.size();
        for (int i = 0; i < 8; ++i) b[table + size_t(y) * 8 + i] = uint8_t(off >> (8 * i));
        put_u32le(b, y);
        put_u32le(b, w * 16u);
        for (int c : order)
            for (uint32_t x = 0; x < w; ++x) put_f32le(b, rgba[(size_t(y) * w + x) * 4 + c]);
    }
    return b;
}

void exr_unzip(const uint8_t *src, size_t n, std::vector<uint8_t> &out, size_t expect, const std::string &path) {
    std::vector<uint8_t> tmp(expect);
    uLongf len = uLongf(expect);
    if (uncompress(tmp.data(), &len, src, uLong(n)) != Z_OK || len != expect)
        throw std::runtime_error("bad ZIP chunk in " + path);
    for (size_t i = 1; i < expect; ++i) tmp[i] = uint8_t(int(tmp[i - 1]) + int(tmp[i]) - 128);
    out.resize(expect);
    size_t half = (expect + 1) / 2;
    for (size_t i = 0; i < expect; ++i) out[i] = (i & 1) ? tmp[half + i / 2] : tmp[i / 2];
}

ImageRGBA decode_exr(const std::vector<uint8_t> &d, const std::string &path) {
    Reader r{d, 0, path};
    if (r.u32le() != kExrMagic) throw std::runtime_error("not an OpenEXR file: " + path);
    uint32_t version = r.u32le();
    if ((version & 0xFF) != 2 || (version & 0x1A00))// tiled, deep or multi-part
        throw std::runtime_error("only single-part scanline OpenEXR files are supported: " + path);
    struct Chan {
        std::string name;
        uint32_t type;
    };
    std::vector<Chan> chans;
    int compression = -1;
    int32_t box[4] = {0, 0, -1, -1};
    int line_order = 0;
    while (true) {
        std::string name = r.str0();
        if (name.empty()) break;
        std::string type = r.str0();
        uint32_t size = r.u32le();
        r.need(size);
        size_t end = r.pos + size;
        if (name == "channels") {
            while (true) {
                std::string cn = r.str0();
                if (cn.empty()) break;
                uint32_t pt = r.u32le();
                r.u32le();
                uint32_t xs = r.u32le(), ys = r.u32le();
                if (xs != 1 || ys != 1) throw std::runtime_error("subsampled EXR channels unsupported: " + path);
                if (pt > 2) throw std::runtime_error("bad EXR pixel type in " + path);
                chans.push_back({cn, pt});
            }
        } else if (name == "compression") {
            if (size < 1) throw std::runtime_error("bad EXR compression attribute in " + path);
            compression = d[r.pos];
        } else if (name == "dataWindow") {
            for (int i = 0; i < 4; ++i) box[i] = r.i32le();
        } else if (name == "lineOrder") {
            if (size < 1) throw std::runtime_error("bad EXR lineOrder attribute in " + path);
            line_order = d[r.pos];
        }
        r.pos = end;
    }
    if (chans.empty()) throw std::runtime_error("EXR without channels: " + path);
    if (compression != 0 && compression != 2 && compression != 3)
        throw std::runtime_error("EXR compression " + std::to_string(compression) +
                                 " unsupported (NONE, ZIPS and ZIP are): " + path);
    (void)line_order;// chunks carry their own y
    int64_t w = int64_t(box[2]) - box[0] + 1, h = int64_t(box[3]) - box[1] + 1;
    check_size(uint64_t(std::max<int64_t>(w, 0)), uint64_t(std::max<int64_t>(h, 0)), path);
    uint32_t lines = compression == 3 ? 16u : 1u;
    uint32_t chunks = uint32_t((h + lines - 1) / lines);
    std::vector<uint64_t> offsets(chunks);
    for (auto &o : offsets) o = r.u64le();
    size_t row_bytes = 0;
    for (const Chan &c : chans) row_bytes += size_t(w) * (c.type == 1 ? 2 : 4);
    ImageRGBA img;
    img.width = uint32_t(w);
    img.height = uint32_t(h);
    img.rgba.assign(size_t(w) * h * 4, 0.f);
    for (size_t i = 0; i < img.rgba.size(); i += 4) img.rgba[i + 3] = 1.f;
    bool has_rgb = false;
    for (const Chan &c : chans) has_rgb |= (c.name == "R" || c.name == "G" || c.name == "B");
    std::vector<uint8_t> buf;
    for (uint32_t k = 0; k < chunks; ++k) {
        if (offsets[k] >= d.size()) throw std::runtime_error("bad EXR chunk offset in " + path);
        Reader cr{d, size_t(offsets[k]), path};
        int32_t y0 = cr.i32le() - box[1];
        uint32_t packed = cr.u32le();
        cr.need(packed);
        if (y0 < 0 || y0 >= h) throw std::runtime_error("bad EXR chunk y in " + path);
        uint32_t nl = uint32_t(std::min<int64_t>(lines, h - y0));
        size_t expect = row_bytes * nl;
        const uint8_t *src = d.data() + cr.pos;
        if (packed == expect) {// stored uncompressed (also allowed for ZIP chunks that do not shrink)
            buf.assign(src, src + expect);
        } else if (compression == 0) {
            throw std::runtime_error("bad EXR chunk size in " + path);
        } else {
            exr_unzip(src, packed, buf, expect, path);
        }
        size_t p = 0;
        for (uint32_t l = 0; l < nl; ++l) {
            size_t row = size_t(y0 + l) * size_t(w);
            for (const Chan &c : chans) {
                int dst = c.name == "R" ? 0 : c.name == "G" ? 1 : c.name == "B" ? 2 : c.name == "A" ? 3
                          : c.name == "Y"                      ? 4
                                                               : -1;
                for (int64_t x = 0; x < w; ++x) {
                    float v;
                    if (c.type == 1) {
                        v = half_to_float(uint16_t(buf[p] | buf[p + 1] << 8));
                        p += 2;
                    } else {
                        uint32_t u = uint32_t(buf[p]) | uint32_t(buf[p + 1]) << 8 | uint32_t(buf[p + 2]) << 16 |
                                     uint32_t(buf[p + 3]) << 24;
                        p += 4;
                        if (c.type == 2) {
                            std::memcpy(&v, &u, 4);
                        } else {
                            v = float(u);
                        }
                    }
                    float *px = &img.rgba[(row + size_t(x)) * 4];
                    if (dst >= 0 && dst < 4) px[dst] = v;
                    if (dst == 4 && !has_rgb) px[0] = px[1] = px[2] = v;
                }
            }
        }
    }
    return img;
}

// ---- Radiance RGBE ----------------------------------------------------------

std::vector<uint8_t> encode_hdr(uint32_t w, uint32_t h, const float *rgba) {
    std::string head = "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y " + std::to_string(h) + " +X " +
                       std::to_string(w) + "\n";
    std::vector<uint8_t> b(head.begin(), head.end());
    b.reserve(b.size() + size_t(w) * h * 4);
    for (size_t i = 0; i < size_t(w) * h; ++i) {
        float r = rgba[i * 4], g = rgba[i * 4 + 1], bl = rgba[i * 4 + 2];
        float v = std::max(r, std::max(g, bl));
        if (!(v >= 1e-32f)) {// also catches NaN
            b.insert(b.end(), {0, 0, 0, 0});
            continue;
        }
        int e;
        float m = std::frexp(v, &e) * 256.f / v;
        b.push_back(uint8_t(std::max(r, 0.f) * m));
        b.push_back(uint8_t(std::max(g, 0.f) * m));
        b.push_back(uint8_t(std::max(bl, 0.f) * m));
        b.push_back(uint8_t(e + 128));
    }
    return b;
}

ImageRGBA decode_hdr(const std::vector<uint8_t> &d, const std::string &path) {
    size_t pos = 0;
    auto line = [&]() {
        std::string s;
        while (pos < d.size() && d[pos] != '\n') s.push_back(char(d[pos++]));
        if (pos >= d.size()) throw std::runtime_error("truncated HDR header in " + path);
        ++pos;
        return s;
    };
    std::string first = line();
    if (first.rfind("#?", 0) != 0) throw std::runtime_error("not a Radiance HDR file: " + path);
    while (true) {
        std::string s = line();
        if (s.empty()) break;
        if (s.rfind("FORMAT=", 0) == 0 && s != "FORMAT=32-bit_rle_rgbe")
            throw std::runtime_error("unsupported HDR format " + s + " in " + path);
    }
    std::istringstream res(line());
    std::string ya, xa;
    long hh = 0, ww = 0;
    res >> ya >> hh >> xa >> ww;
    if (ya != "-Y" || xa != "+X") throw std::runtime_error("unsupported HDR orientation in " + path);
    check_size(uint64_t(std::max(ww, 0L)), uint64_t(std::max(hh, 0L)), path);
    uint32_t w = uint32_t(ww), h = uint32_t(hh);
    ImageRGBA img;
    img.width = w;
    img.height = h;
    img.rgba.resize(size_t(w) * h * 4);
    std::vector<uint8_t> scan(size_t(w) * 4);
    auto need = [&](size_t n) {
        if (pos + n > d.size()) throw std::runtime_error("truncated HDR data in " + path);
    };
    for (uint32_t y = 0; y < h; ++y) {
        need(4);
        bool rle = w >= 8 && w < 32768 && d[pos] == 2 && d[pos + 1] == 2 && ((d[pos + 2] << 8) | d[pos + 3]) == int(w);
        if (rle) {
            pos += 4;
            for (int c = 0; c < 4; ++c) {
                uint32_t x = 0;
                while (x < w) {
                    need(1);
                    uint32_t n = d[pos++];
                    if (n > 128) {
                        n -= 128;
                        need(1);
                        if (x + n > w) throw std::runtime_error("bad HDR run in " + path);
                        uint8_t v = d[pos++];
                        for (uint32_t k = 0; k < n; ++k) scan[size_t(x++) * 4 + c] = v;
                    } else {
                        if (n == 0 || x + n > w) throw std::runtime_error("bad HDR run in " + path);
                        need(n);
                        for (uint32_t k = 0; k < n; ++k) scan[size_t(x++) * 4 + c] = d[pos++];
                    }
                }
            }
        } else {
            need(size_t(w) * 4);
            std::memcpy(scan.data(), d.data() + pos, size_t(w) * 4);
            pos += size_t(w) * 4;
        }
        for (uint32_t x = 0; x < w; ++x) {
            const uint8_t *e = &scan[size_t(x) * 4];
            float *px = &img.rgba[(size_t(y) * w + x) * 4];
            float f = e[3] ? std::ldexp(1.f, int(e[3]) - (128 + 8)) : 0.f;
            px[0] = e[0] * f;
            px[1] = e[1] * f;
            px[2] = e[2] * f;
            px[3] = 1.f;
        }
    }
    return img;
}

// ---- PNG ---------------------------------------------------------------------

void png_chunk(std::vector<uint8_t> &b, const char *type, const std::vector<uint8_t> &data) {
    put_u32be(b, uint32_t(data.size()));
    size_t start = b.size();
    b.insert(b.end(), type, type + 4);
    b.insert(b.end(), data.begin(), data.end());
    uLong crc = crc32(0L, b.data() + start, uInt(b.size() - start));
    put_u32be(b, uint32_t(crc));
}

const uint8_t kPngSig[8] = {137, 80, 78, 71, 13, 10, 26, 10};

uint8_t to_u8(float x) {
    if (!(x > 0.f)) return 0;// NaN and negatives
    if (x >= 1.f) return 255;
    return uint8_t(x * 255.f + 0.5f);
}

std::vector<uint8_t> encode_png(uint32_t w, uint32_t h, const float *rgba) {
    std::vector<uint8_t> b(kPngSig, kPngSig + 8);
    std::vector<uint8_t> ihdr;
    put_u32be(ihdr, w);
    put_u32be(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});// 8-bit RGBA, deflate, adaptive filter set, no interlace
    png_chunk(b, "IHDR", ihdr);
    std::vector<uint8_t> raw;
    raw.reserve(size_t(h) * (1 + size_t(w) * 4));
    for (uint32_t y = 0; y < h; ++y) {
        raw.push_back(0);// filter: none
        for (size_t i = 0; i < size_t(w) * 4; ++i) raw.push_back(to_u8(rgba[size_t(y) * w * 4 + i]));
    }
    uLongf len = compressBound(uLong(raw.size()));
    std::vector<uint8_t> z(len);
    if (compress2(z.data(), &len, raw.data(), uLong(raw.size()), 6) != Z_OK)
        throw std::runtime_error("zlib compression failed");
    z.resize(len);
    png_chunk(b, "IDAT", z);
    png_chunk(b, "IEND", {});
    return b;
}

ImageRGBA decode_png(const std::vector<uint8_t> &d, const std::string &path) {
    if (d.size() < 8 || std::memcmp(d.data(), kPngSig, 8) != 0) throw std::runtime_error("not a PNG file: " + path);
    Reader r{d, 8, path};
    uint32_t w = 0, h = 0, depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> z, palette;
    bool header = false;
    while (true) {
        uint32_t len = r.u32be();
        r.need(4 + size_t(len) + 4);
        std::string type(reinterpret_cast<const char *>(d.data() + r.pos), 4);
        const uint8_t *data = d.data() + r.pos + 4;
        if (uint32_t(crc32(0L, d.data() + r.pos, uInt(4 + len))) !=
            (uint32_t(data[len]) << 24 | uint32_t(data[len + 1]) << 16 | uint32_t(data[len + 2]) << 8 | data[len + 3]))
            throw std::runtime_error("PNG CRC mismatch in " + path);
        if (type == "IHDR") {
            if (len != 13) throw std::runtime_error("bad IHDR in " + path);
            w = uint32_t(data[0]) << 24 | data[1] << 16 | data[2] << 8 | data[3];
            h = uint32_t(data[4]) << 24 | data[5] << 16 | data[6] << 8 | data[7];
            depth = data[8];
            ctype = data[9];
            interlace = data[12];
            header = true;
        } else if (type == "PLTE") {
            palette.assign(data, data + len);
        } else if (type == "IDAT") {
            z.insert(z.end(), data, data + len);
        } else if (type == "IEND") {
            break;
        }
        r.pos += 4 + size_t(len) + 4;
    }
    if (!header) throw std::runtime_error("PNG without IHDR: " + path);
    check_size(w, h, path);
    if (interlace) throw std::runtime_error("interlaced PNG unsupported: " + path);
    int channels = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    if (!channels) throw std::runtime_error("bad PNG colour type in " + path);
    if (!(depth == 8 || (depth == 16 && ctype != 3))) throw std::runtime_error("PNG bit depth must be 8 or 16: " + path);
    size_t bpp = size_t(channels) * depth / 8, stride = size_t(w) * bpp;
    std::vector<uint8_t> raw(size_t(h) * (stride + 1));
    uLongf len = uLongf(raw.size());
    if (uncompress(raw.data(), &len, z.data(), uLong(z.size())) != Z_OK || len != raw.size())
        throw std::runtime_error("bad PNG data in " + path);
    std::vector<uint8_t> px(size_t(h) * stride);
    for (uint32_t y = 0; y < h; ++y) {
        uint8_t f = raw[size_t(y) * (stride + 1)];
        const uint8_t *s = &raw[size_t(y) * (stride + 1) + 1];
        uint8_t *o = &px[size_t(y) * stride];
        const uint8_t *up = y ? &px[size_t(y - 1) * stride] : nullptr;
        for (size_t i = 0; i < stride; ++i) {
            int a = i >= bpp ? o[i - bpp] : 0, b = up ? up[i] : 0, c = (up && i >= bpp) ? up[i - bpp] : 0;
            int v;
            switch (f) {
            case 0: v = s[i]; break;
            case 1: v = s[i] + a; break;
            case 2: v = s[i] + b; break;
            case 3: v = s[i] + ((a + b) >> 1); break;
            case 4: {
                int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
                v = s[i] + ((pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c));
                break;
            }
            default: throw std::runtime_error("bad PNG filter in " + path);
            }
            o[i] = uint8_t(v);
        }
    }
    ImageRGBA img;
    img.width = w;
    img.height = h;
    img.rgba.resize(size_t(w) * h * 4);
    float maxv = depth == 16 ? 65535.f : 255.f;
    for (size_t i = 0; i < size_t(w) * h; ++i) {
        float ch[4];
        for (int c = 0; c < channels; ++c)
            ch[c] = depth == 16 ? float(px[i * bpp + c * 2] << 8 | px[i * bpp + c * 2 + 1]) : float(px[i * bpp + c]);
        float *o = &img.rgba[i * 4];
        if (ctype == 3) {
            size_t k = size_t(ch[0]) * 3;
            if (k + 2 >= palette.size()) throw std::runtime_error("PNG palette index out of range in " + path);
            o[0] = palette[k] / 255.f;
            o[1] = palette[k + 1] / 255.f;
            o[2] = palette[k + 2] / 255.f;
            o[3] = 1.f;
        } else if (channels <= 2) {
            o[0] = o[1] = o[2] = ch[0] / maxv;
            o[3] = channels == 2 ? ch[1] / maxv : 1.f;
        } else {
            o[0] = ch[0] / maxv;
            o[1] = ch[1] / maxv;
            o[2] = ch[2] / maxv;
            o[3] = channels == 4 ? ch[3] / maxv : 1.f;
        }
    }
    return img;
}

// ---- PFM -----------------------------------------------------------------------

std::vector<uint8_t> encode_pfm(uint32_t w, uint32_t h, const float *rgba) {
    std::string head = "PF\n" + std::to_string(w) + " " + std::to_string(h) + "\n-1.0\n";
    std::vector<uint8_t> b(head.begin(), head.end());
    for (uint32_t y = 0; y < h; ++y) {
        uint32_t row = h - 1 - y;// PFM rows run bottom to top
        for (uint32_t x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) put_f32le(b, rgba[(size_t(row) * w + x) * 4 + c]);
    }
    return b;
}

ImageRGBA decode_pfm(const std::vector<uint8_t> &d, const std::string &path) {
    size_t pos = 0;
    auto token = [&]() {
        while (pos < d.size() && std::isspace(d[pos])) ++pos;
        std::string s;
        while (pos < d.size() && !std::isspace(d[pos])) s.push_back(char(d[pos++]));
        return s;
    };
    std::string magic = token();
    if (magic != "PF" && magic != "Pf") throw std::runtime_error("not a PFM file: " + path);
    long iw = std::atol(token().c_str()), ih = std::atol(token().c_str());
    float sc = float(std::atof(token().c_str()));
    ++pos;// the single whitespace byte after the scale
    check_size(uint64_t(std::max(iw, 0L)), uint64_t(std::max(ih, 0L)), path);
    uint32_t w = uint32_t(iw), h = uint32_t(ih);
    int ch = magic == "PF" ? 3 : 1;
    size_t n = size_t(w) * h * ch;
    if (pos + n * 4 > d.size()) throw std::runtime_error("truncated PFM " + path);
    ImageRGBA img;
    img.width = w;
    img.height = h;
    img.rgba.resize(size_t(w) * h * 4);
    for (uint32_t y = 0; y < h; ++y) {
        uint32_t src_row = h - 1 - y;
        for (uint32_t x = 0; x < w; ++x) {
            float *o = &img.rgba[(size_t(y) * w + x) * 4];
            for (int c = 0; c < 3; ++c) {
                const uint8_t *p = &d[pos + ((size_t(src_row) * w + x) * ch + (ch == 3 ? c : 0)) * 4];
                uint32_t u = sc > 0.f ? (uint32_t(p[0]) << 24 | uint32_t(p[1]) << 16 | uint32_t(p[2]) << 8 | p[3])
                                      : (uint32_t(p[3]) << 24 | uint32_t(p[2]) << 16 | uint32_t(p[1]) << 8 | p[0]);
                std::memcpy(&o[c], &u, 4);
            }
            o[3] = 1.f;
        }
    }
    return img;
}

}// namespace

std::string image_extension(const std::string &path) {
    size_t dot = path.find_last_of('.');
    size_t slash = path.find_last_of('/');
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return "";
    std::string e = path.substr(dot + 1);
    for (char &c : e) c = char(std::tolower(static_cast<unsigned char>(c)));
    return e;
}

void write_image(const std::string &path, uint32_t width, uint32_t height, const float *rgba) {
    check_size(width, height, path);
    std::string e = image_extension(path);
    if (e == "exr") {
        write_file(path, encode_exr(width, height, rgba));
    } else if (e == "hdr") {
        write_file(path, encode_hdr(width, height, rgba));
    } else if (e == "png") {
        write_file(path, encode_png(width, height, rgba));
    } else if (e == "pfm") {
        write_file(path, encode_pfm(width, height, rgba));
    } else {
        throw std::runtime_error("unsupported image format ." + e + " (exr, hdr, png, pfm): " + path);
    }
}

ImageRGBA read_image(const std::string &path) {
    std::string e = image_extension(path);
    const bool jpeg = e == "jpg" || e == "jpeg";
    if (e != "exr" && e != "hdr" && e != "png" && e != "pfm" && !jpeg)
        throw std::runtime_error("unsupported image format ." + e + " (exr, hdr, png, pfm, jpg): " + path);
    std::vector<uint8_t> d = read_file(path);
    if (jpeg) return decode_jpeg(d, path);
    if (e == "exr") return decode_exr(d, path);
    if (e == "hdr") return decode_hdr(d, path);
    if (e == "png") return decode_png(d, path);
    return decode_pfm(d, path);
}

}// namespace vs
