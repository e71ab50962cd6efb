// JPEG reader (ITU-T T.81 with the JFIF 1.02 colour convention): baseline
// and extended sequential (SOF0 / SOF1) and progressive (SOF2) frames,
// Huffman entropy coding, 8-bit samples, one component (grey) or three
// (YCbCr, or RGB when an Adobe APP14 marker says transform 0), sampling
// factors 1-4, restart intervals.  Arithmetic coding, lossless,
// hierarchical, 12-bit and four-component (CMYK) files are rejected.
//
// Every scan's coefficients go to per-component arrays first (a progressive
// file refines them scan by scan, T.81 Annex G); when the file ends each
// block is dequantised and inverse-transformed by the separable 8x8 IDCT
// (T.81 A.3.3) in double precision, level-shifted by 128, rounded and
// clamped.  Chroma sampled at half the luma width (and height) is
// reconstructed by the triangle ("fancy") filter of the IJG library — each
// output sample 3/4 of its nearer and 1/4 of its farther input sample per
// halved axis, rounding offsets 8 and 7 alternating — other factors, and
// planes at most two samples wide, by replication; YCbCr converts to RGB by
// the JFIF equations.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "image_io.h"

namespace vs {

namespace {

// natural (row-major) index of the k-th coefficient in zig-zag order
const uint8_t kNatural[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                              12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                              35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                              58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huffman {
    bool present{false};
    int32_t maxcode[18]{};// largest code of each length, -1 if none
    int32_t valptr[17]{};
    int32_t mincode[17]{};
    uint8_t val[256]{};
};

struct Component {
    int id{0}, h{1}, v{1}, tq{0};
    int td{0}, ta{0};
    int bw{0}, bh{0};        // blocks stored (MCU-padded)
    int cw{0}, ch{0};        // samples covering the image: ceil(W * h / hmax)
    std::vector<int16_t> coef;// bw * bh * 64, natural order, not dequantised
    int dc_pred{0};
};

struct Jpeg {
    const std::vector<uint8_t> &d;
    const std::string &path;
    size_t pos{0};
    // bit reader
    uint32_t buf{0};
    int cnt{0};
    bool at_marker{false};

    uint16_t qt[4][64]{};// natural order
    bool qt_present[4]{};
    Huffman dc[4], ac[4];
    std::vector<Component> comp;
    int width{0}, height{0}, hmax{1}, vmax{1}, mcux{0}, mcuy{0};
    bool progressive{false}, frame{false};
    int restart{0};
    int adobe_transform{-1};
    int eobrun{0};

    Jpeg(const std::vector<uint8_t> &data, const std::string &p) : d(data), path(p) {}

    [[noreturn]] void fail(const std::string &what) const {
        throw std::runtime_error("JPEG " + what + ": " + path);
    }
    uint8_t byte() {
        if (pos >= d.size()) fail("truncated");
        return d[pos++];
    }
    int u16() {
        int a = byte();
        return (a << 8) | byte();
    }

    // --- entropy-coded segment bits (F.2.2.5): 0xFF 0x00 is a data 0xFF; a
    // marker ends the data and the reader feeds zeros until the restart
    void fill() {
        while (cnt <= 24) {
            uint32_t b = 0;
            if (!at_marker && pos < d.size()) {
                b = d[pos];
                if (b == 0xFF) {
                    uint8_t nx = pos + 1 < d.size() ? d[pos + 1] : 0xD9;
                    if (nx == 0x00) {
                        pos += 2;
                    } else {
                        at_marker = true;
                        b = 0;
                    }
                } else {
                    ++pos;
                }
            }
            buf |= b << (24 - cnt);
            cnt += 8;
        }
    }
    int bits(int k) {// k in 0..16
        if (k == 0) return 0;
        fill();
        int v = int(buf >> (32 - k));
        buf <<= k;
        cnt -= k;
        return v;
    }
    int bit() { return bits(1); }
    static int extend(int v, int t) {// F.2.2.1
        return t == 0 ? 0 : (v < (1 << (t - 1)) ? v - (1 << t) + 1 : v);
    }
    int decode(const Huffman &h) {// F.2.2.3
        if (!h.present) fail("scan uses an undefined Huffman table");
        fill();
        for (int l = 1; l <= 16; ++l) {
            int32_t c = int32_t(buf >> (32 - l));
            if (c <= h.maxcode[l]) {
                buf <<= l;
                cnt -= l;
                return h.val[h.valptr[l] + c - h.mincode[l]];
            }
        }
        fail("bad Huffman code");
    }
    void reset_bits() {
        buf = 0;
        cnt = 0;
        at_marker = false;
    }

    // --- marker segments
    void read_dqt() {
        int len = u16() - 2;
        while (len > 0) {
            int pq_tq = byte();
            int pq = pq_tq >> 4, tq = pq_tq & 15;
            if (tq > 3) fail("bad quantisation table id");
            for (int k = 0; k < 64; ++k) qt[tq][kNatural[k]] = uint16_t(pq ? u16() : byte());
            qt_present[tq] = true;
            len -= 1 + 64 * (pq ? 2 : 1);
        }
        if (len != 0) fail("bad DQT length");
    }
    void read_dht() {
        int len = u16() - 2;
        while (len > 0) {
            int tc_th = byte();
            int tc = tc_th >> 4, th = tc_th & 15;
            if (tc > 1 || th > 3) fail("bad Huffman table id");
            Huffman &h = tc ? ac[th] : dc[th];
            uint8_t counts[17] = {};
            int total = 0;
            for (int l = 1; l <= 16; ++l) total += counts[l] = byte();
            if (total > 256) fail("bad Huffman table");
            for (int i = 0; i < total; ++i) h.val[i] = byte();
            // canonical codes (C.2): lengths in order, consecutive codes per length
            int code = 0, k = 0;
            for (int l = 1; l <= 16; ++l) {
                h.valptr[l] = k;
                h.mincode[l] = code;
                code += counts[l];
                k += counts[l];
                h.maxcode[l] = counts[l] ? code - 1 : -1;
                code <<= 1;
            }
            h.maxcode[17] = 0x7FFFFFFF;
            h.present = true;
            len -= 17 + total;
        }
        if (len != 0) fail("bad DHT length");
    }
    void read_sof(int marker) {
        if (frame) fail("more than one frame");
        frame = true;
        progressive = marker == 0xC2;
        u16();
        if (byte() != 8) fail("sample precision other than 8 bits unsupported");
        height = u16();
        width = u16();
        int nf = byte();
        if (height == 0) fail("DNL-defined height unsupported");
        if (width == 0) fail("zero width");
        if (nf != 1 && nf != 3) fail("component count " + std::to_string(nf) + " unsupported (1 or 3)");
        comp.resize(nf);
        for (auto &c : comp) {
            c.id = byte();
            int hv = byte();
            c.h = hv >> 4;
            c.v = hv & 15;
            c.tq = byte();
            if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3) fail("bad component parameters");
            hmax = std::max(hmax, c.h);
            vmax = std::max(vmax, c.v);
        }
        if (uint64_t(width) * height > (uint64_t(1) << 28)) fail("image too large");
        mcux = (width + 8 * hmax - 1) / (8 * hmax);
        mcuy = (height + 8 * vmax - 1) / (8 * vmax);
        for (auto &c : comp) {
            if (hmax % c.h || vmax % c.v) fail("sampling factors that do not divide the largest unsupported");
            c.bw = mcux * c.h;
            c.bh = mcuy * c.v;
            c.cw = (width * c.h + hmax - 1) / hmax;
            c.ch = (height * c.v + vmax - 1) / vmax;
            c.coef.assign(size_t(c.bw) * c.bh * 64, 0);
        }
    }

    // --- block decoders (F.2.2 sequential, G.1.2 progressive)
    void block_sequential(Component &c, int16_t *b) {
        int t = decode(dc[c.td]);
        if (t > 11) fail("bad DC magnitude");
        c.dc_pred += extend(bits(t), t);
        b[0] = int16_t(c.dc_pred);
        for (int k = 1; k < 64;) {
            int rs = decode(ac[c.ta]);
            int r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r != 15) break;
                k += 16;
                continue;
            }
            k += r;
            if (k > 63) fail("AC coefficient index out of range");
            b[kNatural[k]] = int16_t(extend(bits(s), s));
            ++k;
        }
    }
    void block_dc_first(Component &c, int16_t *b, int al) {
        int t = decode(dc[c.td]);
        if (t > 11) fail("bad DC magnitude");
        c.dc_pred += extend(bits(t), t);
        b[0] = int16_t(c.dc_pred * (1 << al));
    }
    void block_dc_refine(int16_t *b, int al) {
        if (bit()) b[0] = int16_t(b[0] | (1 << al));
    }
    void block_ac_first(Component &c, int16_t *b, int ss, int se, int al) {
        if (eobrun > 0) {
            --eobrun;
            return;
        }
        for (int k = ss; k <= se;) {
            int rs = decode(ac[c.ta]);
            int r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r < 15) {// EOBr: this block and (2^r - 1 + bits) more end here
                    eobrun = (1 << r) - 1 + bits(r);
                    break;
                }
                k += 16;
                continue;
            }
            k += r;
            if (k > se) fail("AC coefficient index out of range");
            b[kNatural[k]] = int16_t(extend(bits(s), s) * (1 << al));
            ++k;
        }
    }
    void refine_nonzero(int16_t &v, int p1) {
        if (bit() && (v & p1) == 0) v = int16_t(v >= 0 ? v + p1 : v - p1);
    }
    void block_ac_refine(Component &c, int16_t *b, int ss, int se, int al) {
        const int p1 = 1 << al;
        int k = ss;
        if (eobrun == 0) {
            for (; k <= se; ++k) {
                int rs = decode(ac[c.ta]);
                int r = rs >> 4, s = rs & 15;
                int val = 0;
                if (s != 0) {
                    if (s != 1) fail("bad refinement magnitude");
                    val = bit() ? p1 : -p1;
                } else if (r != 15) {
                    eobrun = (1 << r) + bits(r);
                    break;
                }
                // skip r zero-history coefficients (correcting the nonzero
                // ones passed on the way); the next zero one takes val
                for (; k <= se; ++k) {
                    int16_t &v = b[kNatural[k]];
                    if (v != 0) {
                        refine_nonzero(v, p1);
                    } else {
                        if (r == 0) break;
                        --r;
                    }
                }
                if (val != 0) {
                    if (k > se) fail("AC refinement index out of range");
                    b[kNatural[k]] = int16_t(val);
                }
            }
        }
        if (eobrun > 0) {
            for (; k <= se; ++k) {
                int16_t &v = b[kNatural[k]];
                if (v != 0) refine_nonzero(v, p1);
            }
            --eobrun;
        }
    }

    void read_sos() {
        if (!frame) fail("scan before frame header");
        int len = u16();
        int ns = byte();
        if (ns < 1 || ns > 4 || len != 6 + 2 * ns) fail("bad scan header");
        std::vector<Component *> sc;
        for (int i = 0; i < ns; ++i) {
            int id = byte(), t = byte();
            Component *c = nullptr;
            for (auto &x : comp)
                if (x.id == id) c = &x;
            if (!c) fail("scan names an unknown component");
            c->td = t >> 4;
            c->ta = t & 15;
            if (c->td > 3 || c->ta > 3) fail("bad Huffman table selector");
            sc.push_back(c);
        }
        int ss = byte(), se = byte(), a = byte();
        int ah = a >> 4, al = a & 15;
        if (!progressive) {
            ss = 0;
            se = 63;
            ah = al = 0;
        } else {
            if (se > 63 || ss > se || al > 13 || (ss == 0 && se != 0) || (ss > 0 && ns != 1))
                fail("bad progressive scan parameters");
        }
        for (auto *c : sc) c->dc_pred = 0;
        eobrun = 0;
        reset_bits();

        auto block = [&](Component &c, int bx, int by) {
            int16_t *b = &c.coef[(size_t(by) * c.bw + bx) * 64];
            if (!progressive) {
                block_sequential(c, b);
            } else if (ss == 0) {
                if (ah == 0)
                    block_dc_first(c, b, al);
                else
                    block_dc_refine(b, al);
            } else if (ah == 0) {
                block_ac_first(c, b, ss, se, al);
            } else {
                block_ac_refine(c, b, ss, se, al);
            }
        };
        // a scan of one component covers only the blocks holding image
        // samples, one block per MCU (A.2.2); an interleaved scan covers
        // the MCU grid, h x v blocks of each component per MCU (A.2.3)
        int units_x, units_y;
        if (ns == 1) {
            units_x = (sc[0]->cw + 7) / 8;
            units_y = (sc[0]->ch + 7) / 8;
        } else {
            units_x = mcux;
            units_y = mcuy;
        }
        const long total = long(units_x) * units_y;
        for (long u = 0; u < total; ++u) {
            if (restart > 0 && u > 0 && u % restart == 0) next_restart(sc);
            int ux = int(u % units_x), uy = int(u / units_x);
            if (ns == 1) {
                block(*sc[0], ux, uy);
            } else {
                for (auto *c : sc)
                    for (int y = 0; y < c->v; ++y)
                        for (int x = 0; x < c->h; ++x) block(*c, ux * c->h + x, uy * c->v + y);
            }
        }
        // leave pos at the next marker
        reset_bits();
        while (pos + 1 < d.size() && !(d[pos] == 0xFF && d[pos + 1] != 0x00 && !(d[pos + 1] >= 0xD0 && d[pos + 1] <= 0xD7)))
            ++pos;
    }
    void next_restart(std::vector<Component *> &sc) {
        reset_bits();
        // the bit reader stops at the marker; skip fill bytes up to RSTn
        while (pos + 1 < d.size() && !(d[pos] == 0xFF && d[pos + 1] >= 0xD0 && d[pos + 1] <= 0xD7)) ++pos;
        if (pos + 1 >= d.size()) fail("missing restart marker");
        pos += 2;
        for (auto *c : sc) c->dc_pred = 0;
        eobrun = 0;
    }

    void parse() {
        if (d.size() < 4 || d[0] != 0xFF || d[1] != 0xD8) fail("missing SOI");
        pos = 2;
        for (;;) {
            if (pos >= d.size()) fail("missing EOI");
            if (d[pos] != 0xFF) {
                ++pos;// tolerate stray bytes between segments
                continue;
            }
            int m = d[pos + 1 < d.size() ? pos + 1 : pos];
            pos += 2;
            if (m == 0xFF) {// fill byte
                --pos;
                continue;
            }
            if (m == 0xD9) break;
            if (m == 0x01 || (m >= 0xD0 && m <= 0xD7)) continue;
            switch (m) {
                case 0xC0:
                case 0xC1:
                case 0xC2:
                    read_sof(m);
                    break;
                case 0xC3:
                case 0xC5:
                case 0xC6:
                case 0xC7:
                case 0xC9:
                case 0xCA:
                case 0xCB:
                case 0xCD:
                case 0xCE:
                case 0xCF:
                    fail("lossless, hierarchical or arithmetic-coded frames unsupported");
                case 0xC4:
                    read_dht();
                    break;
                case 0xDB:
                    read_dqt();
                    break;
                case 0xDD:
                    u16();
                    restart = u16();
                    break;
                case 0xDA:
                    read_sos();
                    break;
                case 0xEE: {// Adobe APP14: transform flag
                    size_t start = pos;
                    int len = u16();
                    if (len >= 14 && pos + 12 <= d.size() && std::memcmp(&d[pos], "Adobe", 5) == 0)
                        adobe_transform = d[pos + 11];
                    pos = start + len;
                    break;
                }
                default:
                    pos += size_t(u16()) - 2;
            }
        }
        if (!frame) fail("no frame");
    }

    // --- reconstruction
    std::vector<uint8_t> plane(const Component &c) const {
        if (!qt_present[c.tq]) fail("component uses an undefined quantisation table");
        struct Basis {
            double c[8][8];// c[x][u] = C(u) cos((2x+1) u pi / 16) / 2
            Basis() {
                for (int x = 0; x < 8; ++x)
                    for (int u = 0; u < 8; ++u)
                        c[x][u] = (u == 0 ? std::sqrt(0.5) : 1.0) * std::cos((2 * x + 1) * u * M_PI / 16) / 2;
            }
        };
        static const Basis basis;
        const auto &cs = basis.c;
        const int pw = c.bw * 8;
        std::vector<uint8_t> out(size_t(pw) * c.bh * 8);
        const uint16_t *q = qt[c.tq];
        for (int by = 0; by < c.bh; ++by)
            for (int bx = 0; bx < c.bw; ++bx) {
                const int16_t *b = &c.coef[(size_t(by) * c.bw + bx) * 64];
                double f[64], t[64];
                for (int i = 0; i < 64; ++i) f[i] = double(b[i]) * q[i];
                // rows: t[v][x] = sum_u cs[x][u] f[v][u]
                for (int v = 0; v < 8; ++v)
                    for (int x = 0; x < 8; ++x) {
                        double s = 0;
                        for (int u = 0; u < 8; ++u) s += cs[x][u] * f[v * 8 + u];
                        t[v * 8 + x] = s;
                    }
                for (int y = 0; y < 8; ++y)
                    for (int x = 0; x < 8; ++x) {
                        double s = 0;
                        for (int v = 0; v < 8; ++v) s += cs[y][v] * t[v * 8 + x];
                        long r = std::lround(s + 128.0);
                        out[size_t(by * 8 + y) * pw + bx * 8 + x] = uint8_t(std::clamp(r, 0L, 255L));
                    }
            }
        return out;
    }

    // the component's plane at full resolution, width x height
    std::vector<uint8_t> upsampled(const Component &c) const {
        std::vector<uint8_t> p = plane(c);
        const int pw = c.bw * 8;
        const int fx = hmax / c.h, fy = vmax / c.v;
        std::vector<uint8_t> out(size_t(width) * height);
        if (fx == 1 && fy == 1) {
            for (int y = 0; y < height; ++y) std::memcpy(&out[size_t(y) * width], &p[size_t(y) * pw], width);
            return out;
        }
        // planes at most two samples wide are replicated, not filtered
        const bool filter = c.cw > 2;
        if (filter && fx == 2 && fy == 1) {
            // triangle filter along rows: 3/4 nearer + 1/4 farther sample,
            // offsets 1 (left neighbour) and 2 (right); edge samples copy
            const int cw = c.cw;
            for (int y = 0; y < height; ++y) {
                const uint8_t *r = &p[size_t(y) * pw];
                uint8_t *o = &out[size_t(y) * width];
                for (int x = 0; x < width; ++x) {
                    int sx = x >> 1;
                    if ((x & 1) == 0)
                        o[x] = sx == 0 ? r[0] : uint8_t((3 * r[sx] + r[sx - 1] + 1) >> 2);
                    else
                        o[x] = sx >= cw - 1 ? r[cw - 1] : uint8_t((3 * r[sx] + r[sx + 1] + 2) >> 2);
                }
            }
            return out;
        }
        if (filter && fx == 2 && fy == 2) {
            // triangle filter on both axes over the samples that cover the
            // image (cw x ch): column sums 3 * nearer row + farther row, then
            // 3 * nearer + farther column sum, offsets 8 (left) and 7 (right)
            const int cw = c.cw, ch = c.ch;
            std::vector<int> colsum(cw);
            for (int y = 0; y < height; ++y) {
                int sy = y >> 1;
                int near = std::min(sy, ch - 1);
                int far = (y & 1) ? std::min(sy + 1, ch - 1) : std::max(sy - 1, 0);
                for (int x = 0; x < cw; ++x) colsum[x] = 3 * p[size_t(near) * pw + x] + p[size_t(far) * pw + x];
                uint8_t *o = &out[size_t(y) * width];
                for (int x = 0; x < width; ++x) {
                    int sx = x >> 1;
                    int v;
                    if ((x & 1) == 0) {
                        v = sx == 0 ? (colsum[0] * 4 + 8) >> 4 : (colsum[sx] * 3 + colsum[sx - 1] + 8) >> 4;
                    } else {
                        v = sx >= cw - 1 ? (colsum[cw - 1] * 4 + 7) >> 4 : (colsum[sx] * 3 + colsum[sx + 1] + 7) >> 4;
                    }
                    o[x] = uint8_t(v);
                }
            }
            return out;
        }
        for (int y = 0; y < height; ++y)
            for (int x = 0; x < width; ++x) out[size_t(y) * width + x] = p[size_t(y / fy) * pw + x / fx];
        return out;
    }

    ImageRGBA image() const {
        ImageRGBA img;
        img.width = uint32_t(width);
        img.height = uint32_t(height);
        img.rgba.resize(size_t(width) * height * 4);
        const size_t n = size_t(width) * height;
        if (comp.size() == 1) {
            std::vector<uint8_t> y = upsampled(comp[0]);
            for (size_t i = 0; i < n; ++i) {
                float v = y[i] / 255.f;
                img.rgba[i * 4 + 0] = img.rgba[i * 4 + 1] = img.rgba[i * 4 + 2] = v;
                img.rgba[i * 4 + 3] = 1.f;
            }
            return img;
        }
        std::vector<uint8_t> a = upsampled(comp[0]), b = upsampled(comp[1]), c = upsampled(comp[2]);
        const bool rgb = adobe_transform == 0 || (comp[0].id == 'R' && comp[1].id == 'G' && comp[2].id == 'B');
        for (size_t i = 0; i < n; ++i) {
            int r, g, bl;
            if (rgb) {
                r = a[i];
                g = b[i];
                bl = c[i];
            } else {
                const double y = a[i], cb = b[i] - 128.0, cr = c[i] - 128.0;
                r = int(std::floor(y + 1.402 * cr + 0.5));
                g = int(std::floor(y - 0.344136 * cb - 0.714136 * cr + 0.5));
                bl = int(std::floor(y + 1.772 * cb + 0.5));
            }
            img.rgba[i * 4 + 0] = std::clamp(r, 0, 255) / 255.f;
            img.rgba[i * 4 + 1] = std::clamp(g, 0, 255) / 255.f;
            img.rgba[i * 4 + 2] = std::clamp(bl, 0, 255) / 255.f;
            img.rgba[i * 4 + 3] = 1.f;
        }
        return img;
    }
};

}// namespace

ImageRGBA decode_jpeg(const std::vector<uint8_t> &data, const std::string &path) {
    Jpeg j(data, path);
    j.parse();
    return j.image();
}

}// namespace vs
