// CPU oracle, SVGF denoiser and convergence inspector (TEST INFRASTRUCTURE
// ONLY: the tests, the smoke check and the bench's CPU baseline call it as a
// checker; the product path never does).
//
// Restates src/render_core/denoisers/SVGF/{modulator,reproject,
// filter_moments,atrous,svgf}.cpp, the G-buffer gradient of
// src/base/frame_buffer.cpp:124-168 and the VarianceStats update of
// src/render_core/integrators/adaptive/inspector.h:17-33.
//
// Buffers (float32, one record of 8 per pixel, row-major pixels):
//   G-buffer  : nx ny nz fwidth | p_film.x p_film.y depth_gradient linear_depth
//               (PixelGeometry, frame_buffer.h:15-20; a miss keeps depth -1)
//   SVGF data : illum.rgb variance | history moment1 moment2 0
//               (SVGFData, SVGF/utils.h:22-26, with SVGF::history folded in)
//
// The reference runs its passes in place (reproject writes the history it
// reads at a neighbouring pixel, filter_moments and every a-trous pass
// write the records their neighbours read).  That is a data race whose
// result depends on thread order; this restatement (and the HIP kernels)
// take the race-free reading: every pass reads only what the previous pass
// or the previous frame wrote.
//
// Helpers of the absent ocarina submodule are taken at their textbook
// meaning: in_screen(p, res) = 0 <= p < res per axis, luminance = the
// Rec.709 Y weights, lerp(t, a, b) = a + t (b - a), fract(x) = x - floor(x),
// make_int2 of a float truncates toward zero.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../vision_amd/csrc/elementary.h"

namespace {

struct Geo {
    float n[3], fw, pf[2], dg, depth;
};
struct Rec {
    float il[3], var, hist, m1, m2, pad;
};
static_assert(sizeof(Geo) == 32 && sizeof(Rec) == 32, "records are 8 floats");

float lum3(const float *c) { return 0.212671f * c[0] + 0.715160f * c[1] + 0.072169f * c[2]; }
float sat(float x) { return std::min(std::max(x, 0.f), 1.f); }

struct Params {
    uint32_t N;
    float alpha, moments_alpha;
    uint32_t history_limit;
    int moments_radius;
    float sigma_rt, sigma_normal;
    bool reproject, moments;
};

// SVGF::cal_weight (svgf.cpp:69-78)
float cal_weight(float cd, float nd, float sd, const float *cn, const float *nn, float sn, float ci, float ni,
                 float si) {
    const float dn = cn[0] * nn[0] + cn[1] * nn[1] + cn[2] * nn[2];
    const float w_normal = vs::em::pow(sat(dn), sn);
    const float w_depth = sd == 0.f ? 0.f : std::fabs(cd - nd) / sd;
    const float w_illum = std::fabs(ci - ni) / si;
    return vs::em::exp(-(std::max(w_illum, 0.f) + std::max(w_depth, 0.f))) * w_normal;
}

bool in_screen(int x, int y, int W, int H) { return x >= 0 && y >= 0 && x < W && y < H; }

// Reproject::is_valid_reproject (reproject.cpp:14-27)
bool valid_reproject(const Geo &cur, const Geo &prev, int W, int H) {
    const bool inside = in_screen(int(prev.pf[0]), int(prev.pf[1]), W, H);
    const bool z_ok = (std::fabs(cur.depth - prev.depth) / (cur.dg + 1e-2f)) < 10.f;
    const float dx = cur.n[0] - prev.n[0], dy = cur.n[1] - prev.n[1], dz = cur.n[2] - prev.n[2];
    const bool n_ok = (std::sqrt(dx * dx + dy * dy + dz * dz) / (cur.fw + 1e-2f)) < 16.f;
    return inside && z_ok && n_ok;
}

// Reproject::load_prev_data (reproject.cpp:29-125)
bool load_prev(const Geo &cg, const Geo *pg, const Rec *prev, const float *mv, int W, int H, float &history,
               float il[3], float m[2]) {
    const float ppx = cg.pf[0] - mv[0], ppy = cg.pf[1] - mv[1];
    const int qx = int(ppx - 0.5f), qy = int(ppy - 0.5f);
    const bool inside = in_screen(qx, qy, W, H);
    il[0] = il[1] = il[2] = 0.f;
    m[0] = m[1] = 0.f;
    bool valid = false;
    bool v[4] = {false, false, false, false};
    const int ox[4] = {0, 1, 0, 1}, oy[4] = {0, 0, 1, 1};
    for (int i = 0; i < 4; ++i) {
        const int x = qx + ox[i], y = qy + oy[i];
        if (in_screen(x, y, W, H)) {
            v[i] = valid_reproject(cg, pg[size_t(y) * W + x], W, H);
            valid = valid || v[i];
        }
    }
    if (valid) {
        float wsum = 0.f;
        const float fx = ppx - std::floor(ppx), fy = ppy - std::floor(ppy);
        const float w[4] = {(1.f - fx) * (1.f - fy), fx * (1.f - fy), (1.f - fx) * fy, fx * fy};
        for (int i = 0; i < 4; ++i) {
            const int x = qx + ox[i], y = qy + oy[i];
            if (v[i] && in_screen(x, y, W, H)) {
                const Rec &r = prev[size_t(y) * W + x];
                for (int k = 0; k < 3; ++k) il[k] += w[i] * r.il[k];
                m[0] += w[i] * r.m1;
                m[1] += w[i] * r.m2;
                wsum += w[i];
            }
        }
        valid = wsum >= 0.01f;
        for (int k = 0; k < 3; ++k) il[k] = valid ? il[k] / wsum : 0.f;
        m[0] = valid ? m[0] / wsum : 0.f;
        m[1] = valid ? m[1] / wsum : 0.f;
    }
    if (!valid) {
        // foreach_neighbor (frame_buffer.h:49-67) of the 3x3 block around q
        float n = 0.f;
        const int x0 = std::max(0, qx - 1), x1 = std::min(qx + 1, W - 1);
        const int y0 = std::max(0, qy - 1), y1 = std::min(qy + 1, H - 1);
        for (int x = x0; x <= x1; ++x) {
            for (int y = y0; y <= y1; ++y) {
                const size_t j = size_t(y) * W + x;
                if (valid_reproject(cg, pg[j], W, H)) {
                    for (int k = 0; k < 3; ++k) il[k] += prev[j].il[k];
                    m[0] += prev[j].m1;
                    m[1] += prev[j].m2;
                    n += 1.f;
                }
            }
        }
        if (n > 0.f) {
            valid = true;
            for (int k = 0; k < 3; ++k) il[k] /= n;
            m[0] /= n;
            m[1] /= n;
        }
    }
    if (valid && inside) {
        history = prev[size_t(qy) * W + qx].hist;
    } else {
        history = 0.f;
        il[0] = il[1] = il[2] = 0.f;
        m[0] = m[1] = 0.f;
    }
    return valid;
}

// Modulator's demodulate kernel (modulator.cpp:29-37) and the reproject
// kernel (reproject.cpp:127-162) for one pixel
Rec reproject_px(const Params &p, const Geo *gcur, const Geo *gprev, const Rec *prev, const float *mv,
                 const float *rad, const float *alb, const float *emi, int W, int H, int x, int y) {
    const size_t i = size_t(y) * W + x;
    Rec r{};
    for (int k = 0; k < 3; ++k)
        r.il[k] = (rad[4 * i + k] - emi[4 * i + k]) / std::max(alb[4 * i + k], 0.001f);
    if (!p.reproject) return r;
    float history = 0.f, pil[3], pm[2];
    const bool valid = load_prev(gcur[i], gprev, prev, mv + 2 * i, W, H, history, pil, pm);
    history = std::min(float(p.history_limit), valid ? history + 1.0f : 1.0f);
    const float alpha = valid ? std::max(p.alpha, 1.f / history) : 1.f;
    const float malpha = valid ? std::max(p.moments_alpha, 1.f / history) : 1.f;
    float m1 = lum3(r.il), m2 = m1 * m1;
    m1 = pm[0] + malpha * (m1 - pm[0]);
    m2 = pm[1] + malpha * (m2 - pm[1]);
    const float var = std::max(0.f, m2 - m1 * m1);
    for (int k = 0; k < 3; ++k) r.il[k] = pil[k] + alpha * (r.il[k] - pil[k]);
    r.var = var;
    r.m1 = m1;
    r.m2 = m2;
    r.hist = history;
    return r;
}

// FilterMoments (filter_moments.cpp:13-60)
Rec moments_px(const Params &p, const Geo *g, const Rec *in, int W, int H, int x, int y) {
    const size_t i = size_t(y) * W + x;
    Rec r = in[i];
    if (r.hist < 4.f) return r;
    const Geo &cg = g[i];
    if (cg.depth < 0.f) return r;
    const float sigma_depth = std::max(cg.dg, 1e-8f) * 3.f;
    const float cl = lum3(r.il);
    float wsum = 0.f, s[3] = {0.f, 0.f, 0.f}, sm[2] = {0.f, 0.f};
    const int rad = p.moments_radius;
    const int x0 = std::max(0, x - rad), x1 = std::min(x + rad, W - 1);
    const int y0 = std::max(0, y - rad), y1 = std::min(y + rad, H - 1);
    for (int nx = x0; nx <= x1; ++nx) {
        for (int ny = y0; ny <= y1; ++ny) {
            const size_t j = size_t(ny) * W + nx;
            const Rec &q = in[j];
            const float w = cal_weight(cg.depth, g[j].depth, sigma_depth, cg.n, g[j].n, p.sigma_normal, cl, lum3(q.il),
                                       p.sigma_rt);
            wsum += w;
            for (int k = 0; k < 3; ++k) s[k] += q.il[k] * w;
            sm[0] += q.m1 * w;
            sm[1] += q.m2 * w;
        }
    }
    wsum = std::max(wsum, 1e-6f);
    for (int k = 0; k < 3; ++k) r.il[k] = s[k] / wsum;
    sm[0] /= wsum;
    sm[1] /= wsum;
    float var = sm[1] - sm[0] * sm[0];
    var *= 4.f / r.hist;
    r.var = var;
    return r;
}

// AtrousFilter (atrous.cpp:13-93), one pass of step `step`
Rec atrous_px(const Params &p, const Geo *g, const Rec *in, int W, int H, int x, int y, int step) {
    const size_t i = size_t(y) * W + x;
    Rec r = in[i];
    const float cl = lum3(r.il);
    // compute_variance: the 3x3 Gaussian of the variance around the pixel
    float variance = 0.f;
    {
        const float tab[2][2] = {{1.f / 4, 1.f / 8}, {1.f / 8, 1.f / 16}};
        const int x0 = std::max(0, x - 1), x1 = std::min(x + 1, W - 1);
        const int y0 = std::max(0, y - 1), y1 = std::min(y + 1, H - 1);
        for (int nx = x0; nx <= x1; ++nx)
            for (int ny = y0; ny <= y1; ++ny) variance += in[size_t(ny) * W + nx].var * tab[std::abs(nx - x)][std::abs(ny - y)];
    }
    const Geo &cg = g[i];
    if (cg.depth < 0.f) return r;
    const float kw[3] = {1.f, 2.f / 3, 1.f / 6};
    const float sigma_depth = std::max(cg.dg, 1e-8f) * float(step);
    const float sigma_il = p.sigma_rt * std::sqrt(std::max(0.f, 1e-10f + variance));
    float wsum = 1.f;
    float sum[4] = {r.il[0], r.il[1], r.il[2], r.var};
    const int x0 = std::max(0, x - 2), x1 = std::min(x + 2, W - 1);
    const int y0 = std::max(0, y - 2), y1 = std::min(y + 2, H - 1);
    for (int nx = x0; nx <= x1; ++nx) {
        for (int ny = y0; ny <= y1; ++ny) {
            const int ox = x - nx, oy = y - ny;
            if (ox == 0 && oy == 0) continue;
            const float kern = kw[std::abs(ox)] * kw[std::abs(oy)];
            const int tx = x + ox * step, ty = y + oy * step;
            if (!in_screen(tx, ty, W, H)) continue;
            const size_t j = size_t(ty) * W + tx;
            const Rec &q = in[j];
            const float dist = std::sqrt(float(ox) * float(ox) + float(oy) * float(oy));
            const float w = cal_weight(cg.depth, g[j].depth, sigma_depth * dist, cg.n, g[j].n, p.sigma_normal, cl,
                                       lum3(q.il), sigma_il);
            const float iw = w * kern;
            wsum += iw;
            sum[0] += iw * q.il[0];
            sum[1] += iw * q.il[1];
            sum[2] += iw * q.il[2];
            sum[3] += (iw * iw) * q.var;
        }
    }
    r.il[0] = sum[0] / wsum;
    r.il[1] = sum[1] / wsum;
    r.il[2] = sum[2] / wsum;
    r.var = sum[3] / (wsum * wsum);
    return r;
}

Params params_of(const float *pp) {
    Params p;
    p.N = uint32_t(pp[0]);
    p.alpha = pp[1];
    p.moments_alpha = pp[2];
    p.history_limit = uint32_t(pp[3]);
    p.moments_radius = int(pp[4]);
    p.sigma_rt = pp[5];
    p.sigma_normal = pp[6];
    p.reproject = pp[7] != 0.f;
    p.moments = pp[8] != 0.f;
    return p;
}

}// namespace

extern "C" {

// FrameBuffer::compute_gradient over the whole G-buffer (frame_buffer.cpp:124-168):
// fwidth of the normal and the depth gradient from the 3x3 neighbourhood
void or_gbuffer_gradient(float *gbuf, uint32_t W, uint32_t H) {
    Geo *g = reinterpret_cast<Geo *>(gbuf);
    std::vector<Geo> src(g, g + size_t(W) * H);
    for (int y = 0; y < int(H); ++y) {
        for (int x = 0; x < int(W); ++x) {
            const Geo &c = src[size_t(y) * W + x];
            uint32_t xn = 0, yn = 0;
            float ndx[3] = {0.f, 0.f, 0.f}, ndy[3] = {0.f, 0.f, 0.f}, ddx = 0.f, ddy = 0.f;
            const int x0 = std::max(0, x - 1), x1 = std::min(x + 1, int(W) - 1);
            const int y0 = std::max(0, y - 1), y1 = std::min(y + 1, int(H) - 1);
            for (int nx = x0; nx <= x1; ++nx) {
                for (int ny = y0; ny <= y1; ++ny) {
                    const Geo &q = src[size_t(ny) * W + nx];
                    if (x > nx) {
                        ++xn;
                        for (int k = 0; k < 3; ++k) ndx[k] += c.n[k] - q.n[k];
                        ddx += c.depth - q.depth;
                    } else if (nx > x) {
                        ++xn;
                        for (int k = 0; k < 3; ++k) ndx[k] += q.n[k] - c.n[k];
                        ddx += q.depth - c.depth;
                    }
                    if (y > ny) {
                        ++yn;
                        for (int k = 0; k < 3; ++k) ndy[k] += c.n[k] - q.n[k];
                        ddy += c.depth - q.depth;
                    } else if (ny > y) {
                        ++yn;
                        for (int k = 0; k < 3; ++k) ndy[k] += q.n[k] - c.n[k];
                        ddy += q.depth - c.depth;
                    }
                }
            }
            float fw[3];
            for (int k = 0; k < 3; ++k) fw[k] = std::fabs(ndx[k] / float(xn)) + std::fabs(ndy[k] / float(yn));
            Geo &o = g[size_t(y) * W + x];
            o.fw = std::sqrt(fw[0] * fw[0] + fw[1] * fw[1] + fw[2] * fw[2]);
            o.dg = std::fabs(ddx / float(xn)) + std::fabs(ddy / float(yn));
        }
    }
}

// SVGF::dispatch (svgf.cpp:89-106) for one frame: demodulate + reproject,
// filter moments, N a-trous passes of step 1, 2, 4, ..., modulate.
//   params: N, alpha, moments_alpha, history_limit, moments_filter_radius,
//           sigma_rt, sigma_normal, reproject on, moment filter on
//   gcur / gprev: the frame's and the previous frame's G-buffer
//   mv: motion vectors [P][2]; rad: the film's running mean [P][4];
//   alb / emi: G-buffer albedo and emission [P][4]
//   cur (out) / prev: SVGF records of this and the previous frame
//   out: the modulated picture [P][4] (w = 1)
void or_svgf_frame(const float *params, uint32_t W, uint32_t H, const float *gcur, const float *gprev,
                   const float *mv, const float *rad, const float *alb, const float *emi, const float *prev,
                   float *cur, float *out) {
    const Params p = params_of(params);
    const size_t P = size_t(W) * H;
    const Geo *gc = reinterpret_cast<const Geo *>(gcur), *gp = reinterpret_cast<const Geo *>(gprev);
    const Rec *pr = reinterpret_cast<const Rec *>(prev);
    std::vector<Rec> a(P), b(P);
    for (int y = 0; y < int(H); ++y)
        for (int x = 0; x < int(W); ++x)
            a[size_t(y) * W + x] = reproject_px(p, gc, gp, pr, mv, rad, alb, emi, int(W), int(H), x, y);
    if (p.moments) {
        for (int y = 0; y < int(H); ++y)
            for (int x = 0; x < int(W); ++x) b[size_t(y) * W + x] = moments_px(p, gc, a.data(), int(W), int(H), x, y);
        a.swap(b);
    }
    for (uint32_t k = 0; k < p.N; ++k) {
        const int step = 1 << k;
        for (int y = 0; y < int(H); ++y)
            for (int x = 0; x < int(W); ++x)
                b[size_t(y) * W + x] = atrous_px(p, gc, a.data(), int(W), int(H), x, y, step);
        a.swap(b);
    }
    std::memcpy(cur, a.data(), P * sizeof(Rec));
    // Modulator's modulate kernel (modulator.cpp:19-27)
    for (size_t i = 0; i < P; ++i) {
        for (int k = 0; k < 3; ++k) out[4 * i + k] = a[i].il[k] * alb[4 * i + k] + emi[4 * i + k];
        out[4 * i + 3] = 1.f;
    }
}

// VarianceStats::add_sample_impl (inspector.h:22-31) over a sequence of
// samples; stats = (avg, var, N as float bits, 0) per pixel
void or_variance_stats_add(float *stats, const float *x, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        float avg = stats[4 * i], var = stats[4 * i + 1];
        uint32_t N;
        std::memcpy(&N, stats + 4 * i + 2, 4);
        const float xi = x[i];
        const float fN = float(N), fN1 = float(N + 1u);
        const float new_avg = (avg * fN + xi) / fN1;
        const float d = avg - new_avg, e = xi - avg;
        const float new_var = (fN * var + fN * (d * d) + e * e + 2.f * e * d + d * d) / fN1;
        N += 1u;
        stats[4 * i] = new_avg;
        stats[4 * i + 1] = new_var;
        std::memcpy(stats + 4 * i + 2, &N, 4);
    }
}

}// extern "C"
