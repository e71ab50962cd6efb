// CPU restatement of the radiance cache (hash-grid encoding + MLP + relative
// L2 loss + Adam) -- TEST INFRASTRUCTURE ONLY.  The reference has no
// radiance cache (its integrators/sharc directory is a stub); the
// restatement follows the published method (multiresolution hash encoding,
// Mueller et al. 2022; neural radiance caching, Mueller et al. 2021) with
// the sizes and conventions the library documents: 16 levels x 2 features,
// 2^19 entries per level, spatial hash (x * 1) ^ (y * 2654435761) ^
// (z * 805459861), network 32 -> 64 -> 64 -> 64 -> 3 without biases, ReLU
// hidden layers, loss (y - t)^2 / (y^2 + 0.01) averaged over samples and
// channels, Adam (0.9, 0.99, 1e-15).
//
// Two modes: `round` = 1 rounds where the device does (grid read as fp16,
// features / weights / activations / hidden-layer gradients as bf16), so
// results match the GPU up to accumulation order; `round` = 0 is exact
// double-precision arithmetic, the form the finite-difference checks pin.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int L = 16, F = 2, IN = 32, H = 64, OUT = 3;
constexpr uint32_t T = 1u << 19;
constexpr int W1 = 0, W2 = W1 + IN * H, W3 = W2 + H * H, W4 = W3 + H * H, NP = W4 + H * OUT;

struct Cfg {
    uint32_t res[L];
    uint32_t dense_mask;
    float lo[3], inv_ext[3];
};

float bf_round(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

// IEEE binary16 round-to-nearest-even of a float, returned as float
float half_round(float x) {
    if (std::isnan(x) || std::isinf(x)) return x;
    float a = std::fabs(x);
    if (a >= 65520.f) return std::copysign(INFINITY, x);
    if (a < std::ldexp(1.f, -25)) return std::copysign(0.f, x);
    int e;
    std::frexp(a, &e);// a = m 2^e, m in [0.5, 1)
    int exp = e - 1;  // a = 1.xxx 2^exp
    if (exp < -14) exp = -14;
    float q = std::ldexp(1.f, exp - 10);// spacing of halves at this exponent
    float r = std::nearbyint(a / q) * q;// nearbyint: ties to even in the default mode
    return std::copysign(r, x);
}

uint32_t index_of(const Cfg &c, int l, uint32_t x, uint32_t y, uint32_t z) {
    if (c.dense_mask & (1u << l)) {
        uint32_t n = c.res[l] + 1u;
        return x + n * (y + n * z);
    }
    return (x ^ (y * 2654435761u) ^ (z * 805459861u)) & (T - 1u);
}

struct Corner {
    uint32_t idx[8];
    float w[8];
};

Corner corners(const Cfg &c, int l, const float p[3]) {
    float u[3];
    for (int a = 0; a < 3; ++a) u[a] = std::fmin(std::fmax((p[a] - c.lo[a]) * c.inv_ext[a], 0.f), 1.f);
    float s = float(c.res[l]);
    uint32_t ci[3];
    float f[3];
    for (int a = 0; a < 3; ++a) {
        float q = u[a] * s;
        uint32_t v = uint32_t(std::floor(q));
        if (v >= c.res[l]) v = c.res[l] - 1u;
        ci[a] = v;
        f[a] = q - float(v);
    }
    Corner r;
    for (int k = 0; k < 8; ++k) {
        r.idx[k] = index_of(c, l, ci[0] + (k & 1), ci[1] + ((k >> 1) & 1), ci[2] + ((k >> 2) & 1));
        float wx = (k & 1) ? f[0] : 1.f - f[0];
        float wy = (k & 2) ? f[1] : 1.f - f[1];
        float wz = (k & 4) ? f[2] : 1.f - f[2];
        r.w[k] = (wx * wy) * wz;
    }
    return r;
}

struct Net {
    std::vector<double> w;// working copy of the MLP weights (rounded in round mode)
};

void encode(const Cfg &c, const float *grid, const float p[3], bool round, double x[IN]) {
    for (int l = 0; l < L; ++l) {
        Corner cr = corners(c, l, p);
        const float *g = grid + size_t(l) * T * F;
        if (round) {
            float a0 = 0.f, a1 = 0.f;
            for (int k = 0; k < 8; ++k) {
                a0 += cr.w[k] * half_round(g[2 * cr.idx[k]]);
                a1 += cr.w[k] * half_round(g[2 * cr.idx[k] + 1]);
            }
            x[2 * l] = bf_round(a0);
            x[2 * l + 1] = bf_round(a1);
        } else {
            double a0 = 0, a1 = 0;
            for (int k = 0; k < 8; ++k) {
                a0 += double(cr.w[k]) * g[2 * cr.idx[k]];
                a1 += double(cr.w[k]) * g[2 * cr.idx[k] + 1];
            }
            x[2 * l] = a0;
            x[2 * l + 1] = a1;
        }
    }
}

double rd(double v, bool round) { return round ? double(bf_round(float(v))) : v; }

struct Acts {
    double x[IN], h1[H], h2[H], h3[H], y[OUT];
};

void forward(const double *w, const Acts &in, Acts &a, bool round) {
    a = in;
    for (int n = 0; n < H; ++n) {
        double s = 0;
        for (int k = 0; k < IN; ++k) s += a.x[k] * w[W1 + k * H + n];
        a.h1[n] = rd(std::fmax(round ? double(float(s)) : s, 0.0), round);
    }
    for (int n = 0; n < H; ++n) {
        double s = 0;
        for (int k = 0; k < H; ++k) s += a.h1[k] * w[W2 + k * H + n];
        a.h2[n] = rd(std::fmax(round ? double(float(s)) : s, 0.0), round);
    }
    for (int n = 0; n < H; ++n) {
        double s = 0;
        for (int k = 0; k < H; ++k) s += a.h2[k] * w[W3 + k * H + n];
        a.h3[n] = rd(std::fmax(round ? double(float(s)) : s, 0.0), round);
    }
    for (int c = 0; c < OUT; ++c) {
        if (round) {
            float s = 0.f;
            for (int k = 0; k < H; ++k) s += float(a.h3[k]) * float(w[W4 + k * OUT + c]);
            a.y[c] = s;
        } else {
            double s = 0;
            for (int k = 0; k < H; ++k) s += a.h3[k] * w[W4 + k * OUT + c];
            a.y[c] = s;
        }
    }
}

std::vector<double> working_weights(const float *mlp, bool round) {
    std::vector<double> w(NP);
    for (int i = 0; i < NP; ++i) w[i] = (round && i < W4) ? double(bf_round(mlp[i])) : double(mlp[i]);
    return w;
}

}// namespace

extern "C" {

// resolutions of the 16 levels: floor(min_res * b^l), b = (max_res / min_res)^(1/15);
// a level indexes its vertices directly when (res + 1)^3 <= 2^19
void or_nrc_levels(uint32_t min_res, uint32_t max_res, uint32_t *res, uint32_t *dense_mask) {
    double b = std::exp((std::log(double(max_res)) - std::log(double(min_res))) / double(L - 1));
    *dense_mask = 0;
    for (int l = 0; l < L; ++l) {
        res[l] = uint32_t(std::floor(double(min_res) * std::pow(b, double(l)) + 1e-9));
        uint64_t n = uint64_t(res[l]) + 1u;
        if (n * n * n <= T) *dense_mask |= 1u << l;
    }
}

void or_nrc_encode(const void *cfg, const float *grid, const float *pos, uint32_t n, int32_t round, double *out) {
    const Cfg &c = *static_cast<const Cfg *>(cfg);
    for (uint32_t i = 0; i < n; ++i) encode(c, grid, pos + 3 * i, round != 0, out + size_t(i) * IN);
}

void or_nrc_forward(const void *cfg, const float *mlp, const float *grid, const float *pos, uint32_t n, int32_t round,
                    double *out) {
    const Cfg &c = *static_cast<const Cfg *>(cfg);
    std::vector<double> w = working_weights(mlp, round != 0);
    for (uint32_t i = 0; i < n; ++i) {
        Acts in, a;
        encode(c, grid, pos + 3 * i, round != 0, in.x);
        forward(w.data(), in, a, round != 0);
        for (int k = 0; k < OUT; ++k) out[3 * size_t(i) + k] = a.y[k];
    }
}

// samples: n x (p[3], t[3]); mean loss and its gradients w.r.t. the fp32
// MLP masters and the grid (grid_grad may be null)
void or_nrc_backward(const void *cfg, const float *mlp, const float *grid, const float *samples, uint32_t n,
                     int32_t round, double *loss, double *mlp_grad, double *grid_grad) {
    const Cfg &c = *static_cast<const Cfg *>(cfg);
    const bool r = round != 0;
    std::vector<double> w = working_weights(mlp, r);
    std::memset(mlp_grad, 0, sizeof(double) * NP);
    if (grid_grad) std::memset(grid_grad, 0, sizeof(double) * size_t(L) * T * F);
    const double inv = n ? 1.0 / (3.0 * n) : 0.0;
    double ls = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const float *sm = samples + 6 * size_t(i);
        Acts in, a;
        encode(c, grid, sm, r, in.x);
        forward(w.data(), in, a, r);
        double dy[OUT];
        for (int k = 0; k < OUT; ++k) {
            double y = a.y[k], d = y - sm[3 + k], den = y * y + 0.01;
            ls += d * d / den * inv;
            dy[k] = 2.0 * d / den * inv;
        }
        double g3[H], g2[H], g1[H], dx[IN];
        for (int k = 0; k < H; ++k) {
            for (int o = 0; o < OUT; ++o) mlp_grad[W4 + k * OUT + o] += a.h3[k] * dy[o];
            double s = 0;
            for (int o = 0; o < OUT; ++o) s += dy[o] * w[W4 + k * OUT + o];
            g3[k] = rd(a.h3[k] > 0 ? s : 0.0, r);
        }
        for (int k = 0; k < H; ++k)
            for (int m = 0; m < H; ++m) mlp_grad[W3 + k * H + m] += a.h2[k] * g3[m];
        for (int k = 0; k < H; ++k) {
            double s = 0;
            for (int m = 0; m < H; ++m) s += g3[m] * w[W3 + k * H + m];
            g2[k] = rd(a.h2[k] > 0 ? s : 0.0, r);
        }
        for (int k = 0; k < H; ++k)
            for (int m = 0; m < H; ++m) mlp_grad[W2 + k * H + m] += a.h1[k] * g2[m];
        for (int k = 0; k < H; ++k) {
            double s = 0;
            for (int m = 0; m < H; ++m) s += g2[m] * w[W2 + k * H + m];
            g1[k] = rd(a.h1[k] > 0 ? s : 0.0, r);
        }
        for (int k = 0; k < IN; ++k)
            for (int m = 0; m < H; ++m) mlp_grad[W1 + k * H + m] += a.x[k] * g1[m];
        for (int k = 0; k < IN; ++k) {
            double s = 0;
            for (int m = 0; m < H; ++m) s += g1[m] * w[W1 + k * H + m];
            dx[k] = s;
        }
        if (grid_grad) {
            for (int l = 0; l < L; ++l) {
                Corner cr = corners(c, l, sm);
                double *gg = grid_grad + size_t(l) * T * F;
                for (int k = 0; k < 8; ++k) {
                    gg[2 * cr.idx[k]] += double(cr.w[k]) * dx[2 * l];
                    gg[2 * cr.idx[k] + 1] += double(cr.w[k]) * dx[2 * l + 1];
                }
            }
        }
    }
    *loss = ls;
}

// one Adam step (beta1 0.9, beta2 0.99, eps 1e-15) on n parameters, float
// arithmetic in the device's order; gradients are cleared
void or_nrc_adam(float *p, float *m, float *v, float *g, uint32_t n, float lr, float c1, float c2) {
    const float b1 = 0.9f, b2 = 0.99f, eps = 1e-15f;
    for (uint32_t i = 0; i < n; ++i) {
        float gr = g[i];
        m[i] = b1 * m[i] + (1.f - b1) * gr;
        v[i] = b2 * v[i] + (1.f - b2) * gr * gr;
        p[i] -= lr * (m[i] * c1) / (std::sqrt(v[i] * c2) + eps);
        g[i] = 0.f;
    }
}

uint32_t or_nrc_param_count(void) { return NP; }

}// extern "C"
