// Generator of the sRGB -> spectrum sigmoid-polynomial coefficient table
// used by hero-wavelength mode.
//
// The reference consumes this table as data (src/render_core/spectra/srgb2spec.h,
// absent from the reference tree); it is regenerated here with the published
// method of Jakob & Hanika, "A Low-Dimensional Function Space for Efficient
// Spectral Upsampling" (Computer Graphics Forum 38(2), 2019): for every cell
// of a res^3 grid over (max component z, the two other components relative
// to it) find the coefficients of s(l) = sigmoid(c0 l^2 + c1 l + c2) whose
// colour under CIE 1931 and the D65 illuminant matches the target in CIELAB
// (Gauss-Newton, finite-difference Jacobian, warm-started from the
// neighbouring cell).
//
// Layout: float[3][res][res][res][3], [max component][z][y][x][c0, c1, c2],
// coefficients for wavelengths in nm.  The z axis is warped by
// smoothstep(smoothstep(k / (res - 1))).
//
// usage: rgb2spec <cie_xyz_d65.f32> <out.f32> [res]
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr int kCie = 471;// 360..830 nm at 1 nm
constexpr double kLmin = 360.0, kLmax = 830.0;
constexpr int kFine = kCie;// trapezoid quadrature on the 1 nm CIE grid

const double xyz_to_rgb[3][3] = {{3.240479, -1.537150, -0.498535},
                                 {-0.969256, 1.875991, 0.041556},
                                 {0.055648, -0.204043, 1.057311}};
const double rgb_to_xyz[3][3] = {{0.412453, 0.357580, 0.180423},
                                 {0.212671, 0.715160, 0.072169},
                                 {0.019334, 0.119193, 0.950227}};

std::vector<double> g_lambda(kFine);
std::vector<double> g_rgb_tbl[3];
double g_white[3];

double interp(const float *tab, double lambda) {
    double x = lambda - kLmin;
    int off = std::min(std::max(int(std::floor(x)), 0), kCie - 2);
    double w = x - off;
    return (1.0 - w) * tab[off] + w * tab[off + 1];
}

void init_tables(const float *cie) {
    const float *X = cie, *Y = cie + kCie, *Z = cie + 2 * kCie, *I = cie + 3 * kCie;
    double h = (kLmax - kLmin) / double(kFine - 1);
    for (int k = 0; k < 3; ++k) g_rgb_tbl[k].assign(kFine, 0.0);
    double norm = 0.0;
    std::vector<double> w(kFine);
    for (int i = 0; i < kFine; ++i) {
        double weight = (i == 0 || i == kFine - 1) ? 0.5 * h : h;
        w[i] = weight;
        double l = kLmin + i * h;
        g_lambda[i] = l;
        norm += interp(Y, l) * interp(I, l) * weight;
    }
    g_white[0] = g_white[1] = g_white[2] = 0.0;
    for (int i = 0; i < kFine; ++i) {
        double l = g_lambda[i];
        double xyz[3] = {interp(X, l), interp(Y, l), interp(Z, l)};
        double ill = interp(I, l) * w[i] / norm;
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 3; ++j) g_rgb_tbl[k][i] += xyz_to_rgb[k][j] * xyz[j] * ill;
        for (int k = 0; k < 3; ++k) g_white[k] += xyz[k] * ill;
    }
}

void to_lab(double *p) {
    double X = 0, Y = 0, Z = 0;
    for (int j = 0; j < 3; ++j) {
        X += p[j] * rgb_to_xyz[0][j];
        Y += p[j] * rgb_to_xyz[1][j];
        Z += p[j] * rgb_to_xyz[2][j];
    }
    auto f = [](double t) {
        const double d = 6.0 / 29.0;
        return t > d * d * d ? std::cbrt(t) : t / (d * d * 3.0) + 4.0 / 29.0;
    };
    double fy = f(Y / g_white[1]);
    p[0] = 116.0 * fy - 16.0;
    p[1] = 500.0 * (f(X / g_white[0]) - fy);
    p[2] = 200.0 * (fy - f(Z / g_white[2]));
}

void residual(const double *c, const double *rgb_lab, double *r) {
    double out[3] = {0, 0, 0};
    for (int i = 0; i < kFine; ++i) {
        double x = (g_lambda[i] - kLmin) / (kLmax - kLmin);
        double v = (c[0] * x + c[1]) * x + c[2];
        double s = 0.5 + v / (2.0 * std::sqrt(1.0 + v * v));
        for (int j = 0; j < 3; ++j) out[j] += g_rgb_tbl[j][i] * s;
    }
    to_lab(out);
    for (int j = 0; j < 3; ++j) r[j] = rgb_lab[j] - out[j];
}

bool solve3(double A[3][3], double *b) {
    int piv[3] = {0, 1, 2};
    for (int i = 0; i < 3; ++i) {
        int m = i;
        for (int k = i + 1; k < 3; ++k)
            if (std::fabs(A[k][i]) > std::fabs(A[m][i])) m = k;
        if (std::fabs(A[m][i]) < 1e-15) return false;
        std::swap(A[i], A[m]);
        std::swap(b[i], b[m]);
        std::swap(piv[i], piv[m]);
        for (int k = i + 1; k < 3; ++k) {
            double f = A[k][i] / A[i][i];
            for (int j = i; j < 3; ++j) A[k][j] -= f * A[i][j];
            b[k] -= f * b[i];
        }
    }
    for (int i = 2; i >= 0; --i) {
        for (int j = i + 1; j < 3; ++j) b[i] -= A[i][j] * b[j];
        b[i] /= A[i][i];
    }
    return true;
}

void gauss_newton(const double *rgb, double *c) {
    double lab[3] = {rgb[0], rgb[1], rgb[2]};
    to_lab(lab);
    for (int it = 0; it < 15; ++it) {
        double r[3];
        residual(c, lab, r);
        double J[3][3];
        const double eps = 1e-5;
        for (int i = 0; i < 3; ++i) {
            double t[3] = {c[0], c[1], c[2]}, r0[3], r1[3];
            t[i] -= eps;
            residual(t, lab, r0);
            t[i] += 2 * eps;
            residual(t, lab, r1);
            for (int j = 0; j < 3; ++j) J[j][i] = (r1[j] - r0[j]) / (2 * eps);
        }
        double x[3] = {r[0], r[1], r[2]};
        if (!solve3(J, x)) break;
        for (int j = 0; j < 3; ++j) c[j] -= x[j];
        double m = std::max(std::fabs(c[0]), std::max(std::fabs(c[1]), std::fabs(c[2])));
        if (m > 200) {
            for (int j = 0; j < 3; ++j) c[j] *= 200 / m;
        }
        double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        if (rr < 1e-6) break;
    }
}

double smoothstep(double x) { return x * x * (3.0 - 2.0 * x); }

}// namespace

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s cie_xyz_d65.f32 out.f32 [res]\n", argv[0]);
        return 2;
    }
    const int res = argc > 3 ? std::atoi(argv[3]) : 64;
    std::vector<float> cie(4 * kCie);
    FILE *f = std::fopen(argv[1], "rb");
    if (!f || std::fread(cie.data(), 4, cie.size(), f) != cie.size()) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 1;
    }
    std::fclose(f);
    init_tables(cie.data());
    std::vector<double> scale(res);
    for (int k = 0; k < res; ++k) scale[k] = smoothstep(smoothstep(k / double(res - 1)));
    std::vector<float> out(size_t(3) * res * res * res * 3);
    auto work = [&](int l, int j) {
        double y = j / double(res - 1);
        for (int i = 0; i < res; ++i) {
            double x = i / double(res - 1);
            int start = res / 5;
            for (int pass = 0; pass < 2; ++pass) {
                double c[3] = {0, 0, 0};
                int k = start;
                while (pass == 0 ? k < res : k >= 0) {
                    double b = scale[k], rgb[3];
                    rgb[l] = b;
                    rgb[(l + 1) % 3] = x * b;
                    rgb[(l + 2) % 3] = y * b;
                    gauss_newton(rgb, c);
                    const double c0 = kLmin, c1 = 1.0 / (kLmax - kLmin);
                    double A = c[0], B = c[1], C = c[2];
                    size_t idx = ((size_t(l) * res + k) * res + j) * res + i;
                    out[3 * idx + 0] = float(A * c1 * c1);
                    out[3 * idx + 1] = float(B * c1 - 2 * A * c0 * c1 * c1);
                    out[3 * idx + 2] = float(C - B * c0 * c1 + A * c0 * c0 * c1 * c1);
                    k += pass == 0 ? 1 : -1;
                }
            }
        }
    };
    unsigned nth = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    std::vector<std::pair<int, int>> jobs;
    for (int l = 0; l < 3; ++l)
        for (int j = 0; j < res; ++j) jobs.emplace_back(l, j);
    std::atomic_size_t next{0};
    for (unsigned t = 0; t < nth; ++t)
        th.emplace_back([&] {
            while (true) {
                size_t n = next.fetch_add(1);
                if (n >= jobs.size()) break;
                work(jobs[n].first, jobs[n].second);
            }
        });
    for (auto &t : th) t.join();
    FILE *o = std::fopen(argv[2], "wb");
    if (!o || std::fwrite(out.data(), 4, out.size(), o) != out.size()) {
        std::fprintf(stderr, "cannot write %s\n", argv[2]);
        return 1;
    }
    std::fclose(o);
    return 0;
}
