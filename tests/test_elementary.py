"""Accuracy of the library's single-precision elementary functions
(vision_amd/csrc/elementary.h), which the HIP kernels and the oracle share so
that sampling decisions round identically on both sides.

Reference values are numpy's float64 functions (accurate to about one double
ulp, i.e. exact at float32 resolution); errors are in float32 ulps of the
exact value.  A dropped polynomial term, a wrong reduction constant or a wrong
quadrant fails these bounds by orders of magnitude.
"""
import numpy as np
import pytest

from oracle import oracle


def ulps(got, exact):
    exact = np.asarray(exact, np.float64)
    spacing = np.spacing(np.abs(exact).astype(np.float32)).astype(np.float64)
    spacing = np.maximum(spacing, np.float64(np.spacing(np.float32(np.finfo(np.float32).tiny))))
    return np.abs(got.astype(np.float64) - exact) / spacing


def dense(lo, hi, n=400_000, seed=0):
    rng = np.random.default_rng(seed)
    x = np.concatenate([np.linspace(lo, hi, n // 2), rng.uniform(lo, hi, n // 2)])
    return x.astype(np.float32)


@pytest.mark.parametrize("lo,hi", [(-2 * np.pi, 2 * np.pi), (-1e4, 1e4), (-1e-3, 1e-3)])
def test_sin_cos(lo, hi):
    x = dense(lo, hi)
    xd = x.astype(np.float64)
    assert ulps(oracle.elementary("sin", x), np.sin(xd)).max() <= 2.0
    assert ulps(oracle.elementary("cos", x), np.cos(xd)).max() <= 2.0


def test_sin_cos_quadrants_and_limits():
    # the reduction's quadrant table: sin/cos at multiples of pi/2 and the
    # 2 pi theta of the disk warp
    k = np.arange(-8, 9, dtype=np.float64)
    x = (k * np.pi / 2).astype(np.float32)
    s, c = oracle.elementary("sin", x), oracle.elementary("cos", x)
    assert np.allclose(s, np.sin(x.astype(np.float64)), atol=3e-7)
    assert np.allclose(c, np.cos(x.astype(np.float64)), atol=3e-7)
    assert oracle.elementary("sin", np.float32([0.0]))[0] == 0.0
    assert oracle.elementary("cos", np.float32([0.0]))[0] == 1.0
    assert np.isnan(oracle.elementary("sin", np.float32([np.inf, np.nan, 2e5]))).all()
    u = np.linspace(0, 1, 100_001, dtype=np.float32)
    th = (np.float32(2 * np.pi) * u).astype(np.float32)
    s, c = oracle.elementary("sin", th), oracle.elementary("cos", th)
    assert np.abs(s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2 - 1).max() < 4e-7


def test_exp():
    x = dense(-103.0, 88.7)
    assert ulps(oracle.elementary("exp", x), np.exp(x.astype(np.float64))).max() <= 2.0
    x = dense(-1.0, 1.0, seed=1)
    assert ulps(oracle.elementary("exp", x), np.exp(x.astype(np.float64))).max() <= 1.5
    e = oracle.elementary("exp", np.float32([0.0, 89.0, -104.0, np.inf, -np.inf, np.nan, 1.0]))
    assert e[0] == 1.0 and e[1] == np.inf and e[2] == 0.0 and e[3] == np.inf and e[4] == 0.0 and np.isnan(e[5])
    assert abs(float(e[6]) - np.e) <= np.spacing(np.float32(np.e))


def test_log_log1p():
    rng = np.random.default_rng(2)
    bits = rng.integers(1, 0x7F800000, 400_000, dtype=np.uint32)  # every positive finite float, subnormals too
    x = bits.view(np.float32)
    assert ulps(oracle.elementary("log", x), np.log(x.astype(np.float64))).max() <= 2.0
    x = dense(0.5, 2.0, seed=3)
    assert ulps(oracle.elementary("log", x), np.log(x.astype(np.float64))).max() <= 2.0
    sp = oracle.elementary("log", np.float32([0.0, -1.0, np.inf, np.nan, 1.0]))
    assert sp[0] == -np.inf and np.isnan(sp[1]) and sp[2] == np.inf and np.isnan(sp[3]) and sp[4] == 0.0
    x = dense(-0.999, 10.0, seed=4)
    assert ulps(oracle.elementary("log1p", x), np.log1p(x.astype(np.float64))).max() <= 3.0
    x = dense(-1e-5, 1e-5, seed=5)
    assert ulps(oracle.elementary("log1p", x), np.log1p(x.astype(np.float64))).max() <= 3.0


def test_atanh_cosh():
    x = dense(-0.9999, 0.9999)
    assert ulps(oracle.elementary("atanh", x), np.arctanh(x.astype(np.float64))).max() <= 3.0
    # the hero-wavelength warp's argument range
    u = np.linspace(0, 1, 200_001, dtype=np.float32)
    a = (np.float32(0.85691062) - np.float32(1.82750197) * u).astype(np.float32)
    assert ulps(oracle.elementary("atanh", a), np.arctanh(a.astype(np.float64))).max() <= 3.0
    x = dense(-20.0, 20.0, seed=6)
    assert ulps(oracle.elementary("cosh", x), np.cosh(x.astype(np.float64))).max() <= 3.0


def test_acos():
    x = dense(-1.0, 1.0)
    x = np.concatenate([x, np.float32([-1.0, -0.5, 0.0, 0.5, 1.0])])
    assert ulps(oracle.elementary("acos", x), np.arccos(x.astype(np.float64))).max() <= 3.0
    assert np.isnan(oracle.elementary("acos", np.float32([1.5, -2.0, np.nan]))).all()


def test_atan2():
    rng = np.random.default_rng(7)
    y = rng.normal(size=400_000).astype(np.float32)
    x = rng.normal(size=400_000).astype(np.float32)
    got = oracle.elementary("atan2", y, x)
    assert ulps(got, np.arctan2(y.astype(np.float64), x.astype(np.float64))).max() <= 2.5
    # exact axes, signed zeros and infinities as IEEE 754 atan2 has them
    yy = np.float32([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 1.0, 0.0, np.inf, np.inf, -np.inf, 1.0, 1.0])
    xx = np.float32([1.0, 1.0, -1.0, -1.0, 0.0, 0.0, -np.inf, -0.0, 1.0, np.inf, -np.inf, np.inf, 1.0])
    got = oracle.elementary("atan2", yy, xx)
    ref = np.arctan2(yy.astype(np.float64), xx.astype(np.float64))
    assert np.array_equal(np.signbit(got), np.signbit(ref.astype(np.float32)))
    assert (ulps(got, ref) <= 1.0).all()


def test_pow_srgb_domain():
    x = dense(0.0031308, 1.0)
    got = oracle.elementary("pow", x, np.float32(1 / 2.4))
    assert ulps(got, np.power(x.astype(np.float64), np.float64(np.float32(1 / 2.4)))).max() <= 8.0
    assert oracle.elementary("pow", np.float32([0.0]), np.float32(0.5))[0] == 0.0
