"""Polynomial coefficients of the single-precision elementary functions in
vision_amd/csrc/elementary.h.

Each kernel is a Chebyshev (near-minimax) fit, computed with mpmath at 60
significant digits, of the function's remainder after its leading Taylor
terms on the range its argument reduction leaves; the coefficients are then
rounded to float32 and printed as C hex-float literals.

    python vision_amd/gen/fit_elementary.py

The printed block is what elementary.h holds; tests/test_elementary.py checks
the resulting functions against mpmath over their documented domains.
"""
from __future__ import annotations

import struct

import mpmath as mp

mp.mp.dps = 60


def f32(x) -> float:
    return struct.unpack("<f", struct.pack("<f", float(x)))[0]


def lit(x) -> str:
    v = f32(x)
    if v == 0:
        return "0.0f"
    mant, exp = v.hex().split("p")
    mant = mant.rstrip("0").rstrip(".")
    return f"{mant}p{exp}f"


def fit(name, g, a, b, n):
    """coefficients c0..c_{n-1} (lowest degree first) of g on [a, b]"""
    poly, err = mp.chebyfit(g, [a, b], n, error=True)
    coefs = list(reversed(poly))
    print(f"    // {name}: degree {n - 1} on [{mp.nstr(a, 6)}, {mp.nstr(b, 6)}], fit error {mp.nstr(err, 3)}")
    return coefs


def split(c, bits):
    """c = hi + lo with hi holding `bits` significant bits (products k*hi exact)"""
    e = mp.floor(mp.log(abs(c), 2))
    q = mp.mpf(2) ** (e - bits + 1)
    hi = mp.nint(c / q) * q
    lo = c - hi
    return hi, lo


def main():
    pio2 = mp.pi / 2
    # sin r = r + r^3 S(r^2), cos r = 1 - r^2/2 + r^4 C(r^2), |r| <= pi/4
    zmax = (mp.pi / 4) ** 2
    eps = mp.mpf(10) ** -40

    def s_rem(z):
        z = max(z, eps)
        r = mp.sqrt(z)
        return (mp.sin(r) - r) / (r * z)

    def c_rem(z):
        z = max(z, eps)
        r = mp.sqrt(z)
        return (mp.cos(r) - 1 + z / 2) / (z * z)

    S = fit("sin remainder S(z)", s_rem, 0, zmax, 4)
    C = fit("cos remainder C(z)", c_rem, 0, zmax, 4)

    # e^r = 1 + r + r^2 E(r), |r| <= ln2 / 2
    def e_rem(r):
        if abs(r) < eps:
            return mp.mpf(1) / 2
        return (mp.exp(r) - 1 - r) / (r * r)

    E = fit("exp remainder E(r)", e_rem, -mp.log(2) / 2, mp.log(2) / 2, 6)

    # log(1+f) = 2s + s R(s^2), s = f / (2 + f), R(z) = 2 atanh(sqrt z)/sqrt z - 2 = z L(z)
    smax = (mp.sqrt(2) - 1) / (mp.sqrt(2) + 1)

    def l_rem(z):
        z = max(z, eps)
        s = mp.sqrt(z)
        return (2 * mp.atanh(s) / s - 2) / z

    L = fit("log remainder L(z)", l_rem, 0, smax * smax, 4)

    # asin x = x + x^3 A(x^2), |x| <= 1/2
    def a_rem(z):
        z = max(z, eps)
        x = mp.sqrt(z)
        return (mp.asin(x) - x) / (x * z)

    A = fit("asin remainder A(z)", a_rem, 0, mp.mpf(1) / 4, 6)

    # atan a = a + a^3 T(a^2), |a| <= tan(pi/8)
    tmax = mp.tan(mp.pi / 8)

    def t_rem(z):
        z = max(z, eps)
        a = mp.sqrt(z)
        return (mp.atan(a) - a) / (a * z)

    T = fit("atan remainder T(z)", t_rem, 0, tmax * tmax, 6)

    p1 = f32(pio2)
    p2 = f32(pio2 - mp.mpf(p1))
    p3 = f32(pio2 - mp.mpf(p1) - mp.mpf(p2))
    ln2hi, ln2lo = split(mp.log(2), 12)
    pio2lo = pio2 - mp.mpf(f32(pio2))
    pio4lo = mp.pi / 4 - mp.mpf(f32(mp.pi / 4))
    pilo = mp.pi - mp.mpf(f32(mp.pi))
    print(f"    constexpr float kPio2_1 = {lit(p1)}, kPio2_2 = {lit(p2)}, kPio2_3 = {lit(p3)};")
    print(f"    constexpr float kLn2Hi = {lit(ln2hi)}, kLn2Lo = {lit(ln2lo)};")
    print(f"    constexpr float kPio2Hi = {lit(pio2)}, kPio2Lo = {lit(pio2lo)};")
    print(f"    constexpr float kPio4Hi = {lit(mp.pi / 4)}, kPio4Lo = {lit(pio4lo)};")
    print(f"    constexpr float kPiHi = {lit(mp.pi)}, kPiLo = {lit(pilo)};")
    print(f"    constexpr float kTanPi8 = {lit(tmax)};")
    for nm, cs in (("S", S), ("C", C), ("E", E), ("L", L), ("A", A), ("T", T)):
        print(f"    constexpr float k{nm}[{len(cs)}] = {{{', '.join(lit(c) for c in cs)}}};")


if __name__ == "__main__":
    main()
