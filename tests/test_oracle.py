"""Pins the CPU oracle against known answers, closed forms and brute force.

Each check is chosen so that a plausible mistake in the oracle (a dropped
cosine, a wrong pdf, a swapped index, a missing MIS term) fails it.
"""
import math

import numpy as np
import pytest

import vision_amd as va
from oracle import oracle
from vision_amd import scenegen


# --------------------------------------------------------------- sampler
def test_tea_published_known_answers():
    # TEA (Wheeler & Needham 1994), 32 rounds: the oracle's tea() is the
    # reference's tea<N> (src/math/util.h:13-25) with N and the key exposed
    assert oracle.tea_block(0, 0, [0, 0, 0, 0], 32) == (0x41EA3A0A, 0x94BAA940)
    assert oracle.tea_block(0x01234567, 0x89ABCDEF, [0x00112233, 0x44556677, 0x8899AABB, 0xCCDDEEFF],
                            32) == (0x126C6B92, 0xC0653A3E)


def test_lcg_known_outputs():
    # Numerical Recipes' quick generator (a = 1664525, c = 1013904223) from 0
    states, values = oracle.lcg(0, 4)
    assert list(states) == [1013904223, 1196435762, 3519870697, 2868466484]
    # the float is the low 24 bits / 2^24 (util.h:27-34)
    assert values[0] == np.float32((1013904223 & 0xFFFFFF) / 2 ** 24)


def test_sampler_seed_composition():
    # IndependentSampler::start: state = tea(tea(px, py), tea(index, dim))
    px, py, idx, dim = 17, 99, 5, 0
    want = oracle.tea(oracle.tea(px, py), oracle.tea(idx, dim))
    _, states = oracle.sampler_stream(px, py, idx, dim, 1)
    assert states[0] == want
    # streams of different pixels / frames / dims differ
    a, _ = oracle.sampler_stream(1, 2, 3, 0, 8)
    b, _ = oracle.sampler_stream(1, 2, 4, 0, 8)
    c, _ = oracle.sampler_stream(1, 2, 3, 1, 8)
    assert not np.array_equal(a, b) and not np.array_equal(a, c)


def _chi2_ok(counts, expected, alpha_sigma=6.0):
    counts = np.asarray(counts, float).ravel()
    expected = np.asarray(expected, float).ravel()
    keep = expected > 5
    chi2 = np.sum((counts[keep] - expected[keep]) ** 2 / expected[keep])
    dof = keep.sum() - 1
    return chi2 < dof + alpha_sigma * math.sqrt(2 * dof), chi2, dof


# --------------------------------------------------------------- warps
def test_cosine_hemisphere_matches_cos_over_pi():
    rng = np.random.default_rng(1)
    n = 400_000
    w = oracle.cosine_hemisphere(rng.random((n, 2), dtype=np.float32))
    assert np.allclose(np.linalg.norm(w, axis=1), 1, atol=2e-6)
    assert (w[:, 2] >= 0).all()
    # bins in (cos theta, phi); expected mass of a bin = (c1^2 - c0^2)/2 * dphi / pi
    nc, nphi = 16, 16
    c = w[:, 2]
    phi = np.mod(np.arctan2(w[:, 1], w[:, 0]), 2 * np.pi)
    counts, ce, pe = np.histogram2d(c, phi, bins=[nc, nphi], range=[[0, 1], [0, 2 * np.pi]])
    mass = np.outer((ce[1:] ** 2 - ce[:-1] ** 2) / 2, np.diff(pe) / np.pi)
    ok, chi2, dof = _chi2_ok(counts, mass * n)
    assert ok, (chi2, dof)


def test_square_to_triangle_uniform():
    rng = np.random.default_rng(2)
    n = 200_000
    b = oracle.square_to_triangle(rng.random((n, 2), dtype=np.float32))
    assert (b >= 0).all() and (b.sum(1) <= 1 + 1e-6).all()
    # uniform over the triangle: the density of b1 is 2(1 - b1)
    counts, e = np.histogram(b[:, 0], bins=20, range=(0, 1))
    mass = (e[1:] - e[:-1]) * (2 - e[1:] - e[:-1])
    ok, chi2, dof = _chi2_ok(counts, mass * n)
    assert ok, (chi2, dof)


def test_tent_filter_distribution():
    rng = np.random.default_rng(3)
    r = 1.5
    n = 200_000
    x = oracle.sample_tent(rng.random(n, dtype=np.float32), r)
    assert (np.abs(x) <= r).all()
    counts, e = np.histogram(x, bins=30, range=(-r, r))
    cdf = lambda t: np.where(t < 0, (t + r) ** 2 / (2 * r * r), 1 - (r - t) ** 2 / (2 * r * r))
    ok, chi2, dof = _chi2_ok(counts, (cdf(e[1:]) - cdf(e[:-1])) * n)
    assert ok, (chi2, dof)


# --------------------------------------------------------------- alias
@pytest.mark.parametrize("seed,n", [(0, 1), (1, 7), (2, 64), (3, 1000)])
def test_alias_table_reproduces_distribution_exactly(seed, n):
    rng = np.random.default_rng(seed)
    w = (rng.random(n) ** 3 + 1e-3).astype(np.float32)
    prob, alias, integral = oracle.alias_build(w)
    # exact mass per outcome: column j keeps prob[j] of its 1/n, gives the rest to alias[j]
    mass = np.zeros(n)
    for j in range(n):
        mass[j] += prob[j] / n
        mass[alias[j]] += (1 - prob[j]) / n
    assert np.allclose(mass, w / w.sum(), atol=5e-6)
    assert np.isclose(integral, w.astype(np.float64).sum() / n, rtol=1e-6)


def test_alias_sampling_pmf_and_remap():
    rng = np.random.default_rng(4)
    w = np.array([1, 0.5, 3, 0.01, 2], np.float32)
    u = rng.random(300_000, dtype=np.float32)
    idx, urem, pmf = oracle.alias_sample(w, u)
    counts = np.bincount(idx, minlength=len(w))
    ok, chi2, dof = _chi2_ok(counts, w / w.sum() * len(u))
    assert ok, (chi2, dof)
    assert np.allclose(pmf, (w / w.sum())[idx], rtol=1e-5)
    # the remapped u is again uniform on [0, 1)
    c2, _ = np.histogram(urem, bins=20, range=(0, 1))
    ok, chi2, dof = _chi2_ok(c2, np.full(20, len(u) / 20))
    assert ok, (chi2, dof)


# --------------------------------------------------------------- ray offset
def test_offset_ray_origin_moves_off_the_surface():
    rng = np.random.default_rng(5)
    p = (rng.random((5000, 3)) * 20 - 10).astype(np.float32)
    n = rng.normal(size=(5000, 3))
    n = (n / np.linalg.norm(n, axis=1, keepdims=True)).astype(np.float32)
    q = oracle.offset_ray_origin(p, n)
    d = (q.astype(np.float64) - p) * n
    assert (d.sum(1) > 0).all()
    # 256 ulps away from the origin region, n / 65536 inside |p| < 1/32
    assert (np.abs(q - p) <= np.maximum(1e-4 * np.abs(p), 1.6e-5)).all()


# --------------------------------------------------------------- BSDFs
def _hemi_grid(n):
    # midpoint quadrature over (cos theta, phi)
    c = (np.arange(n) + 0.5) / n
    phi = (np.arange(2 * n) + 0.5) / (2 * n) * 2 * np.pi
    C, P = np.meshgrid(c, phi, indexing="ij")
    s = np.sqrt(1 - C ** 2)
    wi = np.stack([s * np.cos(P), s * np.sin(P), C], -1).reshape(-1, 3)
    dw = (1 / n) * (2 * np.pi / (2 * n))
    return wi.astype(np.float32), dw


def test_lambert_albedo_and_pdf_normalization():
    wi, dw = _hemi_grid(64)
    wo = np.tile(np.array([0.3, 0.1, 0.9487], np.float32), (len(wi), 1))
    color = np.array([0.2, 0.5, 0.9], np.float32)
    f, pdf = oracle.eval_diffuse(0, color, 0.0, wo, wi)
    # eval returns f * |cos wi|: its integral is the albedo; the pdf integrates to 1
    assert np.allclose(f.sum(0) * dw, color, rtol=1e-3)
    assert np.isclose(pdf.sum() * dw, 1.0, rtol=1e-3)
    # below the surface: zero pdf and zero f
    f2, pdf2 = oracle.eval_diffuse(0, color, 0.0, wo[:4], -wi[:4])
    assert (pdf2 == 0).all() and (f2 == 0).all()


def test_oren_nayar_reduces_to_lambert_and_conserves_energy():
    wi, dw = _hemi_grid(48)
    color = np.ones(3, np.float32)
    wo = np.tile(np.array([0.6, 0.0, 0.8], np.float32), (len(wi), 1))
    f0, _ = oracle.eval_diffuse(1, color, 0.0, wo, wi)
    fl, _ = oracle.eval_diffuse(0, color, 0.0, wo, wi)
    assert np.allclose(f0, fl, rtol=1e-6)
    for sigma in (0.2, 0.5, 1.0):
        f, _ = oracle.eval_diffuse(1, color, sigma, wo, wi)
        alb = f[:, 0].sum() * dw
        assert 0.5 < alb <= 1.0 + 1e-3, (sigma, alb)


# --------------------------------------------------------------- traversal
def _random_soup(rng, n):
    c = rng.random((n, 1, 3)) * 4 - 2
    tri = c + rng.normal(scale=0.15, size=(n, 3, 3))
    P = tri.reshape(-1, 3).astype(np.float32)
    T = np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)
    return P, T


def _random_rays(rng, n, tmax=np.inf):
    o = rng.random((n, 3)) * 6 - 3
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = 0
    r[:, 4:7] = d
    r[:, 7] = np.float32(tmax) if np.isfinite(tmax) else np.finfo(np.float32).max
    return r


def test_bvh2_equals_brute_force():
    rng = np.random.default_rng(6)
    P, T = _random_soup(rng, 3000)
    R = _random_rays(rng, 4000)
    a = oracle.trace(P, T, R, bvh=False)
    b = oracle.trace(P, T, R, bvh=True)
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2])
    assert (a[0] != 0xFFFFFFFF).mean() > 0.2
    R[:, 7] = 1.0
    a = oracle.trace(P, T, R, any_hit=True, bvh=False)
    b = oracle.trace(P, T, R, any_hit=True, bvh=True)
    assert np.array_equal(a[0] != 0xFFFFFFFF, b[0] != 0xFFFFFFFF)


def test_watertight_shared_edges_never_leak():
    # a closed cube of 12 triangles: every ray from inside must hit
    rng = np.random.default_rng(7)
    ctx = va.Context(-1)
    sc = ctx.load_text(scenegen.furnace_box())
    ex = sc.export()
    R = _random_rays(rng, 20000)
    R[:, 0:3] = rng.random((20000, 3)) * 1.8 - 0.9
    tri, t, bary = oracle.trace(ex["positions"], ex["triangles"], R, bvh=False)
    assert (tri != 0xFFFFFFFF).all()
    # hits aimed exactly at vertices and edges too
    verts = ex["positions"]
    R2 = np.zeros((len(verts), 8), np.float32)
    R2[:, 4:7] = verts
    R2[:, 7] = np.finfo(np.float32).max
    tri2, _, _ = oracle.trace(verts, ex["triangles"], R2, bvh=False)
    assert (tri2 != 0xFFFFFFFF).all()
    sc.close()
    ctx.close()


# --------------------------------------------------------------- integrator closed forms
def _render_scene(scene_dict, spp, window=None, frame_start=0):
    ctx = va.Context(-1)
    sc = ctx.load_text(scene_dict)
    ex = sc.export()
    film = oracle.render(ex, frame_start, spp, window=window)
    sc.close()
    ctx.close()
    return film, ex


@pytest.mark.parametrize("albedo", [(0.5, 0.5, 0.5), (0.8, 0.3, 0.1)])
def test_furnace_box_closed_form(albedo):
    # every inner wall emits 1 and reflects with albedo rho: L = 1 / (1 - rho)
    film, _ = _render_scene(scenegen.furnace_box(albedo=albedo, width=32, height=32), spp=64)
    mean = film[..., :3].mean(axis=(0, 1))
    expect = 1.0 / (1.0 - np.array(albedo))
    assert np.allclose(mean, expect, rtol=0.01), (mean, expect)


def test_furnace_oren_nayar_white_walls_lose_energy_only_slightly():
    film, _ = _render_scene(scenegen.furnace_box(albedo=(0.5, 0.5, 0.5), material_sigma=0.3), spp=32)
    mean = film[..., :3].mean()
    assert 1.8 < mean < 2.02, mean


def test_point_light_single_bounce_closed_form():
    h, I, rho = 0.7, 3.0, 0.6
    light = {"type": "point", "param": {"position": [0.0, h, 0.0], "color": [I, I, I], "scale": 1.0}}
    scene = scenegen.lit_plane(light, albedo=rho, cam_height=2.0, fov=0.5, width=8, height=8)
    film, _ = _render_scene(scene, spp=16)
    # radiance of a Lambert floor under an isotropic point light: rho/pi * I cos(theta) / r^2
    expect = rho / math.pi * I / (h * h)
    got = film[2:6, 2:6, :3].mean()
    assert abs(got / expect - 1) < 2e-3, (got, expect)


def test_spot_light_falloff_closed_form():
    # a spot light aimed straight down: full intensity inside angle - falloff
    h, I, rho = 1.0, 2.0, 0.5
    light = {"type": "spot", "param": {"position": [0.0, h, 0.0], "direction": [0, -1, 0], "angle": 30,
                                       "falloff": 5, "color": [I, I, I], "scale": 1.0}}
    scene = scenegen.lit_plane(light, albedo=rho, cam_height=2.0, fov=0.5, width=8, height=8)
    film, _ = _render_scene(scene, spp=16)
    expect = rho / math.pi * I / (h * h)
    assert abs(film[2:6, 2:6, :3].mean() / expect - 1) < 2e-3


def test_area_light_form_factor_closed_form():
    # square emitter (side a) at height h above the viewed point, facing down.
    # Radiance reflected by a Lambert floor: rho * Le * F, with F the
    # differential-area-to-parallel-rectangle form factor summed over the
    # four corner rectangles (A = B = (a/2)/h).
    a, h, Le, rho = 0.5, 0.8, 4.0, 0.5
    emitter = {"type": "quad", "name": "lamp",
               "param": {"width": 1, "height": 1, "material": "Black",
                         "transform": scenegen.xform(scenegen.translate(0, h, 0) @ scenegen.scale(a, -1, -a)),
                         "emission": {"type": "area", "param": {"color": [Le, Le, Le], "scale": 1}}}}
    # the camera sits below the lamp so the lamp does not hide the floor
    scene = scenegen.lit_plane(None, albedo=rho, cam_height=0.4, fov=0.5, width=8, height=8,
                               extra_shapes=[emitter])
    film, _ = _render_scene(scene, spp=256)
    A = B = (a / 2) / h
    Fc = (A / math.sqrt(1 + A * A) * math.atan(B / math.sqrt(1 + A * A)) +
          B / math.sqrt(1 + B * B) * math.atan(A / math.sqrt(1 + B * B))) / (2 * math.pi)
    expect = rho * Le * 4 * Fc
    got = film[2:6, 2:6, :3].mean()
    assert abs(got / expect - 1) < 0.01, (got, expect)
