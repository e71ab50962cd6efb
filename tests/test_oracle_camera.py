"""Oracle camera-side pieces pinned by closed forms: the fitted-curve pixel
filters (gaussian, mitchell, sinc) and their importance sampling
(src/render_core/filters/*.cpp, fitted_curve.h), and the direct-only bounce
that IlluminationIntegrator::Li adds when max_depth < 2 (integrator.cpp:276-281).
"""
import math

import numpy as np
import pytest

from oracle import oracle
from vision_amd import scenegen

import vision_amd as va


def _axis(cam, x):
    """f(x, 0): the filter along one axis, times its value at 0 on the other."""
    x = np.asarray(x, np.float32)
    return oracle.filter_eval(cam, np.stack([x, np.zeros_like(x)], axis=1))


def _midpoint(a, b, n):
    h = (b - a) / n
    return a + h * (np.arange(n) + 0.5), h


@pytest.mark.parametrize("B,C", [(1 / 3, 1 / 3), (0.0, 0.5), (1.0, 0.0)])
def test_mitchell_kernel_closed_forms(B, C):
    r = 2.0
    cam = oracle.filter_camera("mitchell", r, b=B, c=C)
    k0 = (6 - 2 * B) / 6
    # value at the centre, B/6 at |2x/r| = 1, zero at |2x/r| = 2 (up to the
    # cubic's float rounding) and exactly zero beyond
    assert _axis(cam, [0.0])[0] == pytest.approx(k0 * k0, rel=1e-6)
    assert _axis(cam, [1.0])[0] == pytest.approx(B / 6 * k0, rel=1e-5, abs=1e-7)
    assert np.all(np.abs(_axis(cam, [2.0, -2.0])) < 1e-6)
    assert np.all(_axis(cam, [2.5, -3.0]) == 0)
    # the cubic integrates to one over [-2, 2] for every (B, C)
    x, h = _midpoint(-r, r, 20000)
    integral = _axis(cam, x).astype(np.float64).sum() * h / k0 * (2 / r)
    assert integral == pytest.approx(1.0, abs=1e-5)


@pytest.mark.parametrize("sigma,r", [(1.0, 0.5), (0.5, 1.5), (0.3, 2.0)])
def test_gaussian_filter_closed_forms(sigma, r):
    cam = oracle.filter_camera("gaussian", r, sigma=sigma)
    g = lambda x: math.exp(-x * x / (2 * sigma * sigma)) / math.sqrt(2 * math.pi * sigma * sigma)
    c0 = g(0) - g(r)
    assert _axis(cam, [0.0])[0] == pytest.approx(c0 * c0, rel=1e-5)
    # zero at the radius and beyond, never negative
    assert np.all(_axis(cam, [r, -r, r * 1.5]) == 0)
    assert np.all(_axis(cam, np.linspace(-3 * r, 3 * r, 1001)) >= 0)
    # the truncated, shifted normal density integrates to erf(r / (sigma sqrt 2)) - 2 r g(r)
    x, h = _midpoint(-r, r, 20000)
    integral = _axis(cam, x).astype(np.float64).sum() * h / c0
    assert integral == pytest.approx(math.erf(r / (sigma * math.sqrt(2))) - 2 * r * g(r), rel=2e-5)


@pytest.mark.parametrize("tau,r", [(3.0, 3.0), (2.0, 1.5)])
def test_sinc_filter_closed_forms(tau, r):
    cam = oracle.filter_camera("sinc", r, tau=tau)
    assert _axis(cam, [0.0])[0] == pytest.approx(4.0, rel=1e-6)
    # zeros of sin(pi x) inside the radius, nothing outside it
    ints = np.arange(1, int(r) + 1, dtype=np.float32)
    ints = ints[ints <= r]
    assert np.all(np.abs(_axis(cam, np.concatenate([ints, -ints]))) < 1e-6)
    assert np.all(_axis(cam, [r * 1.01, -r * 1.2]) == 0)
    x = np.float32(0.37)
    ref = 4 * math.sin(math.pi * x) / (math.pi * x) * math.sin(math.pi * x / tau) / (math.pi * x / tau)
    assert _axis(cam, [x])[0] == pytest.approx(ref, rel=1e-5)


def _cams():
    return [oracle.filter_camera("gaussian", 1.5, sigma=0.5), oracle.filter_camera("mitchell", 2.0),
            oracle.filter_camera("sinc", 3.0, tau=3.0), oracle.filter_camera("gaussian", 0.5)]


@pytest.mark.parametrize("cam", _cams(), ids=["gauss", "mitchell", "sinc", "gauss_default"])
def test_filter_table_is_filter_at_cell_centres(cam):
    lut = oracle.filter_table(cam).astype(np.float64)
    r = float(cam[22])
    c = (np.arange(20) + 0.5) / 20 * r
    X, Y = np.meshgrid(c, c)
    f = oracle.filter_eval(cam, np.stack([X.ravel(), Y.ravel()], 1)).reshape(20, 20).astype(np.float64)
    # rows run along y, columns along x; normalised to mean one
    assert lut.mean() == pytest.approx(1.0, rel=1e-5)
    np.testing.assert_allclose(lut * f.mean(), f, rtol=1e-5, atol=1e-7 * np.abs(f).max())


@pytest.mark.parametrize("cam", _cams(), ids=["gauss", "mitchell", "sinc", "gauss_default"])
def test_filter_importance_sampling(cam):
    from scipy import stats
    n = 400_000
    rng = np.random.default_rng(5)
    u = rng.random((n, 2), dtype=np.float32)
    p, w = oracle.filter_sample(cam, u)
    r = float(cam[22])
    lut = oracle.filter_table(cam).astype(np.float64)
    assert np.all(np.abs(p) <= r)
    # |p| lands in grid cells with probability |lut| / sum |lut|, each sign with 1/2
    q = np.minimum((np.abs(p) / r * 20).astype(int), 19)
    cell = q[:, 1] * 20 + q[:, 0]
    expect = np.abs(lut).ravel() / np.abs(lut).sum() * n
    obs = np.bincount(cell, minlength=400)
    keep = expect > 0
    assert obs[~keep].sum() == 0
    chi = ((obs[keep] - expect[keep]) ** 2 / expect[keep]).sum()
    assert stats.chi2.sf(chi, keep.sum() - 1) > 1e-4
    for axis in range(2):
        assert abs((p[:, axis] < 0).mean() - 0.5) < 4 * 0.5 / math.sqrt(n)
    # the weight is the table value over the sampled density: |w| is the
    # constant mean|lut| / mean(lut) and carries the sign of the filter
    np.testing.assert_allclose(np.abs(w), np.abs(lut).mean(), rtol=1e-5)
    np.testing.assert_array_equal(np.sign(w), np.sign(lut.ravel()[cell]))
    # estimates of the filter-weighted integral of a test function: E[w g(p)]
    # = 1/4 * integral over [-1, 1]^2 of lut(q) g(q r) dq, done cell by cell
    g = lambda x, y: 1.0 + x * x - 0.4 * np.cos(y)
    est = (w.astype(np.float64) * g(p[:, 0].astype(np.float64), p[:, 1].astype(np.float64)))
    edges = np.arange(21) / 20 * r
    # cell integrals of x^2 and cos(y) over one quadrant, mirrored by symmetry
    ix2 = (edges[1:] ** 3 - edges[:-1] ** 3) / 3
    icos = np.sin(edges[1:]) - np.sin(edges[:-1])
    h = r / 20
    cell_int = h * h + ix2[None, :] * h - 0.4 * icos[:, None] * h
    exact = (lut * cell_int).sum() / (r * r)
    assert abs(est.mean() - exact) < 5 * est.std() / math.sqrt(n) + 1e-6


def test_box_and_tent_need_no_table():
    assert np.all(oracle.filter_table(oracle.filter_camera("box", 0.5)) == 0)
    assert np.all(oracle.filter_table(oracle.filter_camera("triangle", 0.5)) == 0)


def _direct_only_scene(max_depth, env=None):
    light = {"type": "quad", "name": "lamp",
             "param": {"width": 1, "height": 1, "material": "Black",
                       "transform": scenegen.xform(scenegen.translate(0.2, 0.9, 0.1) @ scenegen.rot_x(180)
                                                   @ scenegen.scale(0.8, 1, 0.5)),
                       "emission": {"type": "area", "param": {"color": [4, 3, 2], "two_sided": False,
                                                              "scale": 1}}}}
    scene = scenegen.lit_plane(None, width=24, height=24, fov=70, cam_height=2.0, extra_shapes=[light])
    scene["integrator"]["param"]["max_depth"] = max_depth
    scene["light_sampler"] = {"type": "uniform", "param": {}}
    return scene


def _oracle_film(scene, spp):
    with va.Context(-1) as ctx:
        ex = ctx.load_text(scene).build_accel().export()
    return oracle.render(ex, 0, spp)


def test_direct_only_bounce_equals_two_bounce_direct_light():
    """A floor lit by an emitter whose own material is black: at max_depth 2
    the second bounce adds only the MIS-weighted emitter hit of the BSDF
    ray, which is what the extra ray of max_depth 1 adds, from the same
    random numbers: the two films are identical sample by sample."""
    one = _oracle_film(_direct_only_scene(1), 4)
    two = _oracle_film(_direct_only_scene(2), 4)
    assert one[..., :3].max() > 0
    np.testing.assert_array_equal(one, two)


def test_direct_only_bounce_needs_both_strategies():
    """With light sampling only (mis_mode 1) there is no extra ray, and
    with both strategies the extra ray adds the BSDF-sampled half of MIS."""
    full = _oracle_film(_direct_only_scene(1), 8)[..., :3]
    s = _direct_only_scene(1)
    s["integrator"]["param"]["mis_mode"] = 1  # MISMode::ELight
    light_only = _oracle_film(s, 8)[..., :3]
    # MIS(both) direct light < light sampling alone in each sample's NEE
    # weight, with the difference made up by the extra ray
    assert full.mean() == pytest.approx(light_only.mean(), rel=0.05)
    assert not np.array_equal(full, light_only)


def test_zero_depth_is_black():
    assert np.all(_oracle_film(_direct_only_scene(0), 2)[..., :3] == 0)


# ------------------------------------------------------------ thin lens
# ThinLensCamera::generate_ray_in_local_space (thin_lens.cpp:34-42): the
# pinhole ray is re-aimed from a point of the lens disk, square_to_disk(u)
# scaled by lens_radius, at the point where it crosses the plane of focus
# (z = focal_distance in camera space).

def _lens_box(radius, focal, filter_radius=0.5, w=32, h=32):
    s = scenegen.cornell_box(w, h, spp=4, max_depth=3)
    s["camera"]["type"] = "thin_lens"
    s["camera"]["param"]["lens_radius"] = radius
    s["camera"]["param"]["focal_distance"] = focal
    s["camera"]["param"]["filter"]["param"]["radius"] = filter_radius
    return s


def _export(scene):
    with va.Context(-1) as ctx:
        return ctx.load_text(scene).build_accel().export()


def _cam_local(cam, o, d):
    """camera-space origins and directions from the world-space rays"""
    m = np.eye(4)
    m[:3, :] = np.asarray(cam[:12], np.float64).reshape(3, 4)
    inv = np.linalg.inv(m)
    oc = (inv[:3, :3] @ o.astype(np.float64).T).T + inv[:3, 3]
    dc = (inv[:3, :3] @ d.astype(np.float64).T).T
    return oc, dc


def test_thin_lens_of_zero_radius_is_the_pinhole():
    pin = scenegen.cornell_box(32, 32, spp=4, max_depth=3)
    lens = _lens_box(0.0, 4.0)
    a, b = _export(pin), _export(lens)
    for px, py in ((3, 5), (16, 16), (31, 0)):
        oa, da, fa = oracle.camera_rays(a["camera"], px, py, 0, 64)
        ob, db, fb = oracle.camera_rays(b["camera"], px, py, 0, 64)
        np.testing.assert_array_equal(oa, ob)
        np.testing.assert_array_equal(da, db)
        np.testing.assert_array_equal(fa, fb)
    np.testing.assert_array_equal(oracle.render(a, 0, 2), oracle.render(b, 0, 2))


@pytest.mark.parametrize("radius,focal", [(0.1, 6.8), (0.3, 4.0)])
def test_thin_lens_rays_meet_on_the_plane_of_focus(radius, focal):
    # zero filter radius: every frame's sensor sample is the pixel centre,
    # so every lens sample's ray passes through the pinhole ray's point at
    # depth focal_distance, and starts on the lens disk (z = 0, |xy| <= r)
    lens = _export(_lens_box(radius, focal, filter_radius=0.0))
    pin = _export(_lens_box(0.0, focal, filter_radius=0.0))
    for px, py in ((2, 30), (16, 16), (29, 7)):
        o, d, pf = oracle.camera_rays(lens["camera"], px, py, 0, 512)
        np.testing.assert_array_equal(pf, np.tile([px + 0.5, py + 0.5], (512, 1)).astype(np.float32))
        oc, dc = _cam_local(lens["camera"], o, d)
        assert np.abs(oc[:, 2]).max() < 1e-5
        rr = np.hypot(oc[:, 0], oc[:, 1])
        assert rr.max() <= radius * (1 + 1e-5) and rr.max() > 0.9 * radius
        p = oc + ((focal - oc[:, 2]) / dc[:, 2])[:, None] * dc
        o0, d0, _ = oracle.camera_rays(pin["camera"], px, py, 0, 1)
        oc0, dc0 = _cam_local(pin["camera"], o0, d0)
        p0 = oc0 + ((focal - oc0[:, 2]) / dc0[:, 2])[:, None] * dc0
        assert np.abs(oc0).max() < 1e-5
        np.testing.assert_allclose(p, np.tile(p0, (512, 1)), atol=2e-5 * focal)


def test_thin_lens_disk_samples_are_uniform():
    from scipy import stats
    radius = 0.25
    lens = _export(_lens_box(radius, 5.0))
    n = 60_000
    o, d, _ = oracle.camera_rays(lens["camera"], 11, 19, 0, n)
    oc, _ = _cam_local(lens["camera"], o, d)
    r2 = (oc[:, 0] ** 2 + oc[:, 1] ** 2) / radius ** 2
    phi = np.mod(np.arctan2(oc[:, 1], oc[:, 0]), 2 * np.pi)
    # equal-area cells: r^2 and phi are uniform for a uniform disk density
    ir = np.minimum((r2 * 10).astype(int), 9)
    ip = np.minimum((phi / (2 * np.pi) * 12).astype(int), 11)
    obs = np.bincount(ir * 12 + ip, minlength=120)
    chi = ((obs - n / 120) ** 2 / (n / 120)).sum()
    assert stats.chi2.sf(chi, 119) > 1e-3, chi
    assert r2.max() <= 1 + 1e-5
