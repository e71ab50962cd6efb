"""The oracle's G-buffer (src/base/frame_buffer.cpp:72-177), SVGF denoiser
(src/render_core/denoisers/SVGF/*.cpp) and convergence statistics
(src/render_core/integrators/adaptive/inspector.h:17-33) pinned by closed
forms: gradients of linear fields and of a step, demodulation followed by
modulation returning the input, temporal accumulation equal to the running
mean and then the exponential average, the a-trous passes equal to a
B-spline convolution with the variance scaled by the kernel's squared
weights, edge-stopping on depth, normal and luminance, the moments filter's
4 / history variance, the population variance of the statistics, adaptive
sampling stopping each pixel at its own count, and the G-buffer of the
Cornell box (wall colours, light radiance, back-wall depth)."""
import numpy as np
import pytest
from scipy import ndimage

import vision_amd as va
from oracle import oracle
from vision_amd import scenegen


def _plane_gbuffer(W, H, depth=2.0, normal=(0.0, 0.0, 1.0), dg=0.0):
    g = np.zeros((H * W, 8), np.float32)
    y, x = np.mgrid[0:H, 0:W]
    g[:, 0:3] = normal
    g[:, 4] = (x + 0.5).ravel()
    g[:, 5] = (y + 0.5).ravel()
    g[:, 6] = dg
    g[:, 7] = depth
    return g


def test_gradient_of_linear_fields():
    W, H = 9, 7
    y, x = np.mgrid[0:H, 0:W]
    g = np.zeros((H * W, 8), np.float32)
    a, b = 0.25, -0.5
    g[:, 7] = (a * x + b * y + 3.0).ravel()
    g[:, 0] = (0.125 * x).ravel()
    g[:, 1] = (0.0625 * y).ravel()
    g[:, 2] = 1.0
    out = oracle.gbuffer_gradient(g, W, H)
    dg = out[:, 6].reshape(H, W)
    fw = out[:, 3].reshape(H, W)
    corner = np.zeros((H, W), bool)
    corner[[0, 0, -1, -1], [0, -1, 0, -1]] = True
    # away from the corners the diagonal neighbours cancel: the slope
    np.testing.assert_allclose(dg[~corner], abs(a) + abs(b), rtol=0, atol=1e-6)
    np.testing.assert_allclose(fw[~corner], np.hypot(0.125, 0.0625), rtol=1e-6)
    # at a corner the one diagonal neighbour counts in both directions:
    # d/dx = (2a + sy b) / 2, d/dy = (2b + sx a) / 2 with the signs of the
    # diagonal offset, and the normal's fwidth grows by half
    for (yy, xx) in [(0, 0), (0, W - 1), (H - 1, 0), (H - 1, W - 1)]:
        sx = 1 if xx == 0 else -1
        sy = 1 if yy == 0 else -1
        ex = abs((2 * a + sx * sy * b) / 2) + abs((2 * b + sx * sy * a) / 2)
        assert dg[yy, xx] == pytest.approx(ex, abs=1e-6)
        assert fw[yy, xx] == pytest.approx(1.5 * np.hypot(0.125, 0.0625), rel=1e-6)
    # the other fields pass through
    np.testing.assert_array_equal(out[:, [0, 1, 2, 4, 5, 7]], g[:, [0, 1, 2, 4, 5, 7]])


def test_gradient_of_a_depth_step_along_x():
    W, H = 10, 6
    y, x = np.mgrid[0:H, 0:W]
    g = np.zeros((H * W, 8), np.float32)
    g[:, 2] = 1.0
    g[:, 7] = (x >= 4).astype(np.float32).ravel()
    out = oracle.gbuffer_gradient(g, W, H)[:, 6].reshape(H, W)
    # the two columns beside the step see half of it: (0 * k + 1 * k) / 2k;
    # in the first and last rows the one-sided vertical difference also
    # catches the diagonal neighbour across the step: 1 / 3 more
    expect = np.zeros((H, W), np.float32)
    expect[:, 3:5] = 0.5
    expect[[0, -1], 3:5] = 0.5 + 1.0 / 3.0
    np.testing.assert_allclose(out, expect, atol=1e-6)


def _frame(params, W, H, gcur, gprev, rad, alb=None, emi=None, prev=None, motion=None):
    P = W * H
    alb = np.ones((P, 4), np.float32) if alb is None else alb
    emi = np.zeros((P, 4), np.float32) if emi is None else emi
    prev = np.zeros((P, 8), np.float32) if prev is None else prev
    motion = np.zeros((P, 2), np.float32) if motion is None else motion
    return oracle.svgf_frame(params, W, H, gcur, gprev, motion, rad, alb, emi, prev)


def test_demodulate_then_modulate_returns_the_radiance():
    W, H = 12, 8
    rng = np.random.default_rng(1)
    P = W * H
    alb = np.zeros((P, 4), np.float32)
    alb[:, :3] = rng.uniform(0.05, 1.0, (P, 3))
    emi = np.zeros((P, 4), np.float32)
    emi[:, :3] = rng.uniform(0.0, 0.5, (P, 3))
    rad = np.zeros((P, 4), np.float32)
    rad[:, :3] = emi[:, :3] + rng.uniform(0.0, 2.0, (P, 3))
    alb[:5] = 0.0  # a miss: the modulated picture keeps only the emission
    g = _plane_gbuffer(W, H)
    p = oracle.svgf_params(N=0, reproject=False, filter_moments=False)
    rec, pic = _frame(p, W, H, g, g, rad, alb, emi)
    np.testing.assert_allclose(pic[5:, :3], rad[5:, :3], rtol=2e-6, atol=1e-7)
    np.testing.assert_array_equal(pic[:5, :3], emi[:5, :3])
    assert np.all(pic[:, 3] == 1.0)
    assert np.all(rec[:, 3:7] == 0.0)  # no reprojection: no variance, history or moments


def test_temporal_accumulation_is_the_running_mean_then_the_exponential_average():
    W, H = 6, 5
    P = W * H
    rng = np.random.default_rng(2)
    c = rng.uniform(0.5, 2.0, 40).astype(np.float32)
    g = _plane_gbuffer(W, H)
    p = oracle.svgf_params(N=0, filter_moments=False, alpha=0.05, moments_alpha=0.2, history_limit=32)
    prev = np.zeros((P, 8), np.float32)
    for t in range(len(c)):
        rad = np.zeros((P, 4), np.float32)
        rad[:, :3] = c[t]
        rec, pic = _frame(p, W, H, g, g, rad, prev=prev)
        prev = rec
        h = t + 1
        assert np.all(rec[:, 4] == min(32, h))
        if h <= 20:  # alpha = 1 / history: the running mean
            il = c[:h].astype(np.float64).mean()
        else:        # then the exponential average of weight 0.05
            il = 0.95 ** (h - 20) * c[:20].astype(np.float64).mean() + \
                sum(0.05 * 0.95 ** (t - k) * float(c[k]) for k in range(20, h))
        np.testing.assert_allclose(rec[:, :3], il, rtol=2e-5)
        np.testing.assert_allclose(pic[:, :3], il, rtol=2e-5)
        if h <= 5:   # moments_alpha = 1 / history: population variance of the frames
            lum = c[:h].astype(np.float64)
            np.testing.assert_allclose(rec[:, 3], lum.var(), rtol=1e-3, atol=2e-6)
            np.testing.assert_allclose(rec[:, 5], lum.mean(), rtol=2e-5)
            np.testing.assert_allclose(rec[:, 6], (lum * lum).mean(), rtol=2e-5)


def test_moments_filter_variance_after_four_frames():
    W, H = 7, 6
    P = W * H
    c = np.array([1.0, 2.0, 0.5, 1.5, 3.0], np.float32)
    g = _plane_gbuffer(W, H)
    p = oracle.svgf_params(N=0, filter_moments=True)
    prev = np.zeros((P, 8), np.float32)
    for t in range(len(c)):
        rad = np.zeros((P, 4), np.float32)
        rad[:, :3] = c[t]
        rec, _ = _frame(p, W, H, g, g, rad, prev=prev)
        prev = rec
        h = t + 1
        lum = c[:h].astype(np.float64)
        v = lum.var()
        if h < 4:  # too little history: the reprojected variance stays
            np.testing.assert_allclose(rec[:, 3], v, rtol=1e-3, atol=2e-6)
        else:      # the spatial estimate of a uniform image, scaled by 4 / history
            np.testing.assert_allclose(rec[:, 3], v * 4.0 / h, rtol=1e-3)


def _bspline(step):
    k = np.array([1 / 6, 2 / 3, 1.0, 2 / 3, 1 / 6])
    h = np.outer(k, k)
    if step == 1:
        return h
    d = np.zeros((4 * step + 1, 4 * step + 1))
    d[::step, ::step] = h
    return d


def test_atrous_passes_are_bspline_convolutions():
    W = H = 24
    P = W * H
    rng = np.random.default_rng(3)
    img = rng.uniform(0.2, 3.0, (H, W)).astype(np.float64)
    rad = np.zeros((P, 4), np.float32)
    rad[:, 0] = img.ravel()
    rad[:, 1] = 2 * img.ravel()
    rad[:, 2] = 0.5
    g = _plane_gbuffer(W, H)
    # infinite luminance sigma, zero normal exponent, flat depth: only the kernel weighs
    p = oracle.svgf_params(N=2, reproject=False, filter_moments=False, sigma_rt=1e30, sigma_normal=0.0)
    _, pic = _frame(p, W, H, g, g, rad)
    ref = img
    for step in (1, 2):
        h = _bspline(step)
        ref = ndimage.correlate(ref, h, mode="constant") / ndimage.correlate(np.ones_like(ref), h, mode="constant")
    inner = np.s_[6:H - 6, 6:W - 6]
    np.testing.assert_allclose(pic[:, 0].reshape(H, W)[inner], ref[inner], rtol=2e-6)
    np.testing.assert_allclose(pic[:, 1].reshape(H, W)[inner], 2 * ref[inner], rtol=2e-6)
    np.testing.assert_allclose(pic[:, 2], 0.5, rtol=2e-6)


def test_atrous_scales_uniform_variance_by_the_squared_weights():
    W = H = 16
    P = W * H
    g = _plane_gbuffer(W, H)
    p = oracle.svgf_params(N=1, filter_moments=False, sigma_normal=0.0)
    prev = np.zeros((P, 8), np.float32)
    for cval in (1.0, 3.0):  # two frames: a uniform variance of 1
        rad = np.zeros((P, 4), np.float32)
        rad[:, :3] = cval
        prev, _ = _frame(p, W, H, g, g, rad, prev=prev)
    h = _bspline(1)
    expect = (h ** 2).sum() / h.sum() ** 2
    var = prev[:, 3].reshape(H, W)
    np.testing.assert_allclose(var[2:-2, 2:-2], 1.0 * expect, rtol=1e-5)
    np.testing.assert_allclose(prev[:, :3], 2.0, rtol=1e-6)


@pytest.mark.parametrize("edge", ["depth", "normal", "luminance"])
def test_edges_stop_the_filter(edge):
    W, H = 16, 12
    P = W * H
    y, x = np.mgrid[0:H, 0:W]
    left = (x < 8).ravel()
    g = _plane_gbuffer(W, H, dg=0.01)
    big = {"sigma_rt": 1e30, "sigma_normal": 30.0}
    if edge == "depth":
        g[~left, 7] = 5.0
        big["sigma_normal"] = 0.0
    elif edge == "normal":
        g[~left, 0:3] = (1.0, 0.0, 0.0)
    else:
        big = {"sigma_rt": 10.0, "sigma_normal": 0.0}
    rad = np.zeros((P, 4), np.float32)
    rad[left, :3] = 1.0
    rad[~left, :3] = 3.0
    p = oracle.svgf_params(N=3, reproject=False, filter_moments=False, **big)
    _, pic = _frame(p, W, H, g, g, rad)
    np.testing.assert_allclose(pic[:, :3], rad[:, :3], rtol=1e-6)
    # without the edge the two sides blend
    if edge != "luminance":
        g2 = _plane_gbuffer(W, H, dg=0.01)
        _, mixed = _frame(p, W, H, g2, g2, rad)
        assert np.abs(mixed[:, 0] - rad[:, 0]).max() > 0.1


def test_variance_stats_are_the_population_variance():
    rng = np.random.default_rng(4)
    x = rng.uniform(0.0, 2.0, (50, 300)).astype(np.float32)
    st = np.zeros((300, 4), np.float32)
    for k in range(len(x)):
        oracle.variance_stats_add(st, x[k])
    np.testing.assert_allclose(st[:, 0], x.astype(np.float64).mean(0), rtol=1e-5)
    np.testing.assert_allclose(st[:, 1], x.astype(np.float64).var(0), rtol=1e-4)
    assert np.all(st[:, 2].view(np.uint32) == 50)


def _export(scene, adaptive=None):
    with va.Context(-1) as ctx:
        sc = ctx.load_text(scene).build_accel()
        if adaptive is not None:
            sc.set_adaptive(True, *adaptive)
        return sc.export()


def test_adaptive_sampling_stops_each_pixel_at_its_own_count():
    scene = scenegen.cornell_box(24, 20, max_depth=3)
    ex = _export(scene, adaptive=(0.5, 4))
    assert ex["render"][9] == 1.0
    P = 24 * 20
    st = np.zeros((P, 4), np.float32)
    film = oracle.render(ex, 0, 24, vstats=st)
    n = st[:, 2].view(np.uint32)
    assert n.min() >= 5 and n.min() < 24  # some pixels converged early
    # every pixel's film is the plain render of its own sample count
    plain = _export(scene)
    for k in np.unique(n):
        ref = oracle.render(plain, 0, int(k))
        sel = (n == k).reshape(20, 24)
        np.testing.assert_array_equal(film[sel], ref[sel])
    # a stopped pixel satisfied the criterion; one still running did not
    stopped = n < 24
    assert np.all(st[stopped, 1] < 0.5 * st[stopped, 0])


def test_cornell_gbuffer_closed_forms():
    W, H = 48, 40
    scene = scenegen.cornell_box(W, H)
    scene["integrator"] = scenegen.restir_integrator(M_light=2)
    with va.Context(-1) as ctx:
        sc = ctx.load_text(scene).build_accel()
        sc.set_denoiser(True)
        ex = sc.export()
    assert ex["restir"][48] == 1.0
    orc = oracle.Restir(ex)
    orc.render(1)
    cur = 1
    geo, alb, emi = orc.geo[cur], orc.alb, orc.emi
    n = geo[:, :3]
    hit = geo[:, 7] > 0
    assert hit.sum() > 1200
    np.testing.assert_allclose(np.linalg.norm(n[hit], axis=1), 1.0, rtol=1e-5)
    # the back wall (normal +z): gray albedo, depth 7.8 along the view axis
    back = hit & (n[:, 2] > 0.999)
    assert back.sum() > 200
    assert np.all(alb[back, :3] == np.float32([0.73, 0.73, 0.73]))
    np.testing.assert_allclose(geo[back, 7], 7.8, rtol=1e-5)
    # the left (red) and right (green) walls
    red = hit & (n[:, 0] > 0.999)
    green = hit & (n[:, 0] < -0.999)
    assert red.sum() > 20 and green.sum() > 20
    assert np.all(alb[red, :3] == np.float32([0.65, 0.05, 0.05]))
    assert np.all(alb[green, :3] == np.float32([0.12, 0.45, 0.15]))
    # the light seen from below: its radiance, a black albedo
    lit = emi[:, 0] > 0
    assert lit.sum() > 5
    assert np.all(emi[lit, :3] == np.float32([17, 12, 4]))
    assert np.all(alb[lit, :3] == 0)
    # flat walls: no depth gradient inside the back wall away from its edges
    bw = back.reshape(H, W)
    inner = bw & ndimage.binary_erosion(bw, np.ones((3, 3)))
    assert np.abs(geo[:, 6].reshape(H, W)[inner]).max() < 1e-4


def test_svgf_smooths_the_cornell_box_back_wall():
    # ReSTIR DI frames of the Cornell box, the film not accumulating (the
    # real-time setting), with the denoiser on: inside the back wall (flat,
    # one albedo) the denoised picture varies much less from pixel to pixel
    # than the frame it was computed from
    W, H = 64, 48
    scene = scenegen.cornell_box(W, H, max_depth=2, light_sampler="power")
    scene["camera"]["param"]["radiance_collector"]["param"]["accumulation"] = 0
    scene["integrator"] = scenegen.restir_integrator(M_light=2, M_bsdf=1)
    with va.Context(-1) as ctx:
        sc = ctx.load_text(scene).build_accel()
        sc.set_denoiser(True)
        ex = sc.export()
    orc = oracle.Restir(ex)
    orc.render(8)
    cur = orc.frame & 1
    n = orc.geo[cur][:, :3]
    back = (orc.geo[cur][:, 7] > 0) & (n[:, 2] > 0.999)
    inner = ndimage.binary_erosion(back.reshape(H, W), np.ones((3, 3)))
    film = orc.film.reshape(H, W, 4)[..., :3]
    pic = orc.picture.reshape(H, W, 4)[..., :3]
    pair = inner[:, 1:] & inner[:, :-1]

    def rough(img):
        return np.abs(np.diff(img, axis=1))[pair].mean()
    print(f"neighbour differences on the back wall: film {rough(film):.4f}, denoised {rough(pic):.4f}")
    assert rough(pic) < 0.5 * rough(film)
    # and it keeps the wall's brightness
    assert abs(pic[inner].mean() / film[inner].mean() - 1) < 0.05
