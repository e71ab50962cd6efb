"""The radiance cache on the GPU against the oracle (oracle/nrc.cpp in its
device-rounded mode): level layout, forward outputs, loss and gradients of
the hand-written MFMA forward/backward kernels, the Adam step; the plain
fallback bit for bit against vs_render; and what the trained cache gains
over the plain path tracer.

Tolerances: the device sums each 16-term MFMA block in its own order and
rounds activations to BF16 (8 bits of mantissa), so a rounding can flip
against the oracle's; forward outputs agree within 2% (relative to the
output scale), gradients within 5% in norm per layer."""
import json

import numpy as np
import pytest

import vision_amd as va
from oracle import oracle
from vision_amd import scenegen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cbox(gpu_ctx):
    sc = gpu_ctx.load_text(json.dumps(scenegen.cornell_box(64, 64, max_depth=8))).build_accel()
    yield sc
    sc.close()


def _oracle_cfg(nrc):
    lv = nrc.levels()
    return oracle.nrc_cfg(lv["res"], lv["dense_mask"], lv["lo"], lv["inv_ext"])


def _random_params(seed):
    rng = np.random.default_rng(seed)
    mlp = np.concatenate([rng.uniform(-a, a, n) for a, n in
                          [(np.sqrt(6 / 96), 32 * 64), (np.sqrt(6 / 128), 64 * 64), (np.sqrt(6 / 128), 64 * 64),
                           (np.sqrt(6 / 67), 64 * 3)]]).astype(np.float32)
    grid = rng.uniform(-1, 1, va.NRC_GRID_PARAMS).astype(np.float32)
    return mlp, grid


def _positions(rng, n):
    return rng.uniform([-1, 0, -1], [1, 2, 1], (n, 3)).astype(np.float32)


def test_levels_match_the_oracle(cbox):
    with va.RadianceCache(cbox) as nrc:
        lv = nrc.levels()
        res, mask = oracle.nrc_levels(16, 2048)
        np.testing.assert_array_equal(lv["res"], res)
        assert lv["dense_mask"] == mask


def test_forward_matches_the_oracle(cbox):
    rng = np.random.default_rng(1)
    with va.RadianceCache(cbox) as nrc:
        mlp, grid = _random_params(2)
        nrc.set_params(mlp, grid)
        pos = _positions(rng, 1000)
        got = nrc.forward(pos)
        ref = oracle.nrc_forward(_oracle_cfg(nrc), mlp, grid, pos, round=True)
        scale = np.abs(ref).max()
        print(f"forward: max |gpu - oracle| = {np.abs(got - ref).max():.3e}, scale {scale:.3e}")
        np.testing.assert_allclose(got, ref, rtol=0, atol=2e-2 * scale)
        assert np.mean(np.abs(got - ref) <= 1e-3 * scale) > 0.9


def test_gradients_match_the_oracle(cbox):
    rng = np.random.default_rng(3)
    with va.RadianceCache(cbox) as nrc:
        mlp, grid = _random_params(4)
        nrc.set_params(mlp, grid)
        pos = _positions(rng, 300)
        t = rng.uniform(0, 2, (300, 3)).astype(np.float32)
        sm = np.concatenate([pos, t], 1)
        loss, gm, gg = nrc.gradients(sm)
        rl, rm, rg = oracle.nrc_backward(_oracle_cfg(nrc), mlp, grid, sm, round=True)
        print(f"loss gpu {loss:.6f} oracle {rl:.6f}")
        assert loss == pytest.approx(rl, rel=2e-2)
        for name, a, b in [("W1", 0, 2048), ("W2", 2048, 6144), ("W3", 6144, 10240), ("W4", 10240, 10432)]:
            err = np.linalg.norm(gm[a:b] - rm[a:b]) / np.linalg.norm(rm[a:b])
            print(f"{name}: relative gradient error {err:.3e}")
            assert err < 1e-5
        touched = rg != 0
        assert np.all(gg[~touched] == 0)
        err = np.linalg.norm(gg[touched] - rg[touched]) / np.linalg.norm(rg[touched])
        print(f"grid: relative gradient error {err:.3e} over {touched.sum()} entries")
        assert err < 1e-5


def test_adam_first_step(cbox):
    rng = np.random.default_rng(5)
    with va.RadianceCache(cbox, learning_rate=1e-3) as nrc:
        mlp, grid = _random_params(6)
        nrc.set_params(mlp, grid)
        pos = _positions(rng, 200)
        sm = np.concatenate([pos, rng.uniform(0, 2, (200, 3)).astype(np.float32)], 1)
        _, rm, rg = oracle.nrc_backward(_oracle_cfg(nrc), mlp, grid, sm, round=True)
        nrc.train(sm, steps=1)
        m2, g2 = nrc.params()
        # the first Adam step moves every parameter with a gradient by lr
        # against it (bias-corrected m / sqrt(v) = sign(g)); parameters with
        # no gradient stay put
        for p0, p1, g in [(mlp, m2, rm), (grid, g2, rg)]:
            clear = np.abs(g) > 1e-3 * np.abs(g).max()
            np.testing.assert_allclose(p1[clear], p0[clear] - 1e-3 * np.sign(g[clear]), atol=2e-6)
            np.testing.assert_array_equal(p1[g == 0], p0[g == 0])


def test_plain_fallback_equals_vs_render(gpu_ctx):
    scene = json.dumps(scenegen.cornell_box(48, 40, max_depth=6))
    a = gpu_ctx.load_text(scene).build_accel()
    b = gpu_ctx.load_text(scene).build_accel()
    a.render(8, fused=True)
    with va.RadianceCache(b) as nrc:
        nrc.render(2, 4, terminate=False, train=False)
    np.testing.assert_array_equal(a.film(0).view(np.uint32), b.film(0).view(np.uint32))
    np.testing.assert_array_equal(a.film(1).view(np.uint32), b.film(1).view(np.uint32))
    a.close()
    b.close()


def test_training_lowers_the_loss(gpu_ctx):
    sc = gpu_ctx.load_text(json.dumps(scenegen.cornell_box(96, 96, max_depth=8))).build_accel()
    with va.RadianceCache(sc, train_paths=4096) as nrc:
        first = nrc.render(1, 1)["last_loss"]
        nrc.render(40, 1)
        last = nrc.render(1, 1)["last_loss"]
        st = nrc.stats()
    print(f"loss {first:.4f} -> {last:.4f}; {st}")
    assert st["queries_last_frame"] > 0.5 * 96 * 96
    assert st["train_samples_last_frame"] > 4096
    assert last < 0.7 * first
    sc.close()


def test_trained_cache_gains_over_the_plain_path_tracer(gpu_ctx):
    """The cbox at 1920x1080, 4 spp x 32 frames, against a 4096-spp plain
    render: at equal samples the cache's film is closer to the reference
    (its paths end at bounce 2 in the cache's mean instead of a noisy
    continuation), and its path kernel costs under 60% of the plain one."""
    scene = json.dumps(scenegen.cornell_box(1920, 1080, max_depth=8))
    ref = gpu_ctx.load_text(scene).build_accel()
    for _ in range(8):
        ref.render(512, fused=True)
    R = ref.film(0)[..., :3].astype(np.float64)
    ref.close()

    def relmse(sc):
        x = sc.film(0)[..., :3].astype(np.float64)
        return float(np.mean((x - R) ** 2 / (R * R + 1e-2)))

    a = gpu_ctx.load_text(scene).build_accel()
    with va.RadianceCache(a) as nrc:
        plain = nrc.render(32, 4, terminate=False, train=False)
        e_plain = relmse(a)
    b = gpu_ctx.load_text(scene).build_accel()
    with va.RadianceCache(b) as nrc:
        nrc.render(8, 4)  # the cache learns on 8 frames first
        b.reset_film()
        cached = nrc.render(32, 4)
        e_cache = relmse(b)
    print(f"relMSE plain {e_plain:.3e} cache {e_cache:.3e}; path kernel {cached['render_ms']:.1f} ms vs "
          f"{plain['render_ms']:.1f} ms; stages {cached}")
    assert e_cache < 0.95 * e_plain
    assert cached["render_ms"] < 0.6 * plain["render_ms"]
    a.close()
    b.close()
