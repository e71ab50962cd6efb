"""The oracle's radiance cache (oracle/nrc.cpp) pinned without the GPU:
level resolutions by their closed form, the encoding by exact trilinear
reproduction of linear functions on the directly indexed levels and by the
published hash on the others, the network against an independent numpy
forward, its backward against central finite differences, the rounded
(device) mode against the exact one within BF16 precision, and Adam's first
step by its closed form (each parameter moves by lr in the direction
opposite its gradient)."""
import numpy as np
import pytest

from oracle import oracle

T = 1 << 19


def _cfg(min_res=16, max_res=2048, lo=(-1.0, 0.0, -1.0), ext=(2.0, 2.0, 2.0)):
    res, mask = oracle.nrc_levels(min_res, max_res)
    return oracle.nrc_cfg(res, mask, lo, [1.0 / e for e in ext]), res, mask


def test_levels_are_geometric():
    res, mask = oracle.nrc_levels(16, 2048)
    assert res[0] == 16 and res[-1] == 2048
    b = (2048 / 16) ** (1 / 15)
    np.testing.assert_array_equal(res, np.floor(16 * b ** np.arange(16) + 1e-9).astype(np.uint32))
    for l in range(16):
        assert bool(mask & (1 << l)) == ((int(res[l]) + 1) ** 3 <= T)
    res, mask = oracle.nrc_levels(4, 4)
    assert np.all(res == 4) and mask == 0xFFFF


def _grid_linear(res, mask, a, b):
    """Grid whose directly indexed levels hold feature0 = a . (vertex / res)
    and feature1 = b . (vertex / res)."""
    g = np.zeros((16, T, 2), np.float32)
    for l in range(16):
        if not mask & (1 << l):
            continue
        n = int(res[l]) + 1
        z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        idx = (x + n * (y + n * z)).reshape(-1)
        v = np.stack([x, y, z], -1).reshape(-1, 3) / res[l]
        g[l, idx, 0] = v @ a
        g[l, idx, 1] = v @ b
    return g.reshape(-1)


def test_encoding_reproduces_linear_functions_on_direct_levels():
    cfg, res, mask = _cfg()
    a, b = np.array([0.3, -0.7, 0.2]), np.array([-0.1, 0.05, 0.9])
    grid = _grid_linear(res, mask, a, b)
    rng = np.random.default_rng(0)
    pos = rng.uniform([-1, 0, -1], [1, 2, 1], (200, 3)).astype(np.float32)
    x = oracle.nrc_encode(cfg, grid, pos, round=False)
    u = (pos.astype(np.float64) - [-1, 0, -1]) / 2.0
    for l in range(16):
        if mask & (1 << l):
            np.testing.assert_allclose(x[:, 2 * l], u @ a, atol=2e-6)
            np.testing.assert_allclose(x[:, 2 * l + 1], u @ b, atol=2e-6)
        else:
            assert np.all(x[:, 2 * l] == 0)


def test_encoding_hash_and_partition_of_unity():
    cfg, res, mask = _cfg()
    # constant grid: every level's blend is that constant (weights sum to 1)
    grid = np.full(16 * T * 2, 0.25, np.float32)
    rng = np.random.default_rng(1)
    pos = rng.uniform([-1, 0, -1], [1, 2, 1], (100, 3)).astype(np.float32)
    np.testing.assert_allclose(oracle.nrc_encode(cfg, grid, pos, round=False), 0.25, atol=1e-7)
    # a position on a grid vertex of a hashed level reads exactly that entry
    l = 15
    assert not mask & (1 << l)
    vx, vy, vz = 1001, 517, 12
    p = np.array([[-1 + 2 * vx / res[l], 2 * vy / res[l], -1 + 2 * vz / res[l]]], np.float32)
    u = (p.astype(np.float64)[0] - [-1, 0, -1]) / 2 * res[l]
    cell = np.minimum(np.floor(u.astype(np.float32)), res[l] - 1).astype(np.int64)
    assert np.all(np.abs(u - cell) < 1e-3)
    h = ((int(cell[0]) * 1) ^ (int(cell[1]) * 2654435761) ^ (int(cell[2]) * 805459861)) & 0xFFFFFFFF & (T - 1)
    g = np.zeros((16, T, 2), np.float32)
    g[l, h] = [3.0, -2.0]
    x = oracle.nrc_encode(cfg, g.reshape(-1), p, round=False)
    np.testing.assert_allclose(x[0, 2 * l:2 * l + 2], [3.0, -2.0], rtol=2e-3)


def _params(seed=0, grid_scale=1.0):
    rng = np.random.default_rng(seed)
    mlp = np.concatenate([rng.uniform(-a, a, n) for a, n in
                          [(np.sqrt(6 / 96), 32 * 64), (np.sqrt(6 / 128), 64 * 64), (np.sqrt(6 / 128), 64 * 64),
                           (np.sqrt(6 / 67), 64 * 3)]]).astype(np.float32)
    grid = rng.uniform(-grid_scale, grid_scale, 16 * T * 2).astype(np.float32)
    return mlp, grid


def _numpy_forward(x, mlp):
    w1 = mlp[:2048].reshape(32, 64).astype(np.float64)
    w2 = mlp[2048:6144].reshape(64, 64).astype(np.float64)
    w3 = mlp[6144:10240].reshape(64, 64).astype(np.float64)
    w4 = mlp[10240:].reshape(64, 3).astype(np.float64)
    h = np.maximum(x @ w1, 0)
    h = np.maximum(h @ w2, 0)
    h = np.maximum(h @ w3, 0)
    return h @ w4


def test_forward_matches_numpy():
    cfg, _, _ = _cfg()
    mlp, grid = _params(2)
    rng = np.random.default_rng(3)
    pos = rng.uniform([-1, 0, -1], [1, 2, 1], (64, 3)).astype(np.float32)
    x = oracle.nrc_encode(cfg, grid, pos, round=False)
    np.testing.assert_allclose(oracle.nrc_forward(cfg, mlp, grid, pos, round=False), _numpy_forward(x, mlp),
                               rtol=1e-12, atol=1e-12)
    # the device-rounded form stays within bf16 precision of the exact one
    r = oracle.nrc_forward(cfg, mlp, grid, pos, round=True)
    e = oracle.nrc_forward(cfg, mlp, grid, pos, round=False)
    assert np.linalg.norm(r - e) / np.linalg.norm(e) < 3e-2


def _samples(rng, n, lo=(-1, 0, -1), hi=(1, 2, 1)):
    p = rng.uniform(lo, hi, (n, 3))
    t = rng.uniform(0, 2, (n, 3))
    return np.concatenate([p, t], 1).astype(np.float32)


def test_backward_matches_finite_differences():
    cfg, _, _ = _cfg(min_res=4, max_res=64)
    mlp, grid = _params(4, grid_scale=0.5)
    rng = np.random.default_rng(5)
    sm = _samples(rng, 8)
    loss, gm, gg = oracle.nrc_backward(cfg, mlp, grid, sm, round=False)

    def loss_at(m, g):
        y = oracle.nrc_forward(cfg, m, g, sm[:, :3], round=False)
        t = sm[:, 3:].astype(np.float64)
        # the loss's denominator holds y constant: its derivative treats
        # 1 / (y^2 + 0.01) as a weight, so compare against that weighted form
        return y, t

    y0, t = loss_at(mlp, grid)
    den = y0 ** 2 + 0.01
    assert loss == pytest.approx(np.mean((y0 - t) ** 2 / den), rel=1e-10)

    def wloss(m, g):
        y = oracle.nrc_forward(cfg, m, g, sm[:, :3], round=False)
        return np.mean((y - t) ** 2 / den)

    idx = list(rng.choice(10432, 24, replace=False)) + [10240, 10300, 10431, 0, 2047]
    for i in idx:
        # float32 parameters: step in float32, difference in float64
        h = 1e-2 * max(1e-3, abs(float(mlp[i])))
        mp, mm = mlp.copy(), mlp.copy()
        mp[i] += np.float32(h)
        mm[i] -= np.float32(h)
        fd = (wloss(mp, grid) - wloss(mm, grid)) / (float(mp[i]) - float(mm[i]))
        assert gm[i] == pytest.approx(fd, rel=2e-3, abs=1e-7), i
    touched = np.flatnonzero(gg)
    assert len(touched) > 100
    for j in rng.choice(touched, 16, replace=False):
        gp, gmi = grid.copy(), grid.copy()
        # a small step: the network is piecewise linear, a large one can cross a ReLU kink
        gp[j] += np.float32(2e-4)
        gmi[j] -= np.float32(2e-4)
        fd = (wloss(mlp, gp) - wloss(mlp, gmi)) / (float(gp[j]) - float(gmi[j]))
        assert gg[j] == pytest.approx(fd, rel=5e-3, abs=1e-6), j


def test_adam_first_step_moves_by_lr_against_the_gradient():
    rng = np.random.default_rng(6)
    p = rng.normal(size=100).astype(np.float32)
    g = rng.normal(size=100).astype(np.float32)
    g[:5] = 0
    m = np.zeros(100, np.float32)
    v = np.zeros(100, np.float32)
    p0 = p.copy()
    oracle.nrc_adam(p, m, v, g.copy(), 0.01, 1)
    np.testing.assert_allclose(p[5:], p0[5:] - 0.01 * np.sign(g[5:]), atol=1e-6)
    np.testing.assert_array_equal(p[:5], p0[:5])
    # second step with the same gradient: m = 0.19 g, v = 0.0199 g^2 -> step lr * sign(g)
    # after bias correction (c1 = 1/0.19, c2 = 1/0.0199)
    p1 = p.copy()
    gg = g.copy()
    oracle.nrc_adam(p, m, v, gg, 0.01, 2)
    np.testing.assert_allclose(p[5:], p1[5:] - 0.01 * np.sign(g[5:]), atol=1e-6)
    assert np.all(gg == 0)
