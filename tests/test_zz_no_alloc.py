"""vs_trace_closest_noalloc (the staging-array variant of vs_trace_closest
that allocates no device memory) against vs_trace_closest on the MI355X:
bit-identical hits on two calls in a row, the second smaller, so the
second reads a staging array that still holds the first call's rays."""
import tempfile

import numpy as np
import pytest

import vision_amd as va
from oracle import oracle

pytestmark = pytest.mark.gpu

MAX_RAYS = 1 << 18


def test_noalloc_trace_equals_trace_closest(gpu_ctx):
    import test_gpu_parity as P
    with tempfile.TemporaryDirectory() as d:
        sc = gpu_ctx.load_text(P._soup_scene(3000, 7, d), base_dir=d)
        sc.build_accel()
    rng = np.random.default_rng(11)
    for n in (MAX_RAYS, 40000):
        print(f"case: noalloc trace n={n}", flush=True)
        rays = P._rays(rng, n)
        a = sc.trace_closest_noalloc(rays)
        b = sc.trace_closest(rays)
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), n
        assert (a["tri"] != 0xFFFFFFFF).mean() > 0.3
    # the smaller call's hits against the oracle's brute force
    ex = sc.export()
    k = slice(0, 5000)
    tri, t, _ = oracle.trace(ex["positions"], ex["triangles"], P._as_oracle(rays[k]), bvh=False)
    assert np.array_equal(a["tri"][k], tri)
    hit = tri != 0xFFFFFFFF
    assert np.array_equal(a["t"][k][hit], t[hit])
    # beyond the staging array's capacity: a clear error, nothing launched
    big = np.zeros(MAX_RAYS + 1, dtype=va.RAY_DTYPE)
    with pytest.raises(va.VisionError, match="at most 262144 rays"):
        sc.trace_closest_noalloc(big)
    assert len(sc.trace_closest_noalloc(big[:0])) == 0
    sc.close()
