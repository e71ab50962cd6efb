"""The wavefront form of the path tracer against the megakernel (MI355X).

Both run path_shade's arithmetic on the same sampler streams and fold
samples into the film in frame order, so their films are bit-identical;
the megakernel is itself checked against the oracle (test_gpu_parity.py,
test_gpu_materials.py).
"""
import numpy as np
import pytest

import vision_amd as va
from vision_amd import scenegen

pytestmark = pytest.mark.gpu


def _scenes(tmp_path):
    yield "cbox", scenegen.cornell_box(64, 48, max_depth=8), "."
    import test_gpu_materials as G
    yield "glass_hero4", G._material_box("glass", "hero4"), "."
    yield "principled_hero3", G._material_box("principled", "hero3"), "."
    yield "procedural_hero", scenegen.procedural_scene(tmp_path, seed=5, n_instances=9, level=3, width=40, height=24,
                                                       spp=5, max_depth=8, spectrum="hero",
                                                       hdri=(256, 128)), str(tmp_path)


@pytest.mark.parametrize("block", [64, 256])
def test_wavefront_equals_megakernel(gpu_ctx, tmp_path, block):
    for name, scene, base in _scenes(tmp_path):
        print(f"case: wavefront {name} block {block}", flush=True)
        a = gpu_ctx.load_text(scene, base)
        a.build_accel()
        a.set_counting(True)
        a.render(5, fused=True)
        sa = a.stats()
        fa, oa = a.film(0), a.film(1)
        b = gpu_ctx.load_text(scene, base)
        b.build_accel()
        b.set_kernel("wavefront", chunk_frames=2, trace_block=block)  # chunks of 2, 2, 1 frames
        b.set_counting(True)
        b.render(3)
        b.render(2)  # progressive continuation across calls
        sb = b.stats()
        fb, ob = b.film(0), b.film(1)
        same = np.array_equal(fa, fb) and np.array_equal(oa, ob)
        print(f"{name}: films identical {same}; rays mega {sa['rays']} wave(last call) {sb['rays']}")
        assert same, name
        a.close()
        b.close()


def test_wavefront_ray_counts_match(gpu_ctx):
    scene = scenegen.cornell_box(64, 64, max_depth=8)
    res = []
    for mode in ("megakernel", "wavefront"):
        s = gpu_ctx.load_text(scene)
        s.build_accel()
        s.set_kernel(mode, chunk_frames=4)
        s.set_counting(True)
        s.render(4)
        st = s.stats()
        res.append((st["rays"], st["node_fetches"], st["tri_fetches"]))
        s.close()
    print(res)
    assert res[0] == res[1]
