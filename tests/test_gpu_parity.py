"""HIP path vs the CPU oracle on the same inputs (MI355X).

Traversal: identical rays give bit-identical (triangle, t, barycentrics).
Rendering: identical sampler streams give per-pixel radiance within 1e-4
relative (the kernels and the oracle share their elementary functions, so
every sampling decision takes the same branch).
"""
import struct
import tempfile
from pathlib import Path

import numpy as np
import pytest

import vision_amd as va
from oracle import oracle
from vision_amd import scenegen

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4
ABS_FLOOR = 1e-6


def _soup_scene(n, seed, d):
    rng = np.random.default_rng(seed)
    # a triangle soup inside the Cornell box ([-1, 1] x [0, 2] x [-1, 1])
    c = rng.random((n, 1, 3)) * 1.6 - 0.8 + np.array([0, 1, 0])
    tri = (c + rng.normal(scale=0.05, size=(n, 3, 3))).reshape(-1, 3).astype(np.float32)
    idx = np.arange(3 * n, dtype=np.uint32)
    fn = Path(d) / "soup.vsmesh"
    with open(fn, "wb") as f:
        f.write(b"VSMESH01" + struct.pack("<III", 3 * n, n, 0) + tri.tobytes() + idx.tobytes())
    scene = scenegen.cornell_box(8, 8)
    scene["shapes"].append({"type": "model", "name": "soup", "param": {"fn": "soup.vsmesh", "material": "Gray"}})
    return scene


def _rays(rng, n, lo=-0.95, hi=0.95):
    # origins inside the box: most rays hit a wall or the soup
    r = np.zeros(n, va.RAY_DTYPE)
    r["org"] = rng.random((n, 3)) * (hi - lo) + lo + np.array([0, 1, 0])
    d = rng.normal(size=(n, 3))
    r["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r["t_min"] = 0
    r["t_max"] = np.finfo(np.float32).max
    return r


def _as_oracle(r):
    out = np.zeros((len(r), 8), np.float32)
    out[:, 0:3] = r["org"]
    out[:, 3] = r["t_min"]
    out[:, 4:7] = r["dir"]
    out[:, 7] = r["t_max"]
    return out


@pytest.mark.parametrize("n_tris,seed", [(1, 0), (50, 1), (20000, 2)])
def test_trace_closest_bit_exact(gpu_ctx, n_tris, seed):
    with tempfile.TemporaryDirectory() as d:
        sc = gpu_ctx.load_text(_soup_scene(n_tris, seed, d), base_dir=d)
        sc.build_accel()
    ex = sc.export()
    rng = np.random.default_rng(100 + seed)
    rays = _rays(rng, 50000)
    hits = sc.trace_closest(rays)
    tri, t, bary = oracle.trace(ex["positions"], ex["triangles"], _as_oracle(rays), bvh=False)
    assert np.array_equal(hits["tri"], tri)
    hit = tri != 0xFFFFFFFF
    assert hit.mean() > 0.3
    assert np.array_equal(hits["t"][hit], t[hit])
    assert np.array_equal(hits["bary"][hit], bary[hit])
    # inst / prim decode back to the scene-global triangle
    assert np.array_equal(ex["instances"][hits["inst"][hit], 2] + hits["prim"][hit], tri[hit])
    assert (hits["inst"][~hit] == 0xFFFFFFFF).all()
    sc.close()


def test_trace_any_and_bounded_rays(gpu_ctx):
    with tempfile.TemporaryDirectory() as d:
        sc = gpu_ctx.load_text(_soup_scene(5000, 3, d), base_dir=d)
        sc.build_accel()
    ex = sc.export()
    rng = np.random.default_rng(7)
    rays = _rays(rng, 40000)
    rays["t_max"] = rng.random(len(rays)) * 3
    rays["t_min"] = rng.random(len(rays)) * 0.2
    occ = sc.trace_any(rays)
    tri, _, _ = oracle.trace(ex["positions"], ex["triangles"], _as_oracle(rays), any_hit=True, bvh=False)
    assert np.array_equal(occ.astype(bool), tri != 0xFFFFFFFF)
    hits = sc.trace_closest(rays)
    tri2, t2, _ = oracle.trace(ex["positions"], ex["triangles"], _as_oracle(rays), bvh=False)
    assert np.array_equal(hits["tri"], tri2)
    ok = tri2 != 0xFFFFFFFF
    assert ((hits["t"][ok] > rays["t_min"][ok]) & (hits["t"][ok] < rays["t_max"][ok])).all()
    sc.close()


def test_watertight_edges_and_vertices(gpu_ctx):
    sc = gpu_ctx.load_text(scenegen.furnace_box())
    sc.build_accel()
    ex = sc.export()
    verts = ex["positions"]
    rng = np.random.default_rng(9)
    rays = np.zeros(len(verts) * 64, va.RAY_DTYPE)
    rays["org"] = rng.random((len(rays), 3)) * 1.6 - 0.8
    # aim exactly at vertices and at edge midpoints
    T = ex["triangles"]
    targets = np.concatenate([verts, 0.5 * (verts[T[:, 0]] + verts[T[:, 1]])])
    rays["dir"] = targets[rng.integers(0, len(targets), len(rays))] - rays["org"]
    rays["t_max"] = np.finfo(np.float32).max
    hits = sc.trace_closest(rays)
    assert (hits["tri"] != 0xFFFFFFFF).all()
    tri, t, _ = oracle.trace(verts, T, _as_oracle(rays), bvh=False)
    assert np.array_equal(hits["tri"], tri) and np.array_equal(hits["t"], t)
    sc.close()


def _compare(gpu, ref):
    g, r = gpu[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    err = np.abs(g - r)
    ok = err <= REL_TOL * np.maximum(np.abs(g), np.abs(r)) + ABS_FLOOR
    pix_ok = ok.all(axis=-1)
    return pix_ok.mean(), err.max()


@pytest.mark.parametrize("light_sampler", ["uniform", "power"])
def test_cornell_box_per_pixel_parity(gpu_ctx, light_sampler):
    scene = scenegen.cornell_box(128, 128, light_sampler=light_sampler)
    sc = gpu_ctx.load_text(scene)
    sc.build_accel()
    sc.render(2, fused=True)
    gpu = sc.film(0)
    ref = oracle.render(sc.export(), 0, 2)
    frac, _ = _compare(gpu, ref)
    print(f"cbox {light_sampler}: {frac * 100:.4f}% of pixels within {REL_TOL} relative")
    assert frac >= 0.999
    assert np.abs(gpu[..., :3].mean() / ref[..., :3].mean() - 1) < 1e-3
    sc.close()


def test_fused_and_per_frame_films_identical(gpu_ctx):
    scene = scenegen.cornell_box(64, 64)
    a = gpu_ctx.load_text(scene)
    a.build_accel()
    a.render(5, fused=True)
    fa = a.film(0)
    b = gpu_ctx.load_text(scene)
    b.build_accel()
    b.render(5, fused=False)
    fb = b.film(0)
    assert np.array_equal(fa, fb)
    # progressive continuation: 3 + 2 frames equal 5 frames
    c = gpu_ctx.load_text(scene)
    c.build_accel()
    c.render(3)
    c.render(2)
    assert np.array_equal(c.film(0), fa)
    assert c.info()["frame_index"] == 5
    for s in (a, b, c):
        s.close()


def test_gpu_furnace_closed_form(gpu_ctx):
    sc = gpu_ctx.load_text(scenegen.furnace_box(albedo=(0.7, 0.5, 0.2), width=64, height=64))
    sc.build_accel()
    sc.render(64)
    mean = sc.film(0)[..., :3].mean(axis=(0, 1))
    expect = 1 / (1 - np.array([0.7, 0.5, 0.2]))
    assert np.allclose(mean, expect, rtol=0.01), (mean, expect)
    sc.close()


def test_point_and_spot_lights_parity(gpu_ctx):
    lights = [{"type": "point", "param": {"position": [0.3, 0.7, 0.1], "color": [3, 2, 1], "scale": 1.0}},
              {"type": "spot", "param": {"position": [0.0, 1.0, 0.0], "direction": [0.1, -1, 0], "angle": 20,
                                         "falloff": 2, "color": [1, 1, 1], "scale": 2.0}}]
    scene = scenegen.lit_plane(lights[0], width=32, height=32, fov=60)
    scene["light_sampler"]["param"]["lights"] = lights
    scene["light_sampler"]["type"] = "power"
    sc = gpu_ctx.load_text(scene)
    sc.build_accel()
    sc.render(4)
    frac, _ = _compare(sc.film(0), oracle.render(sc.export(), 0, 4))
    assert frac >= 0.999
    sc.close()


def test_ray_counters(gpu_ctx):
    sc = gpu_ctx.load_text(scenegen.cornell_box(64, 64))
    sc.build_accel()
    sc.set_counting(True)
    st = sc.render(2)
    assert st["rays"] >= 64 * 64 * 2
    assert st["node_fetches"] >= st["rays"]
    sc.set_counting(False)
    sc.close()


@pytest.mark.parametrize("filt,param", [("gaussian", {"radius": 1.5, "sigma": 0.5}), ("mitchell", {"radius": 2.0}),
                                        ("sinc", {"radius": 3.0, "tau": 3.0})])
def test_fitted_filter_parity(gpu_ctx, filt, param):
    """Filter importance sampling of the fitted-curve filters (fitted_curve.h)."""
    scene = scenegen.cornell_box(96, 96, max_depth=6)
    scene["camera"]["param"]["filter"] = {"type": filt, "param": param}
    sc = gpu_ctx.load_text(scene)
    sc.build_accel()
    sc.render(2)
    gpu, ref = sc.film(0), oracle.render(sc.export(), 0, 2)
    frac, _ = _compare(gpu, ref)
    print(f"{filt}: {frac * 100:.4f}% of pixels within {REL_TOL} relative")
    assert frac >= 0.999
    sc.close()


@pytest.mark.parametrize("kernel", ["megakernel", "wavefront"])
def test_direct_only_bounce_parity(gpu_ctx, kernel):
    """max_depth 1: the extra BSDF-sampled emitter ray (integrator.cpp:276-281),
    in both the megakernel and the wavefront form."""
    scene = scenegen.cornell_box(64, 64, max_depth=1)
    sc = gpu_ctx.load_text(scene)
    sc.build_accel()
    sc.set_kernel(kernel, 2, 256)
    sc.render(3)
    gpu, ref = sc.film(0), oracle.render(sc.export(), 0, 3)
    frac, _ = _compare(gpu, ref)
    assert frac >= 0.999
    sc.close()


@pytest.mark.parametrize("material,spectrum", [("diffuse", "srgb"), ("principled", "srgb"), ("diffuse", "hero4")])
def test_textured_colour_parity(gpu_ctx, tmp_path, material, spectrum):
    """A floor whose colour slot is a PNG checkerboard (ImageNode)."""
    gpu_ctx.write_image(tmp_path / "checker.png", scenegen.checker_rgba(16, 8))
    scene = scenegen.textured_box("checker.png", width=64, height=64, material=material, spectrum=spectrum)
    sc = gpu_ctx.load_text(scene, base_dir=str(tmp_path))
    sc.build_accel()
    sc.render(2, fused=True)
    gpu = sc.film(0)
    ref = oracle.render(sc.export(), 0, 2)
    frac, _ = _compare(gpu, ref)
    print(f"textured {material} {spectrum}: {frac * 100:.4f}% of pixels within {REL_TOL} relative")
    assert frac >= 0.999
    sc.close()


@pytest.mark.parametrize("spectrum", ["srgb", "hero4"])
def test_emission_image_parity(gpu_ctx, tmp_path, spectrum):
    """The ceiling lamp emits a PNG checkerboard (an ImageNode in the area
    light's colour slot)."""
    gpu_ctx.write_image(tmp_path / "lamp.png", scenegen.checker_rgba(8, 4, a=(0.9, 0.5, 0.1), b=(0.2, 0.3, 0.9)))
    scene = scenegen.cornell_box(64, 64, max_depth=4, light_sampler="power")
    lamp = [sh for sh in scene["shapes"] if "emission" in sh["param"]][0]
    lamp["param"]["emission"]["param"]["color"] = {"node": {"type": "image", "param": {"fn": "lamp.png"}}}
    if spectrum != "srgb":
        scene["spectrum"] = {"type": "hero", "param": {"dimension": 4}}
    sc = gpu_ctx.load_text(scene, base_dir=str(tmp_path))
    sc.build_accel()
    sc.render(2, fused=True)
    gpu = sc.film(0)
    ref = oracle.render(sc.export(), 0, 2)
    frac, _ = _compare(gpu, ref)
    print(f"emission image {spectrum}: {frac * 100:.4f}% of pixels within {REL_TOL} relative")
    assert frac >= 0.999
    sc.close()


@pytest.mark.parametrize("mode", ["megakernel", "wavefront"])
@pytest.mark.parametrize("radius,focal", [(0.15, 6.0), (0.4, 5.2)])
def test_thin_lens_parity(gpu_ctx, mode, radius, focal):
    """ThinLensCamera (thin_lens.cpp:34-42): lens samples on the same sampler
    streams as the oracle, in both kernel forms."""
    scene = scenegen.cornell_box(96, 96, max_depth=6)
    scene["camera"]["type"] = "thin_lens"
    scene["camera"]["param"]["lens_radius"] = radius
    scene["camera"]["param"]["focal_distance"] = focal
    print(f"case: thin lens r {radius} focal {focal} {mode}", flush=True)
    sc = gpu_ctx.load_text(scene)
    sc.build_accel()
    sc.set_kernel(mode, chunk_frames=2)
    sc.render(3)
    gpu, ref = sc.film(0), oracle.render(sc.export(), 0, 3)
    frac, _ = _compare(gpu, ref)
    print(f"thin lens {radius}/{focal} {mode}: {frac * 100:.4f}% of pixels within {REL_TOL} relative")
    assert frac >= 0.9999
    # defocus changes the picture: the same scene through the pinhole differs
    scene["camera"]["type"] = "pinhole"
    pin = gpu_ctx.load_text(scene)
    pin.build_accel()
    pin.render(3)
    assert np.abs(pin.film(0)[..., :3] - gpu[..., :3]).max() > 1e-2
    sc.close()
    pin.close()
