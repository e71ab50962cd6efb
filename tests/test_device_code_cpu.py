"""The device path-tracing source, compiled for the host, against the oracle.

tests/devcpu builds the very headers the HIP kernels include (dev_pt.h and
what it pulls in) with g++ against a small host stand-in for the HIP
runtime header, and drives them as the kernels do.  With identical inputs the
result must equal the oracle bit for bit: BSDF values, sampled directions
and whole films.  On the GPU the same source differs from the oracle only
through code generation (the elementary functions are shared,
vision_amd/csrc/elementary.h); this test shows that the arithmetic is the
oracle's.
"""
import ctypes as C
import json
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import vision_amd as va
from oracle import oracle
from vision_amd import build as vbuild
from vision_amd import scenegen

HERE = Path(__file__).resolve().parent
SRC = HERE / "devcpu" / "devcpu.cpp"
BUILD = HERE / "devcpu" / "_build"
LIB = BUILD / "libdevcpu.so"
OBJS = ["data_tables.cpp.o", "json.cpp.o", "scene.cpp.o", "bvh.cpp.o", "tables.cpp.o", "image_io.cpp.o", "jpeg.cpp.o"]

_f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u32 = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")


def _harness():
    deps = [SRC, HERE / "devcpu" / "hip" / "hip_runtime.h", *vbuild.CSRC.glob("*.h")]
    objs = [vbuild.BUILD / o for o in OBJS]
    if not all(o.exists() for o in objs):
        vbuild.build()
    if not LIB.exists() or LIB.stat().st_mtime < max(p.stat().st_mtime for p in deps + objs):
        BUILD.mkdir(parents=True, exist_ok=True)
        data = BUILD / "data"
        if not data.exists():
            try:
                data.symlink_to(vbuild.DATA)
            except FileExistsError:  # another pytest worker made it first
                pass
        tmp = LIB.with_suffix(f".tmp{os.getpid()}")
        cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-w", "-I", str(HERE / "devcpu"),
               "-I", str(vbuild.CSRC), "-o", str(tmp), str(SRC), *map(str, objs), "-lpthread", "-lz"]
        subprocess.run(cmd, check=True, capture_output=True)
        tmp.replace(LIB)
    h = C.CDLL(str(LIB))
    h.dc_bsdf.argtypes = [C.c_int, C.c_int, _u32, _f32, _f32, _f32, C.c_void_p, C.c_uint32, C.c_int, C.c_uint32,
                          _f32, _f32, _f32]
    h.dc_load.restype = C.c_void_p
    h.dc_load.argtypes = [C.c_char_p, C.c_char_p]
    h.dc_set_camera.argtypes = [C.c_void_p, _f32]
    h.dc_render.argtypes = [C.c_void_p, _f32, C.c_uint32, _f32]
    return h


@pytest.fixture(scope="module")
def dc():
    return _harness()


def _dev_bsdf(h, rec, spd, wo, wi, sample, frame, rgb, ns, lam):
    n = len(wo)
    wi_out = np.zeros((n, 3), np.float32)
    f = np.zeros((n, 4), np.float32)
    pdf = np.zeros(n, np.float32)
    h.dc_bsdf(1 if rgb else 0, ns, rec, spd, wo, wi, None if lam is None else lam.ctypes.data, n, sample, frame,
              wi_out, f, pdf)
    return wi_out, f, pdf


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


@pytest.mark.parametrize("spectrum", ["srgb", "hero4"])
def test_device_bsdf_source_equals_oracle(dc, spectrum):
    import test_gpu_materials as G
    rng = np.random.default_rng(8)
    n = 4000
    wo = rng.normal(size=(n, 3)).astype(np.float32)
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    wi = rng.normal(size=(n, 3)).astype(np.float32)
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    lam = None if spectrum == "srgb" else oracle.wavelengths(rng.random(4 * n, dtype=np.float32))[0].reshape(n, 4)
    for mat in G.TEST_MATERIALS:
        with va.Context(-1) as ctx:
            s = ctx.load_text(G._material_box(mat, spectrum)).build_accel()
            ex = s.export()
        rec = oracle.material_table(ex["materials"], len(G._material_box(mat, spectrum)["materials"]) - 1)
        spec = oracle.Spectral(ex)
        rgb, ns = bool(ex["spectrum_rgb"]), 3 if ex["spectrum_rgb"] else ex["spectrum_dim"]
        _, fd, pd = _dev_bsdf(dc, rec, spec.spd, wo, wi, 0, 0, rgb, ns, lam)
        fo, po = oracle.bsdf_eval(spec, rec, wo, wi, lam)
        assert _bits_equal(fd, fo) and _bits_equal(pd, po), mat
        wsd, fsd, psd = _dev_bsdf(dc, rec, spec.spd, wo, wi, 1, 3, rgb, ns, lam)
        wso, fso, pso = oracle.bsdf_sample(spec, rec, wo, frame=3, lam=lam)
        assert _bits_equal(wsd, wso) and _bits_equal(fsd, fso) and _bits_equal(psd, pso), mat


def _dev_render(h, scene, spp, base="."):
    with va.Context(-1) as ctx:
        ex = ctx.load_text(scene, base).build_accel().export()
    p = h.dc_load(json.dumps(scene).encode(), base.encode())
    cam = np.ascontiguousarray(ex["camera"], np.float32)
    h.dc_set_camera(p, cam)
    W, H = int(cam[19]), int(cam[20])
    film = np.zeros((H, W, 4), np.float32)
    h.dc_render(p, np.ascontiguousarray(ex["render"], np.float32), spp, film)
    return film, oracle.render(ex, 0, spp)


@pytest.mark.parametrize("spectrum", ["srgb", "hero4", "hero3"])
def test_device_path_source_equals_oracle_material_boxes(dc, spectrum):
    import test_gpu_materials as G
    for mat in G.TEST_MATERIALS:
        film, ref = _dev_render(dc, G._material_box(mat, spectrum, res=24), 2)
        assert _bits_equal(film, ref), (mat, spectrum)


@pytest.mark.parametrize("spectrum", ["srgb", "hero"])
def test_device_path_source_equals_oracle_procedural(dc, tmp_path, spectrum):
    scene = scenegen.procedural_scene(tmp_path, seed=3, n_instances=9, level=3, width=40, height=24, spp=2,
                                      max_depth=8, spectrum=spectrum, hdri=(256, 128))
    film, ref = _dev_render(dc, scene, 2, str(tmp_path))
    assert _bits_equal(film, ref)


@pytest.mark.parametrize("filt,param", [("gaussian", {"radius": 1.5, "sigma": 0.5}), ("mitchell", {"radius": 2.0}),
                                        ("sinc", {"radius": 3.0, "tau": 3.0}), ("gaussian", {})])
def test_device_path_source_equals_oracle_fitted_filters(dc, filt, param):
    scene = scenegen.cornell_box(24, 24, max_depth=4)
    scene["camera"]["param"]["filter"] = {"type": filt, "param": param}
    film, ref = _dev_render(dc, scene, 3)
    assert _bits_equal(film, ref), filt


@pytest.mark.parametrize("max_depth", [0, 1])
def test_device_path_source_equals_oracle_direct_only(dc, tmp_path, max_depth):
    scene = scenegen.procedural_scene(tmp_path, seed=4, n_instances=9, level=3, width=32, height=20, spp=2,
                                      max_depth=max_depth, spectrum="hero", hdri=(256, 128))
    film, ref = _dev_render(dc, scene, 2, str(tmp_path))
    assert _bits_equal(film, ref)
    assert (film[..., :3].max() > 0) == (max_depth > 0)


def _restir_pair(dc, scene, frames, base=".", cameras=None, fov=19.5):
    """The device ReSTIR source and the oracle over the same frames; returns
    both (film, reservoirs, surfaces)."""
    dc.dc_restir_frame.argtypes = [C.c_void_p] + [_f32] * 3 + [C.c_uint32] + [_f32] * 8
    ctx = va.Context(-1)
    sc = ctx.load_text(scene, base).build_accel()
    ex = sc.export()
    p = dc.dc_load(json.dumps(scene).encode(), base.encode())
    orc = oracle.Restir(ex)
    dev = oracle.Restir(ex)  # the same initial state buffers, advanced by the device source
    for f in range(frames):
        if cameras is not None:
            sc.update_camera(cameras[f], fov)
            ex = sc.export()
        cam = np.ascontiguousarray(ex["camera"], np.float32)
        dc.dc_set_camera(p, cam)
        dc.dc_restir_frame(p, np.ascontiguousarray(ex["restir"], np.float32), dev.prev_w2c, dev.prev_pos, f,
                           dev.rsv, dev.surf, dev.ext, dev.film.reshape(-1),
                           np.ascontiguousarray(ex["render"], np.float32), dev.gi_rsv.reshape(-1),
                           dev.gi_aux.reshape(-1), dev.gi_rad.reshape(-1))
        dev.prev_w2c = np.ascontiguousarray(ex["restir"][20:32], np.float32)
        dev.prev_pos = np.array([cam[3], cam[7], cam[11]], np.float32)
        orc.render(1, scene=ex)
    ctx.close()
    return dev, orc


def _restir_equal(dev, orc):
    assert _bits_equal(dev.surf, orc.surf)
    assert _bits_equal(dev.rsv, orc.rsv)
    assert _bits_equal(dev.film, orc.film)
    # ReSTIR GI state (zeros when its block is closed)
    assert _bits_equal(dev.gi_aux, orc.gi_aux)
    assert _bits_equal(dev.gi_rsv, orc.gi_rsv)
    assert _bits_equal(dev.gi_rad, orc.gi_rad)


@pytest.mark.parametrize("cfg", [dict(sample_num=3, radius=5),
                                 dict(sample_num=4, radius=6, pairwise=True, temporal_mis=True,
                                      extra_spatial_passes=1),
                                 dict(M_light=0, M_bsdf=2, spatial=False)], ids=["default", "pairwise", "bsdf"])
def test_device_restir_source_equals_oracle_cbox(dc, cfg):
    scene = scenegen.cornell_box(24, 20, light_sampler="power")
    scene["integrator"] = scenegen.restir_integrator(**{"M_light": 4, **cfg})
    dev, orc = _restir_pair(dc, scene, 4)
    _restir_equal(dev, orc)
    assert orc.film[..., :3].mean() > 0


def test_device_restir_source_equals_oracle_many_lights_moving_camera(dc):
    scene = scenegen.many_lights_box(28, 20, n_quads=40, n_points=16, seed=2, M_light=8, sample_num=3, radius=6,
                                     history_limit=8, extra_spatial_passes=1)
    cams = scenegen.jitter_path(4, amplitude=0.03, seed=5)
    dev, orc = _restir_pair(dc, scene, 4, cameras=cams)
    _restir_equal(dev, orc)
    # the camera moved: reprojection kept some temporal history
    assert orc.rsv[2, 0, :, 1].max() > 1


def test_device_restir_source_equals_oracle_hero_env(dc, tmp_path):
    scene = scenegen.procedural_scene(tmp_path, seed=6, n_instances=9, level=3, width=28, height=16, spp=1,
                                      max_depth=2, spectrum="hero", hdri=(128, 64))
    scene["integrator"] = scenegen.restir_integrator(M_light=4, M_bsdf=1, sample_num=2, radius=4)
    dev, orc = _restir_pair(dc, scene, 3, str(tmp_path))
    _restir_equal(dev, orc)


def test_device_restir_source_equals_oracle_mirror(dc):
    scene = scenegen.cornell_box(24, 24, light_sampler="power")
    scene["materials"] += [{"type": "mirror", "name": "Mirror", "param": {"color": [0.9, 0.8, 0.6]}},
                           {"type": "mirror", "name": "Rough", "param": {"roughness": 0.4}}]
    for sh in scene["shapes"]:
        if sh["name"] == "short_box":
            sh["param"]["material"] = "Mirror"
        if sh["name"] == "tall_box":
            sh["param"]["material"] = "Rough"
    scene["integrator"] = scenegen.restir_integrator(M_light=4, M_bsdf=1, sample_num=2, radius=4)
    dev, orc = _restir_pair(dc, scene, 3)
    _restir_equal(dev, orc)


def test_device_restir_source_equals_oracle_near_specular(dc):
    scene = scenegen.cornell_box(24, 24, light_sampler="power")
    scene["materials"] += [{"type": "metal", "name": "Mirror", "param": {"material_name": "Cu", "roughness": 0.01}},
                           {"type": "glass", "name": "Clear", "param": {"ior": 1.5, "roughness": 0.01}}]
    for sh in scene["shapes"]:
        if sh["name"] == "short_box":
            sh["param"]["material"] = "Mirror"
        if sh["name"] == "tall_box":
            sh["param"]["material"] = "Clear"
    scene["integrator"] = scenegen.restir_integrator(M_light=4, M_bsdf=1, sample_num=2, radius=4)
    dev, orc = _restir_pair(dc, scene, 3)
    _restir_equal(dev, orc)
    # near-specular surfaces were replaced by what they reflect
    flags = orc.surf[orc.frame & 1, 0, :, 3].view(np.uint32)
    assert np.any((flags >> 16) == 1)


@pytest.mark.parametrize("material,spectrum", [("diffuse", "srgb"), ("principled", "srgb"), ("diffuse", "hero4")])
def test_device_path_source_equals_oracle_textured(dc, tmp_path, material, spectrum):
    with va.Context(-1) as c:
        c.write_image(tmp_path / "checker.png", scenegen.checker_rgba(16, 8))
    scene = scenegen.textured_box("checker.png", width=24, height=24, material=material, spectrum=spectrum)
    film, ref = _dev_render(dc, scene, 2, str(tmp_path))
    assert _bits_equal(film, ref), (material, spectrum)


@pytest.mark.parametrize("spectrum", ["srgb", "hero4"])
def test_device_path_source_equals_oracle_emission_image(dc, tmp_path, spectrum):
    with va.Context(-1) as c:
        c.write_image(tmp_path / "lamp.png", scenegen.checker_rgba(8, 4, a=(0.9, 0.5, 0.1), b=(0.2, 0.3, 0.9)))
    scene = scenegen.cornell_box(24, 24, max_depth=4, light_sampler="power")
    lamp = [sh for sh in scene["shapes"] if "emission" in sh["param"]][0]
    lamp["param"]["emission"]["param"]["color"] = {"node": {"type": "image", "param": {"fn": "lamp.png"}}}
    if spectrum != "srgb":
        scene["spectrum"] = {"type": "hero", "param": {"dimension": 4}}
    film, ref = _dev_render(dc, scene, 2, str(tmp_path))
    assert _bits_equal(film, ref), spectrum
    scene["integrator"] = scenegen.restir_integrator(M_light=4, M_bsdf=1, sample_num=2, radius=4)
    dev, orc = _restir_pair(dc, scene, 3, str(tmp_path))
    _restir_equal(dev, orc)


# ---------------------------------------------------------------- ReSTIR GI
def _gi_scene(w, h, max_depth=4, temporal=True, spatial=True, point_lights=False, **gi_extra):
    s = scenegen.cornell_box(w, h, max_depth=max_depth, light_sampler="power")
    if point_lights:
        s["shapes"] = [x for x in s["shapes"] if x["name"] != "light"]
        s["light_sampler"]["param"]["lights"] = [
            {"type": "point", "param": {"position": [0.2, 1.7, 0.3], "color": [6, 5, 4], "scale": 1.0}},
            {"type": "spot", "param": {"position": [-0.5, 1.6, -0.4], "color": [3, 4, 5], "scale": 1.0,
                                       "direction": [0.2, -1, 0.1], "angle": 50, "falloff": 10}}]
    gi = {"open": True, "max_age": 6, "temporal": {"open": temporal, "history_limit": 8, "radius": 4, "N": 5},
          "spatial": {"open": spatial, "sample_num": 3, "radius": 4}, **gi_extra}
    s["integrator"] = scenegen.restir_integrator(M_light=4, M_bsdf=1, sample_num=2, radius=4, indirect=gi,
                                                 max_depth=max_depth)
    return s


@pytest.mark.parametrize("cfg", [dict(), dict(point_lights=True), dict(temporal=False, spatial=False),
                                 dict(max_depth=2), dict(max_depth=1), dict(max_depth=6, spatial=False)],
                         ids=["area", "point_spot", "no_reuse", "depth2", "depth1", "depth6"])
def test_device_restir_gi_source_equals_oracle(dc, cfg):
    scene = _gi_scene(22, 18, **cfg)
    dev, orc = _restir_pair(dc, scene, 4)
    _restir_equal(dev, orc)
    # GI ran: the pass-through reservoirs hold samples with light
    assert (orc.gi_rsv[2, 3, :, :3].sum(-1) > 0).mean() > (0.2 if cfg.get("max_depth", 4) > 1 else -1)


def test_device_restir_gi_source_equals_oracle_moving_camera_near_specular(dc):
    scene = _gi_scene(24, 20)
    scene["materials"] += [{"type": "metal", "name": "Mirror", "param": {"material_name": "Cu", "roughness": 0.01}},
                           {"type": "metal", "name": "Rough", "param": {"material_name": "Al", "roughness": 0.3}}]
    for sh in scene["shapes"]:
        if sh["name"] == "short_box":
            sh["param"]["material"] = "Mirror"
        if sh["name"] == "tall_box":
            sh["param"]["material"] = "Rough"
    cams = scenegen.jitter_path(4, amplitude=0.03, seed=9)
    dev, orc = _restir_pair(dc, scene, 4, cameras=cams)
    _restir_equal(dev, orc)
    # temporal reuse kept history somewhere
    assert orc.gi_rsv[2, 0, :, 1].max() > 1
