"""Host-side logic on the CPU: C-ABI exports, scene-file parsing, BVH8 layout."""
import ctypes as C
import json
import math
import re
from pathlib import Path

import numpy as np
import pytest

import vision_amd as va
from vision_amd import scenegen

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = Path(__file__).resolve().parent / "golden"


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "vision_amd.h").read_text()
    names = set(re.findall(r"\b(vs_[a-z_0-9]+)\s*\(", header))
    assert len(names) >= 20
    lib = va.lib()
    for n in sorted(names):
        assert hasattr(lib, n), n
        assert n in va.SIGNATURES, n


def test_ray_and_hit_record_layout():
    assert C.sizeof(va.Ray) == 32 and va.RAY_DTYPE.itemsize == 32
    assert C.sizeof(va.Hit) == 24 and va.HIT_DTYPE.itemsize == 24


def test_host_context_refuses_device_work(host_ctx):
    sc = host_ctx.load_text(scenegen.cornell_box(16, 16))
    sc.build_accel()
    with pytest.raises(va.VisionError) as e:
        sc.render(1)
    assert e.value.status == 6
    sc.close()


def test_cornell_box_resolves_to_expected_geometry(host_ctx):
    sc = host_ctx.load_text(scenegen.cornell_box())
    info = sc.info()
    assert info["num_triangles"] == 5 * 2 + 2 * 12 + 2
    assert info["num_instances"] == 8 and info["num_lights"] == 1
    assert (info["width"], info["height"], info["max_depth"], info["output_spp"]) == (256, 256, 8, 16)
    ex = sc.export()
    P = ex["positions"]
    # walls span [-1, 1] x [0, 2] x [-1, 1]
    assert np.allclose(P.min(0), [-1, 0, -1], atol=1e-5) and np.allclose(P.max(0), [1, 2, 1], atol=1e-5)
    # the light: 0.47 x 0.38 quad at y = 1.98, facing down, radiance (17, 12, 4)
    light = ex["lights"][0]
    assert light[0] == 0  # area
    inst = light[2]
    tris = ex["triangles"][ex["tri_instance"] == inst]
    lp = P[tris.reshape(-1)]
    assert np.allclose(lp[:, 1], 1.98, atol=1e-6)
    assert np.isclose(np.ptp(lp[:, 0]), 0.47, atol=1e-6) and np.isclose(np.ptp(lp[:, 2]), 0.38, atol=1e-6)
    v0, v1, v2 = P[tris[0]]
    n = np.cross(v0 - v2, v1 - v2)
    assert n[1] < 0
    lf = light.view(np.float32)
    # Light ctor normalizes the colour by its max and folds it into scale
    assert np.allclose(lf[8:11] * lf[11], [17, 12, 4], rtol=1e-6) and lf[8:11].max() == 1.0
    # camera: at (0, 1, 6.8) looking down -z (c2w rows [R | t])
    cam = ex["camera"]
    assert np.allclose(cam[[3, 7, 11]], [0, 1, 6.8], atol=1e-6)
    R = cam[:12].reshape(3, 4)[:, :3]
    assert np.allclose(R @ [0, 0, 1], [0, 0, -1], atol=1e-6)
    assert np.isclose(cam[16], math.tan(math.radians(19.5) / 2), rtol=1e-6)
    sc.close()


def test_reference_cbox_json_parses(host_ctx):
    # the reference's cbox.json (res/render_scene/cbox/cbox.json) loads
    # unchanged: its comment syntax, the thin-lens camera, the rt integrator,
    # a mix material and the materials this build does not render (metallic,
    # plastic, substrate), which no shape uses
    text = (GOLDEN / "cbox_reference.json").read_text()
    sc = host_ctx.load_text(text)
    ex = sc.export()
    info = sc.info()
    assert info["num_lights"] == 1 and (info["width"], info["height"]) == (1024, 1024)
    # floor quad spans x, z in [-1, 1] at y = 0 (cbox.json shape_0 matrix)
    fl = ex["positions"][ex["triangles"][ex["tri_instance"] == 0].reshape(-1)]
    assert np.allclose(np.abs(fl[:, [0, 2]]), 1, atol=1e-5) and np.allclose(fl[:, 1], 0, atol=1e-5)
    lt = ex["lights"][0].view(np.float32)
    assert np.allclose(lt[8:11] * lt[11], [17, 12, 4])
    # the look_at camera of cbox.json: position (0, 1, 6.8) toward (0, 1, 0)
    cam = ex["camera"]
    assert np.allclose(cam[[3, 7, 11]], [0, 1, 6.8], atol=1e-5)
    sc.close()
    # with the path tracer, the file's "mix" material is packed: components
    # after the scene's own materials, the record pointing at them (under
    # "rt" it is listed but not rendered)
    doc = json.loads(re.sub(r"//[^\n]*", "", text))
    doc["integrator"] = {"type": "pt", "param": {"max_depth": 8}}
    names = [m["name"] for m in doc["materials"]]
    sc = host_ctx.load_text(json.dumps(doc))
    ex = sc.export()
    mix = ex["materials"][names.index("mix")]
    assert mix[0] == 7 and mix[2] >= len(names) and mix[3] == mix[2] + 1
    assert np.float32(0.1) == mix[4:5].view(np.float32)[0]
    assert ex["materials"][mix[2]][0] == 5 and ex["materials"][mix[3]][0] == 0  # principled, diffuse
    sc.close()
    # a shape that uses an unsupported material is refused, naming it
    doc["shapes"][0]["param"]["material"] = "plastic"
    with pytest.raises(va.VisionError) as e:
        host_ctx.load_text(json.dumps(doc))
    assert "plastic" in str(e.value)


def test_json_comments_and_errors(host_ctx):
    scene = scenegen.cornell_box(8, 8)
    text = "// leading comment\n/* block */" + json.dumps(scene, indent=1).replace(
        '"shapes": [', '"shapes": [ // trailing\n', 1)
    sc = host_ctx.load_text(text)
    assert sc.info()["num_triangles"] == 36
    sc.close()
    with pytest.raises(va.VisionError) as e:
        host_ctx.load_text('{"shapes": [ }')
    assert e.value.status == 2
    bad = scenegen.cornell_box(8, 8)
    bad["materials"][0]["type"] = "plastic"
    with pytest.raises(va.VisionError) as e:
        host_ctx.load_text(bad)
    assert e.value.status == 3


def test_defaults_follow_reference_descriptors(host_ctx):
    # IlluminationIntegrator defaults: max_depth 16, min_depth 5, rr 1 (integrator.cpp:60-67);
    # Sensor fov 20 (sensor.cpp:22); RadianceCollector resolution 1280x720
    scene = scenegen.cornell_box()
    scene["integrator"] = {"type": "pt"}
    scene["camera"]["param"].pop("fov_y")
    scene["camera"]["param"]["radiance_collector"] = {"type": "normal"}
    sc = host_ctx.load_text(scene)
    ex = sc.export()
    assert list(ex["render"][:3]) == [16, 5, 1]
    assert (sc.info()["width"], sc.info()["height"]) == (1280, 720)
    assert np.isclose(ex["camera"][16], math.tan(math.radians(20) / 2))
    sc.close()


def _decode_bvh(nodes):
    """Decode 80-byte BVH8 nodes into child boxes (float64) and metadata."""
    n = len(nodes)
    w = nodes.astype(np.uint64)
    origin = nodes[:, 0:3].view(np.float32).astype(np.float64)
    e = np.stack([(w[:, 3] >> s) & 0xFF for s in (0, 8, 16)], 1).astype(np.int64) - 127
    step = np.ldexp(1.0, e)
    imask = (w[:, 3] >> 24) & 0xFF

    def bytes8(word0):
        lo, hi = w[:, word0], w[:, word0 + 1]
        return np.stack([((lo if s < 4 else hi) >> (8 * (s & 3))) & 0xFF for s in range(8)], 1).astype(np.float64)

    qlo = np.stack([bytes8(8), bytes8(10), bytes8(12)], 2)
    qhi = np.stack([bytes8(14), bytes8(16), bytes8(18)], 2)
    meta = bytes8(6).astype(np.int64)
    lo = origin[:, None, :] + qlo * step[:, None, :]
    hi = origin[:, None, :] + qhi * step[:, None, :]
    return lo, hi, imask, meta


def test_bvh8_layout_is_complete_and_conservative(host_ctx):
    rng = np.random.default_rng(11)
    # a procedural mesh file (vsmesh) of 5000 random triangles
    import tempfile
    import struct
    n = 5000
    c = rng.random((n, 1, 3)) * 10 - 5
    tri = (c + rng.normal(scale=0.2, size=(n, 3, 3))).reshape(-1, 3).astype(np.float32)
    idx = np.arange(3 * n, dtype=np.uint32)
    with tempfile.TemporaryDirectory() as d:
        fn = Path(d) / "soup.vsmesh"
        with open(fn, "wb") as f:
            f.write(b"VSMESH01" + struct.pack("<III", 3 * n, n, 0) + tri.tobytes() + idx.tobytes())
        scene = scenegen.cornell_box(8, 8)
        scene["shapes"].append({"type": "model", "name": "soup",
                                "param": {"fn": "soup.vsmesh", "material": "Gray"}})
        sc = host_ctx.load_text(scene, base_dir=d)
        sc.build_accel()
    nodes, ltris = sc.bvh()
    ex = sc.export()
    P, T = ex["positions"], ex["triangles"]
    ids = ltris[:, 3].view(np.uint32)
    assert sorted(ids.tolist()) == list(range(len(T)))
    # the leaf triangle records hold the scene's vertices
    assert np.array_equal(ltris[:, 0:3], P[T[ids, 0]]) and np.array_equal(ltris[:, 8:11], P[T[ids, 2]])
    lo, hi, imask, meta = _decode_bvh(nodes)
    info = sc.info()
    assert info["bvh_nodes"] == len(nodes) and info["bvh_depth"] >= 2
    child_base = nodes[:, 4].astype(np.int64)
    tri_base = nodes[:, 5].astype(np.int64)

    def subtree_tris(i, out):
        for s in range(8):
            if (imask[i] >> s) & 1:
                ci = child_base[i] + bin(int(imask[i]) & ((1 << s) - 1)).count("1")
                subtree_tris(ci, out)
            elif meta[i, s]:
                cnt, off = meta[i, s] >> 5, meta[i, s] & 31
                out.extend(range(tri_base[i] + off, tri_base[i] + off + cnt))
        return out

    for i in range(len(nodes)):
        for s in range(8):
            if (imask[i] >> s) & 1:
                ci = child_base[i] + bin(int(imask[i]) & ((1 << s) - 1)).count("1")
                leaf_ids = subtree_tris(ci, [])
            elif meta[i, s]:
                cnt, off = meta[i, s] >> 5, meta[i, s] & 31
                assert 1 <= cnt <= 3
                leaf_ids = list(range(tri_base[i] + off, tri_base[i] + off + cnt))
            else:
                assert (lo[i, s] > hi[i, s]).all()  # empty slot never hits
                continue
            v = ltris[leaf_ids].reshape(-1, 4)[:, :3].astype(np.float64)
            assert (v >= lo[i, s] - 1e-12).all() and (v <= hi[i, s] + 1e-12).all()
    sc.close()
