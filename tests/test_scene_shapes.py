"""Scene descriptors beyond plain meshes: the UV sphere shape
(src/render_core/shapes/sphere.cpp), an area light listed in the light
sampler with its own quad (AreaLight::init_geometry,
src/render_core/lights/area.cpp:57-68), a spherical light of constant
colour (SphericalMap::prepare, spherical.cpp:198-211); and every scene file
the reference ships either loads or fails with an error, never a crash."""
import json
import math
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import vision_amd as va
from oracle import oracle
from vision_amd import scenegen

REPO = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def ctx():
    with va.Context(-1) as c:
        yield c


def _sphere_scene(radius, sub_div):
    sc = scenegen.lit_plane({"type": "point", "param": {"color": [1, 1, 1], "position": [0, 3, 0]}})
    sc["shapes"] = [{"type": "sphere", "name": "ball",
                     "param": {"radius": radius, "sub_div": sub_div, "material": "Floor"}}]
    return sc


@pytest.mark.parametrize("sub_div", [2, 5, 60])
def test_sphere_mesh_is_a_closed_latitude_longitude_tessellation(ctx, sub_div):
    r = 0.75
    ex = ctx.load_text(json.dumps(_sphere_scene(r, sub_div))).export()
    P, N, UV, T = ex["positions"], ex["normals"], ex["uvs"], ex["triangles"].astype(np.int64)
    phi = 2 * sub_div
    assert len(P) == 2 + (sub_div - 1) * phi
    assert len(T) == 2 * phi + (sub_div - 2) * 2 * phi
    np.testing.assert_allclose(np.linalg.norm(P, axis=1), r, rtol=1e-6)
    np.testing.assert_allclose(N, P / r, atol=1e-6)
    # v runs from the north pole (0) to the south pole (1) with the polar angle
    np.testing.assert_allclose(UV[1:-1, 1], np.arccos(np.clip(P[1:-1, 1] / r, -1, 1)) / np.pi, atol=1e-5)
    # watertight: every edge is shared by exactly two triangles, once each way
    edges = {}
    for a, b, c in T:
        for e in ((a, b), (b, c), (c, a)):
            edges[e] = edges.get(e, 0) + 1
    assert all(n == 1 for n in edges.values())
    assert all((b, a) in edges for a, b in edges)
    # one consistent winding over the whole surface (all normals point the same way)
    g = np.cross(P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]])
    side = np.sign(np.einsum("ij,ij->i", g, P[T].mean(axis=1)))
    assert np.all(side == side[0])
    # the enclosed volume approaches the sphere's from below
    vol = abs(np.einsum("ij,ij->i", P[T[:, 0]], np.cross(P[T[:, 1]], P[T[:, 2]])).sum()) / 6
    assert vol < 4 / 3 * np.pi * r ** 3
    if sub_div == 60:
        assert vol > 0.99 * 4 / 3 * np.pi * r ** 3


def test_sphere_rejects_degenerate_subdivision(ctx):
    with pytest.raises(va.VisionError, match="sub_div"):
        ctx.load_text(json.dumps(_sphere_scene(1.0, 1)))


def test_light_list_area_light_brings_its_quad(ctx):
    """An area light in the light sampler's list with width / height / o2w:
    the same form factor closed form as the emissive-shape version."""
    a, h, Le, rho = 0.5, 0.8, 4.0, 0.5
    lamp = {"type": "area", "param": {"color": [Le, Le, Le], "scale": 1, "width": a, "height": a,
                                      "o2w": scenegen.xform(scenegen.translate(0, h, 0) @ scenegen.scale(1, -1, -1))}}
    scene = scenegen.lit_plane(lamp, albedo=rho, cam_height=0.4, fov=0.5, width=8, height=8)
    sc = ctx.load_text(json.dumps(scene))
    ex = sc.export()
    assert ex["instances"].shape[0] == 2 and ex["lights"].shape[0] == 1
    film = oracle.render(ex, 0, 256)
    sc.close()
    A = B = (a / 2) / h
    Fc = (A / math.sqrt(1 + A * A) * math.atan(B / math.sqrt(1 + A * A)) +
          B / math.sqrt(1 + B * B) * math.atan(A / math.sqrt(1 + B * B))) / (2 * math.pi)
    got = film[2:6, 2:6, :3].mean()
    assert abs(got / (rho * Le * 4 * Fc) - 1) < 0.01, got


def test_constant_colour_environment(ctx):
    """A spherical light with a constant colour over an open plane: the
    floor reflects rho * L (the whole upper hemisphere is sky)."""
    L, scale, rho = (0.6, 0.3, 0.15), 2.0, 0.5
    env = {"type": "spherical", "param": {"color": list(L), "scale": scale}}
    scene = scenegen.lit_plane(env, albedo=rho, width=8, height=8)
    sc = ctx.load_text(json.dumps(scene))
    ex = sc.export()
    assert ex["env_rgb"].shape == (1, 1, 3)
    film = oracle.render(ex, 0, 256)
    sc.close()
    got = film[..., :3].mean(axis=(0, 1))
    np.testing.assert_allclose(got, rho * scale * np.array(L), rtol=0.02)


REF_RES = Path("/root/reference/res")


@pytest.mark.skipif(not REF_RES.is_dir(), reason="the reference's scene files are not on this machine")
def test_reference_scene_files_load_or_fail_cleanly():
    """Each scene file under the reference's res/ loads, or raises a
    VisionError (unsupported feature, missing mesh or image): the loader
    never crashes.  Run in a child process so a crash fails the test."""
    code = r'''
import sys, pathlib, vision_amd as va
root = pathlib.Path(sys.argv[1])
ok = bad = 0
with va.Context(-1) as ctx:
    for p in sorted(root.rglob("*.json")):
        if "precompute" in str(p):
            continue
        try:
            ctx.load_text(p.read_text(errors="replace"), str(p.parent)).close()
            ok += 1
        except va.VisionError:
            bad += 1
print(ok, bad)
'''
    r = subprocess.run([sys.executable, "-c", code, str(REF_RES)], capture_output=True, text=True, cwd=REPO,
                       timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    ok, bad = map(int, r.stdout.split())
    print(f"{ok} scene files load, {bad} fail with an error")
    assert ok + bad >= 40 and ok >= 2
