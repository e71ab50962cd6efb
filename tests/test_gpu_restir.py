"""ReSTIR DI on the GPU (restir.hip) against the oracle on the same frames.

The kernels run the very source the CPU harness checks bit for bit
(test_device_code_cpu.py); on the GPU the device math library rounds
sin/cos/sqrt differently in the last bit, which can move a spatial or
temporal neighbour draw across a pixel boundary.  Per frame that touches a
few pixels, and the reservoirs carry it forward, so the bar is: nearly all
pixels of the film within 1e-4 relative, nearly all reservoirs holding the
same light sample, and the image means within 0.5%.
"""
import numpy as np
import pytest

import vision_amd as va
from oracle import oracle
from vision_amd import scenegen

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4
ABS_FLOOR = 1e-6


def _frames(gpu_ctx, scene, frames, cameras=None, fov=19.5):
    sc = gpu_ctx.load_text(scene)
    sc.build_accel()
    orc = oracle.Restir(sc.export())
    for f in range(frames):
        if cameras is not None:
            sc.update_camera(cameras[f], fov)
        sc.render(1)
        orc.render(1, scene=sc.export())
    film = sc.film(0)
    rsv, surf = sc.restir_state()
    sc.close()
    return film, rsv, surf, orc


def _check(film, rsv, surf, orc, pix_frac=0.995, sample_frac=0.99):
    g, r = film[..., :3].astype(np.float64), orc.film[..., :3].astype(np.float64)
    ok = (np.abs(g - r) <= REL_TOL * np.maximum(np.abs(g), np.abs(r)) + ABS_FLOOR).all(axis=-1)
    print(f"film: {ok.mean() * 100:.3f}% of pixels within {REL_TOL}; "
          f"means {g.mean():.6f} / {r.mean():.6f}")
    assert ok.mean() >= pix_frac
    assert abs(g.mean() / r.mean() - 1) < 5e-3
    cur = orc.frame & 1
    same_tri = np.mean(surf[cur, 0, :, 0].view(np.uint32) == orc.surf[cur, 0, :, 0].view(np.uint32))
    same_light = np.mean(rsv[cur, 1, :, 0].view(np.uint32) == orc.rsv[cur, 1, :, 0].view(np.uint32))
    print(f"surfaces {same_tri * 100:.3f}% same triangle, reservoirs {same_light * 100:.3f}% same light")
    assert same_tri >= 0.999
    assert same_light >= sample_frac


@pytest.mark.parametrize("cfg", [dict(sample_num=3, radius=8),
                                 dict(sample_num=4, radius=8, pairwise=True, temporal_mis=True,
                                      extra_spatial_passes=1)], ids=["default", "pairwise"])
def test_restir_cbox_parity(gpu_ctx, cfg):
    scene = scenegen.cornell_box(96, 96, light_sampler="power")
    scene["integrator"] = scenegen.restir_integrator(M_light=8, M_bsdf=1, history_limit=10, **cfg)
    _check(*_frames(gpu_ctx, scene, 4))


def test_restir_many_lights_moving_camera_parity(gpu_ctx):
    scene = scenegen.many_lights_box(96, 64, n_quads=400, n_points=32, seed=3, M_light=16, sample_num=5,
                                     radius=12, history_limit=20, extra_spatial_passes=1)
    cams = scenegen.jitter_path(4, amplitude=0.01, seed=7)
    _check(*_frames(gpu_ctx, scene, 4, cameras=cams), pix_frac=0.999, sample_frac=0.999)


def test_restir_full_hd_frames_are_finite(gpu_ctx):
    """The C3 workload's shape: 1080p, 10 000 emissive triangles and 256
    point / spot lights, two frames with a moving camera."""
    scene = scenegen.many_lights_box(1920, 1080, seed=0)
    sc = gpu_ctx.load_text(scene)
    sc.build_accel()
    for c2w in scenegen.jitter_path(2, seed=1):
        sc.update_camera(c2w, 19.5)
        st = sc.render(1)
    film = sc.film(0)
    assert np.isfinite(film).all()
    assert film[..., :3].mean() > 0
    assert st["launches"] == 3
    sc.close()


# ---------------------------------------------------------------- ReSTIR GI
def _gi_scene(w, h, max_depth=4, point_lights=False, temporal=True, spatial=True):
    s = scenegen.cornell_box(w, h, max_depth=max_depth, light_sampler="power")
    if point_lights:
        s["shapes"] = [x for x in s["shapes"] if x["name"] != "light"]
        s["light_sampler"]["param"]["lights"] = [
            {"type": "point", "param": {"position": [0.2, 1.7, 0.3], "color": [6, 5, 4], "scale": 1.0}},
            {"type": "spot", "param": {"position": [-0.5, 1.6, -0.4], "color": [3, 4, 5], "scale": 1.0,
                                       "direction": [0.2, -1, 0.1], "angle": 50, "falloff": 10}}]
    gi = {"open": True, "max_age": 6, "temporal": {"open": temporal, "history_limit": 8, "radius": 4, "N": 5},
          "spatial": {"open": spatial, "sample_num": 3, "radius": 6}}
    s["integrator"] = scenegen.restir_integrator(M_light=4, M_bsdf=1, sample_num=2, radius=6, indirect=gi,
                                                 max_depth=max_depth)
    return s


@pytest.mark.parametrize("cfg", [dict(), dict(point_lights=True, max_depth=6), dict(max_depth=2)],
                         ids=["area", "point_spot_depth6", "depth2"])
def test_restir_gi_cbox_parity(gpu_ctx, cfg):
    """ReSTIR DI + GI frames (indirect.cpp) against the oracle: the film,
    the GI reservoirs' sample points and the indirect radiance."""
    scene = _gi_scene(96, 80, **cfg)
    sc = gpu_ctx.load_text(scene)
    sc.build_accel()
    orc = oracle.Restir(sc.export())
    cams = scenegen.jitter_path(4, amplitude=0.01, seed=11)
    for f in range(4):
        sc.update_camera(cams[f], 19.5)
        st = sc.render(1)
        orc.render(1, scene=sc.export())
    film = sc.film(0)
    rsv, surf = sc.restir_state()
    gi_rsv, gi_rad = sc.restir_gi_state()
    sc.close()
    assert st["launches"] == 4
    _check(film, rsv, surf, orc, pix_frac=0.999, sample_frac=0.999)
    cur = orc.frame & 1
    same_pos = np.all(gi_rsv[cur, 1, :, :3] == orc.gi_rsv[cur, 1, :, :3], axis=-1).mean()
    g, r = gi_rad[:, :3].astype(np.float64), orc.gi_rad[:, :3].astype(np.float64)
    ok = (np.abs(g - r) <= REL_TOL * np.maximum(np.abs(g), np.abs(r)) + ABS_FLOOR).all(-1).mean()
    print(f"GI: {same_pos * 100:.3f}% same sample point, {ok * 100:.3f}% indirect radiance within {REL_TOL}; "
          f"indirect mean {g.mean():.5f} / {r.mean():.5f}")
    assert same_pos >= 0.999
    assert ok >= 0.999
    assert r.mean() > 0
