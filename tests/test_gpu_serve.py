"""Batched serving on the GPU: every job's film, rendered in a mix of scenes,
resolutions, sample counts and spectrum modes that arrive while others run,
is bit-identical to the plain render of the same job alone (vs_render of
the scene at that resolution and spp); admission keeps within the HBM
budget; graph replay, direct launches and every tile shape agree."""
import copy
import json
import time

import numpy as np
import pytest

import vision_amd as va
from vision_amd import scenegen

pytestmark = pytest.mark.gpu


def _scenes(tmp):
    """Three small scenes (JSON, base directory): the Cornell box (diffuse
    only), the procedural scene with all four material families and an HDRI,
    and the Cornell box in hero-wavelength mode."""
    cbox = scenegen.cornell_box(64, 64, light_sampler="power")
    proc = scenegen.procedural_scene(tmp, seed=5, n_instances=4, level=2, width=64, height=48, spp=1,
                                     max_depth=8, spectrum="srgb", hdri=(64, 32))
    hero = copy.deepcopy(cbox)
    hero["spectrum"] = {"type": "hero", "param": {"dimension": 4}}
    return {"cbox": (cbox, "."), "proc": (proc, str(tmp)), "hero": (hero, ".")}


def _plain(ctx, scene, w, h, spp, max_depth=0, spectral=False, dim=4):
    sj = copy.deepcopy(scene[0])
    if spectral:
        sj["spectrum"] = {"type": "hero", "param": {"dimension": dim}}
    else:
        sj["spectrum"] = {"type": "srgb"}
    sc = ctx.load_text(json.dumps(sj), base_dir=scene[1]).build_accel()
    sc.set_resolution(w, h)
    if max_depth:
        p = sj["integrator"]["param"]
        sc.set_integrator(max_depth, p.get("min_depth", 5), p.get("rr_threshold", 1.0))
    sc.render(spp, fused=True)
    rt, out = sc.film(0), sc.film(1)
    sc.close()
    return rt, out


@pytest.fixture(scope="module")
def served(gpu_ctx, tmp_path_factory):
    js = _scenes(tmp_path_factory.mktemp("serve"))
    loaded = {k: gpu_ctx.load_text(json.dumps(v[0]), base_dir=v[1]).build_accel() for k, v in js.items()}
    return js, loaded


JOBS = [  # scene, width, height, spp, max_depth, spectral, dim
    ("cbox", 64, 64, 4, 0, False, 0),
    ("cbox", 96, 40, 7, 0, False, 0),
    ("proc", 80, 60, 3, 0, False, 0),
    ("cbox", 33, 65, 1, 3, False, 0),
    ("proc", 48, 48, 5, 6, True, 4),
    ("cbox", 50, 30, 9, 0, True, 3),
    ("hero", 64, 32, 2, 0, True, 4),
    ("hero", 40, 40, 6, 0, False, 0),
]


def _run(gpu_ctx, served, jobs, staggered=False, routes=(), **cfg):
    js, loaded = served
    with va.Server(gpu_ctx, **cfg) as srv:
        idx = {k: srv.add_scene(v) for k, v in loaded.items()}
        for name, spectral in routes:
            srv.set_route(idx[name], spectral, True)
        ids = []
        for n, (name, w, h, spp, md, spectral, dim) in enumerate(jobs):
            ids.append(srv.submit(idx[name], w, h, spp, max_depth=md, spectral=spectral, spectrum_dim=dim))
            if staggered and n % 3 == 2:
                time.sleep(0.01)  # later jobs arrive while earlier ones run
        srv.drain(60000)
        res = []
        for jid, (name, w, h, spp, md, spectral, dim) in zip(ids, jobs):
            info = srv.info(jid)
            if info["state"] != "done":
                srv.wait(jid)  # raises with the job's error
            assert info["state"] == "done"
            lane = (name, spectral) in routes or 0 < cfg.get("lane_min_samples", 0) <= w * h * spp
            assert info["route"] == ("wavefront" if lane else "megakernel")
            res.append((srv.result(jid, 0, (h, w)), srv.result(jid, 1, (h, w))))
            srv.release(jid)
        stats = srv.stats()
    return res, stats


def test_each_job_matches_its_plain_render(gpu_ctx, served):
    js, _ = served
    res, stats = _run(gpu_ctx, served, JOBS, staggered=True)
    for (name, w, h, spp, md, spectral, dim), (rt, out) in zip(JOBS, res):
        print(f"job {name} {w}x{h} spp={spp} depth={md} spectral={spectral}/{dim}", flush=True)
        prt, pout = _plain(gpu_ctx, js[name], w, h, spp, md, spectral, dim or 4)
        np.testing.assert_array_equal(rt.view(np.uint32), prt.view(np.uint32))
        np.testing.assert_array_equal(out.view(np.uint32), pout.view(np.uint32))
    assert stats["jobs_done"] == len(JOBS)
    assert stats["graph_launches"] >= 1
    assert 0 < stats["latency_p50_ms"] <= stats["latency_p99_ms"] <= stats["latency_max_ms"]
    assert stats["samples_done"] == sum(w * h * spp for _, w, h, spp, *_ in JOBS)


@pytest.mark.parametrize("cfg", [dict(use_graphs=False), dict(tile_w=16, waves_per_cu=1),
                                 dict(tile_w=64, step_ms=1.0), dict(tile_w=32, max_jobs=2),
                                 dict(min_waves_per_eu=6)],
                         ids=["direct", "tile16-wpc1", "tile64-short-steps", "tile32-two-slots", "regs6"])
def test_same_films_under_every_launch_shape(gpu_ctx, served, cfg):
    base, _ = _run(gpu_ctx, served, JOBS[:5])
    got, stats = _run(gpu_ctx, served, JOBS[:5], **cfg)
    for (a, b), (c, d) in zip(base, got):
        np.testing.assert_array_equal(a.view(np.uint32), c.view(np.uint32))
        np.testing.assert_array_equal(b.view(np.uint32), d.view(np.uint32))
    if "max_jobs" in cfg:
        assert stats["max_active"] <= cfg["max_jobs"]


def test_admission_keeps_within_the_hbm_budget(gpu_ctx, served):
    # room for two 64x64 films (32 B per pixel each) at a time
    budget = 2 * 64 * 64 * 32
    jobs = [("cbox", 64, 64, 3, 0, False, 0)] * 6
    res, stats = _run(gpu_ctx, served, jobs, hbm_budget=budget)
    assert stats["jobs_done"] == 6
    assert stats["max_active"] <= 2
    assert stats["hbm_peak"] <= budget
    for rt, _ in res[1:]:
        np.testing.assert_array_equal(rt.view(np.uint32), res[0][0].view(np.uint32))
    with va.Server(gpu_ctx, hbm_budget=budget) as srv:
        i = srv.add_scene(served[1]["cbox"])
        with pytest.raises(va.VisionError) as e:
            srv.submit(i, 128, 128, 1)  # needs 512 KiB > budget
        assert e.value.status == 5  # out of memory


def test_rejects_bad_jobs(gpu_ctx, served):
    with va.Server(gpu_ctx) as srv:
        i = srv.add_scene(served[1]["cbox"])
        for bad in [dict(width=0), dict(spp=0), dict(spectrum_dim=5, spectral=True)]:
            kw = dict(scene=i, width=16, height=16, spp=1)
            kw.update(bad)
            with pytest.raises(va.VisionError):
                srv.submit(**kw)
        with pytest.raises(va.VisionError):
            srv.submit(7, 16, 16, 1)  # unknown scene


def test_autotuner_picks_a_measured_shape(gpu_ctx, served):
    with va.Server(gpu_ctx) as srv:
        i = srv.add_scene(served[1]["cbox"])
        probe = [dict(scene=i, width=64, height=64, spp=2), dict(scene=i, width=128, height=32, spp=1)]
        r = srv.autotune(probe)
        assert len(r["trials"]) >= 4
        assert [(x["scene"], x["spectral"]) for x in r["routes"]] == [(i, False)]
        rt = r["routes"][0]
        # the lane is chosen only when it is more than twice as fast alone
        assert srv.route(i, False) == ("wavefront" if 2 * rt["wavefront_ms"] < rt["megakernel_ms"] else "megakernel")
        best = max(r["trials"], key=lambda t: t["msamples_per_s"])
        assert (r["tile_w"], r["waves_per_cu"], r["min_waves_per_eu"]) == (best["tile_w"], best["waves_per_cu"],
                                                                           best["min_waves_per_eu"])
        cfg = srv.config()
        assert (cfg["tile_w"], cfg["waves_per_cu"], cfg["min_waves_per_eu"]) == (r["tile_w"], r["waves_per_cu"],
                                                                                 r["min_waves_per_eu"])
        # the server still renders correctly under the chosen shape
        j = srv.submit(i, 64, 64, 2)
        srv.wait(j)
        film = srv.result(j, 0, (64, 64))
    rt, _ = _plain(gpu_ctx, served[0]["cbox"], 64, 64, 2)
    np.testing.assert_array_equal(film.view(np.uint32), rt.view(np.uint32))


def test_wavefront_lane_matches_the_plain_render(gpu_ctx, served):
    """Jobs routed to the wavefront lane render beside the persistent steps
    and give the same films."""
    base, _ = _run(gpu_ctx, served, JOBS)
    got, stats = _run(gpu_ctx, served, JOBS, staggered=True, routes=(("cbox", False), ("proc", True)))
    for (a, b), (c, d) in zip(base, got):
        np.testing.assert_array_equal(a.view(np.uint32), c.view(np.uint32))
        np.testing.assert_array_equal(b.view(np.uint32), d.view(np.uint32))
    assert stats["jobs_done"] == len(JOBS)


def test_large_jobs_go_to_the_lanes_and_match(gpu_ctx, served):
    """Jobs of lane_min_samples samples or more render on the wavefront
    lanes (several at once) and give the same films; the lane statistics
    count their samples, stage times and, in a counting run, their rays."""
    base, _ = _run(gpu_ctx, served, JOBS)
    lane_min = 64 * 64 * 2
    got, stats = _run(gpu_ctx, served, JOBS, lane_min_samples=lane_min)
    for (a, b), (c, d) in zip(base, got):
        np.testing.assert_array_equal(a.view(np.uint32), c.view(np.uint32))
        np.testing.assert_array_equal(b.view(np.uint32), d.view(np.uint32))
    big = sum(w * h * spp for _, w, h, spp, *_ in JOBS if w * h * spp >= lane_min)
    assert 0 < big < stats["samples_done"]
    assert stats["lane_samples"] == big
    assert all(ms > 0 for ms in stats["lane_stage_ms"]) and all(n > 0 for n in stats["lane_stage_launches"])
    _, cnt = _run(gpu_ctx, served, JOBS, lane_min_samples=lane_min, count_rays=True)
    assert cnt["lane_rays"][0] >= big  # at least one camera ray per sample
    assert cnt["lane_node_fetches"][0] > cnt["lane_rays"][0] > 0 and cnt["lane_tri_fetches"][0] > 0
    assert cnt["lane_rays"][1] > 0 and cnt["lane_node_fetches"][1] > 0  # shadow rays of the lit scenes
