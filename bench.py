#!/usr/bin/env python3
"""Benchmark: rendering throughput on one MI355X.

Workloads:
  serve (default)       batched serving against the plain per-job form on
                        the same seeded list of 1000 jobs over 8 resident
                        Cornell-box scenes of 0.1-2M triangles (resolutions
                        256^2..1920x1080, 1-64 spp, depth 8-16, sRGB or
                        hero-4): throughput with every job queued at once
                        (the timed step), then p50/p99 latency under Poisson
                        arrivals at 0.8 of the plain form's capacity.
  procedural            plain path tracing, 1920x1080, 64 spp, depth 16,
                        hero-4 spectral, 1.31M triangles + HDRI; a step is
                        one render job (film reset + all 64 frames).
  cbox                  the 256x256 Cornell box, 16 spp, depth 8, sRGB.
  restir                ReSTIR DI: the Cornell box lit by 10 000 emissive
                        triangles and 256 point / spot lights, 1920x1080,
                        1 spp per frame x 64 frames along a seeded camera
                        jitter path; a step is one 64-frame job.
Scene, BVH and film are resident in HBM before the timed region.  Prints one
JSON line.

The roofline entry is for the dominant kernel (the one with the most
HIP-event time in the timed region), with algorithmic bytes from a counting
run of the same job divided by that kernel's event time.  Traversal bytes
per ray: BVH8 node fetches x 80 B + leaf-triangle fetches x 48 B + 48 B ray
in / hit out.  Per-unit record bytes (the records the kernel reads and
writes, cache hits included):
  megakernel PT: traversal + 32 B film read / write per pixel;
  persistent serving kernel: traversal + 32 B film per sample (the first
      --count-jobs jobs, routed as in the timed run, are run once timed and
      once counting; jobs of --lane-min-samples samples or more go to the
      wavefront lanes, whose stages are accounted as below, and the dominant
      kernel is the one of the persistent kernel and the three lane stages
      with the most event time);
  wavefront closest-hit: traversal;
  wavefront shading: SHADE_BYTES per shaded path (path state in / out,
      hit, queues, triangle, vertices, instance, material, light sample)
      + 48 B per shadow ray written;
  wavefront any-hit: traversal + 32 B per shadow ray (contribution, radiance);
  ReSTIR shader0: traversal + 4 x 48 B records (surface and reservoir
      written, previous frame's read) + 124 B surface gathers + 128 B
      material + CANDIDATE_BYTES per light candidate, per pixel;
  ReSTIR shader1: traversal + records (pass-through, current reservoir,
      own and neighbour surfaces and reservoirs) + film, per pixel.
`traffic` is the HBM bytes the same kernel moved in rocprofv3 PMC runs of
the workload (profiles/hbm_traffic.json: FETCH_SIZE x 2.0 + WRITE_SIZE x
1.0, the factors a plain 256 MiB copy reads back), null where no PMC run
covers the kernel and workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Msamples/s on 1 MI355X at relMSE≤1e-3 (batched vs plain); p50/p99 job latency"
HBM_PEAK_GBS = 8000.0
NODE_BYTES, TRI_BYTES, RAY_BYTES, FILM_BYTES = 80, 48, 48, 32
RECORD_BYTES = 48      # ReSTIR surface / reservoir record
SURFACE_GATHER = 124   # triangle indices 12 + positions 36 + normals 36 + uvs 24 + instance 16
MATERIAL_BYTES = 128
CANDIDATE_BYTES = 184  # light selection alias 8 + func 4, light record 112, triangle alias 12, triangle 48
SHADE_BYTES = 240 + 16 + 8 + SURFACE_GATHER + MATERIAL_BYTES + CANDIDATE_BYTES


SCENE_CACHE = ROOT / "vision_amd" / "_scenes"
TRAFFIC = ROOT / "profiles" / "hbm_traffic.json"


def measured_traffic(key: str):
    """HBM bytes a kernel moved per sample or per launch in rocprofv3 PMC runs
    of this workload (FETCH_SIZE and WRITE_SIZE, each scaled by what the same
    counter reads for a plain 256 MiB copy); None where no run covers it."""
    try:
        return json.loads(TRAFFIC.read_text())["kernels"].get(key)
    except (OSError, ValueError, KeyError):
        return None


def workload(name: str):
    """(scene dict, base_dir, spp, config) of a named workload."""
    from vision_amd import scenegen
    if name == "cbox":
        spp = 16
        return scenegen.cornell_box(256, 256, spp=spp, max_depth=8), ".", spp, {
            "workload": "cornell box 256x256, 16 spp, max depth 8, sRGB, box filter (plain megakernel PT)",
            "resolution": [256, 256], "spp": spp, "max_depth": 8, "triangles": 36, "spectrum": "srgb"}
    if name == "procedural":
        spp = 64
        out = SCENE_CACHE / "procedural_s0"
        scene = scenegen.procedural_scene(out, seed=0, n_instances=64, level=5, width=1920, height=1080, spp=spp,
                                          max_depth=16, spectrum="hero", dimension=4, hdri=(2048, 1024))
        return scene, str(out), spp, {
            "workload": "procedural: 64 displaced level-5 icospheres (1,310,720 tris) + 2-tri ground, "
                        "matte/metal(Cu,Au,Al GGX r U[0.05,0.5])/glass(Cauchy 1.5, B 0.004)/Disney, "
                        "2048x1024 sky+sun HDRI (peak 5e4); 1920x1080, 64 spp, max depth 16, "
                        "hero spectral (4 wavelengths), box filter (plain megakernel PT)",
            "resolution": [1920, 1080], "spp": spp, "max_depth": 16, "triangles": 1310722,
            "spectrum": "hero4"}
    if name in ("restir", "restir_pairwise"):
        frames = 64
        scene = scenegen.many_lights_box(1920, 1080, seed=0)
        pairwise = name == "restir_pairwise"
        scene["integrator"]["param"]["direct"]["pairwise"] = pairwise
        return scene, ".", frames, {
            "workload": "ReSTIR DI: Cornell-box geometry with 10,000 emissive triangles (5,000 quads, radiance "
                        "log-uniform [1,100]) + 256 point/spot lights, power light sampler; 32 light + 1 BSDF "
                        "candidates, temporal reuse (history 20), 5 spatial neighbours within 30 px in 2 passes "
                        f"({'pairwise MIS' if pairwise else 'constant 1/M combine, the default'}); "
                        "1920x1080, 1 spp/frame x 64 frames along a seeded camera jitter path, sRGB, fp32 "
                        "reservoirs",
            "resolution": [1920, 1080], "frames": frames, "M_light": 32, "M_bsdf": 1, "spatial_passes": 2,
            "spatial_neighbours": 5, "history_limit": 20, "emissive_triangles": 10000, "point_spot_lights": 256,
            "spectrum": "srgb"}
    raise SystemExit(f"unknown workload {name}")


def host_cpu() -> dict:
    """Threads the CPU baseline uses and what the host has: the GPU box gives
    each GPU a share of 16 host threads, whatever nproc shows."""
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    return {"cores": max(1, min(16, nproc)), "nproc": nproc, "cpu_model": model,
            "cores_note": "16 threads: the host-CPU share per GPU on the GPU box"}


def cpu_baseline(scene, spp: int, budget_s: float = 15.0) -> dict:
    """The oracle (scalar C++ port of the reference path) on the host cores
    (at most 16, this machine's share per GPU), over a bounded band of rows
    of the same job; its per-call setup (BVH2 and light tables) is timed
    separately and left out."""
    from oracle import oracle
    ex = scene.export()
    W, H = int(ex["camera"][19]), int(ex["camera"][20])
    hc = host_cpu()
    cores = hc["cores"]
    t0 = time.perf_counter()
    oracle.render(ex, 0, spp, window=(0, 0, 0, 0), threads=cores)
    setup = time.perf_counter() - t0
    rows = 1
    while True:
        y0 = (H - rows) // 2
        t0 = time.perf_counter()
        oracle.render(ex, 0, spp, window=(0, y0, W, y0 + rows), threads=cores)
        t = max(time.perf_counter() - t0 - setup, 1e-6)
        if t > budget_s / 4 or rows >= H:
            break
        rows = min(H, rows * 4)
    samples = W * rows * spp
    return {"value": round(samples / t / 1e6, 4), "unit": "Msamples/s", **hc, "kind": "port",
            "sample": f"{rows} centre rows x {W} px x {spp} spp of the same job ({samples} samples, {t:.2f} s "
                      f"after {setup:.2f} s setup)"}


def restir_cpu_baseline(scene, frames: int = 2) -> dict:
    """The oracle's ReSTIR DI frames on the host cores (at most 16) over the
    full image: the first `frames` frames of the same job."""
    from oracle import oracle
    from vision_amd import scenegen
    hc = host_cpu()
    cores = hc["cores"]
    r = oracle.Restir(scene.export(), threads=cores)
    cams = scenegen.jitter_path(frames, seed=1)
    t0 = time.perf_counter()
    for c2w in cams:
        scene.update_camera(c2w, 19.5)
        r.render(1, scene=scene.export())
    t = time.perf_counter() - t0
    samples = r.W * r.H * frames
    return {"value": round(samples / t / 1e6, 4), "unit": "Msamples/s", **hc, "kind": "port",
            "sample": f"the first {frames} frames of the job at {r.W}x{r.H} ({samples} samples, {t:.2f} s, "
                      f"per-frame BVH2 / light-table setup included)"}


SERVE_RES = ((256, 256), (512, 512), (1024, 1024), (1920, 1080))
SERVE_SPP = (1, 4, 16, 64)


def serve_jobs(n: int, n_scenes: int, seed: int = 0) -> list:
    """The serving mix: scene uniform over the resident scenes, resolution
    from SERVE_RES, spp from SERVE_SPP, max depth uniform in [8, 16], sRGB or
    hero-4 spectral with equal odds."""
    import numpy as np
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(n):
        w, h = SERVE_RES[int(rng.integers(0, len(SERVE_RES)))]
        jobs.append(dict(scene=int(rng.integers(0, n_scenes)), width=w, height=h,
                         spp=SERVE_SPP[int(rng.integers(0, len(SERVE_SPP)))], max_depth=int(rng.integers(8, 17)),
                         spectral=bool(rng.integers(0, 2))))
    return jobs


def plain_job(sc, job: dict) -> float:
    """One job the plain way: switch the scene to the job's spectrum mode,
    resolution and depth, render all its frames in one megakernel launch,
    download the output film.  Returns its service time in seconds."""
    t0 = time.perf_counter()
    sc.set_spectrum(job["spectral"], 4)
    sc.set_resolution(job["width"], job["height"])
    sc.set_integrator(job["max_depth"], 5, 1.0)
    sc.render(job["spp"], fused=True)
    sc.film(1)
    return time.perf_counter() - t0


def lindley(arrivals, service) -> list:
    """Latencies of a first-come first-served single server (one job on the
    device at a time) from arrival and service times."""
    free, out = 0.0, []
    for a, s in zip(arrivals, service):
        start = max(a, free)
        free = start + s
        out.append(free - a)
    return out


def percentile(xs, q: float) -> float:
    ys = sorted(xs)
    r = max(1, int(-(-q * len(ys) // 1)))
    return ys[min(len(ys), r) - 1]


def serve_main(args) -> None:
    """Batched serving against the plain per-job form on the same job list:
    throughput with every job queued at once, then latency under Poisson
    arrivals at `load` times the plain form's capacity."""
    import threading

    import numpy as np
    import vision_amd as va
    from vision_amd import scenegen

    out = SCENE_CACHE / "serving"
    ctx = va.Context(0)
    scenes, build_s, tris = [], 0.0, []
    for k, t in enumerate(scenegen.SERVING_TRIANGLES):
        sd = scenegen.serving_scene(out, t, seed=k)
        sc = ctx.load_text(sd, str(out))
        t0 = time.perf_counter()
        sc.build_accel()
        build_s += time.perf_counter() - t0
        tris.append(sc.info()["num_triangles"])
        scenes.append(sc)
        print(f"scene {k}: {tris[-1]} triangles", file=sys.stderr, flush=True)
    # a step is `jobs_per_step` jobs of the seeded mix; the timed region
    # queues all `steps` steps' jobs at once (continuous batching across step
    # boundaries) and drains the device at both ends
    jobs = serve_jobs(args.steps * args.jobs_per_step, len(scenes), seed=args.seed)
    warm_jobs = serve_jobs(max(args.warmup, 1) * args.jobs_per_step, len(scenes), seed=args.seed + 200)
    total_samples = sum(j["width"] * j["height"] * j["spp"] for j in jobs)

    # plain form: one job at a time, in order (a few untimed jobs first)
    for j in warm_jobs[:8]:
        plain_job(scenes[j["scene"]], j)
    service = []
    t0 = time.perf_counter()
    for n, j in enumerate(jobs):
        service.append(plain_job(scenes[j["scene"]], j))
        if n % 100 == 99:
            print(f"plain: {n + 1} jobs", file=sys.stderr, flush=True)
    plain_wall = time.perf_counter() - t0
    plain_msps = total_samples / plain_wall / 1e6
    # what filling the GPU can gain over this plain form: every job at the
    # best rate a plain job of its scene and spectrum mode reached (a large
    # job keeps the GPU full, so that rate stands for the kernel's own
    # throughput), against the plain form's measured service times.  The
    # batched form can exceed it only through a faster kernel.
    best = {}
    for j, t in zip(jobs, service):
        key = (j["scene"], j["spectral"])
        best[key] = max(best.get(key, 0.0), j["width"] * j["height"] * j["spp"] / t)
    ideal_s = sum(j["width"] * j["height"] * j["spp"] / best[(j["scene"], j["spectral"])] for j in jobs)
    batching_bound = sum(service) / ideal_s
    small = sum(j["width"] * j["height"] * j["spp"] for j in jobs if j["width"] * j["height"] * j["spp"] < 1 << 20)
    # plain films of a few jobs for the batched-vs-plain comparison
    check = list(range(min(args.check_jobs, len(jobs))))
    plain_films = {}
    for i in check:
        plain_job(scenes[jobs[i]["scene"]], jobs[i])
        plain_films[i] = scenes[jobs[i]["scene"]].film(0)

    srv = va.Server(ctx, max_jobs=args.max_jobs, step_ms=args.step_ms, tile_w=args.tile_w,
                    waves_per_cu=args.waves_per_cu, min_waves_per_eu=args.min_waves_per_eu,
                    lane_min_samples=args.lane_min_samples)
    for sc in scenes:
        srv.add_scene(sc)

    lane_total = [0]  # lane samples over every batch of this process (normalises PMC sums)

    def run_batch(js, arrivals=None):
        """Submit js (at their arrival offsets, or all at once) and wait for
        all; returns the server statistics."""
        srv.reset_stats()
        ids = [None] * len(js)

        def submitter():
            t_start = time.perf_counter()
            for i, j in enumerate(js):
                if arrivals is not None:
                    d = t_start + arrivals[i] - time.perf_counter()
                    if d > 0:
                        time.sleep(d)
                ids[i] = srv.submit(j["scene"], j["width"], j["height"], j["spp"], max_depth=j["max_depth"],
                                    spectral=j["spectral"], spectrum_dim=4)

        th = threading.Thread(target=submitter)
        th.start()
        done = 0
        for i in range(len(js)):
            while ids[i] is None:
                time.sleep(0.0005)
            srv.wait(ids[i])
            done += 1
            if done % 100 == 0:
                print(f"batched: {done} jobs", file=sys.stderr, flush=True)
        th.join()
        st = srv.stats()
        lane_total[0] += st["lane_samples"]
        return ids, st

    # warm-up steps (graph capture, first-step rate estimate)
    wids, _ = run_batch(warm_jobs)
    for i in wids:
        srv.release(i)
    tune = None
    if args.autotune:
        # launch shape picked on a separate probe list of the same mix
        probe = serve_jobs(args.autotune, len(scenes), seed=args.seed + 100)
        tune = srv.autotune(probe)
        print(f"autotune: {tune}", file=sys.stderr, flush=True)
    # throughput: every job queued at once
    ids, st_sat = run_batch(jobs)
    relmse = 0.0
    identical = 0
    for i in check:
        j = jobs[i]
        film = srv.result(ids[i], 0, (j["height"], j["width"]))
        ref = plain_films[i]
        identical += int(np.array_equal(film.view(np.uint32), ref.view(np.uint32)))
        x, r = film[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
        relmse = max(relmse, float(np.mean((x - r) ** 2 / (r * r + 1e-2))))
    for i in ids:
        srv.release(i)
    batched_msps = st_sat["msamples_per_s"]
    # latency: Poisson arrivals at load x the plain capacity
    mean_service = sum(service) / len(service)
    lam = args.load / mean_service
    rng = np.random.default_rng(args.seed + 1)
    arrivals = np.cumsum(rng.exponential(1.0 / lam, len(jobs))).tolist()
    arrivals = [a - arrivals[0] for a in arrivals]
    ids, st_lat = run_batch(jobs, arrivals)
    for i in ids:
        srv.release(i)
    plain_lat = lindley(arrivals, service)
    # roofline of the dominant kernel: the first count_jobs jobs, routed as in
    # the timed run, once timed (the persistent steps' event time and the
    # lanes' per-stage event time) and once counting rays and BVH fetches
    lane_pairs = [(r["scene"], r["spectral"]) for r in (tune or {}).get("routes", []) if r["route"] == "wavefront"]
    sub = jobs[:args.count_jobs]
    ids, st_time = run_batch(sub)
    for i in ids:
        srv.release(i)
    srv.set_config(count_rays=True)
    ids, st_cnt = run_batch(sub)
    for i in ids:
        srv.release(i)
    srv.set_config(count_rays=False)
    sub_samples = sum(j["width"] * j["height"] * j["spp"] for j in sub)
    lane_samples = st_cnt["lane_samples"]
    pers_samples = sub_samples - lane_samples
    stage_names = ["trace_queue_kernel (wavefront lanes' closest-hit traversal)",
                   "shade_kernel (wavefront lanes' shading)", "shadow_kernel (wavefront lanes' any-hit traversal)"]
    cands = [("serve_kernel (persistent multi-job megakernel)", st_time["busy_ms"])]
    cands += [(stage_names[k], st_time["lane_stage_ms"][k]) for k in range(3)]
    kernel, kernel_ms = max(cands, key=lambda c: c[1])
    rays_total = st_cnt["rays"] + sum(st_cnt["lane_rays"])
    rays_per_sample = rays_total / sub_samples
    if kernel.startswith("serve_kernel"):
        alg_bytes = (NODE_BYTES * st_cnt["node_fetches"] + TRI_BYTES * st_cnt["tri_fetches"] +
                     RAY_BYTES * st_cnt["rays"] + FILM_BYTES * pers_samples)
        unit_samples, launches = pers_samples, None
        tr = measured_traffic("serve_kernel@serving_mix") if not lane_samples else None
    else:
        k = stage_names.index(kernel)
        trav = lambda t: (NODE_BYTES * st_cnt["lane_node_fetches"][t] + TRI_BYTES * st_cnt["lane_tri_fetches"][t] +
                          RAY_BYTES * st_cnt["lane_rays"][t])
        alg_bytes = [trav(0), SHADE_BYTES * st_cnt["lane_rays"][0] + RAY_BYTES * st_cnt["lane_rays"][1],
                     trav(1) + 32 * st_cnt["lane_rays"][1]][k]
        unit_samples, launches = lane_samples, st_time["lane_stage_launches"][k]
        tr = measured_traffic(kernel.split()[0] + "@serving_mix_lanes")
    achieved = alg_bytes / (kernel_ms / 1e3) / 1e9
    line = {
        "metric": METRIC,
        "value": round(batched_msps, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(st_sat["wall_ms"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"batched serving: {len(jobs)} jobs over 8 resident Cornell-box scenes of "
                        f"{min(tris):,}-{max(tris):,} triangles (displaced level-5 icospheres, diffuse); "
                        "resolution from {256^2, 512^2, 1024^2, 1920x1080}, spp from {1, 4, 16, 64}, max depth "
                        "U[8, 16], sRGB or hero-4 spectral; persistent multi-job kernels in captured hipGraphs, "
                        f"HBM-budget admission; a step is {args.jobs_per_step} jobs of the mix, the timed "
                        f"{args.steps} steps' jobs queued at once",
            "jobs_per_step": args.jobs_per_step,
            "jobs": len(jobs), "samples": total_samples, "triangles": tris, "seed": args.seed,
            "max_jobs": args.max_jobs, "step_ms": args.step_ms, **{k: srv.config()[k] for k in
                                                                    ("tile_w", "waves_per_cu",
                                                                              "min_waves_per_eu")}},
        "autotune": tune,
        "plain_msamples_per_s": round(plain_msps, 3),
        "batched_over_plain": round(batched_msps / plain_msps, 3),
        "occupancy_gain_bound": round(batching_bound, 3),
        "occupancy_gain_note": ("what filling the GPU can gain over the plain form: its service time over the time "
                                "of every job at the best plain rate of its scene and spectrum mode; jobs under "
                                f"1 M samples carry {100.0 * small / total_samples:.2f}% of the samples"),
        "relmse_batched_vs_plain": relmse,
        "bit_identical_jobs_checked": f"{identical}/{len(check)}",
        "latency": {
            "arrivals": "Poisson", "rate_jobs_per_s": round(lam, 3), "load_vs_plain_capacity": args.load,
            "batched_p50_ms": round(st_lat["latency_p50_ms"], 2), "batched_p99_ms": round(st_lat["latency_p99_ms"], 2),
            "plain_p50_ms": round(percentile(plain_lat, 0.5) * 1e3, 2),
            "plain_p99_ms": round(percentile(plain_lat, 0.99) * 1e3, 2),
            "plain_from": "measured per-job service times, first-come first-served on one device",
            "saturated_batched_p50_ms": round(st_sat["latency_p50_ms"], 2),
            "saturated_batched_p99_ms": round(st_sat["latency_p99_ms"], 2)},
        "mrays_per_s": round(rays_per_sample * batched_msps, 2),
        "rays_per_sample": round(rays_per_sample, 3),
        "bvh_build_s": round(build_s, 3),
        "steps_run": st_sat["steps"], "graph_launches": st_sat["graph_launches"],
        "busy_ms_persistent_steps": round(st_sat["busy_ms"], 3),
        "max_active_jobs": st_sat["max_active"], "hbm_peak_bytes": st_sat["hbm_peak"],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": round(tr["hbm_bytes_per_sample"] * unit_samples) if tr else None,
                     "traffic_unit": "HBM bytes over the counting jobs, the same unit as the algorithmic bytes",
                     "kernel": kernel, "kernel_ms": round(kernel_ms, 3), "kernel_jobs": len(sub),
                     "launches": launches,
                     "avg_launch_ms": round(kernel_ms / launches, 4) if launches else None,
                     "algorithmic_bytes": int(alg_bytes),
                     "algorithmic_bytes_per_sample": round(alg_bytes / max(unit_samples, 1), 1),
                     "hbm_bytes_per_sample": tr["hbm_bytes_per_sample"] if tr else None,
                     "candidates_ms": {name.split()[0]: round(ms, 3) for name, ms in cands},
                     "note": ("the lanes' stages run on three streams beside each other and the persistent "
                              "kernel; a lane launch's event time includes the time it shares the device, so "
                              "achieved is a lower bound on the stage's rate alone") if lane_samples else None},
        "lane_min_samples": args.lane_min_samples,
        "lane_sample_share": round(st_sat["lane_samples"] / max(st_sat["samples_done"], 1), 4),
        "lane_samples_all_batches": lane_total[0],
        "wavefront_lane_routes": lane_pairs,
    }
    if not args.no_cpu:
        # a 1920x1080 x 4 spp, depth 12, sRGB job on the smallest scene; the
        # oracle renders a band of its rows
        sc = scenes[0]
        sc.set_spectrum(False, 4)
        sc.set_resolution(1920, 1080)
        sc.set_integrator(12, 5, 1.0)
        line["cpu_baseline"] = cpu_baseline(sc, 4, budget_s=40.0)
        line["cpu_baseline"]["sample"] += f"; scene 0 ({tris[0]} triangles), 1920x1080 x 4 spp job, depth 12, sRGB"
    srv.close()
    for sc in scenes:
        sc.close()
    ctx.close()
    if not args.no_configs:
        # the launch shape the serving run used (autotuned or given)
        shape = {k: line["config"][k] for k in ("tile_w", "waves_per_cu", "min_waves_per_eu")}
        shape["lane_min_samples"] = args.lane_min_samples
        line["mixed_material"] = serve_mixed(args, shape)
        line["configs"] = other_configs(args)
    print(json.dumps(line))


def _compact(l: dict) -> dict:
    keep = ("value", "unit", "steps", "warmup", "ms_per_step", "mrays_per_s", "rays_per_sample", "bvh_build_s",
            "relmse_vs_4096spp", "relmse_detail", "relmse_at", "relmse_reference", "stages_ms_per_step", "note")
    out = {k: l[k] for k in keep if k in l}
    out["workload"] = l["config"]["workload"] + (f"; {l['config']['kernel']}" if "kernel" in l["config"] else "")
    rf = l["roofline"]
    out["roofline"] = {k: rf[k] for k in ("kernel", "achieved", "frac", "avg_launch_ms", "traffic", "note") if k in rf}
    return out


def serve_mixed(args, shape: dict, n_jobs: int = 100) -> dict:
    """Serving over mixed-material scenes (the procedural generator's matte /
    metal / glass / Disney icospheres under the sky HDRI, 102k-2.0M
    triangles: the every-material kernels): a seeded list of n_jobs jobs of
    the serving mix rendered the plain way one at a time, then all queued at
    once through the server at the launch shape `shape` (tile_w,
    waves_per_cu, min_waves_per_eu); two jobs' films compared bit for bit."""
    import numpy as np
    import vision_amd as va
    from vision_amd import scenegen
    t_start = time.perf_counter()
    out = SCENE_CACHE / "procedural_s0"
    ctx = va.Context(0)
    scenes, tris = [], []
    for n in (5, 17, 40, 98):
        sd = scenegen.procedural_scene(out, seed=0, n_instances=n, level=5, width=256, height=256, spp=1,
                                       max_depth=16, spectrum="hero", dimension=4, hdri=(2048, 1024))
        sc = ctx.load_text(sd, str(out))
        sc.build_accel()
        sc.set_kernel("megakernel")
        scenes.append(sc)
        tris.append(sc.info()["num_triangles"])
    jobs = serve_jobs(n_jobs, len(scenes), seed=args.seed + 300)
    total = sum(j["width"] * j["height"] * j["spp"] for j in jobs)
    for j in serve_jobs(8, len(scenes), seed=args.seed + 301):
        plain_job(scenes[j["scene"]], j)
    t0 = time.perf_counter()
    for j in jobs:
        plain_job(scenes[j["scene"]], j)
    plain_msps = total / (time.perf_counter() - t0) / 1e6
    check = {}
    for i in range(2):
        plain_job(scenes[jobs[i]["scene"]], jobs[i])
        check[i] = scenes[jobs[i]["scene"]].film(0)
    srv = va.Server(ctx, max_jobs=args.max_jobs, step_ms=args.step_ms, **shape)
    for sc in scenes:
        srv.add_scene(sc)

    def batch(js):
        srv.reset_stats()
        ids = [srv.submit(j["scene"], j["width"], j["height"], j["spp"], max_depth=j["max_depth"],
                          spectral=j["spectral"], spectrum_dim=4) for j in js]
        for i in ids:
            srv.wait(i)
        return ids, srv.stats()

    ids, _ = batch(serve_jobs(16, len(scenes), seed=args.seed + 302))
    for i in ids:
        srv.release(i)
    ids, st = batch(jobs)
    identical = 0
    for i, ref in check.items():
        j = jobs[i]
        film = srv.result(ids[i], 0, (j["height"], j["width"]))
        identical += int(np.array_equal(film.view(np.uint32), ref.view(np.uint32)))
    for i in ids:
        srv.release(i)
    srv.close()
    for sc in scenes:
        sc.close()
    ctx.close()
    res = {"value": round(st["msamples_per_s"], 3), "unit": "Msamples/s", "plain_msamples_per_s": round(plain_msps, 3),
           "batched_over_plain": round(st["msamples_per_s"] / plain_msps, 3),
           "bit_identical_jobs_checked": f"{identical}/{len(check)}", "jobs": n_jobs, "samples": total,
           "triangles": tris, **shape,
           "workload": "mixed-material serving: jobs of the serving mix (256^2..1920x1080, 1-64 spp, depth "
                       "U[8, 16], sRGB or hero-4) over 4 procedural scenes (displaced level-5 icospheres of "
                       "matte / metal / glass / Disney under the sky + sun HDRI) of the listed triangle counts"}
    print(f"mixed-material serving: {res['value']} Msamples/s, plain {res['plain_msamples_per_s']} "
          f"({time.perf_counter() - t_start:.1f} s)", file=sys.stderr, flush=True)
    return res


def other_configs(args) -> dict:
    """The plain workloads beside the serving headline, each a few timed
    jobs plus relMSE against a 4096-spp plain render: the Cornell box 256^2 x
    16 spp (megakernel), the procedural 1080p x 64 spp hero-4 scene
    (wavefront form) and ReSTIR DI over 64 jitter-path frames at 1080p, with
    the reference's default constant 1/M combine and with pairwise MIS (the
    constant combine darkens where neighbours' samples lack support at the
    pixel: one-sided emitters, spot cones; pairwise MIS does not)."""
    out = {}
    for key, name, steps, form in (("C1_cbox", "cbox", 5, "megakernel"), ("C2_procedural", "procedural", 2, "wavefront"),
                                   ("C3_restir", "restir", 2, "wavefront"),
                                   ("C3_restir_pairwise", "restir_pairwise", 2, "wavefront")):
        t0 = time.perf_counter()
        l = plain_workload(name, steps, 1, form, args.chunk_frames, args.trace_block, cpu=False)
        out[key] = _compact(l)
        print(f"config {key}: {out[key]['value']} Msamples/s, relMSE {out[key].get('relmse_vs_4096spp')} "
              f"({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    return out


def nrc_main(args) -> None:
    """The radiance cache on the Cornell box at 1920x1080, 4 spp per frame x
    128 frames, depth 8: paths stop into the cache at bounce 2, 65 536
    training paths per frame, 4 Adam steps of 65 536 samples.  Against the
    plain path tracer on the same frames; both judged by relMSE against a
    4096-spp plain render."""
    import numpy as np
    import vision_amd as va
    from vision_amd import scenegen

    W, H, spp, frames = 1920, 1080, 4, args.nrc_frames
    base = "."
    if args.nrc_scene == "procedural":
        # the C2 scene in sRGB mode at depth 16 (glass / metal / Disney under
        # the sky HDRI: long, costly paths)
        base = str(SCENE_CACHE / "procedural_s0")
        scene_d = scenegen.procedural_scene(base, seed=0, n_instances=64, level=5, width=W, height=H, spp=spp,
                                            max_depth=16, spectrum="srgb", hdri=(2048, 1024))
        label, depth = "procedural scene (64 icospheres, 1.31M triangles, sky HDRI), sRGB", 16
    else:
        scene_d = scenegen.cornell_box(W, H, max_depth=8)
        label, depth = "Cornell box", 8
    scene = json.dumps(scene_d)
    ctx = va.Context(0)

    def relmse(x, r):
        x, r = x[..., :3].astype(np.float64), r[..., :3].astype(np.float64)
        return float(np.mean((x - r) ** 2 / (r * r + 1e-2)))

    ref = ctx.load_text(scene, base).build_accel()
    left = args.nrc_ref_spp
    while left > 0:
        ref.render(min(512, left), fused=True)
        left -= min(512, left)
    R = ref.film(0)
    ref.close()
    plain = ctx.load_text(scene, base).build_accel()
    cached = ctx.load_text(scene, base).build_accel()
    with va.RadianceCache(plain) as p0, va.RadianceCache(cached) as nrc:
        for _ in range(args.warmup):
            p0.render(1, spp, terminate=False, train=False)
            nrc.render(1, spp)
        plain.reset_film()
        cached.reset_film()
        t0 = time.perf_counter()
        sp = p0.render(frames, spp, terminate=False, train=False)
        plain.film(0)
        t_plain = time.perf_counter() - t0
        t0 = time.perf_counter()
        sn = nrc.render(frames, spp)
        cached.film(0)
        t_nrc = time.perf_counter() - t0
        e_plain, e_nrc = relmse(plain.film(0), R), relmse(cached.film(0), R)
        # the plain path tracer given the cache's wall time
        frames_eq = max(1, int(round(frames * t_nrc / t_plain)))
        plain.reset_film()
        p0.render(frames_eq, spp, terminate=False, train=False)
        e_plain_eq = relmse(plain.film(0), R)
        samples = W * H * spp * frames
        # MFMA work of the cache per frame: forward on the queries, forward
        # + backward (3x) on the training batches
        f_fwd = 2 * (32 * 64 + 64 * 64 + 64 * 64 + 64 * 3)
        q = sn["queries_last_frame"]
        trained = 4 * min(sn["train_samples_last_frame"], 65536)
        infer_tflops = f_fwd * q * frames / (sn["infer_ms"] / 1e3) / 1e12
        line = {
            "metric": METRIC, "value": round(samples / t_nrc / 1e6, 3), "unit": "Msamples/s", "n_gpus": 1,
            "steps": 1, "warmup": args.warmup, "ms_per_step": round(t_nrc * 1e3, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (MLP) / f32", "data": "synthetic",
            "config": {"workload": f"radiance cache: {label} 1920x1080, {spp} spp/frame x {frames} frames, "
                                   f"depth {depth}, paths terminate into the cache at bounce 2; hash grid 16 levels x 2^19 x "
                                   "2 fp16, MLP 32-64-64-64-3 ReLU on bf16 MFMA, Adam lr 1e-2; 65 536 training "
                                   "paths/frame, 4 steps x 65 536 samples",
                       "reference_spp": args.nrc_ref_spp},
            "plain_msamples_per_s": round(samples / t_plain / 1e6, 3),
            "relmse_cache": e_nrc, "relmse_plain": e_plain,
            "relmse_plain_equal_time": e_plain_eq, "plain_frames_equal_time": frames_eq,
            "stage_ms": {k: round(sn[k], 2) for k in ("render_ms", "train_paths_ms", "infer_ms", "train_ms",
                                                      "accumulate_ms")},
            "plain_render_ms": round(sp["render_ms"], 2),
            "queries_per_frame": q, "train_samples_per_frame": sn["train_samples_last_frame"],
            "last_loss": sn["last_loss"],
            "roofline": {"bound": "mfma", "achieved": round(infer_tflops, 3), "peak": 2500.0, "unit": "TFLOP/s",
                         "frac": round(infer_tflops / 2500.0, 5), "traffic": None,
                         "kernel": "nrc_forward_kernel (hash-grid encoding + bf16 MLP inference)",
                         "flops_per_query": f_fwd, "trained_samples_per_frame": trained},
        }
    plain.close()
    cached.close()
    ctx.close()
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="serve",
                    choices=["serve", "procedural", "cbox", "restir", "restir_pairwise", "nrc"])
    ap.add_argument("--nrc-frames", type=int, default=128, help="nrc: frames of 4 spp")
    ap.add_argument("--nrc-ref-spp", type=int, default=4096, help="nrc: spp of the plain reference render")
    ap.add_argument("--nrc-scene", default="cbox", choices=["cbox", "procedural"], help="nrc: scene")
    ap.add_argument("--jobs-per-step", type=int, default=50, help="serve: jobs in one step of the list")
    ap.add_argument("--no-configs", action="store_true",
                    help="serve: skip the mixed-material serving and the cbox / procedural / restir keys")
    ap.add_argument("--seed", type=int, default=0, help="serve: job-list seed")
    ap.add_argument("--load", type=float, default=0.8, help="serve: Poisson rate / plain capacity")
    ap.add_argument("--max-jobs", type=int, default=64, help="serve: jobs admitted at once")
    ap.add_argument("--step-ms", type=float, default=16.0)
    ap.add_argument("--tile-w", type=int, default=8)
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--min-waves-per-eu", type=int, default=0, help="serve: kernel register budget (0 or 6)")
    ap.add_argument("--lane-min-samples", type=int, default=1_000_000,
                    help="serve: jobs of at least this many samples render on the wavefront lanes (0: by route only)")
    ap.add_argument("--check-jobs", type=int, default=16, help="serve: jobs compared with their plain film")
    ap.add_argument("--count-jobs", type=int, default=100, help="serve: jobs in the ray-counting run")
    ap.add_argument("--autotune", type=int, default=48, help="serve: probe jobs for the launch-shape autotuner "
                                                             "(0: keep --tile-w / --waves-per-cu)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-quality", action="store_true", help="skip the relMSE against 4096-spp references")
    ap.add_argument("--kernel", default="wavefront", choices=["megakernel", "wavefront"])
    # wavefront launch shape, measured on the procedural job: frames per chunk
    # 8 / 16 / 32 / 64 gave 477 / 514 / 531 / 538 Msamples/s at traversal
    # blocks of 64 threads, which beat 128 and 256 at every chunk size
    ap.add_argument("--chunk-frames", type=int, default=64)
    ap.add_argument("--trace-block", type=int, default=64)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 256 MiB device to device three times (the reference point that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profiling run to bytes)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        x = torch.ones(64 * 1024 * 1024, device="cuda", dtype=torch.float32)
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": x.numel() * 4, "copies": 3}))
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU scaling is not part of it", file=sys.stderr)
        sys.exit(2)
    if args.workload == "serve":
        serve_main(args)
        return
    if args.workload == "nrc":
        nrc_main(args)
        return

    line = plain_workload(args.workload, args.steps, args.warmup, args.kernel, args.chunk_frames, args.trace_block,
                          cpu=not args.no_cpu, quality=not args.no_quality)
    print(json.dumps(line))


def relmse(x, ref) -> float:
    """mean over pixels and channels of (x - ref)^2 / (ref^2 + 1e-2)"""
    import numpy as np
    x, r = x[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    return float(np.mean((x - r) ** 2 / (r * r + 1e-2)))


def relmse_detail(x, ref) -> dict:
    """How the relMSE is spread over pixels: its median per pixel, the mean
    without the 0.1% of pixels with the largest error, that 0.1%'s share of
    the total, and the largest per-pixel term (a few fireflies of a bright
    HDRI sun or emitter can carry the mean); the film's mean over the
    reference's, over all pixels and without that 0.1% (a biased estimator
    shows in the second, a few fireflies only in the first)."""
    import numpy as np
    x, r = x[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    e = ((x - r) ** 2 / (r * r + 1e-2)).mean(axis=-1).ravel()
    order = np.sort(e)
    k = max(1, int(len(e) * 0.001))
    keep = e < order[-k]
    xm, rm = x.mean(axis=-1).ravel(), r.mean(axis=-1).ravel()
    return {"median": float(np.median(e)), "mean_without_top_0.1pct": float(order[:-k].mean()),
            "top_0.1pct_share": float(order[-k:].sum() / max(e.sum(), 1e-300)), "max_pixel": float(order[-1]),
            "mean_over_reference": float(x.mean() / max(r.mean(), 1e-300)),
            "mean_over_reference_without_top_0.1pct": float(xm[keep].mean() / max(rm[keep].mean(), 1e-300))}


def reference_film(ctx, scene_desc, base_dir, spp: int, resolution=None, pt_depth=None, cameras=None,
                   fov_y: float = 19.5):
    """A plain path-traced film of `spp` samples per pixel (megakernel, in
    launches of at most 512 frames), optionally at another resolution or as
    a path tracer of the given max depth (for an "rt" scene: its direct-light
    estimate is the PT image of max depth 1).  With `cameras` (camera-to-world
    matrices) the spp are split evenly over them and the films averaged: the
    reference of a job whose camera follows that path."""
    import copy
    import numpy as np
    d = copy.deepcopy(scene_desc)
    if pt_depth is not None:
        d["integrator"] = {"type": "pt", "param": {"max_depth": pt_depth, "min_depth": 5, "rr_threshold": 1}}
    sc = ctx.load_text(d, base_dir)
    sc.build_accel()
    if resolution:
        sc.set_resolution(*resolution)
    sc.set_kernel("megakernel")

    def render(n):
        sc.reset_film()
        left = n
        while left > 0:
            sc.render(min(512, left), fused=True)
            left -= min(512, left)
        return sc.film(0).astype(np.float64)

    if cameras is None:
        film = render(spp)
    else:
        per = max(1, spp // len(cameras))
        film = 0.0
        for c2w in cameras:
            sc.update_camera(c2w, fov_y)
            film = film + render(per)
        film = film / len(cameras)
    sc.close()
    return np.asarray(film, np.float32)


def plain_workload(name: str, steps: int, warmup: int, kernel_form: str, chunk_frames: int, trace_block: int,
                   cpu: bool = True, quality: bool = True, ref_spp: int = 4096) -> dict:
    """One of the plain (non-serving) workloads: `steps` timed jobs after
    `warmup` untimed ones, a counting run for the roofline's algorithmic
    bytes, relMSE against a 4096-spp plain render, the CPU baseline."""
    import numpy as np

    import vision_amd as va
    from vision_amd import scenegen
    scene_desc, base_dir, spp, config = workload(name)
    restir = name.startswith("restir")
    ctx = va.Context(0)
    sc = ctx.load_text(scene_desc, base_dir)
    t0 = time.perf_counter()
    sc.build_accel()
    build_s = time.perf_counter() - t0
    if not restir:
        sc.set_kernel(kernel_form, chunk_frames, trace_block)
        config["kernel"] = kernel_form + (f" ({chunk_frames} frames per chunk)" if kernel_form == "wavefront"
                                          else "")
    info = sc.info()
    W, H = info["width"], info["height"]
    cams = scenegen.jitter_path(spp, seed=1) if restir else None

    def step():
        """One job; returns the summed stats of its render calls."""
        sc.reset_film()
        if not restir:
            return sc.render(spp, fused=True)
        tot = None
        for c2w in cams:
            sc.update_camera(c2w, 19.5)
            st = sc.render(1)
            if tot is None:
                tot = st
            else:
                for k, v in st.items():
                    tot[k] = [a + b for a, b in zip(tot[k], v)] if isinstance(v, list) else tot[k] + v
        return tot

    for _ in range(warmup):
        step()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(step())
    wall = time.perf_counter() - t0
    ms_per_step = wall / steps * 1e3
    samples = W * H * spp
    msps = samples * steps / wall / 1e6

    # counting run of the same job: rays, node and triangle fetches
    sc.set_counting(True)
    cst = step()
    sc.set_counting(False)
    trav = lambda k: (NODE_BYTES * cst["stage_node_fetches"][k] + TRI_BYTES * cst["stage_tri_fetches"][k] +
                      RAY_BYTES * cst["stage_rays"][k])
    avg_kernel_s = sum(s["kernel_ms"] for s in stats) / len(stats) / 1e3
    staged = restir or kernel_form == "wavefront"
    if staged:
        stage_s = [sum(s["stage_ms"][k] for s in stats) / len(stats) / 1e3 for k in range(3)]
        k = max(range(3), key=lambda i: stage_s[i])
        px = W * H * spp
        if restir:
            r = scene_desc["integrator"]["param"]["direct"]
            n_sp = r["spatial"]["sample_num"]
            names = ["initial_kernel (shader0: candidates, visibility, temporal reuse)",
                     "spatial_kernel (extra spatial pass)", "final_kernel (shader1: spatial reuse, shading)"]
            unit = [4 * RECORD_BYTES + SURFACE_GATHER + MATERIAL_BYTES + r["M_light"] * CANDIDATE_BYTES,
                    (2 + 2 * n_sp) * RECORD_BYTES + n_sp * RECORD_BYTES,
                    (3 + 3 * n_sp) * RECORD_BYTES + SURFACE_GATHER + MATERIAL_BYTES + FILM_BYTES]
            alg_bytes = (trav(k) if k != 1 else 0) + unit[k] * px
        else:
            names = ["trace_queue_kernel (wavefront closest-hit traversal)", "shade_kernel (wavefront shading)",
                     "shadow_kernel (wavefront any-hit traversal)"]
            if k == 1:
                alg_bytes = SHADE_BYTES * cst["stage_rays"][0] + RAY_BYTES * cst["stage_rays"][2]
            elif k == 0:
                alg_bytes = trav(0)
            else:
                alg_bytes = trav(2) + 32 * cst["stage_rays"][2]
        kernel, kernel_s, launches = names[k], stage_s[k], cst["stage_launches"][k]
    else:
        kernel_s = avg_kernel_s
        launches = cst["launches"]
        alg_bytes = (NODE_BYTES * cst["node_fetches"] + TRI_BYTES * cst["tri_fetches"] + RAY_BYTES * cst["rays"] +
                     FILM_BYTES * W * H)
        kernel = "pt_kernel (megakernel)"
    achieved = alg_bytes / kernel_s / 1e9
    mrays = cst["rays"] / avg_kernel_s / 1e6
    tag = {"procedural": "procedural_1080p_64spp_hero4", "restir": "restir_many_lights_1080p_64frames"}.get(name)
    tr = measured_traffic(kernel.split()[0] + "@" + tag) if tag else None
    if tr and staged and not restir and (tr.get("chunk_frames"), tr.get("trace_block")) != (chunk_frames, trace_block):
        tr = None  # measured at another launch shape: bytes per launch differ
    film = sc.film(0)

    line = {
        "metric": METRIC,
        "value": round(msps, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": config,
        "mrays_per_s": round(mrays, 2),
        "rays_per_sample": round(cst["rays"] / samples, 3),
        "kernel_ms_per_step": round(avg_kernel_s * 1e3, 3),
        "bvh_build_s": round(build_s, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": tr["hbm_bytes_per_launch"] if tr else None, "kernel": kernel,
                     "launches_per_step": int(launches),
                     "avg_launch_ms": round(kernel_s / max(launches, 1) * 1e3, 4),
                     "algorithmic_bytes_per_launch": int(alg_bytes / max(launches, 1)),
                     "kernel_ms_per_step": round(kernel_s * 1e3, 3)},
    }
    if staged:
        line["stages_ms_per_step"] = [round(x * 1e3, 3) for x in stage_s]
    if name == "restir_pairwise":
        line["note"] = ("the reference's pairwise combine weights a neighbour's sample by its target at the "
                        "neighbour's own pixel and normalises by it (ReSTIR/direct.cpp:248, 283; restated as is), "
                        "so where the targets at the two pixels differ by orders of magnitude the estimate explodes")
    if restir:
        line["roofline"]["note"] = ("algorithmic bytes count each light candidate's 184 B gather (32 per pixel) "
                                    "from light tables that stay in L2, so achieved can pass the HBM peak; "
                                    "traffic is what reached HBM")
    if quality:
        # relMSE of the job's film against a 4096-spp plain render of the same
        # scene; the 1080p workloads are judged at 480x270 (same scene, spp and
        # camera path), where the reference costs 0.5 G samples
        if name == "cbox":
            ref = reference_film(ctx, scene_desc, base_dir, ref_spp)
            line["relmse_vs_4096spp"] = relmse(film, ref)
            line["relmse_detail"] = relmse_detail(film, ref)
            line["relmse_at"] = f"{W}x{H}, the timed job's film"
        else:
            rw, rh = 480, 270
            if restir:
                sc.set_resolution(rw, rh)
                step()
                test = sc.film(0)
                ref = reference_film(ctx, scene_desc, base_dir, ref_spp, (rw, rh), pt_depth=1, cameras=cams)
                line["relmse_reference"] = ("plain PT of max depth 1 (emission + direct light, the quantity ReSTIR DI "
                                            "estimates), 64 spp at each of the job's 64 jitter-path cameras")
            else:
                sc.set_resolution(rw, rh)
                sc.reset_film()
                sc.render(spp, fused=True)
                test = sc.film(0)
                ref = reference_film(ctx, scene_desc, base_dir, ref_spp, (rw, rh))
            line["relmse_vs_4096spp"] = relmse(test, ref)
            line["relmse_detail"] = relmse_detail(test, ref)
            line["relmse_at"] = f"{rw}x{rh}, same scene and {spp} {'frames' if restir else 'spp'}"
            if os.environ.get("VS_BENCH_DUMP"):
                np.savez_compressed(Path(os.environ["VS_BENCH_DUMP"]) / f"relmse_{name}.npz", test=test, ref=ref)
            sc.set_resolution(W, H)
    if cpu:
        line["cpu_baseline"] = restir_cpu_baseline(sc) if restir else cpu_baseline(sc, spp)
    sc.close()
    ctx.close()
    return line


if __name__ == "__main__":
    main()
