#!/usr/bin/env python3
"""Benchmark: megakernel path tracing throughput on one MI355X.

A step is one complete render job of the workload (film reset + all spp
frames of the megakernel over the full film), with the scene, BVH and film
already resident in HBM.  Prints one JSON line.

The roofline entry is for the dominant kernel (the PT megakernel): its
algorithmic bytes come from a counting run of the same job (BVH8 node
fetches x 80 B + leaf-triangle fetches x 48 B + rays x 48 B for ray in /
hit out + 32 B film read/write per pixel), divided by the kernel's average
HIP-event duration in the timed region.
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


SCENE_CACHE = ROOT / "vision_amd" / "_scenes"


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
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(scene, spp: int, budget_s: float = 15.0) -> dict:
    """The oracle (scalar C++ port of the reference path) on the host cores
    (at most 16, this machine's share per GPU), over a bounded band of rows
    of the same job; its per-call setup (BVH2 and light tables) is timed
    separately and left out."""
    from oracle import oracle
    ex = scene.export()
    W, H = int(ex["camera"][19]), int(ex["camera"][20])
    cores = max(1, min(16, os.cpu_count() or 1))
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
    return {"value": round(samples / t / 1e6, 4), "unit": "Msamples/s", "cores": cores, "kind": "port",
            "sample": f"{rows} centre rows x {W} px x {spp} spp of the same job ({samples} samples, {t:.2f} s "
                      f"after {setup:.2f} s setup)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="procedural", choices=["procedural", "cbox"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--kernel", default="megakernel", choices=["megakernel", "wavefront"])
    ap.add_argument("--chunk-frames", type=int, default=8)
    ap.add_argument("--trace-block", type=int, default=128)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU scaling is not part of it", file=sys.stderr)
        sys.exit(2)

    import vision_amd as va
    scene_desc, base_dir, spp, config = workload(args.workload)
    ctx = va.Context(0)
    sc = ctx.load_text(scene_desc, base_dir)
    t0 = time.perf_counter()
    sc.build_accel()
    build_s = time.perf_counter() - t0
    sc.set_kernel(args.kernel, args.chunk_frames, args.trace_block)
    info = sc.info()
    W, H = info["width"], info["height"]

    def step():
        sc.reset_film()
        return sc.render(spp, fused=True)

    for _ in range(args.warmup):
        step()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = step()
        kernel_ms.append(st["kernel_ms"])
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    samples = W * H * spp
    msps = samples * args.steps / wall / 1e6

    # counting run of the same job: rays, node and triangle fetches
    sc.set_counting(True)
    sc.reset_film()
    cst = sc.render(spp, fused=True)
    sc.set_counting(False)
    alg_bytes = (NODE_BYTES * cst["node_fetches"] + TRI_BYTES * cst["tri_fetches"] +
                 RAY_BYTES * cst["rays"] + FILM_BYTES * W * H)
    avg_kernel_s = sum(kernel_ms) / len(kernel_ms) / 1e3
    achieved = alg_bytes / avg_kernel_s / 1e9
    mrays = cst["rays"] / avg_kernel_s / 1e6

    line = {
        "metric": METRIC,
        "value": round(msps, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
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
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_launch": int(alg_bytes),
                     "node_fetches": int(cst["node_fetches"]), "tri_fetches": int(cst["tri_fetches"]),
                     "rays": int(cst["rays"])},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(sc, spp)
    sc.close()
    ctx.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
