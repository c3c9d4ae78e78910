"""Benchmark: BDPT paths/s and Mrays/s of the render core on one MI355X.

Workload (one "step" = one full render):
  HDRI-lit torus knot: 100,000 triangles, glTF metallic-roughness with
  1024^2 roughness/metallic maps (RGBA8, the reference's texture format),
  2048x1024 fp32 HDRI (log-normal sky + 0.5 deg sun), 512x512, 64 spp,
  depth 8 (B = b = q = 8), Philox seed 0, reference estimator semantics.
  Scene commit (mesh preprocessing, BVH build, upload) is outside the timed
  region, as the reference times BDPTRenderer::render only (main.cpp:499-507).

Prints one JSON line.  value = paths/s over the timed steps (inputs resident
in HBM); roofline = the path kernel's algorithmic bytes (rays x B_ray from the
oracle's canonical binary SAH BVH over the same rays) per launch / its mean
HIP-event duration, against 8 TB/s; cpu_baseline = the CPU oracle (float,
reference estimator, same Philox streams) on a band of rows of the same
workload, all threads it was given.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

import __graft_entry__ as ge

HBM_PEAK_GBS = 8000.0
B_RAY_IN, B_RAY_OUT, B_NODE, B_TRI = 32, 16, 64, 48


def cpu_baseline(ref, desc, view, params, cam_desc, budget_s, threads):
    """Times the oracle on whole rows of the same workload, growing the band
    until about budget_s of CPU work; returns paths/s and the per-ray counts."""
    rs = ref.Scene()
    desc.apply(rs, ref.pbr_desc)
    p = ref.default_params(**params)
    rows, row0 = 1, view.height // 2
    tot_t, tot_samples, tot_rays, tot_nodes, tot_tris = 0.0, 0, 0, 0, 0
    while True:
        t0 = time.perf_counter()
        _, st = rs.render(p, cam_desc, row0=row0, rows=rows, threads=threads)
        dt = time.perf_counter() - t0
        tot_t += dt
        tot_samples += st["samples"]
        tot_rays += st["rays"]
        tot_nodes += st["nodes_visited"]
        tot_tris += st["tris_tested"]
        row0 += rows
        if tot_t >= budget_s or row0 >= view.height:
            break
        rows = max(1, min(view.height - row0, int(rows * max(1.5, (budget_s - tot_t) / max(dt, 1e-3)))))
    return {
        "paths_per_s": tot_samples / tot_t,
        "seconds": tot_t,
        "rows": row0 - view.height // 2,
        "samples": tot_samples,
        "rays_per_path": tot_rays / max(tot_samples, 1),
        "nodes_per_ray": tot_nodes / max(tot_rays, 1),
        "tris_per_ray": tot_tris / max(tot_rays, 1),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="hdri_torus_knot", choices=["hdri_torus_knot", "plain_bdpt_icosphere", "large_instanced"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (no multi-GPU path in this build)", file=sys.stderr)
        sys.exit(2)

    pkg = ge.load_package()
    wl = ge.load_workloads()
    if args.workload == "hdri_torus_knot":
        desc, view, params = wl.hdri_torus_knot()
        cfg = {"workload": "hdri_torus_knot", "triangles": 100000, "resolution": [512, 512], "spp": 64, "depth": 8,
               "textures": "2 x 1024^2 maps (RGBA8)", "hdri": [2048, 1024], "estimator": "reference"}
    elif args.workload == "plain_bdpt_icosphere":
        desc, view, params = wl.plain_bdpt_icosphere()
        cfg = {"workload": "plain_bdpt_icosphere", "triangles": 5120, "resolution": [256, 256], "spp": 16, "depth": 5,
               "estimator": "reference"}
    else:
        desc, view, params = wl.large_instanced()
        cfg = {"workload": "large_instanced", "triangles": desc.triangles(), "resolution": [1920, 1080], "spp": 16, "depth": 8,
               "estimator": "reference"}

    t0 = time.perf_counter()
    scene = pkg.Scene()
    desc.apply(scene, pkg.pbr_desc)
    commit_s = time.perf_counter() - t0
    cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    r = pkg.Renderer(pkg.default_params(**params))
    r.set_scene(scene)
    print(f"# device {pkg.device_info()} scene commit {commit_s:.2f}s bvh {scene.bvh_info()}", file=sys.stderr, flush=True)

    for i in range(args.warmup):
        r.render(cam)
        print(f"# warmup {i} done", file=sys.stderr, flush=True)
    pkg.device_synchronize()
    kernel_ms, rays, samples = [], 0, 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        r.render(cam)
        st = r.stats()
        kernel_ms.append(st["kernel_ms"])
        rays += st["rays"]
        samples += st["samples"]
    pkg.device_synchronize()
    elapsed = time.perf_counter() - t0
    print(f"# timed {args.steps} steps in {elapsed:.3f}s", file=sys.stderr, flush=True)
    paths_per_s = samples / elapsed
    mrays = rays / elapsed / 1e6
    rays_per_launch = rays / max(args.steps, 1)
    k_ms = float(np.mean(kernel_ms))

    cpu = None
    b_ray = None
    if not args.no_cpu:
        ref = ge.load_oracle()
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(ref, desc, view, params, cam.desc(), args.cpu_seconds, threads)
        b_ray = B_RAY_IN + B_RAY_OUT + B_NODE * cpu["nodes_per_ray"] + B_TRI * cpu["tris_per_ray"]
        print(f"# cpu baseline {cpu}", file=sys.stderr, flush=True)

    roofline = None
    if b_ray is not None:
        achieved = rays_per_launch * b_ray / (k_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "b_ray": round(b_ray, 1), "rays_per_launch": int(rays_per_launch), "kernel_ms": round(k_ms, 3)}
    line = {
        "metric": "BDPT paths/s (Mrays/s beside it) on 1 GPU", "value": round(paths_per_s, 1), "unit": "paths/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": cfg, "mrays_per_s": round(mrays, 1), "rays_per_path": round(rays / max(samples, 1), 3),
        "roofline": roofline,
        "cpu_baseline": None if cpu is None else {
            "value": round(cpu["paths_per_s"], 1), "unit": "paths/s", "cores": min(16, os.cpu_count() or 1), "kind": "port",
            "sample": f"{cpu['rows']} rows x {view.width} px x {params['samples_per_pixel']} spp ({cpu['samples']} paths) of the same "
                      f"workload in {cpu['seconds']:.1f}s, float oracle, reference estimator"},
    }
    if cpu is not None:
        line["gpu_over_cpu"] = round(paths_per_s / cpu["paths_per_s"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
