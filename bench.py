"""Benchmark: BDPT paths/s and Mrays/s of the render core on one MI355X.

Default workload (one "step" = one full render), large_instanced:
  200 instances of their own randomly displaced ellipsoid / box meshes of
  ~10k triangles (2.0M unique triangles), 8 materials (OpenMATERIAL complex
  IOR mirrors and glTF metallic-roughness, one with an RGBA8 metallic-
  roughness map), four point lights plus a 2048x1024 fp32 HDRI (log-normal
  sky + 0.5 deg sun), 1920x1080, 16 spp, depth 8 (B = b = q = 8), Philox
  seed 0, reference estimator semantics.  Scene commit (mesh preprocessing,
  BVH build, upload) is outside the timed region, as the reference times
  BDPTRenderer::render only (main.cpp:499-507).

Prints one JSON line.  value = paths/s over the timed steps (inputs resident
in HBM); roofline = the traversal kernel's algorithmic bytes (rays x B_ray
from the oracle's canonical binary SAH BVH over the same rays) per render /
the summed HIP-event duration of its launches, against 8 TB/s, with the HBM
bytes the counters measured beside it; cpu_baseline = the CPU oracle (float,
reference estimator, same Philox streams) on a band of rows of the same
workload: median of 3 runs on the host's CPU share, plus one 1-thread run;
equal_rmse = GPU and CPU images scored against the FP64 oracle at 16x spp.

--workload hdri_torus_knot: 100k-triangle torus knot, 512x512, 64 spp.
--workload batched_lightbox prints the batched form's line instead (jobs/s of
256 lightbox jobs in one launch, beside sequential plain launches).
--workload material_fit prints the adjoint form's line.
--copy-probe GIB only runs a plain copy, the calibration kernel for the
FETCH_SIZE / WRITE_SIZE counters behind roofline.traffic.
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


def host_cpu():
    """The host's CPUs as this process sees them: nproc, the affinity mask,
    the cgroup CPU quota, the CPU model, and the thread count the baseline
    uses = the CPUs this process may run on (affinity, capped by the cgroup
    quota and by OMP_NUM_THREADS, the host's declared CPU share)."""
    nproc = os.cpu_count() or 1
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    share = aff
    if quota is not None:
        share = min(share, max(1, int(quota)))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        share = min(share, int(omp))
    return {"nproc": nproc, "affinity": aff, "cgroup_quota_cpus": quota, "omp_num_threads": omp, "model": model,
            "threads": share}


def _cpu_pass(ref, rs, params, cam_desc, view, rows, threads, bands=8):
    """One CPU run: `bands` bands of `rows` whole rows each, evenly spaced
    from the top of the image to the bottom, at the workload's spp and seed
    (the same samples the GPU renders for those rows).  Returns (seconds,
    summed stats)."""
    p = ref.default_params(**params)
    bands = max(1, min(bands, view.height // max(rows, 1)))
    tot = {"samples": 0, "rays": 0, "nodes_visited": 0, "tris_tested": 0}
    dt = 0.0
    for b in range(bands):
        row0 = min(view.height - rows, (2 * b + 1) * view.height // (2 * bands) - rows // 2)
        t0 = time.perf_counter()
        _, st = rs.render(p, cam_desc, row0=max(0, row0), rows=rows, threads=threads)
        dt += time.perf_counter() - t0
        for k in tot:
            tot[k] += st[k]
    return dt, tot, bands


def cpu_baseline(ref, desc, view, params, cam_desc, budget_s, threads):
    """The oracle (float, same estimator and Philox streams) on the host's
    CPUs, the reference's row scheduler (an atomic row counter over
    std::thread workers, BDPTRenderer.cpp:349-393).  Bounded sample: 8 bands
    of whole rows spread evenly over the image, the band height sized from a
    calibration pass so that one run takes about budget_s / 4 (at least 4
    rows per thread in a band); that sample is rendered 3 times on `threads`
    threads (median reported) and 8 bands holding 1/threads as many rows once
    on 1 thread.  Also returns the per-ray
    counts of the oracle's canonical binary BVH over the measured rays."""
    rs = ref.Scene()
    desc.apply(rs, ref.pbr_desc)
    # calibration: one band of 4 rows per thread through the middle (the row
    # scheduler needs several rows per thread to keep every thread busy)
    cal = min(view.height, 4 * threads)
    dt, _, _ = _cpu_pass(ref, rs, params, cam_desc, view, cal, threads, bands=1)
    per_run = budget_s / 4.0
    total_rows = max(1, int(per_run / max(dt / cal, 1e-6)))
    rows = max(min(view.height, 4 * threads), total_rows // 8)
    rows = min(rows, view.height // 8 if view.height >= 64 else view.height)
    runs, tot = [], {"samples": 0, "rays": 0, "nodes_visited": 0, "tris_tested": 0}
    for _ in range(3):
        dt, st, nb = _cpu_pass(ref, rs, params, cam_desc, view, rows, threads)
        runs.append(st["samples"] / dt)
        for k in tot:
            tot[k] += st[k]
    rows1 = max(1, int(round(nb * rows / max(threads, 1) / 8)))   # about the same seconds on one thread
    dt1, st1, nb1 = _cpu_pass(ref, rs, params, cam_desc, view, rows1, 1)
    return {
        "paths_per_s": float(np.median(runs)),
        "runs_paths_per_s": [round(x, 1) for x in runs],
        "bands": [nb, rows],
        "samples": int(tot["samples"] // 3),
        "seconds": round(float(sum(tot["samples"] / 3 / x for x in runs)), 2),
        "one_thread_paths_per_s": st1["samples"] / dt1,
        "one_thread_bands": [nb1, rows1],
        "one_thread_seconds": round(dt1, 2),
        "rays_per_path": tot["rays"] / max(tot["samples"], 1),
        "nodes_per_ray": tot["nodes_visited"] / max(tot["rays"], 1),
        "tris_per_ray": tot["tris_tested"] / max(tot["rays"], 1),
    }


def rmse_check(ref, desc, view, params, cam_desc, gpu_img, rows, threads, plain_img=None):
    """Equal-RMSE comparison on a band of `rows` rows through the image centre:
    the GPU image and the float oracle at the same spp and seed, each scored
    against the FP64 oracle at 16x spp with an independent seed.  plain_img:
    the GPU image of the reference estimator at the same spp, scored too when
    the measured estimator is an extension."""
    rs = ref.Scene()
    desc.apply(rs, ref.pbr_desc)
    row0 = max(0, view.height // 2 - rows // 2)
    p = ref.default_params(**params)
    t0 = time.perf_counter()
    cpu, _ = rs.render(p, cam_desc, row0=row0, rows=rows, threads=threads)
    cpu_s = time.perf_counter() - t0
    p64 = ref.default_params(**dict(params, samples_per_pixel=16 * params["samples_per_pixel"], seed=params.get("seed", 0) + 1000003))
    t0 = time.perf_counter()
    hi, _ = rs.render(p64, cam_desc, fp64=True, row0=row0, rows=rows, threads=threads)
    ref_s = time.perf_counter() - t0
    band = slice(row0, row0 + rows)
    g = gpu_img[band, :, :3].astype(np.float64)
    c = cpu[band, :, :3].astype(np.float64)
    h = hi[band, :, :3].astype(np.float64)
    rg = float(np.sqrt(np.mean((g - h) ** 2)))
    rc = float(np.sqrt(np.mean((c - h) ** 2)))
    out = {"rows": [row0, row0 + rows], "spp": params["samples_per_pixel"], "reference_spp": 16 * params["samples_per_pixel"],
           "reference": "FP64 oracle, independent seed", "rmse_gpu": rg, "rmse_cpu": rc, "ratio": rg / rc if rc > 0 else None,
           "cpu_band_s": round(cpu_s, 2), "reference_s": round(ref_s, 2)}
    if plain_img is not None:
        rp = float(np.sqrt(np.mean((plain_img[band, :, :3].astype(np.float64) - h) ** 2)))
        out["rmse_gpu_reference_estimator"] = rp
        out["rmse_reduction"] = rp / rg if rg > 0 else None
    return out


def copy_probe(gib):
    """A plain device copy of known size, the calibration kernel for the
    FETCH_SIZE / WRITE_SIZE counters (each launch reads and writes gib GiB)."""
    import torch
    n = int(gib * (1 << 30)) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_per_launch": n * 4, "launches": 3}), flush=True)


def measured_traffic(workload):
    """HBM bytes per launch of the traversal kernel (the launches of one
    render) from the committed rocprofv3 --pmc summary of this workload
    (FETCH_SIZE + WRITE_SIZE, calibrated against a plain copy) and the file
    it came from, or (None, None) when there is none."""
    rel = os.path.join("profiles", f"trace_kernel_{workload}_pmc.txt")
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), rel)
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1]), rel
    except OSError:
        pass
    return None, None


BUILD_MODE = 0   # set from --build


def scene_of(pkg, desc):
    s = pkg.Scene()
    s.set_build_mode(BUILD_MODE)
    desc.apply(s, pkg.pbr_desc)
    return s


def bench_batched(pkg, wl, args):
    """Batched validation: 256 lightbox jobs (32 assets x 8 views) in one
    persistent graph-replayed launch, autotuned; against the same jobs as
    sequential plain renders.  value = jobs/s of the batch (films read back)."""
    assets, jobs = wl.batched_lightbox()
    t0 = time.perf_counter()
    scenes = [scene_of(pkg, d) for d in assets]
    commit_s = time.perf_counter() - t0
    cams = [pkg.Camera(v.width, v.height, v.position, v.forward, v.up, v.yfov) for _, v, _ in jobs]
    prm = [pkg.default_params(**p) for _, _, p in jobs]
    b = pkg.Batch()
    for (a, _, _), cam, p in zip(jobs, cams, prm):
        b.add_job(scenes[a], cam, p)
    print(f"# {len(jobs)} jobs, {len(assets)} assets, commit {commit_s:.1f}s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    tune_s = None
    if not args.fixed_tuning:
        best = b.autotune(reps=1)
        tune_s = time.perf_counter() - t0
        print(f"# autotune {best} in {tune_s:.1f}s", file=sys.stderr, flush=True)
    else:
        # the configuration the autotuner chose on this workload (8 streams, graph)
        b.set_tuning(streams=8, use_graph=True)
        best = b.tuning()
        print(f"# fixed tuning {best}", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        b.run()
    pkg.device_synchronize()
    t0 = time.perf_counter()
    kms, rays, samples = [], 0, 0
    for i in range(args.steps):
        b.run()
        for j in range(len(jobs)):
            b.film(j)
        st = b.stats()
        kms.append(st["kernel_ms"])
        rays += st["rays"]
        samples += st["samples"]
        print(f"# batch step {i}: {st['kernel_ms']:.1f} ms kernel", file=sys.stderr, flush=True)
    pkg.device_synchronize()
    el = time.perf_counter() - t0
    # sequential plain launches of the same jobs (one renderer, re-pointed at
    # each job's scene and parameters, film read back)
    r = pkg.Renderer(prm[0])
    r.set_scene(scenes[jobs[0][0]])
    r.render(cams[0])
    pkg.device_synchronize()
    t0 = time.perf_counter()
    for (a, _, _), cam, p in zip(jobs, cams, prm):
        r.set_parameters(p)
        r.set_scene(scenes[a])
        r.render(cam)
    pkg.device_synchronize()
    seq = time.perf_counter() - t0
    print(f"# sequential {len(jobs)} renders in {seq:.2f}s", file=sys.stderr, flush=True)
    jobs_s = len(jobs) * args.steps / el
    return {
        "commit_s": round(commit_s, 3), "build": "device LBVH" if BUILD_MODE else "host SAH",
        "jobs_per_s_incl_commit": round(len(jobs) / (commit_s + el / args.steps), 3),
        "metric": "batched validation jobs/s on 1 GPU (vs sequential plain launches)", "value": round(jobs_s, 3),
        "unit": "jobs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched_lightbox", "jobs": len(jobs), "assets": len(assets), "tris": "1k-200k log-uniform",
                   "resolution": "128^2-512^2", "spp": "8-64", "depth": "4-8", "tuning": best},
        "autotune_s": None if tune_s is None else round(tune_s, 2),
        "sequential_jobs_per_s": round(len(jobs) / seq, 3), "batch_over_sequential": round(jobs_s / (len(jobs) / seq), 3),
        "paths_per_s": round(samples / el, 1), "mrays_per_s": round(rays / el / 1e6, 1),
        "kernel_ms": round(float(np.mean(kms)), 3),
    }


def bench_material_fit(pkg, wl, args):
    """Differentiable material fit: forward render + path-replay backward of an
    L2 image loss against a target rendered with hidden parameters (seed
    5000) from the initial ones (seed 5001).  value = (forward + backward)
    paths/s; a short Adam fit shows the loss falling; --fd N checks N texels
    (seed 5002) and the OM (n, k) against central finite differences of the
    FP64 oracle's detached functional (h = 1e-3)."""
    import importlib.util
    import torch
    spec = importlib.util.spec_from_file_location(
        "omv_amd.torch_adjoint", os.path.join(os.path.dirname(os.path.abspath(__file__)), "openmaterial-validation_amd", "torch_adjoint.py"))
    ta = importlib.util.module_from_spec(spec)
    sys.modules["omv_amd.torch_adjoint"] = ta
    spec.loader.exec_module(ta)
    dt, di, view, params, ids = wl.material_fit()
    dev = torch.device("cuda")
    st, si = pkg.Scene(), pkg.Scene()
    dt.apply(st, pkg.pbr_desc)
    di.apply(si, pkg.pbr_desc)
    cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    rt = pkg.Renderer(pkg.default_params(**params))
    rt.set_scene(st)
    T = torch.from_numpy(rt.render(cam)).to(dev)
    r = pkg.Renderer(pkg.default_params(**params))
    r.set_scene(si)
    fit = ta.MaterialFit(si, r, cam, ids["base_color"], ids["metal_rough"], ids["om"])
    base = torch.tensor(di.textures[0][1], device=dev, requires_grad=True)
    mr = torch.tensor(di.textures[1][1], device=dev, requires_grad=True)
    n = torch.tensor(np.asarray(di.materials[1][1], np.float32), device=dev, requires_grad=True)
    k = torch.tensor(np.asarray(di.materials[1][2], np.float32), device=dev, requires_grad=True)

    def step():
        for t in (base, mr, n, k):
            t.grad = None
        img = fit(base, mr, n, k)
        loss = ((img[..., :3] - T[..., :3]) ** 2).sum()
        loss.backward()
        return float(loss)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    # forward and backward device times on their own
    fwd_ms, bwd_ms = [], []
    d = torch.ones_like(T)
    for _ in range(3):
        r.render(cam)
        fwd_ms.append(r.stats()["render_ms"])
        r.backward(cam, d.data_ptr())
        bwd_ms.append(r.stats()["kernel_ms"])
    samples = view.width * view.height * params["samples_per_pixel"]
    # a short fit: Adam on all parameters, clamped to the valid ranges
    opt = torch.optim.Adam([base, mr, n, k], lr=0.02)
    losses = []
    for it in range(args.fit_steps):
        opt.zero_grad()
        img = fit(base, mr, n, k)
        loss = ((img[..., :3] - T[..., :3]) ** 2).sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            base.clamp_(0.0, 1.0); mr.clamp_(0.0, 1.0); n.clamp_(0.05, 3.0); k.clamp_(0.0, 10.0)
        losses.append(float(loss))
    line = {
        "metric": "differentiable material fit: (forward + backward) paths/s on 1 GPU", "value": round(samples * args.steps / el, 1),
        "unit": "paths/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "material_fit", "resolution": [view.width, view.height], "spp": params["samples_per_pixel"],
                   "depth": params["max_path_length"], "textures": "512^2 fp32 base colour + metallic-roughness", "om_sphere": True,
                   "sampling": "detached"},
        "forward_ms": round(float(np.median(fwd_ms)), 3), "backward_ms": round(float(np.median(bwd_ms)), 3),
        "fit_loss": [round(losses[0], 3), round(losses[-1], 3)] if losses else None, "fit_steps": args.fit_steps,
    }
    if args.fd > 0:
        line["fd_check"] = fd_check(pkg, args.fd, di, view, params, T.cpu().numpy().astype(np.float64))
    return line


def fd_check(pkg, count, di, view, params, T):
    """The GPU adjoint vs central differences (h = 1e-3) of the FP64 detached
    functional on the device's own paths (omv_renderer_render_detached), on
    `count` texels drawn with seed 5002 among those the loss sees and on the
    OM (n, k); loss = sum (I - T)^2 over the detached film I."""
    import importlib
    import torch
    gc = importlib.import_module("omv_amd.gradcheck")
    dev = torch.device("cuda")
    s0, se = pkg.Scene(), pkg.Scene()
    di.apply(s0, pkg.pbr_desc)
    di.apply(se, pkg.pbr_desc)
    cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    r = pkg.Renderer(pkg.default_params(**params))
    r.set_scene(s0)
    I0 = r.render_detached(cam, se)
    d_dev = torch.from_numpy(gc.d_loss(I0, T)).to(dev)
    tex = di.textures[0][1].shape[0]
    gb = torch.zeros((tex, tex, 4), device=dev)
    gmr = torch.zeros((tex, tex, 4), device=dev)
    gom = torch.zeros(6, device=dev)
    torch.cuda.synchronize()
    r.backward(cam, d_dev.data_ptr(), textures=[(0, gb.data_ptr()), (1, gmr.data_ptr())], om_materials=[1], om_grads_ptr=gom.data_ptr())
    torch.cuda.synchronize()
    grads = {0: gb.cpu().numpy().astype(np.float64), 1: gmr.cpu().numpy().astype(np.float64)}
    textures = {0: np.array(di.textures[0][1], np.float32), 1: np.array(di.textures[1][1], np.float32)}

    def near_clamp(t, y, x, c, texs):
        if not (t == 1 and c == 1):
            return False
        img = texs[t][..., c]
        h, w = img.shape
        return float(img[np.ix_([(y + d) % h for d in (-1, 0, 1)], [(x + d) % w for d in (-1, 0, 1)])].max()) > 0.998

    picks = gc.pick_texels(grads, count, 5002, ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2)), textures, avoid=near_clamp)
    t0 = time.perf_counter()
    checks = gc.central_differences(r, cam, se, T, textures, picks, om=(1, di.materials[1][1], di.materials[1][2]), h=1e-3)
    rel = gc.relative_errors(checks, grads, gom.cpu().numpy().astype(np.float64))
    for key, adj, fd, e in rel:
        print(f"# fd {key}: adjoint {adj:.7g} fd {fd:.7g} rel {e:.3g}", file=sys.stderr, flush=True)
    r_tex = np.array([e[3] for e in rel if e[0][0] == "tex"])
    r_om = np.array([e[3] for e in rel if e[0][0] == "om"])
    return {"texels": len(picks), "seed": 5002, "h": 1e-3, "reference": "central differences of the FP64 detached functional on the device's paths",
            "median_rel": float(np.median(r_tex)), "p90_rel": float(np.quantile(r_tex, 0.9)), "max_rel": float(r_tex.max()),
            "within_1e-3": float(np.mean(r_tex <= 1e-3)), "om_max_rel": float(r_om.max()), "seconds": round(time.perf_counter() - t0, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=16.0,
                    help="CPU baseline budget: 3 runs of about a quarter of it on all threads + one 1-thread run")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=float, default=0.0, metavar="GIB",
                    help="only run a plain device copy of GIB GiB (counter calibration)")
    ap.add_argument("--workload", default="large_instanced",
                    choices=["hdri_torus_knot", "plain_bdpt_icosphere", "plain_bdpt_icosphere_om_rough", "large_instanced",
                             "batched_lightbox", "material_fit"])
    ap.add_argument("--fit-steps", type=int, default=20, help="material_fit: Adam steps after the timed region")
    ap.add_argument("--rmse", type=int, default=16, metavar="ROWS",
                    help="equal-RMSE check of GPU and CPU images against the FP64 oracle at 16x spp on ROWS rows "
                         "(0: off; skipped with --no-cpu)")
    ap.add_argument("--fd", type=int, default=0, help="material_fit: finite-difference check on this many texels")
    ap.add_argument("--fixed-tuning", action="store_true",
                    help="batched_lightbox: skip the autotuner and use 8 streams + graph")
    ap.add_argument("--build", default="host", choices=["host", "device"],
                    help="per-mesh BVH build at commit: host binned SAH (default) or device linear BVH")
    ap.add_argument("--estimator", default="reference", choices=["reference", "nee"],
                    help="hdri_torus_knot: the reference estimator, or the HDRI next-event-estimation extension")
    args = ap.parse_args()
    global BUILD_MODE
    BUILD_MODE = 1 if args.build == "device" else 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (no multi-GPU path in this build)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return

    # torch (device buffers for the film) is loaded before the render core's
    # library so that both use one HIP runtime initialised by torch
    import torch  # noqa: F401
    pkg = ge.load_package()
    wl = ge.load_workloads()
    if args.workload == "batched_lightbox":
        print(json.dumps(bench_batched(pkg, wl, args)), flush=True)
        return
    if args.workload == "material_fit":
        print(json.dumps(bench_material_fit(pkg, wl, args)), flush=True)
        return
    if args.workload == "hdri_torus_knot":
        desc, view, params = wl.hdri_torus_knot()
        cfg = {"workload": "hdri_torus_knot", "triangles": 100000, "resolution": [512, 512], "spp": 64, "depth": 8,
               "textures": "2 x 1024^2 maps (RGBA8)", "hdri": [2048, 1024], "estimator": "reference"}
        if args.estimator == "nee":
            params = dict(params, extensions=params.get("extensions", 0) | 2)
            cfg["estimator"] = "HDRI next-event estimation + MIS extension"
    elif args.workload.startswith("plain_bdpt_icosphere"):
        desc, view, params = wl.plain_bdpt_icosphere()
        rough = args.workload.endswith("om_rough")
        if rough:
            params = dict(params, extensions=1, om_roughness=0.3)
        cfg = {"workload": args.workload, "triangles": 5120, "resolution": [256, 256], "spp": 16, "depth": 5,
               "estimator": "OM rough conductor extension (r=0.3)" if rough else "reference"}
    else:
        desc, view, params = wl.large_instanced()
        cfg = {"workload": "large_instanced", "triangles": desc.triangles(), "resolution": [1920, 1080], "spp": 16, "depth": 8,
               "estimator": "reference"}

    if BUILD_MODE:
        cfg["bvh_build"] = "device LBVH"
    t0 = time.perf_counter()
    scene = scene_of(pkg, desc)
    commit_s = time.perf_counter() - t0
    cam = pkg.Camera(view.width, view.height, view.position, view.forward, view.up, view.yfov)
    r = pkg.Renderer(pkg.default_params(**params))
    r.set_scene(scene)
    print(f"# device {pkg.device_info()} scene commit {commit_s:.2f}s bvh {scene.bvh_info()}", file=sys.stderr, flush=True)

    # the film stays in HBM (omv_renderer_render_device): value is the render's
    # throughput with inputs and output resident on the device; the rate with
    # the film handed over to host memory is measured beside it
    film_dev = torch.empty(view.width * view.height * 4, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    for i in range(args.warmup):
        r.render_device(cam, film_dev.data_ptr())
        print(f"# warmup {i} done", file=sys.stderr, flush=True)
    pkg.device_synchronize()
    kernel_ms, trace_ms, rays, samples, nodes, tris = [], [], 0, 0, 0, 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        r.render_device(cam, film_dev.data_ptr())
        st = r.stats()
        kernel_ms.append(st["kernel_ms"])
        trace_ms.append(st["trace_ms"])
        rays += st["rays"]
        samples += st["samples"]
        nodes += st["nodes_visited"]
        tris += st["tris_tested"]
    pkg.device_synchronize()
    elapsed = time.perf_counter() - t0
    print(f"# timed {args.steps} steps in {elapsed:.3f}s; kernel ms {[round(k, 3) for k in kernel_ms]}; "
          f"traversal ms {[round(k, 3) for k in trace_ms]}; "
          f"device BVH4: {nodes / max(rays, 1):.2f} nodes/ray, {tris / max(rays, 1):.2f} tris/ray", file=sys.stderr, flush=True)
    paths_per_s = samples / elapsed
    mrays = rays / elapsed / 1e6
    # the same renders with the film read back into the camera (PCIe-inclusive)
    n_host = max(1, min(args.steps, 5))
    r.render(cam)
    pkg.device_synchronize()
    t0 = time.perf_counter()
    for i in range(n_host):
        r.render(cam)
    pkg.device_synchronize()
    host_paths_per_s = view.width * view.height * params["samples_per_pixel"] * n_host / (time.perf_counter() - t0)
    rays_per_launch = rays / max(args.steps, 1)
    k_ms = float(np.mean(kernel_ms))
    t_ms = float(np.mean(trace_ms))

    cpu = None
    b_ray = None
    hc = host_cpu()
    threads = hc["threads"]
    if not args.no_cpu:
        ref = ge.load_oracle()
        cpu = cpu_baseline(ref, desc, view, params, cam.desc(), args.cpu_seconds, threads)
        b_ray = B_RAY_IN + B_RAY_OUT + B_NODE * cpu["nodes_per_ray"] + B_TRI * cpu["tris_per_ray"]
        print(f"# cpu baseline {cpu} on {hc}", file=sys.stderr, flush=True)

    roofline = None
    if b_ray is not None:
        # dominant kernel: the traversal kernel (its launches summed per render);
        # the per-sample schedule has a single path kernel
        dom_ms = t_ms if t_ms > 0 else k_ms
        achieved = rays_per_launch * b_ray / (dom_ms * 1e-3) / 1e9
        traffic, traffic_src = measured_traffic(args.workload)
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "achieved_model": "cold-cache algorithmic bytes: rays x B_ray, B_ray = 32 + 16 + 64 x nodes + 48 x tris "
                                      "of the oracle's canonical binary SAH BVH over the CPU sample's rays",
                    "kernel": "traversal launches of one render: omv_wf_trace<k>, omv_wf_trace_fused, and the tail launch "
                              "omv_wf_tail (which also shades the late rounds)" if t_ms > 0 else "omv_bdpt_kernel",
                    "b_ray": round(b_ray, 1), "rays_per_render": int(rays_per_launch), "kernel_ms": round(dom_ms, 3),
                    "render_kernels_ms": round(k_ms, 3)}
        # the same model over the device's own BVH4 (128-byte nodes, 48-byte
        # triangle records) and its measured visit counts
        b_dev = B_RAY_IN + B_RAY_OUT + 128.0 * nodes / max(rays, 1) + B_TRI * tris / max(rays, 1)
        roofline["b_ray_device_bvh4"] = round(b_dev, 1)
        roofline["achieved_device_bvh4"] = round(rays_per_launch * b_dev / (dom_ms * 1e-3) / 1e9, 1)
        if achieved > HBM_PEAK_GBS:
            roofline["frac_note"] = ("the cache-cold model's bytes exceed what HBM could deliver in this time: repeated node and "
                                     "triangle fetches are served by L2 and the Infinity Cache, so frac > 1 measures that reuse; "
                                     "measured_hbm_frac is the kernel's share of HBM bandwidth")
        if traffic:
            # what the memory system actually moved (rocprofv3 counters): the
            # traversal's working set sits in L2 and the Infinity Cache, so this
            # rate, not the model's, says how far the kernel is from HBM-bound
            roofline["traffic_source"] = traffic_src
            roofline["measured_hbm_gbs"] = round(traffic / (dom_ms * 1e-3) / 1e9, 1)
            roofline["measured_hbm_frac"] = round(traffic / (dom_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
            roofline["limiter"] = "dependent node-fetch latency and lane divergence (measured HBM rate well below peak)"
    line = {
        "metric": "BDPT paths/s (Mrays/s beside it) on 1 GPU", "value": round(paths_per_s, 1), "unit": "paths/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": cfg, "mrays_per_s": round(mrays, 1), "rays_per_path": round(rays / max(samples, 1), 3),
        "film": "device-resident (omv_renderer_render_device)",
        "pcie_inclusive_paths_per_s": round(host_paths_per_s, 1),
        "roofline": roofline,
        "cpu_baseline": None if cpu is None else {
            "value": round(cpu["paths_per_s"], 1), "unit": "paths/s", "cores": threads, "kind": "port",
            "sample": f"{cpu['bands'][0]} bands of {cpu['bands'][1]} rows spread over the image x {view.width} px x "
                      f"{params['samples_per_pixel']} spp ({cpu['samples']} paths) of the same workload, median of 3 runs on {threads} threads "
                      f"({cpu['seconds']:.1f}s in all); float oracle, {cfg['estimator']} estimator, same Philox streams; "
                      f"reference-semantics CPU restatement (Embree unavailable offline)",
            "runs": cpu["runs_paths_per_s"], "one_thread": round(cpu["one_thread_paths_per_s"], 1),
            "one_thread_sample": f"{cpu['one_thread_bands'][0]} bands of {cpu['one_thread_bands'][1]} rows, {cpu['one_thread_seconds']}s",
            "nproc": hc["nproc"], "affinity_cpus": hc["affinity"], "cgroup_quota_cpus": hc["cgroup_quota_cpus"],
            "omp_num_threads": hc["omp_num_threads"], "cpu_model": hc["model"]},
    }
    if cpu is not None:
        line["gpu_over_cpu"] = round(paths_per_s / cpu["paths_per_s"], 1)
        # upper bound on an all-core CPU: the 1-thread rate scaled linearly to nproc
        line["gpu_over_cpu_nproc_linear"] = round(paths_per_s / (cpu["one_thread_paths_per_s"] * hc["nproc"]), 2)
    if args.rmse > 0 and not args.no_cpu:
        ref = ge.load_oracle()
        plain = None
        if params.get("extensions", 0) & 2:   # NEE estimates the same image: score the reference estimator beside it
            rp = pkg.Renderer(pkg.default_params(**dict(params, extensions=params["extensions"] & ~2)))
            rp.set_scene(scene)
            plain = rp.render(cam)
        line["equal_rmse"] = rmse_check(ref, desc, view, params, cam.desc(), r.render(cam), args.rmse, threads, plain)
        print(f"# equal rmse {line['equal_rmse']}", file=sys.stderr, flush=True)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
