#!/usr/bin/env python3
"""Benchmark of the AdaptiveKSearchSurface hot path on one MI355X.

Workload: terrestrial-scan-like room, 4,000,000 float32 points ray-cast from a scanner at the
origin (3200 x 1250 angular grid) into a 20 x 15 x 4 m room with 40 boxes/cylinders, radial
noise 2 mm; kn=50, ki=20, kd=10, flip point = scanner origin, voxel 0.05 m.  The SDF query
vertices are the PointsetGrid lattice of that cloud (cells touched by points, extruded by one
cell, all cell corners), built on the GPU by the library.

A step = calculateNormals (adaptive kNN + PCA + flip) + interpolateNormals + distances at all
query vertices, inputs resident in HBM.  value = points / (t_normals + t_interpolate) in
Mpts/s; the SDF rate is reported beside it.  The roofline entry covers the dominant kernel
(the kNN tile kernel of the normals over its adaptive-round launches), timed with hipEvents
on the library's stream.  The line also carries the other configurations measured in the same
run (`forms`: the 32M-point outdoor cloud, the 64-scan project in the batched form, and the
differentiation form; `--no-forms` skips them, `--form NAME` runs one alone).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(w, gpu_normals, verts, sample_pts: int, sample_q: int, reps: int = 3):
    """The oracle (CPU restatement of the reference path, OpenMP over all host cores) on a
    bounded sample of the same workload; every phase is the median of `reps` runs."""
    import oracle

    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    rng = np.random.default_rng(0)
    t = time.perf_counter()
    o = oracle.Surface(w.points)
    t_build = time.perf_counter() - t
    sub = np.sort(rng.choice(len(w.points), sample_pts, replace=False))
    qs = verts[np.sort(rng.choice(len(verts), min(sample_q, len(verts)), replace=False))]

    def med(fn):
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t)
        return float(np.median(ts))

    t_norm = med(lambda: o.normals(w.kn, 3, w.flip, subset=sub))
    t_int = med(lambda: o.interpolate(gpu_normals, w.ki, subset=sub))
    t_sdf = med(lambda: o.distance(gpu_normals, w.kd, qs))
    return {
        "mpts": sample_pts / (t_norm + t_int) / 1e6,
        "sdf_mqps": len(qs) / t_sdf / 1e6,
        "cores": oracle.num_threads(),
        "t_build_s": t_build,
        "t_norm_s": t_norm,
        "t_interp_s": t_int,
        "t_sdf_s": t_sdf,
        "sample": f"{sample_pts} random points of the {len(w.points)}-point room (normals kn={w.kn} + "
                  f"interpolation ki={w.ki}, kd-tree over all points) and {len(qs)} random lattice vertices "
                  f"(SDF kd={w.kd}); median of {reps} runs per phase",
    }


def plain_small_leg(with_cpu: bool, reps: int = 5):
    """The reference's README workload shape: 300,000 points, kn 10, ki 10, least-squares plane
    fit (nem 0), flip point at the origin; normals + interpolation on the GPU (median of reps)
    and, beside it, the oracle on all 300,000 points (median of 3)."""
    import torch

    from lvr2_amd import HipSurface, scenes

    w = scenes.plain_small()
    s = HipSurface(torch.from_numpy(w.points).to("cuda:0"))
    s.setKn(w.kn)
    s.setKi(w.ki)
    s.setMethod("LS")
    s.setFlippoint(*w.flip)
    for _ in range(2):
        s.calculateNormals()
        s.interpolateNormals()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        s.calculateNormals()
        s.interpolateNormals()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t)
    n = len(w.points)
    out = {"workload": f"plain scene {n} pts kn{w.kn} ki{w.ki} least-squares plane (nem 0)",
           "gpu_mpts": round(n / float(np.median(ts)) / 1e6, 2),
           "reference_readme_mpts": {"estimate": 2.47, "estimate_and_interpolate": 1.80,
                                     "note": "README.md:95-124, 289,252-point scan.pts on an unstated 16-thread CPU"}}
    if with_cpu:
        import oracle

        o = oracle.Surface(w.points)
        tc = []
        for _ in range(3):
            t = time.perf_counter()
            n32, _, _ = o.normals(w.kn, 0, w.flip)
            o.interpolate(n32, w.ki)
            tc.append(time.perf_counter() - t)
        out["cpu_mpts"] = round(n / float(np.median(tc)) / 1e6, 3)
        out["cpu_cores"] = oracle.num_threads()
    return out


def bench_batched(args):
    """The batched form on the scan project: 64 scans of the 200 x 200 m outdoor scene, each from
    its own scanner pose, sizes log-uniform in [50,000, 2,000,000] (about 30M points), per-scan
    flip points, kn 50 ki 20.  Compares one handle per scan (the reference's per-scan loop,
    ScanProjectUtils.cpp:361-410) with one batched handle, eager and replayed from a hipGraph
    (LBVH build included in every variant)."""
    import torch

    from lvr2_amd import HipSurface, scenes

    log("[bench] generating the scan project")
    sp = scenes.scan_project_outdoor(args.scans)
    n = len(sp.points)
    dev = torch.device("cuda:0")
    pts = torch.from_numpy(sp.points).to(dev)
    out = torch.empty_like(pts)
    offs = np.concatenate([[0], np.cumsum(sp.sizes)])
    log(f"[bench] {len(sp.sizes)} scans, {n} points")

    def per_scan():
        for s_ in range(len(sp.sizes)):
            h = HipSurface(pts[offs[s_]:offs[s_ + 1]])
            h.setKn(sp.kn)
            h.setKi(sp.ki)
            h.setFlippoint(*sp.flips[s_])
            h.calculateNormals()
            h.interpolateNormals()
            h.getNormals(out[offs[s_]:offs[s_ + 1]])
            h.freeGPU()

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps

    t_loop = timed(per_scan, 1)
    ref = out.clone()
    log(f"[bench] per-scan handles: {t_loop * 1e3:.1f} ms")
    b = HipSurface(pts, scan_sizes=sp.sizes)
    b.setKn(sp.kn)
    b.setKi(sp.ki)
    b.setScanFlippoints(sp.flips)

    def eager():
        b.rebuild()
        b.calculateNormals()
        b.interpolateNormals()
        b.getNormals(out)

    t_eager = timed(eager, args.steps)
    same = bool(torch.equal(out, ref))
    # normals of unit length: the largest angle between the batched and the per-scan normals
    cos = torch.clamp(torch.abs((out.double() * ref.double()).sum(1)), max=1.0)
    max_angle = float(torch.acos(cos).max())
    log(f"[bench] batched eager: {t_eager * 1e3:.1f} ms (bit-equal to per-scan: {same}; "
        f"largest angle to the per-scan normals {max_angle:.2e} rad)")
    tune = b.autotune()
    log(f"[bench] autotuned launch shape: {tune}")
    t_tuned = timed(eager, args.steps)
    log(f"[bench] batched eager, autotuned: {t_tuned * 1e3:.1f} ms")
    b.captureGraph(out, rebuild=True)

    def replay():
        b.launchGraph()
        b.synchronize()

    t_graph = timed(replay, args.steps)
    log(f"[bench] batched graph replay: {t_graph * 1e3:.1f} ms")
    b.freeGPU()
    return {
        "metric": "batched normals Mpts/s (LBVH build + adaptive kNN/PCA + interpolation, all scans)",
        "form": "batched", "unit": "Mpts/s", "n_gpus": 1, "points": n, "scans": len(sp.sizes),
        "per_scan_handles": round(n / t_loop / 1e6, 2), "batched_eager": round(n / t_eager / 1e6, 2),
        "batched_eager_autotuned": round(n / t_tuned / 1e6, 2),
        "batched_graph": round(n / t_graph / 1e6, 2), "graph_vs_per_scan": round(t_loop / t_graph, 2),
        "ms": {"per_scan": round(t_loop * 1e3, 1), "eager": round(t_eager * 1e3, 1),
               "eager_autotuned": round(t_tuned * 1e3, 1), "graph": round(t_graph * 1e3, 1)},
        "autotuned": tune, "eager_bit_equal_to_per_scan": same, "eager_max_angle_to_per_scan_rad": max_angle,
        "data": "synthetic (seeded ray-cast scans of the outdoor scene)",
    }


def bench_adjoint(args):
    """The differentiation form: 1,000,000-point noisy sphere + torus, kn 30 kd 10, loss = sum of
    squared Hoppe SDF at 500,000 band queries; forward alone vs forward + adjoint, then 20
    gradient-descent denoising steps."""
    import torch

    from lvr2_amd import scenes
    from lvr2_amd.autograd import hoppe_sdf

    w, q = scenes.sdf_denoise()
    dev = torch.device("cuda:0")
    qt = torch.from_numpy(q).to(dev)
    x0 = torch.from_numpy(w.points).to(dev)

    def fwd(x):
        return hoppe_sdf(x, qt, kn=w.kn, kd=w.kd, flip=w.flip, voxel=w.voxel)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps

    t_f = timed(lambda: fwd(x0), args.steps)

    def fb():
        x = x0.detach().requires_grad_(True)
        (fwd(x) ** 2).sum().backward()
        return x.grad

    t_fb = timed(fb, args.steps)
    losses = []
    x = x0.clone()
    l0 = float((fwd(x) ** 2).sum())
    lr = 0.2 * w.voxel ** 2 / max(l0, 1e-30) * len(q)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(20):
        x = x.detach().requires_grad_(True)
        loss = (fwd(x) ** 2).sum()
        loss.backward()
        losses.append(float(loss.detach()))
        with torch.no_grad():
            x = x - lr * x.grad
    torch.cuda.synchronize()
    t_steps = (time.perf_counter() - t) / 20
    log(f"[bench] adjoint: forward {t_f * 1e3:.1f} ms, forward+backward {t_fb * 1e3:.1f} ms, step {t_steps * 1e3:.1f} ms")
    return {
        "metric": "SDF adjoint: forward + backward ms (1M points, 500k queries) vs forward alone",
        "form": "adjoint", "unit": "ms", "n_gpus": 1, "points": len(w.points), "queries": len(q),
        "forward_ms": round(t_f * 1e3, 2), "forward_backward_ms": round(t_fb * 1e3, 2),
        "backward_over_forward": round((t_fb - t_f) / t_f, 2), "denoise_step_ms": round(t_steps * 1e3, 2),
        "loss_first": losses[0], "loss_last": losses[-1], "steps": 20,
        "data": "synthetic (seeded noisy sphere + torus)",
    }


def bench_large(args):
    """The large single cloud: 32,000,000 points ray-cast from the centre of the 200 x 200 x 30 m
    outdoor scene (Perlin-noise terrain, 500 box buildings), 3 mm noise, kn 100, ki 30, kd 10,
    voxel 0.1: LBVH build, normals + interpolation, lattice and SDF, each timed."""
    import torch

    from lvr2_amd import HipSurface, scenes

    log("[bench] generating the outdoor scan")
    w = scenes.large_outdoor(args.points or 32_000_000)
    n = len(w.points)
    t_pts = torch.from_numpy(w.points).cuda()

    def sync_time(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) * 1e3, r

    HipSurface(t_pts).freeGPU()  # warm-up
    t_build, s = sync_time(lambda: HipSurface(t_pts))
    s.setKn(w.kn)
    s.setKi(w.ki)
    s.setKd(w.kd)
    s.setFlippoint(*w.flip)
    s.calculateNormals()
    s.interpolateNormals()
    t_norm, _ = sync_time(lambda: (s.calculateNormals(), s.interpolateNormals()))
    t_lat, (v, _, _) = sync_time(lambda: s.grid(w.voxel, True, device=True))
    s.distances(v, w.voxel)
    t_sdf, _ = sync_time(lambda: s.distances(v, w.voxel))
    rounds = [int(x) for x in s.normalsStats()]
    log(f"[bench] large: build {t_build:.1f} ms, normals+interp {t_norm:.1f} ms, lattice {t_lat:.1f} ms, "
        f"sdf {t_sdf:.1f} ms for {v.shape[0]} vertices")
    s.freeGPU()
    return {
        "metric": f"large cloud normals + interpolation Mpts/s ({n} points, kn {w.kn})", "form": "large",
        "unit": "Mpts/s", "n_gpus": 1, "points": n, "value": round(n / t_norm * 1e3 / 1e6, 2),
        "sdf_mqps": round(v.shape[0] / t_sdf * 1e3 / 1e6, 2), "vertices": int(v.shape[0]),
        "ms": {"build": round(t_build, 1), "normals_interp": round(t_norm, 1), "lattice": round(t_lat, 1),
               "sdf": round(t_sdf, 1)}, "adaptive_rounds": rounds,
        "data": "synthetic (seeded ray-cast outdoor scene: Perlin terrain + 500 buildings, 200 x 200 x 30 m)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-az", type=int, default=3200)
    ap.add_argument("--n-el", type=int, default=1250)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-sample", type=int, default=200_000)
    ap.add_argument("--cpu-sample-q", type=int, default=1_000_000)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain copy kernel (y = x * 1) over MIB MiB three times: the known-size "
                         "reference for HBM counter scaling")
    ap.add_argument("--form", choices=("plain", "batched", "adjoint", "large"), default="plain",
                    help="plain: the room-scan headline line; batched / adjoint: the two other forms; "
                         "large: the 32M-point cloud")
    ap.add_argument("--scans", type=int, default=64)
    ap.add_argument("--points", type=int, default=0, help="points of the large form (default 32,000,000)")
    ap.add_argument("--no-forms", action="store_true",
                    help="the plain line only: skip the large, batched and adjoint summaries it carries")
    args = ap.parse_args()
    forms = {"batched": bench_batched, "adjoint": bench_adjoint, "large": bench_large}
    if args.form != "plain" and args.gpus == 1:
        print(json.dumps(forms[args.form](args)), flush=True)
        return
    if args.copy_calibration:
        import torch

        x = torch.ones(args.copy_calibration << 18, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        log(f"[bench] copy calibration: 3 x (read {args.copy_calibration} MiB + write {args.copy_calibration} MiB)")
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every box here has one GPU)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from lvr2_amd import HipSurface, scenes

    log("[bench] generating scene")
    w = scenes.room(args.n_az, args.n_el)
    n = len(w.points)
    dev = torch.device("cuda:0")
    pts_t = torch.from_numpy(w.points).to(dev)
    st = torch.cuda.current_stream(dev)
    log(f"[bench] {n} points; building LBVH")
    torch.cuda.synchronize()
    t = time.perf_counter()
    s = HipSurface(pts_t)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t) * 1e3
    s.grid(w.voxel, True, device=True)  # warm-up of the lattice kernels
    torch.cuda.synchronize()
    t = time.perf_counter()
    verts_t, _, cell_idx = s.grid(w.voxel, True, device=True)
    torch.cuda.synchronize()
    lattice_ms = (time.perf_counter() - t) * 1e3
    m = int(verts_t.shape[0])
    ncells = int(cell_idx.shape[0])
    log(f"[bench] {m} query vertices ({ncells} cells, lattice {lattice_ms:.1f} ms)")
    s.setKn(w.kn)
    s.setKi(w.ki)
    s.setKd(w.kd)
    s.setFlippoint(*w.flip)
    out = (torch.empty(m, dtype=torch.float32, device=dev), torch.empty(m, dtype=torch.float32, device=dev),
           torch.empty(m, dtype=torch.uint8, device=dev))

    def step(evs=None):
        if evs:
            evs[0].record(st)
        s.calculateNormals()
        if evs:
            evs[1].record(st)
        s.interpolateNormals()
        if evs:
            evs[2].record(st)
        s.distances(verts_t, w.voxel, out)
        if evs:
            evs[3].record(st)

    for i in range(args.warmup):
        step()
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    # the LBVH build alone (device-resident points; the rebuild keeps the handle's buffers)
    rb = []
    for _ in range(4):
        torch.cuda.synchronize()
        t = time.perf_counter()
        s.rebuild()
        torch.cuda.synchronize()
        rb.append((time.perf_counter() - t) * 1e3)
    rebuild_ms = float(np.median(rb[1:]))
    log(f"[bench] LBVH rebuild {rebuild_ms:.2f} ms")
    torch.cuda.synchronize()
    s.setProfiling(True)
    s.timings(reset=True)
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(evs[i])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    log(f"[bench] {args.steps} timed steps done in {wall:.3f} s")
    t_norm = sum(e[0].elapsed_time(e[1]) for e in evs) / args.steps
    t_int = sum(e[1].elapsed_time(e[2]) for e in evs) / args.steps
    t_sdf = sum(e[2].elapsed_time(e[3]) for e in evs) / args.steps
    ms_step = sum(e[0].elapsed_time(e[3]) for e in evs) / args.steps
    tm = s.timings()
    kss = s.kernelStats()
    ks, kw = kss["sdf"], kss["merged_rounds"]
    if ks["tiles"]:
        log(f"[bench] SDF search: {ks['resolved'] / args.steps:.0f} vertices one lane each, "
            f"{ks['refine'] / args.steps:.0f} left to the wave pass; per vertex {ks['candidates'] / max(ks['resolved'], 1):.1f} "
            f"points scanned, {ks['pops'] / max(ks['resolved'], 1):.1f} nodes expanded; slowest lane per wave "
            f"{16 * ks['attempts'] / ks['tiles']:.1f} points; {ks['cy_total'] / ks['tiles']:.0f} cycles per wave")
    if kw["tiles"]:
        log(f"[bench] SDF wave pass: {kw['resolved'] / args.steps:.0f} vertices, {kw['refine'] / args.steps:.0f} left "
            f"to the tile kernel; per vertex {kw['candidates'] / kw['tiles']:.0f} points, {kw['pops'] / kw['tiles']:.0f} "
            f"nodes, {kw['cy_total'] / kw['tiles']:.0f} cycles")
    rounds = s.normalsStats()
    mpts = n / ((t_norm + t_int) / 1e3) / 1e6
    sdf_mqps = m / (t_sdf / 1e3) / 1e6

    # the mesh after the path (not in the step): marching cubes on the step's distances
    mesh = {}
    for dec in ("PMC", "MC"):
        s.mesh(out[0], out[2], dec, device=True)
        torch.cuda.synchronize()
        ts = []
        for _ in range(3):
            t = time.perf_counter()
            mv, mf = s.mesh(out[0], out[2], dec, device=True)
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t) * 1e3)
        mesh[dec] = {"ms": round(float(np.median(ts)), 3), "vertices": int(mv.shape[0]), "faces": int(mf.shape[0])}
    log(f"[bench] mesh: PMC {mesh['PMC']['ms']:.2f} ms, MC {mesh['MC']['ms']:.2f} ms, {mesh['MC']['faces']} faces")

    # roofline of the dominant kernel: the kNN tile kernel of the normals (fused bounding-box
    # test + moments), over its five adaptive-round launches.  Algorithmic bytes of a launch at
    # k: per query point 12 (point) + 12 k (neighbours) + 12 (normal) + 4 ki (ki list, first
    # round only), the per-unit figure of the fused kNN + PCA step.
    lanes = s.tuning()["lanes_per_query"]
    ms_first, n_first = tm["normals_first"]
    ms_retry, n_retry = tm["normals_retry"]
    launches = n_first + n_retry
    per_step = [int(x) for x in rounds]
    bytes_step = 0
    for r, nr in enumerate(per_step):
        k_r = w.kn << (r + 1)
        bytes_step += nr * (12 + 12 * k_r + 12) + (n * 4 * w.ki if r == 0 else 0)
    launches_per_step = len(per_step)
    avg_ms = (ms_first + ms_retry) / max(launches, 1)
    achieved = bytes_step / launches_per_step / (avg_ms / 1e3) / 1e9
    roofline = {
        "bound": "hbm",
        "limiter": "instruction issue and LDS / L2 latency (the kernel is occupancy-bound: three waves per SIMD, "
                   "set by LDS and registers); measured HBM traffic is a small fraction of the algorithmic bytes",
        "kernel": f"knn_tile_kernel<{lanes}, SinkNormals<{lanes}> | SinkMoments<{lanes}>> (the five adaptive-round launches)",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "avg_launch_ms": round(avg_ms, 4),
        "bytes_per_launch": int(bytes_step / launches_per_step),
        "units_per_launch": int(sum(per_step) / launches_per_step),
        "sum_k_per_step": int(sum(nr * (w.kn << (r + 1)) for r, nr in enumerate(per_step))),
    }
    traffic_file = os.path.join(ROOT, "profiles", "traffic_knn_normals_room_4m.json")
    if os.path.exists(traffic_file):
        try:
            roofline["traffic"] = json.load(open(traffic_file)).get("hbm_bytes_per_launch")
            roofline["traffic_source"] = ("profiles/traffic_knn_normals_room_4m.md: rocprofv3 --pmc FETCH_SIZE and "
                                          "WRITE_SIZE of this kernel on this workload, scaled by a copy kernel")
        except (OSError, ValueError):
            pass

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle, bounded sample)")
        gn = s.getNormals()
        cpu = cpu_baseline(w, gn, verts_t.cpu().numpy(), args.cpu_sample, args.cpu_sample_q)
        log(f"[bench] CPU baseline done: {cpu['mpts']:.3f} Mpts/s, {cpu['sdf_mqps']:.3f} M q/s on {cpu['cores']} threads")

    c1 = plain_small_leg(not args.no_cpu)
    log(f"[bench] plain scene: GPU {c1['gpu_mpts']} Mpts/s" + (f", CPU {c1['cpu_mpts']} Mpts/s" if "cpu_mpts" in c1
                                                                 else ""))

    line = {
        "metric": "Mpts/s kNN+PCA normals and M SDF queries/s at kn=50, vs host CPU; % HBM peak",
        "value": round(mpts, 2),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ray-cast room scan)",
        "config": {"workload": f"room scan {n} pts kn50 ki20 kd10 voxel 0.05 m", "points": n, "queries": m,
                   "kn": w.kn, "ki": w.ki, "kd": w.kd, "voxel": w.voxel},
        "sdf_mqps": round(sdf_mqps, 2),
        "phase_ms": {"normals": round(t_norm, 3), "interpolate": round(t_int, 3), "sdf": round(t_sdf, 3),
                     "build": round(build_ms, 2), "lbvh_rebuild": round(rebuild_ms, 3),
                     "lattice": round(lattice_ms, 2)},
        "kernel_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in tm.items() if v[1]},
        "adaptive_rounds": [int(x) for x in rounds],
        "mesh": mesh,
        "plain_small": c1,
        "roofline": roofline,
        "cpu_baseline": None if cpu is None else {
            "value": round(cpu["mpts"], 4), "unit": "Mpts/s", "cores": cpu["cores"], "kind": "port",
            "sample": cpu["sample"], "sdf_mqps": round(cpu["sdf_mqps"], 4),
        },
    }
    if cpu:
        line["speedup_vs_cpu"] = {"normals": round(mpts / cpu["mpts"], 1), "sdf": round(sdf_mqps / cpu["sdf_mqps"], 1)}
    if not args.no_forms:
        # the other configurations, measured in the same run: the 32M-point outdoor cloud, the
        # 64-scan project (batched form) and the differentiation form; a failure is recorded, not
        # raised, so that the line above is always printed
        s.freeGPU()
        del pts_t, verts_t, out
        torch.cuda.empty_cache()
        line["forms"] = {}
        for name in ("large", "batched", "adjoint"):
            try:
                log(f"[bench] form {name}")
                line["forms"][name] = forms[name](args)
            except Exception as e:  # noqa: BLE001 -- recorded in the line
                line["forms"][name] = {"error": f"{type(e).__name__}: {e}"}
            torch.cuda.empty_cache()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
