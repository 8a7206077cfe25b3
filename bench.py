#!/usr/bin/env python3
"""t-SNE iterations/s at 1M points in 2D on one MI355X.

Workload (the single-job configuration): 1,000,000 points from a 100-cluster
Gaussian mixture in 50 dimensions (centres N(0, 5^2 I), within N(0, I),
cluster sizes Multinomial(n, Dirichlet(1)), seed 201); exact k = 90 KNN
(computed on the device as an untimed fixture); perplexity 30; 2D; FFT grid
1024 x 1024 (the backend's rule at this size); schedule 250 early-exaggeration
iterations (12) then 500 normal ones, learning rate n/12, momentum 0.8.

A step is one optimizer iteration over all points.  P, the layout and all
optimizer state are resident in HBM before timing starts; the timed region is
exactly --steps iterations (the first 250 with early exaggeration), bracketed
by stream synchronisation.  After timing, the per-kernel durations are read
with hipEvents on the plan's own stream (fastembedr_hip_tsne_plan_profile)
for the roofline figure, and the reference CPU iteration (float32 port,
oracle/cpu_port.c) is timed on a bounded sample on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=750)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--k", type=int, default=90)
    ap.add_argument("--profile-iters", type=int, default=20)
    ap.add_argument("--cpu-iters", type=int, default=160,
                    help="iterations of the CPU reference algorithm in the baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, cpu_count)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-json", default=None,
                    help="rocprofv3-derived HBM bytes per launch of the dominant kernel")
    ap.add_argument("--quality", action="store_true",
                    help="also report trustworthiness / kNN preservation on a 2000-point sample")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="run a plain device copy of this many MiB (3 times) so hardware byte "
                         "counters can be scaled against a known transfer")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build measures one GPU only; multi-GPU runs are not "
                                   "part of it (use --gpus 1)", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    import fastembedr_amd as fe
    from fastembedr_amd import workloads

    if not fe.available():
        raise SystemExit("no gfx950 device visible to the HIP backend")
    if args.copy_calibration_mb:
        a = torch.ones(args.copy_calibration_mb * (1 << 18), dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        log(f"[bench] copy calibration: 3 x {args.copy_calibration_mb} MiB read + write")
    n, k = args.n, args.k
    t0 = time.perf_counter()
    x, labels, centres = workloads.gaussian_mixture(n, 50, 100, seed=201, dirichlet=True)
    log(f"[bench] data {x.shape} in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    knn_i, knn_d = workloads.mixture_knn_device(x, labels, centres, k)
    torch.cuda.synchronize()
    log(f"[bench] device KNN k={k} in {time.perf_counter() - t0:.1f}s")
    y0 = workloads.centred_normal_init(n, seed=202)
    t0 = time.perf_counter()
    plan = fe.TsnePlan(knn_i, knn_d, perplexity=30.0, grid_size=0, Y_init=y0, device=True,
                       n=n, k=k, index_offset=0)
    setup_s = time.perf_counter() - t0
    info = plan.info()
    del knn_i, knn_d
    torch.cuda.empty_cache()
    log(f"[bench] plan: {info} (P build + init {setup_s:.2f}s)")

    lr = n / 12.0
    early = dict(exaggeration=12.0, momentum=0.8, learning_rate=lr)
    normal = dict(exaggeration=1.0, momentum=0.8, learning_rate=lr)
    if args.warmup:
        plan.run(args.warmup, **early)
        plan.sync()
    log("[bench] warmup done")
    steps = args.steps
    n_early = min(steps, 250)
    torch.cuda.synchronize()
    plan.sync()
    t0 = time.perf_counter()
    plan.run(n_early, **early)
    if steps > n_early:
        plan.run(steps - n_early, **normal)
    plan.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / steps
    its = steps / elapsed
    log(f"[bench] {steps} steps in {elapsed:.3f}s -> {its:.1f} it/s")

    prof = plan.profile(args.profile_iters, **normal)
    log(f"[bench] per-kernel ms: {prof}")
    g = info["grid_size"]
    f = 2 * g
    nnz_dir = n * k
    # algorithmic bytes per launch of each kernel
    algo = {
        "splat": 32 * n + 16 * g * g,
        "conv_rows_fwd": 16 * g * g + 2 * 8 * g * f + 8 * (g + 1) * g,
        "conv_cols": 4 * 8 * g * f,
        "conv_rows_inv": 2 * 8 * g * f + 16 * g * g,
        "attract": 8 * nnz_dir + 28 * n,
        "update": 72 * n + 16 * g * g,
    }
    dominant = max(algo, key=lambda kk: prof[kk])
    achieved = algo[dominant] / (prof[dominant] * 1e-3) / 1e9
    traffic = None
    if args.traffic_json and os.path.exists(args.traffic_json):
        with open(args.traffic_json) as fh:
            tj = json.load(fh)
        traffic = tj.get(dominant)
    b_iter = 8 * nnz_dir + 116 * n + 28 * g * g + 144 * f * f
    layout = plan.layout()
    finite = bool(np.all(np.isfinite(layout)))

    cpu = None
    if not args.no_cpu_baseline:
        import oracle

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        row_ptr, col, val = plan.affinities()
        _, sec = oracle.port_tsne_iterate(row_ptr, col, val, layout, args.cpu_iters, exag=1.0,
                                          learning_rate=lr, momentum=0.8, grid=0,
                                          n_threads=threads)
        cpu = {"value": round(args.cpu_iters / sec, 4), "unit": "iterations/s", "cores": threads,
               "kind": "port",
               "sample": f"{args.cpu_iters} iterations of the float32 reference CPU algorithm "
                         f"(oracle/cpu_port.c: serial splat, radix-2 FFT at the CPU grid rule "
                         f"G={256 if n >= 50000 else 128}, CSR attraction, update) on the same "
                         f"1M-point P and layout, {threads} threads, {sec:.1f}s"}
        log(f"[bench] cpu baseline {cpu}")
    quality = None
    if args.quality:
        import oracle

        rng = np.random.default_rng(7)
        s = rng.choice(n, 2000, replace=False)
        m = oracle.structure_metrics(x[s].astype(np.float64), layout[s].astype(np.float64), 15)
        quality = {"trustworthiness_k15": float(m[0]), "knn_preservation_k15": float(m[2])}
    plan.close()

    out = {
        "metric": "t-SNE iters/s @1M pts 2D (frac of HBM roofline); batched jobs/s + p99 latency",
        "value": round(its, 3),
        "unit": "iterations/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"single-job t-SNE: {n} points, 100-cluster 50-d Gaussian mixture "
                               f"(Dirichlet(1) sizes), exact k={k} KNN, perplexity 30, 2D, "
                               f"FFT grid {g}x{g}",
                   "n": n, "k": k, "grid": g, "nnz_sym": info["nnz"],
                   "schedule": "250 early (12) + 500 normal", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": dominant,
                     "kernel_ms": round(prof[dominant], 5),
                     "algorithmic_bytes": algo[dominant]},
        "iteration_roofline": {"algorithmic_bytes": b_iter,
                               "achieved": round(b_iter / (ms_per_step * 1e-3) / 1e9, 1),
                               "frac": round(b_iter / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "kernel_ms": {kk: round(v, 5) for kk, v in prof.items()},
        "setup_s": round(setup_s, 3),
        "layout_finite": finite,
        "cpu_baseline": cpu,
    }
    if quality:
        out["quality"] = quality
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
