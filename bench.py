#!/usr/bin/env python3
"""t-SNE iterations/s at 1M points in 2D on one MI355X.

Workload (the single-job configuration): 1,000,000 points from a 100-cluster
Gaussian mixture in 50 dimensions (centres N(0, 5^2 I), within N(0, I),
cluster sizes Multinomial(n, Dirichlet(1)), seed 201); exact k = 90 KNN
(computed by the backend's exact KNN producer, timed and reported apart from
the iterations, with an FP64 brute-force check of a sample of rows);
perplexity 30; 2D; FFT grid
1024 x 1024 (the backend's rule at this size); schedule 250 early-exaggeration
iterations (12) then 500 normal ones, learning rate n/12, momentum 0.8.

A step is one optimizer iteration over all points.  P, the layout and all
optimizer state are resident in HBM before timing starts; the timed region is
exactly --steps iterations (the first 250 with early exaggeration), bracketed
by stream synchronisation.  After timing, the per-kernel durations are read
with hipEvents on the plan's own stream (fastembedr_hip_tsne_plan_profile)
for the roofline figure, and the reference CPU iteration (the reference's own
functions from oracle/_ref, else the float32 port oracle/cpu_port.c) is timed
on bounded samples at the host's thread count and at 4 threads.

Quality gate: a second plan from the same start runs the full schedule
(250 + 500) whatever --steps is, the CPU port runs the same job end to end
(its own float32 P build, the CPU grid rule G = 256, 750 iterations), and
trustworthiness and kNN preservation (k = 15) of both layouts on the same
5000-row sample must agree within 0.01.
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
                    help="iterations of the CPU reference algorithm in the baseline sample at "
                         "the host's thread count (a quarter of them at 4 threads)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0: the host's thread count (nproc: the CPUs this process may run on)")
    ap.add_argument("--no-quality-gate", action="store_true",
                    help="skip the quality leg: the full 250 + 500 schedule on the device and the "
                         "CPU port run end to end at the same size, compared on a 5000-row sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-json", default=None,
                    help="rocprofv3-derived HBM bytes per launch of the dominant kernel")
    ap.add_argument("--umap-graph-k", type=int, default=15,
                    help="build the UMAP fuzzy graph of the workload's first k neighbours on the "
                         "device (timed beside the CPU restatement on a 100k-point slice); 0 "
                         "skips it")
    ap.add_argument("--umap-quality-n", type=int, default=50_000,
                    help="points of the UMAP quality gate (device mode 0 against the reference "
                         "CPU optimizer, 500 epochs); 0 skips it")
    ap.add_argument("--knn-check-rows", type=int, default=32,
                    help="query rows of the workload's exact KNN recomputed by the FP64 brute "
                         "force on the CPU (timed, compared); 0 skips it")
    ap.add_argument("--quality-sample", type=int, default=5000,
                    help="points of the exact structure-metric sample of the final layout "
                         "(device metric, timed beside the CPU restatement); 0 skips it")
    ap.add_argument("--batched-jobs", type=int, default=256,
                    help="jobs of the batched workload (0: skip it)")
    ap.add_argument("--umap-batched-jobs", type=int, default=64,
                    help="jobs of the batched UMAP workload (0: skip it)")
    ap.add_argument("--config1", type=int, default=1,
                    help="run the 2000-point case (exact and G=128 routes; 0: skip it)")
    ap.add_argument("--served-requests", type=int, default=96,
                    help="requests of the served workload (0: skip it)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="run a plain device copy of this many MiB (3 times) so hardware byte "
                         "counters can be scaled against a known transfer")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


SCHEDULE = dict(early=250, normal=500, early_exaggeration=12.0, momentum=0.8)


def run_plan_alone(fe, ki, kd, n, k, y0, shape=None):
    """One job through the plain single-job path: P build, 250 + 500
    iterations, layout read back (with an autotuned shape when given)."""
    shape = shape or {}
    plan = fe.TsnePlan(ki, kd, perplexity=30.0, grid_size=shape.get("grid_size", 0), Y_init=y0,
                       device=True, n=n, k=k, index_offset=0,
                       attraction=shape.get("attraction", 0),
                       pair_workgroups=shape.get("pair_workgroups", 0))
    lr = n / SCHEDULE["early_exaggeration"]
    plan.run(SCHEDULE["early"], SCHEDULE["early_exaggeration"], SCHEDULE["momentum"], lr)
    plan.run(SCHEDULE["normal"], 1.0, SCHEDULE["momentum"], lr)
    out = plan.layout()
    plan.close()
    return out


def bench_batched(fe, workloads, count, k=90):
    """Batched t-SNE: `count` jobs of 2k-100k points (log-uniform), each a
    5-50-cluster mixture in 30-50 dimensions, k = 90, perplexity 30, 750
    iterations.  Times all jobs run one after another through the plain
    single-job path, then all jobs admitted into one job-table batch (graph-
    captured iterations), from KNN resident on the device to every layout on
    the host."""
    import torch

    specs = workloads.batch_job_specs(count, seed=301)
    jobs = []
    t0 = time.perf_counter()
    for sp in specs:
        ki, kd = workloads.job_knn_device(sp, k)
        y0 = workloads.centred_normal_init(sp["n"], seed=sp["seed"] + 1)
        jobs.append((ki, kd, sp["n"], y0))
    torch.cuda.synchronize()
    total_points = sum(j[2] for j in jobs)
    log(f"[batched] {count} jobs, {total_points} points, KNN in {time.perf_counter() - t0:.1f}s")

    t0 = time.perf_counter()
    alone = [run_plan_alone(fe, ki, kd, n, k, y0) for ki, kd, n, y0 in jobs]
    t_seq = time.perf_counter() - t0
    log(f"[batched] sequential plain path: {t_seq:.2f}s")

    from fastembedr_amd.batch import TsneBatch

    t0 = time.perf_counter()
    b = TsneBatch(max_jobs=count)
    slots = [b.add(ki, kd, perplexity=30.0, Y_init=y0, device=True, n=n, k=k, index_offset=0,
                   early_exaggeration_iter=SCHEDULE["early"], n_iter=SCHEDULE["normal"],
                   early_exaggeration=SCHEDULE["early_exaggeration"],
                   initial_momentum=SCHEDULE["momentum"], final_momentum=SCHEDULE["momentum"])
             for ki, kd, n, y0 in jobs]
    t_admit = time.perf_counter() - t0
    b.run_to_completion(chunk=50)
    outs = [b.layout(s) for s in slots]
    t_batch = time.perf_counter() - t0
    b.close()
    same = sum(int(np.array_equal(a, o)) for a, o in zip(alone, outs))
    log(f"[batched] batched: {t_batch:.2f}s (admission {t_admit:.2f}s); "
        f"{same}/{count} layouts bit-identical to the plain path")
    tuned = bench_batched_tuned(fe, workloads, specs, jobs, alone, t_seq, k)
    return {"workload": f"{count} t-SNE jobs, n log-uniform in [2k, 100k] ({total_points} points), "
                        f"5-50-cluster mixtures in 30-50 dims, k={k}, perplexity 30, 750 iterations",
            "jobs": count, "points": total_points,
            "batched_jobs_per_s": round(count / t_batch, 3),
            "sequential_jobs_per_s": round(count / t_seq, 3),
            "speedup_vs_sequential": round(t_seq / t_batch, 3),
            "batched_s": round(t_batch, 3), "sequential_s": round(t_seq, 3),
            "identical_to_plain": same, "autotuned": tuned}


def bench_batched_tuned(fe, workloads, specs, jobs, alone, t_seq, k, check_jobs=16):
    """The same jobs with the autotuner's shape per size bucket (FFT grid kept
    at the smallest size that holds the grid rule's layout quality, attraction
    form and workgroups by timing): tuning time, batched jobs/s against the
    untuned sequential plain path, tuned jobs alone = tuned jobs batched on a
    sample, and the layout quality of tuned against plain jobs on that
    sample (trustworthiness and kNN preservation, k = 15, exact ranks on 2000
    rows per job)."""
    from fastembedr_amd.batch import TsneBatch
    from fastembedr_amd.quality import exact_structure_metrics
    from fastembedr_amd.tune import Autotuner

    count = len(jobs)
    tuner = Autotuner(k=k, perplexity=30.0)
    t0 = time.perf_counter()
    tuner.tune(sizes=[n for _, _, n, _ in jobs])
    t_tune = time.perf_counter() - t0
    log(f"[batched] autotuned {len(tuner.table)} buckets in {t_tune:.1f}s: " + ", ".join(
        f"{e['bucket']}: G={e['grid_size']} (rule {e['grid_rule']}), attraction "
        f"{e['attraction']}, workgroups {e['pair_workgroups']}" for e in tuner.table.values()))
    shapes = [tuner.shape_for(n) for _, _, n, _ in jobs]
    t0 = time.perf_counter()
    b = TsneBatch(max_jobs=count)
    slots = [b.add(ki, kd, perplexity=30.0, Y_init=y0, device=True, n=n, k=k, index_offset=0,
                   early_exaggeration_iter=SCHEDULE["early"], n_iter=SCHEDULE["normal"],
                   early_exaggeration=SCHEDULE["early_exaggeration"],
                   initial_momentum=SCHEDULE["momentum"], final_momentum=SCHEDULE["momentum"],
                   **sh)
             for (ki, kd, n, y0), sh in zip(jobs, shapes)]
    b.run_to_completion(chunk=50)
    outs = [b.layout(s) for s in slots]
    t_batch = time.perf_counter() - t0
    b.close()
    pick = np.linspace(0, count - 1, min(check_jobs, count)).astype(int)
    same = 0
    dq = []
    rng = np.random.default_rng(311)
    for i in pick:
        ki, kd, n, y0 = jobs[i]
        same += int(np.array_equal(run_plan_alone(fe, ki, kd, n, k, y0, shapes[i]), outs[i]))
        sp = specs[i]
        x, _, _ = workloads.gaussian_mixture(sp["n"], sp["d"], sp["clusters"], seed=sp["seed"],
                                             dirichlet=True)
        rows = np.sort(rng.choice(n, size=min(2000, n), replace=False))
        qt = exact_structure_metrics(x[rows], outs[i][rows], [15])[0]
        qp = exact_structure_metrics(x[rows], alone[i][rows], [15])[0]
        dq.append((qt[0] - qp[0], qt[2] - qp[2]))
    dq = np.array(dq)
    log(f"[batched] autotuned batch: {t_batch:.2f}s; {same}/{len(pick)} equal to tuned alone; "
        f"quality delta trust min {dq[:, 0].min():+.4f}, kNN min {dq[:, 1].min():+.4f}")
    return {"tune_s": round(t_tune, 2),
            "table": [{k2: e[k2] for k2 in ("bucket", "n", "grid_size", "grid_rule", "attraction",
                                             "pair_workgroups", "grid_trials", "shape_trials")}
                      for e in tuner.table.values()],
            "batched_jobs_per_s": round(count / t_batch, 3),
            "batched_s": round(t_batch, 3),
            "speedup_vs_sequential": round(t_seq / t_batch, 3),
            "checked_jobs": int(len(pick)), "identical_to_tuned_alone": same,
            "quality_delta_vs_plain": {
                "trustworthiness_mean": round(float(dq[:, 0].mean()), 5),
                "trustworthiness_min": round(float(dq[:, 0].min()), 5),
                "knn_preservation_mean": round(float(dq[:, 1].mean()), 5),
                "knn_preservation_min": round(float(dq[:, 1].min()), 5)}}


def bench_umap_batched(workloads, count, k=15, epochs=200, n_lo=2000, n_hi=100000,
                       forms=("auto",)):
    """Batched UMAP: `count` jobs of n_lo-n_hi points (log-uniform), each a
    5-50-cluster mixture in 30-50 dimensions, exact k = 15 KNN, the default
    fuzzy graph, `epochs` epochs in optimizer_mode 1, uniform [-10, 10]
    starts.  Times every job alone (graph build + epochs, one after another)
    and all jobs in one UMAP batch per form (auto: the library's choice;
    persistent / launches: small jobs in the persistent kernel or not), from
    host KNN to every layout on the host; counts the layouts that are the
    same bits both ways."""
    import os

    from fastembedr_amd.umap import FuzzyGraph, UmapBatch

    specs = workloads.batch_job_specs(count, seed=331, n_lo=n_lo, n_hi=n_hi)
    jobs = []
    t0 = time.perf_counter()
    for sp in specs:
        ki, kd = workloads.job_knn_device(sp, k)
        y0 = np.random.default_rng(sp["seed"] + 5).uniform(-10, 10, size=(sp["n"], 2)).astype(
            np.float32)
        jobs.append((ki.cpu().numpy().T, kd.cpu().numpy().T, y0, sp["seed"] & 0xFFFF))
        del ki, kd
    points = sum(len(j[0]) for j in jobs)
    log(f"[umap batched] {count} jobs, {points} points, KNN in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    alone = []
    for idx, dist, y0, seed in jobs:
        g = FuzzyGraph(idx, dist)
        alone.append(g.optimize(y0, epochs, seed=seed, optimizer_mode=1))
        g.close()
    t_seq = time.perf_counter() - t0
    res = {"workload": f"{count} UMAP jobs, n log-uniform in [{n_lo}, {n_hi}] ({points} "
                       f"points), 5-50-cluster mixtures in 30-50 dims, exact k={k} fuzzy graph, "
                       f"{epochs} epochs, optimizer_mode 1",
           "jobs": count, "points": points, "sequential_jobs_per_s": round(count / t_seq, 3),
           "sequential_s": round(t_seq, 3)}
    env_of = {"auto": None, "persistent": "1", "launches": "0"}
    for form in forms:
        old_env = os.environ.pop("FASTEMBEDR_HIP_UMAP_PERSISTENT", None)
        if env_of[form] is not None:
            os.environ["FASTEMBEDR_HIP_UMAP_PERSISTENT"] = env_of[form]
        try:
            t0 = time.perf_counter()
            b = UmapBatch(max_jobs=count)
            slots = [b.add(idx, dist, y0, n_epochs=epochs, seed=seed)
                     for idx, dist, y0, seed in jobs]
            b.run_to_completion(chunk=50)
            outs = [b.layout(sl) for sl in slots]
            t_batch = time.perf_counter() - t0
            b.close()
        finally:
            os.environ.pop("FASTEMBEDR_HIP_UMAP_PERSISTENT", None)
            if old_env is not None:
                os.environ["FASTEMBEDR_HIP_UMAP_PERSISTENT"] = old_env
        same = sum(int(np.array_equal(a, o)) for a, o in zip(alone, outs))
        log(f"[umap batched] {count} jobs [{n_lo}, {n_hi}]: sequential {t_seq:.2f}s, batched "
            f"({form}) {t_batch:.2f}s, {same}/{count} identical")
        key = "" if form == "auto" else f"_{form}"
        res.update({f"batched_jobs_per_s{key}": round(count / t_batch, 3),
                    f"speedup_vs_sequential{key}": round(t_seq / t_batch, 3),
                    f"batched_s{key}": round(t_batch, 3), f"identical_to_alone{key}": same})
    return res


def bench_config1(fe, workloads, cpu):
    """The reference's CPU-runnable case: 2000 points, 10 clusters of 200 in
    50 dims, exact k = 90 KNN, perplexity 30, 250 early (12) + 750 normal
    iterations from a centred N(0, 1e-4^2) start (seed 102).  The reference
    routes n <= 3000 to exact pair repulsion; both that route and the FFT grid
    at G = 128 run on the device (wall time from host KNN to host layout),
    beside the reference CPU algorithm on the host (the CPU port) when `cpu`;
    quality is the exact metrics at k = 15 on all 2000 points."""
    from fastembedr_amd.quality import exact_structure_metrics

    x, _, _ = workloads.gaussian_mixture(2000, 50, 10, seed=101)
    ki, kd = workloads.exact_knn_tensors(x, 90)
    idx = np.ascontiguousarray(ki.T.cpu().numpy())
    dist = np.ascontiguousarray(kd.T.cpu().numpy())
    y0 = np.random.default_rng(102).normal(scale=1e-4, size=(2000, 2)).astype(np.float32)
    y0 -= y0.mean(0)
    out = {"workload": "2000 points, 10 clusters x 200, 50-d, exact k=90 KNN, perplexity 30, "
                       "250 early (12) + 750 normal, centred N(0, 1e-4^2) start"}
    for route, grid in (("exact", -1), ("fft_g128", 128)):
        method = dict(negative_gradient_method="exact") if grid < 0 else dict(grid_size=grid)
        fe.tsne_knn(idx, dist, perplexity=30.0, Y_init=y0, early_exaggeration_iter=5, n_iter=5,
                    **method)  # warm-up (kernels loaded)
        t0 = time.perf_counter()
        y = fe.tsne_knn(idx, dist, perplexity=30.0, Y_init=y0, early_exaggeration_iter=250,
                        n_iter=750, **method)
        dev_s = time.perf_counter() - t0
        m = exact_structure_metrics(x, y, [15])[0]
        r = {"device_s": round(dev_s, 4), "device_iters_per_s": round(1000 / dev_s, 1),
             "trustworthiness": round(float(m[0]), 5), "knn_preservation": round(float(m[2]), 5)}
        if cpu:
            import oracle

            threads = os.cpu_count() or 1
            t0 = time.perf_counter()
            c, aff_s, opt_s = oracle.port_tsne_run(idx, dist, y0, perplexity=30.0, early_iter=250,
                                                   n_iter=750, early_exag=12.0, exag=1.0,
                                                   learning_rate=2000 / 12.0, mom0=0.8, mom1=0.8,
                                                   max_step=5.0, grid=grid, n_threads=threads)
            cpu_s = time.perf_counter() - t0
            mc = exact_structure_metrics(x, c, [15])[0]
            r["cpu"] = {"kind": "port (oracle/cpu_port.c, the reference CPU algorithm)",
                        "threads": threads, "seconds": round(cpu_s, 3),
                        "affinity_s": round(aff_s, 3), "optimise_s": round(opt_s, 3),
                        "trustworthiness": round(float(mc[0]), 5),
                        "knn_preservation": round(float(mc[2]), 5)}
            r["speedup_vs_cpu"] = round(cpu_s / dev_s, 1)
            r["delta_trustworthiness"] = round(float(m[0] - mc[0]), 5)
            r["delta_knn_preservation"] = round(float(m[2] - mc[2]), 5)
        out[route] = r
        log(f"[config1] {route}: {r}")
    return out


def bench_served(fe, workloads, n_requests, k=90):
    """Served stream: a 500k-point reference embedding resident on the device;
    Poisson arrivals of transform queries (1k-50k points, 50-d, k = 90 to the
    reference) mixed with 10k-200k-point t-SNE jobs; admission by HBM budget;
    latency is arrival to result on the host."""
    import torch

    from fastembedr_amd import serve

    rng = np.random.default_rng(401)
    n_ref = 500_000
    xr, lr_, centres = workloads.gaussian_mixture(n_ref, 50, 50, seed=402, dirichlet=True)
    weights = np.bincount(lr_, minlength=len(centres)) / n_ref
    ki, kd = workloads.exact_knn_tensors(xr, k)
    ref_layout = run_plan_alone(fe, ki, kd, n_ref, k, workloads.centred_normal_init(n_ref, 403))
    del ki, kd
    torch.cuda.empty_cache()
    reqs = []
    t = 0.0
    # 80% transforms, 20% t-SNE jobs; about 24 requests/s keeps the device
    # busy most of the time (a job of the mean size takes ~0.15 s alone)
    rate = 24.0  # requests per second
    for rid in range(n_requests):
        t += rng.exponential(1.0 / rate)
        if rng.random() < 0.2:
            sp = dict(n=int(workloads.log_uniform_sizes(1, 10_000, 200_000, rng)[0]),
                      d=int(rng.integers(30, 51)), clusters=int(rng.integers(5, 51)),
                      seed=int(rng.integers(1 << 30)))
            ji, jd = workloads.job_knn_device(sp, k)
            # host n x k views of the column-major KNN (Fortran order: passed as is)
            payload = dict(indices=ji.cpu().numpy().T, distances=jd.cpu().numpy().T,
                           index_offset=0, perplexity=30.0,
                           Y_init=workloads.centred_normal_init(sp["n"], sp["seed"]),
                           early_exaggeration_iter=SCHEDULE["early"], n_iter=SCHEDULE["normal"])
            del ji, jd
            reqs.append(serve.Request(rid, "embed", t, payload))
        else:
            m = int(workloads.log_uniform_sizes(1, 1_000, 50_000, rng)[0])
            xq, lq = workloads.sample_mixture(m, centres, weights, rng)
            qi, qd = workloads.knn_to_reference(xq, xr, k)
            reqs.append(serve.Request(rid, "transform", t, dict(indices=np.asfortranarray(qi),
                                                                distances=np.asfortranarray(qd))))
    torch.cuda.synchronize()
    n_embed = sum(r.kind == "embed" for r in reqs)
    log(f"[served] {n_requests} requests ({n_embed} embed), trace span {t:.2f}s")
    free, _ = torch.cuda.mem_get_info()
    budget = int(0.5 * free)
    srv = serve.Server(reference_layout=ref_layout, max_jobs=64, hbm_budget_bytes=budget)
    stats = srv.run(reqs, realtime=True)
    sess = srv.session.stats()
    srv.close()
    summ = stats.summary()
    lat_e = [stats.latencies[r.rid] for r in reqs if r.kind == "embed"]
    lat_t = [stats.latencies[r.rid] for r in reqs if r.kind == "transform"]
    log(f"[served] {summ}")
    return {"workload": f"{n_requests} Poisson requests at {rate}/s: transform_tsne of 1k-50k "
                        f"queries against a resident {n_ref}-point reference (k={k}) and "
                        f"10k-200k-point t-SNE jobs (750 iterations), HBM budget {budget >> 20} MiB",
            "engine": "native session (fastembedr_hip_session_*): one engine thread, 25-iteration "
                      "batch ticks, transforms on 2 streams with pooled workspace "
                      f"({sess['pool_bytes'] >> 20} MiB of the budget)",
            "engine_ticks": sess["ticks"],
            **summ,
            "embed_p50_ms": round(1e3 * float(np.percentile(lat_e, 50)), 3) if lat_e else None,
            "embed_p99_ms": round(1e3 * float(np.percentile(lat_e, 99)), 3) if lat_e else None,
            "transform_p50_ms": round(1e3 * float(np.percentile(lat_t, 50)), 3) if lat_t else None,
            "transform_p99_ms": round(1e3 * float(np.percentile(lat_t, 99)), 3) if lat_t else None}


MFMA_F32_PEAK_TFLOPS = 157.3


def host_cpus():
    """The host's CPUs as this process sees them: nproc (the CPUs it may run
    on), the cgroup CPU quota where one is set, and lscpu's summary."""
    import subprocess

    info = {"nproc": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count()}
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            quota, period = fh.read().split()
        if quota != "max":
            info["cgroup_cpus"] = round(int(quota) / int(period), 2)
    except (OSError, ValueError):
        pass
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keep = ("Architecture", "CPU(s)", "Model name", "Socket(s)", "Core(s) per socket",
                "Thread(s) per core", "NUMA node(s)", "CPU max MHz")
        for line in out.splitlines():
            key, _, val = line.partition(":")
            if key.strip() in keep:
                info["lscpu " + key.strip()] = val.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline(row_ptr, col, val, layout, iters, threads, lr, n):
    """The reference CPU optimizer's iteration rate on the same P and layout:
    the reference's own functions (oracle/_ref: compute_gradient_fft_grid_f,
    apply_open_tsne_update_f, zero_mean_f at its CPU grid rule) where they are
    built, else the float32 port (oracle/cpu_port.c); timed at the host's
    thread count and at 4 threads, on bounded samples of iterations."""
    import oracle
    from oracle import ref as oref

    host = host_cpus()
    # nproc, or the cgroup's CPU quota where that is smaller
    nthreads = threads or max(1, min(host["nproc"], int(host.get("cgroup_cpus", host["nproc"]))))
    use_ref = oref.lib() is not None

    def rate(its, th):
        if use_ref:
            _, _, _, sec = oref.tsne_iterate(row_ptr, col, val, layout, its, 1.0, lr, 0.8,
                                             grid=0, threads=th)
        else:
            _, sec = oracle.port_tsne_iterate(row_ptr, col, val, layout, its, exag=1.0,
                                              learning_rate=lr, momentum=0.8, grid=0,
                                              n_threads=th)
        return its / sec, sec

    g_cpu = 256 if n >= 50000 else 128
    what = ("the reference's own CPU iteration (oracle/_ref, built from its sources)" if use_ref
            else "the float32 port of the reference CPU iteration (oracle/cpu_port.c)")
    v, sec = rate(iters, nthreads)
    q = max(1, iters // 4)
    v4, sec4 = rate(q, 4)
    return {"value": round(v, 4), "unit": "iterations/s", "cores": nthreads,
            "kind": "reference" if use_ref else "port",
            "sample": f"{iters} normal-phase iterations of {what}: FFT grid at the CPU rule "
                      f"G={g_cpu}, attraction, update, on the same {n}-point P and layout, "
                      f"{nthreads} threads, {sec:.1f}s",
            "at_4_threads": {"value": round(v4, 4), "cores": 4, "iterations": q,
                             "seconds": round(sec4, 2)},
            "host": host}


def bench_knn(x, k, cpu, check_rows):
    """Exact KNN of the workload on the device (the producer feeding the
    t-SNE plan), with the FP64 brute force (oracle) timed on a sample of
    query rows beside it and compared with the device rows."""
    import torch

    from fastembedr_amd import knn

    n, p = x.shape
    xs = np.ascontiguousarray(x[:2048].T, dtype=np.float32)
    wi = torch.empty((k, 2048), dtype=torch.int32, device="cuda")
    wd = torch.empty((k, 2048), dtype=torch.float32, device="cuda")
    knn.exact_knn_device(torch.from_numpy(xs).cuda(), 2048, p, k, wi, wd)  # module load, LDS attributes
    xd = torch.from_numpy(np.ascontiguousarray(x.T, dtype=np.float32)).cuda()
    ki = torch.empty((k, n), dtype=torch.int32, device="cuda")
    kd = torch.empty((k, n), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    scanned = knn.exact_knn_device(xd, n, p, k, ki, kd)
    sec = time.perf_counter() - t0
    del xd
    padded = (p + 7) // 8 * 8
    tflops = 2.0 * n * n * padded / sec / 1e12
    out = {"n": int(n), "p": int(p), "k": int(k), "seconds": round(sec, 4),
           "points_per_s": round(n / sec, 1), "scanned_rows": int(scanned),
           "mfma": {"bound": "mfma", "achieved": round(tflops, 2), "peak": MFMA_F32_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(tflops / MFMA_F32_PEAK_TFLOPS, 4),
                    "flops": f"2 n^2 P, P = {padded} (columns padded to a multiple of 8)"}}
    if cpu and check_rows > 0:
        import oracle

        rows = np.linspace(0, n - 1, check_rows).astype(np.int32)
        t0 = time.perf_counter()
        oi, od = oracle.exact_knn(x.astype(np.float32), k, queries=rows)
        cpu_s = time.perf_counter() - t0
        gi = ki[:, torch.from_numpy(rows.astype(np.int64)).cuda()].T.cpu().numpy()
        gd = kd[:, torch.from_numpy(rows.astype(np.int64)).cuda()].T.cpu().numpy()
        out.update({"cpu_sample_rows": int(check_rows), "cpu_s": round(cpu_s, 2),
                    "cpu_points_per_s": round(check_rows / cpu_s, 2), "cpu_cores": 1,
                    "cpu_kind": "oracle FP64 brute force, 1 thread",
                    "sample_matches_cpu": bool(np.array_equal(gi, oi) and np.array_equal(gd, od))})
    log(f"[bench] exact knn {out}")
    return ki, kd, out


def bench_umap_graph(fe, knn_i, knn_d, n, kk, cpu):
    """The UMAP fuzzy graph of the first kk neighbours of the workload's KNN:
    device build from the device-resident KNN (whole graph), and the CPU
    restatement (oracle) beside the device build on the first 100k points."""
    import torch

    from fastembedr_amd import umap

    warm = umap.FuzzyGraph(knn_i, knn_d, device=True, n=n, k=knn_i.shape[0], n_cols=kk,
                           index_offset=0)
    warm.close()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g = umap.FuzzyGraph(knn_i, knn_d, device=True, n=n, k=knn_i.shape[0], n_cols=kk,
                        index_offset=0)
    dev_s = time.perf_counter() - t0
    out = {"n": int(n), "k": int(kk), "nnz": g.nnz, "device_ms": round(1e3 * dev_s, 2),
           "device_points_per_s": round(n / dev_s, 1)}
    # the UMAP epochs on the resident graph (U[-10, 10]^2 start, seed 301; 500
    # epochs, 5 negatives per positive, seed 302; the deterministic mode)
    n_epochs = 500
    eps = g.export()["epochs_per_sample"].astype(np.float64)
    init = np.random.default_rng(301).uniform(-10.0, 10.0, size=(n, 2)).astype(np.float32)
    g.optimize(init, 2, negative_sample_rate=5, seed=302)  # warm-up
    pos = float(np.floor(n_epochs / eps).sum())  # positive samples over the run
    epoch_bytes = 16 * g.nnz + 16 * n
    modes = {}
    # mode 0: the reference's default for the GPU path (per-edge atomic
    # updates); mode 1: the deterministic order-free delta form
    for mode in (0, 1):
        t0 = time.perf_counter()
        y_umap = g.optimize(init, n_epochs, negative_sample_rate=5, learning_rate=1.0, seed=302,
                            optimizer_mode=mode)
        ep_s = time.perf_counter() - t0
        modes[mode] = {"optimizer_mode": mode, "epochs": n_epochs, "seconds": round(ep_s, 4),
                       "epochs_per_s": round(n_epochs / ep_s, 1),
                       "positive_samples_per_s": round(pos / ep_s, 1),
                       "sampled_edges_per_s": round(pos * 6 / ep_s, 1),
                       "roofline": {"bound": "hbm", "bytes_per_epoch": int(epoch_bytes),
                                    "achieved_gbs": round(epoch_bytes * n_epochs / ep_s / 1e9, 1),
                                    "frac": round(epoch_bytes * n_epochs / ep_s / 1e9 /
                                                  HBM_PEAK_GBS, 4)},
                       "layout_finite": bool(np.isfinite(y_umap).all()),
                       "note": "host start layout in and layout out included"}
    out["epochs"] = modes[1]
    out["epochs_mode0"] = modes[0]
    g.close()
    if cpu:
        import oracle

        m = min(n, 100_000)
        hi = knn_i[:kk, :m].T.cpu().numpy()
        hd = knn_d[:kk, :m].T.cpu().numpy().astype(np.float64)
        t0 = time.perf_counter()
        ref = oracle.umap_fuzzy_graph(hi, hd, index_offset=0)
        cpu_s = time.perf_counter() - t0
        small = umap.FuzzyGraph(hi, hd, index_offset=0).export()
        same = (np.array_equal(small["offsets"], ref["offsets"]) and
                np.array_equal(small["neighbors"], ref["neighbors"]) and
                float(np.abs(small["weights"] - ref["weights"]).max(initial=0.0)) <= 2.4e-7)
        out.update({"cpu_sample_points": int(m), "cpu_ms": round(1e3 * cpu_s, 1),
                    "cpu_points_per_s": round(m / cpu_s, 1), "cpu_cores": 1,
                    "cpu_kind": "oracle restatement of umap_graph_csr_float_cpp, 1 thread",
                    "sample_matches_cpu": bool(same)})
        # the reference CPU optimizer (optimize_layout_csr) on the 100k-point
        # slice's graph: its hogwild edge loop at the reference's thread cap
        # (effective_cpu_threads: at most 4), and its serial form on 1 thread
        out["epochs_cpu"] = umap_cpu_baseline(ref["offsets"], ref["neighbors"], ref["weights"], m)
    log(f"[bench] umap graph {out}")
    return out


def umap_cpu_baseline(offsets, neighbors, weights, m, n_epochs=500, sample_epochs=20):
    """Positive samples/s of the reference CPU UMAP optimizer on a graph of m
    points over a bounded number of epochs of the 500-epoch schedule: the
    reference's own hogwild edge loop (oracle/_ref) where it is built, else its
    restatement (oracle/oracle_umap_cpu.c), at the reference's cap of 4
    threads; and the serial form (restatement) on 1 thread."""
    import oracle
    from oracle import ref as oref

    rows, tails, eps = oracle.umap_cpu_active(offsets, neighbors, weights, n_epochs)
    ys = np.random.default_rng(301).uniform(-10.0, 10.0, size=(m, 2)).astype(np.float32)
    pos = float(np.floor(sample_epochs / eps.astype(np.float64)).sum())
    threads = min(4, len(os.sched_getaffinity(0)))
    if oref.lib() is not None:
        _, sec = oref.umap_hogwild(rows, tails, eps, ys, sample_epochs, nsr=5, seed=302,
                                   threads=threads)
        kind = "reference"
    else:
        _, _, sec = oracle.umap_cpu_optimize(rows, tails, eps, ys, sample_epochs, nsr=5, seed=302,
                                             n_threads=threads, form=2)
        kind = "port"
    _, _, sec1 = oracle.umap_cpu_optimize(rows, tails, eps, ys, sample_epochs, nsr=5, seed=302,
                                          n_threads=1, form=0)
    return {"value": round(pos / sec, 1), "unit": "positive samples/s", "cores": threads,
            "kind": kind,
            "sample": f"{sample_epochs} epochs (schedule of {n_epochs}) of the reference CPU "
                      f"optimizer's hogwild form on the {m}-point graph, {threads} threads "
                      f"(the reference caps its UMAP threads at 4), {sec:.2f}s",
            "serial_1_thread": {"value": round(pos / sec1, 1), "seconds": round(sec1, 2),
                                "kind": "port (serial form)"}}


def umap_quality_gate(fe, workloads, n, k=15, n_epochs=500, sample=5000):
    """GPU UMAP in the reference's GPU default (optimizer mode 0) against the
    reference CPU optimizer (hogwild edge loop at its 4-thread cap) on one
    n-point draw of the single-job generator: same fuzzy graph, same start;
    trustworthiness and kNN preservation (k = 15) on a shared sample."""
    import oracle
    from oracle import ref as oref
    from fastembedr_amd import knn, quality as fq, umap

    x, _, _ = workloads.gaussian_mixture(n, 50, 100, seed=311, dirichlet=True)
    ki, kd = knn.exact_knn(x, k)
    ki, kd = np.ascontiguousarray(ki.T), np.ascontiguousarray(kd.T)
    g = umap.FuzzyGraph(ki, kd, index_offset=0)
    gr = g.export()
    y0 = np.random.default_rng(301).uniform(-10, 10, size=(n, 2)).astype(np.float32)
    t0 = time.perf_counter()
    dev = g.optimize(y0, n_epochs, negative_sample_rate=5, seed=302, optimizer_mode=0)
    dev_s = time.perf_counter() - t0
    g.close()
    rows, tails, eps = oracle.umap_cpu_active(gr["offsets"], gr["neighbors"], gr["weights"],
                                              n_epochs)
    threads = min(4, len(os.sched_getaffinity(0)))
    if oref.lib() is not None:
        cpu, cpu_s = oref.umap_hogwild(rows, tails, eps, y0, n_epochs, nsr=5, seed=302,
                                       threads=threads)
        kind = "reference hogwild edge loop (oracle/_ref)"
    else:
        cpu, _, cpu_s = oracle.umap_cpu_optimize(rows, tails, eps, y0, n_epochs, nsr=5, seed=302,
                                                 n_threads=threads, form=2)
        kind = "restated hogwild form (oracle/oracle_umap_cpu.c)"
    s = np.sort(np.random.default_rng(7).choice(n, min(sample, n), replace=False))
    hx = x[s].astype(np.float64)
    mg = fq.exact_structure_metrics(hx, dev[s].astype(np.float64), [15])[0]
    mc = fq.exact_structure_metrics(hx, cpu[s].astype(np.float64), [15])[0]
    out = {"n": n, "k": k, "epochs": n_epochs, "sample": int(len(s)), "metric_k": 15,
           "device": {"optimizer_mode": 0, "seconds": round(dev_s, 3),
                      "trustworthiness": float(mg[0]), "knn_preservation": float(mg[2])},
           "cpu": {"kind": kind, "threads": threads, "seconds": round(cpu_s, 3),
                   "trustworthiness": float(mc[0]), "knn_preservation": float(mc[2])},
           "delta_trustworthiness": float(mg[0] - mc[0]),
           "delta_knn_preservation": float(mg[2] - mc[2]), "tolerance": 0.01,
           "pass": bool(abs(mg[0] - mc[0]) <= 0.01 and abs(mg[2] - mc[2]) <= 0.01)}
    log(f"[bench] umap quality gate {out}")
    return out


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
    knn_i, knn_d, knn_report = bench_knn(x, k, not args.no_cpu_baseline, args.knn_check_rows)
    y0 = workloads.centred_normal_init(n, seed=202)
    t0 = time.perf_counter()
    plan = fe.TsnePlan(knn_i, knn_d, perplexity=30.0, grid_size=0, Y_init=y0, device=True,
                       n=n, k=k, index_offset=0)
    setup_s = time.perf_counter() - t0
    gate_plan = host_knn = None
    if not args.no_quality_gate:
        # the quality leg's own plan: the full schedule from the same start,
        # whatever --steps and --warmup are
        gate_plan = fe.TsnePlan(knn_i, knn_d, perplexity=30.0, grid_size=0, Y_init=y0,
                                device=True, n=n, k=k, index_offset=0)
        if not args.no_cpu_baseline:
            host_knn = (np.ascontiguousarray(knn_i.T.cpu().numpy()),
                        np.ascontiguousarray(knn_d.T.cpu().numpy()))
    info = plan.info()
    attr_info = plan.attraction_info()
    log(f"[bench] attraction {attr_info}")
    umap_graph = None
    if args.umap_graph_k > 0:
        umap_graph = bench_umap_graph(fe, knn_i, knn_d, n, min(args.umap_graph_k, k),
                                      not args.no_cpu_baseline)
    # high-dimensional neighbours for the structure score of the final layout
    score_idx = (knn_i[:30].T.cpu().numpy() + 1).astype(np.int32) if args.quality_sample > 0 else None
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
    # instantiate the iteration graphs of the timed run before timing it (set-up,
    # like the plan itself: every timed iteration still runs on the device)
    plan.prepare(n_early, **early)
    if steps > n_early:
        plan.prepare(steps - n_early, **normal, after=n_early)
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
    if args.traffic_json is None and n == 1_000_000:
        # rocprofv3 FETCH_SIZE + WRITE_SIZE of this workload, calibrated
        # against a plain copy (profiles/attract_pairs_1M_counters.md)
        args.traffic_json = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                                         "traffic_1M.json")
    if args.traffic_json and os.path.exists(args.traffic_json):
        with open(args.traffic_json) as fh:
            tj = json.load(fh)
        traffic = tj.get(dominant)
    b_iter = 8 * nnz_dir + 116 * n + 28 * g * g + 144 * f * f
    layout = plan.layout()
    finite = bool(np.all(np.isfinite(layout)))

    cpu = None
    if not args.no_cpu_baseline:
        row_ptr, col, val = plan.affinities()
        cpu = cpu_baseline(row_ptr, col, val, layout, args.cpu_iters, args.cpu_threads, lr, n)
        del row_ptr, col, val
        log(f"[bench] cpu baseline {cpu}")
    gate = None
    if gate_plan is not None:
        t0 = time.perf_counter()
        gate_plan.run(SCHEDULE["early"], SCHEDULE["early_exaggeration"], SCHEDULE["momentum"], lr)
        gate_plan.run(SCHEDULE["normal"], 1.0, SCHEDULE["momentum"], lr)
        gate_plan.sync()
        gate = {"schedule": "250 early (12) + 500 normal, learning rate n/12, momentum 0.8",
                "device_s": round(time.perf_counter() - t0, 3)}
        # the quality figures below describe the layout of the full schedule
        layout = gate_plan.layout()
        gate_plan.close()
        finite = finite and bool(np.all(np.isfinite(layout)))
        if host_knn is not None:
            import oracle

            threads = cpu["cores"]
            t0 = time.perf_counter()
            y_cpu, aff_s, opt_s = oracle.port_tsne_run(
                host_knn[0], host_knn[1], y0, perplexity=30.0, early_iter=SCHEDULE["early"],
                n_iter=SCHEDULE["normal"], early_exag=SCHEDULE["early_exaggeration"], exag=1.0,
                learning_rate=lr, mom0=SCHEDULE["momentum"], mom1=SCHEDULE["momentum"],
                max_step=5.0, grid=0, n_threads=threads)
            gate["cpu"] = {"kind": "port (oracle/cpu_port.c: its own float32 P build, then the "
                                   "reference CPU optimizer at its grid rule G="
                                   f"{256 if n >= 50000 else 128})",
                           "threads": threads, "affinity_s": round(aff_s, 3),
                           "optimise_s": round(opt_s, 3),
                           "wall_s": round(time.perf_counter() - t0, 3)}
            host_knn = None
            log(f"[bench] quality gate CPU run {gate['cpu']}")
    quality = None
    if args.quality_sample > 0:
        from fastembedr_amd import quality as fq

        qs = min(args.quality_sample, n, 5000)
        rng = np.random.default_rng(7)
        sidx = np.sort(rng.choice(n, qs, replace=False))
        hs, ls = x[sidx].astype(np.float64), layout[sidx].astype(np.float64)
        fq.exact_structure_metrics(hs[:64], ls[:64], [15])  # warm-up (module load, LDS attribute)
        t0 = time.perf_counter()
        m = fq.exact_structure_metrics(hs, ls, [15])[0]
        dev_s = time.perf_counter() - t0
        quality = {"sample": int(qs), "k": 15, "trustworthiness": float(m[0]),
                   "continuity": float(m[1]), "knn_preservation": float(m[2]),
                   "mean_neighbor_rank_error": float(m[3]), "device_ms": round(1e3 * dev_s, 3)}
        if gate is not None and "cpu" in gate:
            mc = fq.exact_structure_metrics(hs, y_cpu[sidx].astype(np.float64), [15])[0]
            gate["sample"] = int(qs)
            gate["k"] = 15
            gate["device"] = {"trustworthiness": float(m[0]), "knn_preservation": float(m[2])}
            gate["cpu"].update({"trustworthiness": float(mc[0]), "knn_preservation": float(mc[2])})
            gate["delta_trustworthiness"] = float(m[0] - mc[0])
            gate["delta_knn_preservation"] = float(m[2] - mc[2])
            gate["tolerance"] = 0.01
            gate["pass"] = bool(abs(m[0] - mc[0]) <= 0.01 and abs(m[2] - mc[2]) <= 0.01)
            log(f"[bench] quality gate {gate}")
        if not args.no_cpu_baseline:
            import oracle

            t0 = time.perf_counter()
            mc = oracle.structure_metrics(hs, ls, 15)
            quality["cpu_ms"] = round(1e3 * (time.perf_counter() - t0), 1)
            quality["cpu_kind"] = "oracle restatement of exact_structure_metrics_cpp, 1 thread"
            quality["identical_to_cpu"] = bool(np.array_equal(np.asarray(mc), m))
        # the KNN structure score of the whole layout for the same rows
        # (30 high-dimensional neighbours, preserve_k 15, cluster labels)
        lay64 = layout.astype(np.float64)
        labs = (np.asarray(labels) + 1).astype(np.int32)
        keep = (sidx + 1).astype(np.int32)
        fq.knn_structure_score(lay64, score_idx, keep[:8], 15, labs, 100)  # warm-up
        t0 = time.perf_counter()
        sc = fq.knn_structure_score(lay64, score_idx, keep, 15, labs, 100)
        score = {"rows": int(qs), "n": int(n), "preserve_k": 15, "neighbours": 30,
                 **{kk: float(v) for kk, v in sc.items()},
                 "device_ms": round(1e3 * (time.perf_counter() - t0), 3)}
        if not args.no_cpu_baseline:
            import oracle

            few = keep[:50]
            t0 = time.perf_counter()
            ref = oracle.knn_structure_score(lay64, score_idx, few, 15, labs, 100)
            score["cpu_rows"] = int(len(few))
            score["cpu_ms"] = round(1e3 * (time.perf_counter() - t0), 1)
            score["cpu_kind"] = "oracle restatement of knn_structure_score_cpp, 1 thread"
            got = fq.knn_structure_totals(lay64, score_idx, few, 15, labs, 100)
            score["sample_matches_cpu"] = bool(np.array_equal(got, ref))
        quality["structure_score"] = score
        log(f"[bench] quality {quality}")
    plan.close()
    torch.cuda.empty_cache()

    umap_gate = None
    if args.umap_quality_n > 0 and not args.no_cpu_baseline:
        umap_gate = umap_quality_gate(fe, workloads, args.umap_quality_n)
    batched = bench_batched(fe, workloads, args.batched_jobs) if args.batched_jobs > 0 else None
    torch.cuda.empty_cache()
    config1 = bench_config1(fe, workloads, not args.no_cpu_baseline) if args.config1 else None
    umap_batched = None
    if args.umap_batched_jobs > 0:
        umap_batched = bench_umap_batched(workloads, args.umap_batched_jobs)
        # many small jobs: the persistent kernel against the job-table launches
        umap_batched["small_jobs"] = bench_umap_batched(
            workloads, 8 * args.umap_batched_jobs, n_lo=1000, n_hi=6000,
            forms=("persistent", "launches"))
    served = bench_served(fe, workloads, args.served_requests) if args.served_requests > 0 else None

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
                   "attraction": attr_info,
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
        "quality_gate": gate,
    }
    out["knn"] = knn_report
    if quality:
        out["quality"] = quality
    if umap_graph:
        out["umap_graph"] = umap_graph
    if umap_gate:
        out["umap_quality_gate"] = umap_gate
    out["batched"] = batched
    out["config1"] = config1
    out["umap_batched"] = umap_batched
    out["served"] = served
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
