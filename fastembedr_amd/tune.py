"""Autotuner of the batched t-SNE form: per job-size bucket, the FFT grid and
the attraction kernel's launch shape, measured on this device and cached.

The reference fixes both: its CUDA grid rule gives 256 / 512 / 1024 nodes a
side by size (embedding_cuda_kernels.cpp:5610-5627), its CPU rule 128 / 256
(tsne_neighbors.cpp:1923-1940), and its attraction is one kernel.  Here:

* Grid.  For each bucket a representative job (a Gaussian mixture of the
  bucket's middle size, exact KNN) runs the whole schedule at the CUDA
  rule's grid and at each smaller candidate down to the CPU rule's.  The
  tuner keeps the smallest grid whose final layout's trustworthiness and
  kNN preservation (k = 15, exact ranks on a sample) are within
  ``quality_tol`` of the rule's layout, so a tuned job embeds as well as the
  reference's own rules allow while its FFT passes cost a quarter or a
  sixteenth.
* Launch shape.  With that grid, a batch of copies of the representative job
  is timed for ``time_iters`` iterations with each attraction form (every
  pair once, or segment-blocked with LDS windows) and workgroup count of the
  pair form; the fastest is kept.  The workgroup count changes no bit; the
  forms add the same terms in different orders (agreeing to about 1e-5).

A tuned job is fully described by its shape (grid, attraction, workgroups):
run alone through ``TsnePlan`` with that shape it gives the bits it gives in
a batch.  The table is a JSON file keyed by device, k and perplexity;
``shape_for(n)`` returns a job's choices, the rule's grid outside the tuned
buckets.
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np

from . import _lib

# bucket edges in points; a job of n points falls in [edge_i, edge_i+1)
BUCKET_EDGES = (2000, 5000, 10000, 20000, 50000, 100000, 200000)
ATTRACTION_PAIRS, ATTRACTION_STAGED = 3, 2


def cuda_rule_grid(n):
    """The reference CUDA grid rule (the library's default)."""
    return int(_lib.lib().fastembedr_hip_opentsne_fft_grid_size(int(n)))


def cpu_rule_grid(n):
    """The reference CPU grid rule (tsne_neighbors.cpp:1923-1940)."""
    return 128 if n < 50000 else 256


def bucket_of(n, edges=BUCKET_EDGES):
    for b in range(len(edges) - 1):
        if edges[b] <= n < edges[b + 1]:
            return b
    return -1


def _device_name():
    try:
        import torch

        if torch.cuda.is_available():
            return torch.cuda.get_device_properties(0).gcnArchName
    except Exception:
        pass
    return "unknown"


def _representative(n, k, seed):
    from . import knn, workloads

    x, _, _ = workloads.gaussian_mixture(n, 40, 20, seed=seed, dirichlet=True)
    idx, dist = knn.exact_knn(x, k)  # (k, n): column-major n x k
    return x, idx.T, dist.T, workloads.centred_normal_init(n, seed + 1)


def run_alone(idx, dist, y0, shape, perplexity=30.0, early=250, normal=500,
              early_exaggeration=12.0):
    """One job alone through a ``TsnePlan`` with a tuned shape and the
    reference schedule (learning rate n / 12, momentum 0.8 in both phases,
    the batch's defaults); returns the centred layout."""
    from .tsne import TsnePlan

    n = len(idx)
    plan = TsnePlan(idx, dist, perplexity=perplexity, grid_size=shape["grid_size"], Y_init=y0,
                    attraction=shape.get("attraction", 0),
                    pair_workgroups=shape.get("pair_workgroups", 0))
    lr = n / early_exaggeration
    if early:
        plan.run(early, exaggeration=early_exaggeration, momentum=0.8, learning_rate=lr)
    if normal:
        plan.run(normal, exaggeration=1.0, momentum=0.8, learning_rate=lr)
    out = plan.layout()
    plan.close()
    return out


def _quality(x, y, sample, seed):
    from .quality import exact_structure_metrics

    rng = np.random.default_rng(seed)
    rows = np.sort(rng.choice(len(x), size=min(sample, len(x)), replace=False))
    m = exact_structure_metrics(x[rows], y[rows], [15])[0]
    return float(m[0]), float(m[2])  # trustworthiness, kNN preservation


class Autotuner:
    """Per-bucket table of (grid, attraction form, pair workgroups)."""

    def __init__(self, k=90, perplexity=30.0, cache_path=None, quality_tol=0.003,
                 quality_sample=3000, batch_copies=8, time_iters=50, edges=BUCKET_EDGES):
        self.k = int(k)
        self.perplexity = float(perplexity)
        self.quality_tol = float(quality_tol)
        self.quality_sample = int(quality_sample)
        self.batch_copies = int(batch_copies)
        self.time_iters = int(time_iters)
        self.edges = tuple(edges)
        self.cache_path = cache_path
        self.table = {}
        self.log = []
        if cache_path and os.path.exists(cache_path):
            self._load()

    # ---- cache ----
    def _key(self):
        return f"{_device_name()}|k={self.k}|perplexity={self.perplexity:g}|edges={self.edges}"

    def _load(self):
        with open(self.cache_path) as fh:
            data = json.load(fh)
        if data.get("key") == self._key():
            self.table = {int(b): v for b, v in data["buckets"].items()}

    def save(self):
        if not self.cache_path:
            return
        tmp = self.cache_path + ".tmp"
        with open(tmp, "w") as fh:
            json.dump({"key": self._key(), "buckets": {str(b): v for b, v in self.table.items()}},
                      fh, indent=1)
        os.replace(tmp, self.cache_path)

    # ---- lookup ----
    def shape_for(self, n):
        b = bucket_of(n, self.edges)
        if b in self.table:
            e = self.table[b]
            return {"grid_size": e["grid_size"], "attraction": e["attraction"],
                    "pair_workgroups": e["pair_workgroups"]}
        return {"grid_size": 0, "attraction": 0, "pair_workgroups": 0}

    # ---- tuning ----
    def tune(self, sizes=None, seed=900):
        """Tune every bucket that holds one of ``sizes`` (all buckets when
        None) and is not cached yet; returns the table."""
        buckets = range(len(self.edges) - 1) if sizes is None else sorted(
            {bucket_of(int(n), self.edges) for n in sizes} - {-1})
        for b in buckets:
            if b in self.table:
                continue
            self.table[b] = self._tune_bucket(b, seed + 17 * b)
            self.save()
        return self.table

    def _tune_bucket(self, b, seed):
        lo, hi = self.edges[b], self.edges[b + 1]
        n = int(round(math.sqrt(lo * hi)))
        x, idx, dist, y0 = _representative(n, self.k, seed)
        t0 = time.perf_counter()
        # grid: the smallest candidate whose layout keeps the rule's quality
        g_rule = cuda_rule_grid(n)
        g_min = min(cpu_rule_grid(n), g_rule)
        cands = []
        g = g_rule
        while g >= g_min:
            cands.append(g)
            g //= 2
        ref_q = None
        chosen = g_rule
        grid_trials = []
        for g in cands:
            y = run_alone(idx, dist, y0, {"grid_size": g}, self.perplexity)
            tq, kq = _quality(x, y, self.quality_sample, seed + 1)
            grid_trials.append({"grid_size": g, "trustworthiness": round(tq, 5),
                                "knn_preservation": round(kq, 5)})
            if ref_q is None:
                ref_q = (tq, kq)
                continue
            if tq >= ref_q[0] - self.quality_tol and kq >= ref_q[1] - self.quality_tol:
                chosen = g
            else:
                break
        # launch shape: a batch of copies of the job, timed per candidate
        from .batch import TsneBatch

        cus = _compute_units()
        # the pair form at its rule (one workgroup per CU), two and four per
        # CU, and the segment-blocked form
        shapes = [(ATTRACTION_PAIRS, 0), (ATTRACTION_PAIRS, 2 * cus), (ATTRACTION_PAIRS, 4 * cus),
                  (ATTRACTION_STAGED, 0)]
        timings = []
        for att, wg in shapes:
            bt = TsneBatch(max_jobs=self.batch_copies)
            try:
                for c in range(self.batch_copies):
                    bt.add(idx, dist, perplexity=self.perplexity, Y_init=y0, grid_size=chosen,
                           early_exaggeration_iter=self.time_iters, n_iter=self.time_iters,
                           attraction=att, pair_workgroups=wg)
                bt.step(25)
                bt.status()
                t = time.perf_counter()
                bt.step(self.time_iters)
                bt.status()
                dt = (time.perf_counter() - t) / self.time_iters
            finally:
                bt.close()
            timings.append({"attraction": att, "pair_workgroups": wg,
                            "ms_per_batched_iteration": round(1e3 * dt, 4)})
        best = min(timings, key=lambda r: r["ms_per_batched_iteration"])
        entry = {"bucket": [lo, hi], "n": n, "grid_size": chosen, "grid_rule": g_rule,
                 "attraction": best["attraction"], "pair_workgroups": best["pair_workgroups"],
                 "grid_trials": grid_trials, "shape_trials": timings,
                 "tune_s": round(time.perf_counter() - t0, 2)}
        self.log.append(entry)
        return entry


def _compute_units():
    try:
        import torch

        if torch.cuda.is_available():
            return int(torch.cuda.get_device_properties(0).multi_processor_count)
    except Exception:
        pass
    return 256


__all__ = ["Autotuner", "BUCKET_EDGES", "bucket_of", "cpu_rule_grid", "cuda_rule_grid",
           "run_alone"]
