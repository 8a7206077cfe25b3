"""Synthetic workloads of the measurement plan (seeded Gaussian mixtures and
their exact KNN graphs), used by bench.py and the GPU tests.

* ``gaussian_mixture``: centres N(0, 5^2 I), within-cluster N(0, I), cluster
  sizes equal or Multinomial(n, Dirichlet(1)), numpy PCG64 with a fixed seed.
* ``exact_knn_tensors`` / ``knn_to_reference``: exact Euclidean k-NN (self
  excluded) and exact query k-NN by the backend's own exact KNN producer
  (``knn.py``: FP64-exact neighbours with ties by index, checked bit for bit
  against the FP64 brute force of the oracle in ``tests/test_knn_gpu.py``).
"""
from __future__ import annotations

import numpy as np


def gaussian_mixture(n, d, n_clusters, seed, center_sd=5.0, dirichlet=False):
    rng = np.random.default_rng(seed)
    centres = rng.normal(0.0, center_sd, size=(n_clusters, d))
    if dirichlet:
        sizes = rng.multinomial(n, rng.dirichlet(np.ones(n_clusters)))
    else:
        sizes = np.full(n_clusters, n // n_clusters)
        sizes[: n - sizes.sum()] += 1
    labels = np.repeat(np.arange(n_clusters), sizes)
    x = (centres[labels] + rng.normal(size=(n, d))).astype(np.float32)
    return x, labels, centres


def centred_normal_init(n, seed, sd=1e-4):
    """Y_init = centred N(0, sd^2), float32, shared by CPU and GPU runs."""
    y = np.random.default_rng(seed).normal(scale=sd, size=(n, 2))
    y -= y.mean(0)
    return y.astype(np.float32)


def exact_knn_tensors(x, k, device="cuda"):
    """Exact k-NN of the rows of ``x`` by the backend's exact KNN producer
    (``knn.exact_knn_device``: FP64-exact neighbours, ties by index, the row
    itself excluded).  Returns (indices, distances) as torch tensors of shape
    (k, n): the column-major n x k layout of the C-ABI (element (i, j) at
    j*n + i), int32 0-based and float32 Euclidean."""
    import torch

    from . import knn

    n, p = x.shape
    xd = torch.from_numpy(np.ascontiguousarray(np.asarray(x, np.float32).T)).to(device)
    ki = torch.empty((k, n), dtype=torch.int32, device=device)
    kd = torch.empty((k, n), dtype=torch.float32, device=device)
    torch.cuda.synchronize()
    knn.exact_knn_device(xd, n, p, k, ki, kd)
    return ki, kd


def log_uniform_sizes(count, lo, hi, rng):
    """`count` integers log-uniform in [lo, hi]."""
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=count)).astype(np.int64)


def batch_job_specs(count, seed, n_lo=2000, n_hi=100000):
    """Independent t-SNE jobs of uneven size: n log-uniform in [n_lo, n_hi],
    each a 5-50-cluster Gaussian mixture in 30-50 dimensions."""
    rng = np.random.default_rng(seed)
    sizes = log_uniform_sizes(count, n_lo, n_hi, rng)
    return [dict(n=int(n), d=int(rng.integers(30, 51)), clusters=int(rng.integers(5, 51)),
                 seed=int(rng.integers(1 << 30))) for n in sizes]


def job_knn_device(spec, k, device="cuda"):
    """Data and exact device KNN of one job spec (see batch_job_specs)."""
    x, _, _ = gaussian_mixture(spec["n"], spec["d"], spec["clusters"], seed=spec["seed"],
                               dirichlet=True)
    return exact_knn_tensors(x, k, device=device)


def knn_to_reference(xq, xr, k):
    """Exact k-NN of query points among reference points (queries are not part
    of the reference) by the backend's exact query KNN.  Returns host
    (indices, distances) of shape (nq, k), int32 0-based and float32."""
    from . import knn

    qi, qd = knn.exact_query_knn(xr, xq, k)
    return np.ascontiguousarray(qi.T), np.ascontiguousarray(qd.T)


def sample_mixture(n, centres, weights, rng):
    """n new points of a Gaussian mixture with the given centres, unit
    within-cluster variance and cluster probabilities; returns (x, labels)."""
    labels = rng.choice(len(centres), size=n, p=weights)
    x = (centres[labels] + rng.normal(size=(n, centres.shape[1]))).astype(np.float32)
    return x, labels
