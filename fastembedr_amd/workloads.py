"""Synthetic workloads of the measurement plan (seeded Gaussian mixtures and
their exact KNN graphs), used by bench.py and the GPU tests.

* ``gaussian_mixture``: centres N(0, 5^2 I), within-cluster N(0, I), cluster
  sizes equal or Multinomial(n, Dirichlet(1)), numpy PCG64 with a fixed seed.
* ``mixture_knn_device``: exact Euclidean k-NN (self excluded) computed on the
  GPU with torch, one cluster at a time.  Candidates for a query are its own
  cluster plus, when the cluster holds fewer than k+1 points or when another
  centre lies within ``margin`` of the nearest one, every cluster whose centre
  is that close.  With centres ~50 apart and within-cluster radii ~10 in 50
  dimensions no other point can be nearer; tests check the result against a
  brute-force search on a smaller mixture.
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


def mixture_knn_device(x, labels, centres, k, device="cuda", chunk=4096, margin=20.0):
    """Exact k-NN of a Gaussian mixture on the GPU.

    Returns (indices, distances) as torch tensors of shape (k, n), i.e. the
    column-major n x k layout of the C-ABI (element (i, j) at j*n + i), int32
    0-based and float32 Euclidean.
    """
    import torch

    torch.backends.cuda.matmul.allow_tf32 = False
    n = len(x)
    xt = torch.from_numpy(np.ascontiguousarray(x)).to(device)
    lab = np.asarray(labels)
    order = np.argsort(lab, kind="stable")
    bounds = np.searchsorted(lab[order], np.arange(len(centres) + 1))
    cd = np.sqrt(((centres[:, None, :] - centres[None, :, :]) ** 2).sum(-1))
    out_i = torch.empty((k, n), dtype=torch.int32, device=device)
    out_d = torch.empty((k, n), dtype=torch.float32, device=device)
    sq = (xt * xt).sum(1)
    for c in range(len(centres)):
        members = order[bounds[c]:bounds[c + 1]]
        if len(members) == 0:
            continue
        others = np.argsort(cd[c])
        others = others[others != c]
        cand_clusters = [c]
        nearest = cd[c, others[0]] if len(others) else np.inf
        for o in others:
            if cd[c, o] <= nearest + margin or sum(bounds[q + 1] - bounds[q] for q in cand_clusters) < k + 1:
                cand_clusters.append(o)
            else:
                break
        cand = np.concatenate([order[bounds[q]:bounds[q + 1]] for q in cand_clusters])
        if len(cand) < k + 1:
            raise ValueError("mixture too small for the requested k")
        cand_t = torch.from_numpy(cand.astype(np.int64)).to(device)
        xc = xt[cand_t]
        sqc = sq[cand_t]
        mem_t = torch.from_numpy(members.astype(np.int64)).to(device)
        for s in range(0, len(members), chunk):
            q = mem_t[s:s + chunk]
            xq = xt[q]
            d2 = sq[q][:, None] + sqc[None, :] - 2.0 * (xq @ xc.T)
            d2[cand_t[None, :] == q[:, None]] = float("inf")
            kk = min(k + 8, len(cand) - 1)
            top = torch.topk(d2, kk, dim=1, largest=False).indices
            nb = cand_t[top]
            # exact distances of the short list, then the final ordering
            diff = xq[:, None, :] - xt[nb]
            de = torch.sqrt((diff * diff).sum(-1))
            de = torch.where(nb == q[:, None], torch.full_like(de, float("inf")), de)
            srt = torch.sort(de, dim=1, stable=True)
            sel = srt.indices[:, :k]
            out_i[:, q] = torch.gather(nb, 1, sel).to(torch.int32).T
            out_d[:, q] = srt.values[:, :k].T
    return out_i, out_d


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
    x, labels, centres = gaussian_mixture(spec["n"], spec["d"], spec["clusters"], seed=spec["seed"],
                                          dirichlet=True)
    return mixture_knn_device(x, labels, centres, k, device=device)


def knn_to_reference(xq, lq, xr, lr, centres, k, device="cuda", chunk=4096, margin=20.0):
    """Exact Euclidean k-NN of query points among reference points of the same
    mixture (queries are not part of the reference), on the GPU.  Returns host
    (indices, distances) of shape (nq, k), int32 0-based and float32."""
    import torch

    torch.backends.cuda.matmul.allow_tf32 = False
    xr_t = torch.from_numpy(np.ascontiguousarray(xr)).to(device)
    sqr = (xr_t * xr_t).sum(1)
    order = np.argsort(lr, kind="stable")
    bounds = np.searchsorted(lr[order], np.arange(len(centres) + 1))
    cd = np.sqrt(((centres[:, None, :] - centres[None, :, :]) ** 2).sum(-1))
    out_i = np.zeros((len(xq), k), np.int32)
    out_d = np.zeros((len(xq), k), np.float32)
    for c in np.unique(lq):
        qsel = np.nonzero(lq == c)[0]
        others = np.argsort(cd[c])
        cand_clusters = []
        nearest = cd[c, others[1]] if len(others) > 1 else np.inf
        for o in others:
            if o == c or cd[c, o] <= nearest + margin or \
                    sum(bounds[q + 1] - bounds[q] for q in cand_clusters) < k:
                cand_clusters.append(o)
            else:
                break
        cand = np.concatenate([order[bounds[q]:bounds[q + 1]] for q in cand_clusters])
        if len(cand) < k:
            raise ValueError("reference too small for the requested k")
        cand_t = torch.from_numpy(cand.astype(np.int64)).to(device)
        xc, sqc = xr_t[cand_t], sqr[cand_t]
        for s in range(0, len(qsel), chunk):
            qs = qsel[s:s + chunk]
            xq_t = torch.from_numpy(np.ascontiguousarray(xq[qs])).to(device)
            d2 = (xq_t * xq_t).sum(1)[:, None] + sqc[None, :] - 2.0 * (xq_t @ xc.T)
            top = torch.topk(d2, min(k + 8, len(cand)), dim=1, largest=False).indices
            nb = cand_t[top]
            diff = xq_t[:, None, :] - xr_t[nb]
            de = torch.sqrt((diff * diff).sum(-1))
            srt = torch.sort(de, dim=1, stable=True)
            out_i[qs] = torch.gather(nb, 1, srt.indices[:, :k]).to(torch.int32).cpu().numpy()
            out_d[qs] = srt.values[:, :k].cpu().numpy()
    return out_i, out_d


def sample_mixture(n, centres, weights, rng):
    """n new points of a Gaussian mixture with the given centres, unit
    within-cluster variance and cluster probabilities; returns (x, labels)."""
    labels = rng.choice(len(centres), size=n, p=weights)
    x = (centres[labels] + rng.normal(size=(n, centres.shape[1]))).astype(np.float32)
    return x, labels
