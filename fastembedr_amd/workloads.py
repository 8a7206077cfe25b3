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
