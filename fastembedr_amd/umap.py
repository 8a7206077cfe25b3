"""UMAP SGD epochs on the HIP backend.

``optimize_coo`` mirrors the reference's CUDA COO optimizer entry
(fastembedr_cuda_umap_optimize_coo, src/embedding_cuda_kernels.cpp:5483-5608,
reached from umap_cuda_optimize_csr_cpp / umap_cuda_optimize_coo_impl,
src/embedding_cuda_bridge.cpp:345-378, src/embedding_cuda_impl.cpp:
2048-2107).  ``csr_to_coo`` does the bridge's CSR -> COO expansion with
epochs_per_sample = max_w / max(w, 1e-6) (:3019-3042; CPU
attach_csr_epoch_schedule, src/fast_knn_umap.cpp:1264-1282).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import _fp, _ip, check, ptr

# find_ab_params for min_dist = 0.01, spread = 1 (src/fast_knn_umap.cpp:437-497)
A_MIN_DIST_001 = 1.895605865596314
B_MIN_DIST_001 = 0.8006377738365004


def csr_to_coo(row_ptr, col, weights):
    """Expand a CSR fuzzy graph to COO heads/tails/weights/epochs_per_sample."""
    row_ptr = np.asarray(row_ptr)
    n = len(row_ptr) - 1
    heads = np.repeat(np.arange(n, dtype=np.int32), np.diff(row_ptr))
    tails = np.asarray(col, dtype=np.int32)
    w = np.asarray(weights, dtype=np.float32)
    max_w = float(w.max()) if len(w) else 1.0
    eps = (np.float32(max_w) / np.maximum(w, np.float32(1e-6))).astype(np.float32)
    return heads, tails, w, eps


def optimize_coo(heads, tails, weights, epochs_per_sample, init, n_epochs,
                 negative_sample_rate=5, learning_rate=1.0, a=A_MIN_DIST_001, b=B_MIN_DIST_001,
                 repulsion_strength=1.0, seed=0, optimizer_mode=1):
    """Run `n_epochs` UMAP epochs; returns the float32 n x 2 layout."""
    h = np.ascontiguousarray(heads, dtype=np.int32)
    t = np.ascontiguousarray(tails, dtype=np.int32)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    e = np.ascontiguousarray(epochs_per_sample, dtype=np.float32)
    y = np.ascontiguousarray(init, dtype=np.float32)
    if not (len(h) == len(t) == len(w) == len(e)):
        raise ValueError("COO arrays must have the same length")
    n = y.shape[0]
    out = np.zeros_like(y)
    rc = _lib.lib().fastembedr_hip_umap_optimize_coo(
        ptr(h, _ip), ptr(t, _ip), ptr(w, _fp), ptr(e, _fp), ptr(y, _fp), n, len(h), int(n_epochs),
        int(negative_sample_rate), float(learning_rate), float(a), float(b),
        float(repulsion_strength), int(seed) & 0xFFFFFFFF, int(optimizer_mode), ptr(out, _fp))
    check(rc, "HIP COO UMAP optimizer")
    return out


def schedule(heads, tails, epochs_per_sample, n, epoch, negative_sample_rate=5, seed=0,
             max_neg=None):
    """(n_pos, n_neg, negative ids) of one epoch as the device computes them."""
    h = np.ascontiguousarray(heads, dtype=np.int32)
    t = np.ascontiguousarray(tails, dtype=np.int32)
    e = np.ascontiguousarray(epochs_per_sample, dtype=np.float32)
    m = len(h)
    mn = int(max_neg if max_neg is not None else 2 * max(negative_sample_rate, 1))
    npos = np.zeros(m, np.int32)
    nneg = np.zeros(m, np.int32)
    ids = np.zeros((m, max(mn, 1)), np.int32)
    rc = _lib.lib().fastembedr_hip_umap_schedule(
        ptr(h, _ip), ptr(t, _ip), ptr(e, _fp), m, int(n), int(epoch), int(negative_sample_rate),
        int(seed) & 0xFFFFFFFF, mn, ptr(npos, _ip), ptr(nneg, _ip), ptr(ids, _ip))
    check(rc, "HIP UMAP schedule")
    return npos, nneg, ids[:, :mn]
