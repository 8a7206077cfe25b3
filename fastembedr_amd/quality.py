"""Embedding quality on the HIP backend.

``exact_structure_metrics`` mirrors the reference's
``exact_structure_metrics_cpp(high, low, requested_k, n_threads)``
(src/embedding_utils.cpp:1582-1752): for a sample of at most 5000
observations, trustworthiness, continuity, kNN preservation and mean
neighbour-rank error from exact FP64 ranks in both spaces (ties by index),
one row per requested k.  The ranks are computed on the GPU and the values
equal the reference's bit for bit.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import _dp, _ip, check, ptr

COLUMNS = ("trustworthiness", "continuity", "knn_preservation", "mean_neighbor_rank_error")


def exact_structure_metrics(high, low, requested_k):
    """(len(requested_k), 4) array, columns ``COLUMNS``; high and low are
    n x p arrays of the same row count (rows are observations)."""
    h = np.asarray(high, dtype=np.float64)
    lo = np.asarray(low, dtype=np.float64)
    if h.ndim != 2 or lo.ndim != 2:
        raise ValueError("high and low must be matrices")
    if lo.shape[0] != h.shape[0]:
        raise ValueError("high and low must have the same row count")
    ks = np.ascontiguousarray(np.atleast_1d(np.asarray(requested_k)).astype(np.int32))
    n = h.shape[0]
    # column-major, as the reference's R matrices
    hc = np.asfortranarray(h).ravel(order="F")
    lc = np.asfortranarray(lo).ravel(order="F")
    out = np.zeros(4 * max(len(ks), 1), np.float64)
    check(_lib.lib().fastembedr_hip_exact_structure_metrics(
        ptr(hc, _dp), int(n), int(h.shape[1]), ptr(lc, _dp), int(lo.shape[1]), ptr(ks, _ip),
        int(len(ks)), ptr(out, _dp)), "exact structure metrics")
    return out.reshape(4, len(ks)).T.copy()
