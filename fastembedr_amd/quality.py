"""Embedding quality on the HIP backend.

``exact_structure_metrics`` mirrors the reference's
``exact_structure_metrics_cpp(high, low, requested_k, n_threads)``
(src/embedding_utils.cpp:1582-1752): for a sample of at most 5000
observations, trustworthiness, continuity, kNN preservation and mean
neighbour-rank error from exact FP64 ranks in both spaces (ties by index),
one row per requested k.  The ranks are computed on the GPU and the values
equal the reference's bit for bit.

``knn_structure_score`` mirrors ``knn_structure_score_cuda_impl(layout,
indices, keep, preserve_k, labels, n_label_levels)``
(src/embedding_cuda_impl.cpp:1143-1196): kNN preservation, local
trustworthiness and continuity of a 2-d layout against the high-dimensional
KNN list, their mean, and the embedding's kNN label accuracy.
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


STRUCTURE_SCORE_NAMES = ("knn_preservation", "local_trustworthiness", "local_continuity",
                         "structure_score", "embedding_knn_accuracy")


def knn_structure_score(layout, indices, keep, preserve_k, labels=None, n_label_levels=0):
    """Dict of ``STRUCTURE_SCORE_NAMES`` (NaN when no row is scored).

    layout: n x 2; indices: n x k (or len(keep) x k) 1-based high-dimensional
    neighbours; keep: 1-based rows to score; labels: n integer codes in
    1..n_label_levels (other values: unlabelled)."""
    y = np.asarray(layout, dtype=np.float64)
    if y.ndim != 2 or y.shape[1] != 2:
        raise ValueError("HIP structure scoring currently supports two-dimensional embeddings.")
    n = y.shape[0]
    ind = np.asarray(indices)
    if ind.ndim != 2:
        raise ValueError("indices must be a matrix")
    kp = np.ascontiguousarray(np.atleast_1d(keep), dtype=np.int32)
    compact = ind.shape[0] == len(kp)
    if ind.shape[0] != n and not compact:
        raise ValueError("indices row count must match layout row count or keep length")
    lab = None
    if labels is not None and len(labels) != 0:
        if len(labels) != n:
            raise ValueError("labels length must match layout row count")
        lab = np.ascontiguousarray(labels, dtype=np.int32)
    nan = {name: float("nan") for name in STRUCTURE_SCORE_NAMES}
    if len(kp) == 0:
        return nan
    yc = np.asfortranarray(y).ravel(order="F")
    ic = np.asfortranarray(ind.astype(np.int32)).ravel(order="F")
    tot = np.zeros(6)
    levels = int(n_label_levels) if lab is not None else 0
    check(_lib.lib().fastembedr_hip_knn_structure_score(
        ptr(yc, _dp), ptr(ic, _ip), ptr(kp, _ip), None if lab is None else ptr(lab, _ip), int(n),
        int(ind.shape[0]), int(ind.shape[1]), int(preserve_k), int(len(kp)), int(compact),
        levels, ptr(tot, _dp)), "structure score")
    scored = int(tot[5])
    if scored == 0:
        return nan
    pres, trust, cont = tot[0] / scored, tot[1] / scored, tot[2] / scored
    return {"knn_preservation": pres, "local_trustworthiness": trust, "local_continuity": cont,
            "structure_score": (pres + trust + cont) / 3.0,
            "embedding_knn_accuracy": tot[3] / tot[4] if tot[4] > 0 else float("nan")}


def knn_structure_totals(layout, indices, keep, preserve_k, labels=None, n_label_levels=0):
    """The six raw totals of the C-ABI (for checks against the oracle)."""
    y = np.asfortranarray(np.asarray(layout, dtype=np.float64)).ravel(order="F")
    ind = np.asarray(indices)
    ic = np.asfortranarray(ind.astype(np.int32)).ravel(order="F")
    kp = np.ascontiguousarray(np.atleast_1d(keep), dtype=np.int32)
    lab = None if labels is None else np.ascontiguousarray(labels, dtype=np.int32)
    tot = np.zeros(6)
    check(_lib.lib().fastembedr_hip_knn_structure_score(
        ptr(y, _dp), ptr(ic, _ip), ptr(kp, _ip), None if lab is None else ptr(lab, _ip),
        int(len(y) // 2), int(ind.shape[0]), int(ind.shape[1]), int(preserve_k), int(len(kp)),
        int(ind.shape[0] == len(kp)), int(n_label_levels) if lab is not None else 0,
        ptr(tot, _dp)), "structure score")
    return tot
