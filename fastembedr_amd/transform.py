"""Fixed-reference t-SNE transform on the HIP backend.

``transform_tsne`` mirrors the reference's public ``transform_tsne()``
(R/transform_tsne.R:58-104) for a precomputed query KNN, with its defaults
(perplexity 5, 250 iterations, no early exaggeration, learning rate 0.1,
exaggeration 1.5, momentum 0.8, max_grad_norm 0.25) and its choice of
n_negatives (all reference points up to 4096, else min(256, n_ref);
:201-238).  ``TransformReference`` keeps a reference layout resident on the
device for repeated requests.
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib
from ._lib import _dp, _fp, _ip, check, ptr

_INIT = {"median": 0, "weighted": 1, "random": 2}


def _controls(n_ref, n_negatives, threshold):
    if n_negatives is None:
        n_negatives = n_ref if n_ref <= threshold else min(256, n_ref)
    n_negatives = int(n_negatives)
    if n_negatives < 1:
        raise ValueError("`n_negatives` must be NULL or a positive integer.")
    return n_negatives


def _knn(knn, indices, distances):
    if knn is not None:
        indices, distances = knn["indices"], knn["distances"]
    idx = np.asarray(indices)
    dist = np.asarray(distances)
    if idx.ndim != 2 or idx.shape != dist.shape:
        raise ValueError("KNN `indices` and `distances` must have the same dimensions.")
    return idx, dist


def _offset(idx, n_ref):
    """resolve_reference_index_offset (src/tsne_neighbors.cpp:1367-1382)."""
    return 1 if idx.size and idx.min() >= 1 and idx.max() <= n_ref else 0


def transform_tsne(reference_layout, knn=None, indices=None, distances=None, perplexity=5.0,
                   initialization="median", Y_init=None, n_iter=250, early_exaggeration_iter=0,
                   learning_rate=0.1, early_exaggeration=4.0, exaggeration=1.5,
                   initial_momentum=0.8, final_momentum=0.8, max_grad_norm=0.25,
                   max_step_norm=math.inf, n_negatives=None, exact_repulsion_threshold=4096,
                   seed=4, backend="hip"):
    """Place new points into a fixed reference embedding; returns float32 n x 2."""
    if backend != "hip":
        raise ValueError("this package implements backend = \"hip\" only")
    ref = np.ascontiguousarray(reference_layout, dtype=np.float32)
    if ref.ndim != 2 or ref.shape[1] != 2 or len(ref) < 1:
        raise ValueError("`reference_layout` must be a non-empty n x 2 matrix.")
    if initialization not in _INIT:
        raise ValueError("`initialization` must be 'median', 'weighted', or 'random'.")
    idx, dist = _knn(knn, indices, distances)
    nq, k = idx.shape
    nneg = _controls(len(ref), n_negatives, exact_repulsion_threshold)
    init = None
    if Y_init is not None:
        init = np.ascontiguousarray(Y_init, dtype=np.float32)
        if init.shape != (nq, 2):
            raise ValueError("`Y_init` has the wrong shape.")
    ci = np.asfortranarray(idx, dtype=np.int32)
    cd = np.asfortranarray(dist, dtype=np.float64)
    out = np.zeros((nq, 2), np.float32)
    rc = _lib.lib().fastembedr_hip_tsne_transform(
        ptr(ref, _fp), len(ref), ptr(ci, _ip), ptr(cd, _dp), nq, k, _offset(idx, len(ref)),
        ptr(init, _fp) if init is not None else None, _INIT[initialization], float(perplexity),
        int(n_iter), int(early_exaggeration_iter), float(learning_rate), float(early_exaggeration),
        float(exaggeration), float(initial_momentum), float(final_momentum), float(max_grad_norm),
        float(max_step_norm), nneg, int(exact_repulsion_threshold), int(seed), ptr(out, _fp))
    check(rc, "HIP t-SNE transform")
    return out


class TransformReference:
    """A reference layout resident on the device for repeated transforms."""

    def __init__(self, reference_layout):
        ref = np.ascontiguousarray(reference_layout, dtype=np.float32)
        self.n_ref = len(ref)
        self._h = _lib.ctypes.c_void_p()
        check(_lib.lib().fastembedr_hip_transform_reference_create(
            ptr(ref, _fp), self.n_ref, _lib.ctypes.byref(self._h)), "transform reference")

    def transform(self, indices, distances, perplexity=5.0, initialization="median", Y_init=None,
                  n_iter=250, early_exaggeration_iter=0, learning_rate=0.1, early_exaggeration=4.0,
                  exaggeration=1.5, initial_momentum=0.8, final_momentum=0.8, max_grad_norm=0.25,
                  max_step_norm=math.inf, n_negatives=None, exact_repulsion_threshold=4096,
                  seed=4):
        idx, dist = _knn(None, indices, distances)
        nq, k = idx.shape
        nneg = _controls(self.n_ref, n_negatives, exact_repulsion_threshold)
        init = None if Y_init is None else np.ascontiguousarray(Y_init, dtype=np.float32)
        ci = np.asfortranarray(idx, dtype=np.int32)
        cd = np.asfortranarray(dist, dtype=np.float32)
        out = np.zeros((nq, 2), np.float32)
        rc = _lib.lib().fastembedr_hip_transform_reference_run(
            self._h, ptr(ci, _ip), ptr(cd, _fp), nq, k, _offset(idx, self.n_ref),
            ptr(init, _fp) if init is not None else None, _INIT[initialization],
            float(perplexity), int(n_iter), int(early_exaggeration_iter), float(learning_rate),
            float(early_exaggeration), float(exaggeration), float(initial_momentum),
            float(final_momentum), float(max_grad_norm), float(max_step_norm), nneg,
            int(exact_repulsion_threshold), int(seed), ptr(out, _fp))
        check(rc, "HIP t-SNE transform (resident reference)")
        return out

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.lib().fastembedr_hip_transform_reference_destroy(self._h)
            self._h = _lib.ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
