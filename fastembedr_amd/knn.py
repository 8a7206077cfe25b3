"""Exact Euclidean k-nearest neighbours on the HIP backend.

``exact_knn`` is the producer behind the reference's exact native route
(``native_cuda_knn_impl(data, k, method = "exact", metric = "euclidean")``,
src/native_knn_cuda_impl.cpp:520-700, with ``finalize_cuvs_knn_kernel``,
src/embedding_cuda_kernels.cpp:1428-1480): the k nearest other rows of every
row, as the column-major n x k list that the t-SNE plans and the UMAP graph
read from device memory (the reference's GPU-resident list contract,
src/embedding_cuda_impl.cpp:2435-2523).  The neighbours are exact (FP64
squared distances of the float32 rows, ties by the smaller index).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check

MAX_K = 256


def exact_knn(data, k, index_base=0):
    """Exact k-NN of the rows of ``data`` (n x p).

    Returns (indices, distances) as (k, n) arrays: the column-major n x k
    layout (element (i, j) at j * n + i), int32 with ``index_base`` (0 or 1)
    and float32 Euclidean distances, the row itself excluded.
    """
    x = np.asarray(data, dtype=np.float32)
    if x.ndim != 2:
        raise ValueError("`data` must be a matrix")
    n, p = x.shape
    col = np.asfortranarray(x).ravel(order="F")
    idx = np.zeros((max(k, 0), n), dtype=np.int32)
    dist = np.zeros((max(k, 0), n), dtype=np.float32)
    scanned = ctypes.c_int(0)
    check(_lib.lib().fastembedr_hip_exact_knn(
        col.ctypes.data, int(n), int(p), int(k), 0, int(index_base), idx.ctypes.data,
        dist.ctypes.data, 0, ctypes.byref(scanned)), "exact KNN")
    return idx, dist


def exact_knn_device(data, n, p, k, indices, distances, index_base=0):
    """Device form: ``data`` is a device buffer holding the column-major
    n x p float32 matrix (a torch tensor of shape (p, n), or a pointer);
    ``indices`` (int32) and ``distances`` (float32) are device buffers of
    k * n elements written column-major n x k.  Returns the number of rows
    settled by the full FP64 scan."""
    from .umap import _device_ptr

    scanned = ctypes.c_int(0)
    check(_lib.lib().fastembedr_hip_exact_knn(
        _device_ptr(data), int(n), int(p), int(k), 1, int(index_base), _device_ptr(indices),
        _device_ptr(distances), 1, ctypes.byref(scanned)), "exact KNN")
    return scanned.value


def exact_query_knn(data, query, k, index_base=0):
    """Exact k-NN of the rows of ``query`` (m x p) among the rows of ``data``
    (n x p), no row dropped (the reference's exact query route).

    Returns (indices, distances) as (k, m) arrays: column-major m x k,
    indices into ``data`` with ``index_base``, float32 Euclidean distances.
    """
    x = np.asarray(data, dtype=np.float32)
    q = np.asarray(query, dtype=np.float32)
    if x.ndim != 2 or q.ndim != 2 or x.shape[1] != q.shape[1]:
        raise ValueError("`data` and `query` must have at least one row and matching columns.")
    n, p = x.shape
    m = q.shape[0]
    xc = np.asfortranarray(x).ravel(order="F")
    qc = np.asfortranarray(q).ravel(order="F")
    idx = np.zeros((max(k, 0), m), dtype=np.int32)
    dist = np.zeros((max(k, 0), m), dtype=np.float32)
    scanned = ctypes.c_int(0)
    check(_lib.lib().fastembedr_hip_exact_query_knn(
        xc.ctypes.data, int(n), qc.ctypes.data, int(m), int(p), int(k), 0, int(index_base),
        idx.ctypes.data, dist.ctypes.data, 0, ctypes.byref(scanned)), "exact query KNN")
    return idx, dist
