"""CPU oracle for the embedding-optimizer path (test infrastructure only).

This package is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it; the product package
``fastembedr_amd`` never does.

It wraps ``liboracle.so`` built from:

* ``oracle.c`` -- FP64 restatement of the t-SNE affinity, force, FFT-grid and
  update functions of ``src/tsne_neighbors.cpp`` (each C function cites the
  lines it follows) plus the exact rank metrics of
  ``src/embedding_utils.cpp:1582-1752``;
* ``oracle_umap.c`` -- the UMAP edge-sampled SGD epoch and its schedule/hash
  (``src/embedding_cuda_kernels.cpp:253-274,364-407,3044-3154``);
* ``oracle_umap_cpu.c`` -- the reference CPU UMAP optimizer
  (``optimize_layout_csr``, ``src/fast_knn_umap.cpp:3088-3558``): hogwild,
  delta and serial forms, used as the UMAP CPU baseline;
* ``oracle_transform.c`` -- the fixed-reference transform
  (``src/tsne_neighbors.cpp:2124-2373``);
* ``oracle_knn.c`` -- exact Euclidean k-nearest neighbours by FP64 brute
  force (what the reference's exact route, ``src/native_knn_cuda_impl.cpp``
  with ``finalize_cuvs_knn_kernel``, approximates in float32);
* ``cpu_port.c`` -- the float32, multi-threaded restatement of the reference
  CPU optimiser used as the timed CPU baseline.

``oracle.ref`` loads ``_ref/libref.so``, the reference's own C++ for the
functions that need only the standard library (``build_ref.sh``), which pins
the restatements above where it covers them.

Every array crossing this boundary is C-contiguous; neighbour indices are
0-based and row-major (n x k).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_u32p = ctypes.POINTER(ctypes.c_uint32)


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    c_int, c_long, c_double = ctypes.c_int, ctypes.c_long, ctypes.c_double
    L.or_row_probabilities.argtypes = [_dp, c_int, c_double, _dp]
    L.or_build_p.argtypes = [c_int, c_int, _ip, _dp, c_double, _ip, _ip, _dp, _dp]
    L.or_build_p.restype = c_long
    L.or_symmetrize.argtypes = [c_int, c_int, _ip, _dp, _ip, _ip, _dp]
    L.or_symmetrize.restype = c_long
    L.or_attractive.argtypes = [c_int, _ip, _ip, _dp, _dp, c_double, _dp]
    L.or_exact_repulsive.argtypes = [c_int, _dp, _dp]
    L.or_exact_repulsive.restype = c_double
    L.or_sum_q.argtypes = [c_int, _dp]
    L.or_sum_q.restype = c_double
    L.or_kl.argtypes = [c_int, _ip, _ip, _dp, _dp]
    L.or_kl.restype = c_double
    L.or_fft_repulsive.argtypes = [c_int, _dp, c_int, _dp, _dp, _dp, _dp]
    L.or_fft2.argtypes = [_dp, _dp, c_int, c_int]
    L.or_grid_geometry.argtypes = [c_int, _dp, c_int, _dp]
    L.or_direct_convolution.argtypes = [c_int, c_double, _dp, c_int, _dp]
    L.or_update.argtypes = [c_int, _dp, _dp, _dp, _dp, c_double, c_double, c_double, c_double]
    L.or_zero_mean.argtypes = [c_int, _dp]
    L.or_grid_size_cpu.argtypes = [c_int]
    L.or_grid_size_gpu.argtypes = [c_int]
    L.or_tsne_optimize.argtypes = [c_int, _ip, _ip, _dp, _dp, c_int, c_int, c_int, c_double,
                                   c_double, c_double, c_double, c_double, c_double, c_double,
                                   c_double]
    L.or_structure_metrics.argtypes = [c_int, c_int, _dp, c_int, _dp, c_int, _dp]
    L.port_tsne_run.argtypes = [c_int, c_int, _ip, _fp, _fp, c_double, c_int, c_int, c_double,
                                c_double, c_double, c_double, c_double, c_double, c_double,
                                c_int, c_int, _fp, _dp]
    L.port_tsne_iterate.argtypes = [c_int, _ip, _ip, _fp, _fp, c_int, c_double, c_double,
                                    c_double, c_double, c_double, c_int, c_int, _dp]
    if hasattr(L, "or_umap_epoch"):
        _declare_umap(L)
    if hasattr(L, "or_umap_fuzzy_graph"):
        _declare_umap_graph(L)
    if hasattr(L, "or_transform"):
        _declare_transform(L)
    if hasattr(L, "or_umap_cpu_optimize"):
        ll = ctypes.c_longlong
        L.or_umap_cpu_active.argtypes = [c_int, _ip, _ip, _fp, c_int, c_int, _ip, _ip, _fp]
        L.or_umap_cpu_active.restype = ll
        L.or_umap_cpu_optimize.argtypes = [ll, _ip, _ip, _fp, c_int, _fp, c_int, c_int,
                                           c_double, c_double, c_double, c_double, c_int, c_int,
                                           c_int, _dp]
        L.or_umap_cpu_optimize.restype = c_int
        L.or_umap_cpu_det_vertex.argtypes = [c_int, c_int, c_int, ll, c_int]
        L.or_umap_cpu_tau_vertices.argtypes = [c_int, c_int, ll, c_int, c_int, c_int, _ip]
        L.or_umap_cpu_pow.argtypes = [c_double, c_double]
        L.or_umap_cpu_pow.restype = c_double
        L.port_row_probabilities_f.argtypes = [_fp, c_int, c_int, c_double, _fp]
    if hasattr(L, "or_exact_knn"):
        L.or_exact_knn.argtypes = [_fp, c_int, c_int, c_int, _ip, c_int, _ip, _fp]
        L.or_exact_knn.restype = c_int
        L.or_knn_structure_score.argtypes = [_dp, c_int, c_int, _ip, c_int, c_int, _ip, c_int,
                                             c_int, c_int, _ip, c_int, _dp]
        L.or_knn_structure_score.restype = c_int
        L.or_exact_query_knn.argtypes = [_fp, c_int, _fp, c_int, c_int, c_int, _ip, _fp]
        L.or_exact_query_knn.restype = c_int


def _declare_umap(L):
    c_int, c_float, c_uint, c_double = ctypes.c_int, ctypes.c_float, ctypes.c_uint, ctypes.c_double
    L.or_umap_mix_uint.argtypes = [c_uint]
    L.or_umap_mix_uint.restype = c_uint
    L.or_umap_negative_vertex.argtypes = [c_int, c_uint, c_int, c_int, c_int, c_int]
    L.or_umap_negative_vertex.restype = c_int
    L.or_umap_n_positive.argtypes = [c_int, c_float]
    L.or_umap_n_positive.restype = c_int
    L.or_umap_n_negative.argtypes = [c_int, c_float, c_int]
    L.or_umap_n_negative.restype = c_int
    L.or_umap_fast_pow.argtypes = [c_float, c_float]
    L.or_umap_fast_pow.restype = c_float
    L.or_umap_epoch.argtypes = [_ip, _ip, _fp, c_int, _dp, c_int, c_int, c_int, c_int, c_double,
                                c_double, c_double, c_double, c_uint, c_int, _fp]
    L.or_umap_optimize.argtypes = [_ip, _ip, _fp, c_int, c_int, c_int, c_int, c_double, c_double,
                                   c_double, c_double, c_uint, c_int, _fp]
    L.or_umap_schedule.argtypes = [_ip, _ip, _fp, c_int, c_int, c_int, c_int, c_uint, c_int, _ip,
                                   _ip, _ip]


def _declare_umap_graph(L):
    c_int, c_ll = ctypes.c_int, ctypes.c_longlong
    L.or_umap_fuzzy_graph.argtypes = [_ip, _fp, c_int, c_int, c_int, c_int, c_int, _ip, _ip, _fp,
                                      _fp, _fp, _fp, _fp, c_ll]
    L.or_umap_fuzzy_graph.restype = c_int


def _declare_transform(L):
    c_int, c_double = ctypes.c_int, ctypes.c_double
    L.or_transform.argtypes = [c_int, _dp, c_int, c_int, _ip, _dp, _dp, c_int, c_double, c_int,
                               c_int, c_double, c_double, c_double, c_double, c_double,
                               c_double, c_double, c_int, c_int, c_int, _dp]
    L.or_transform.restype = c_int
    L.or_mix_uint32.argtypes = [ctypes.c_uint32]
    L.or_mix_uint32.restype = ctypes.c_uint32
    L.or_uniform_index.argtypes = [_u32p, c_int]
    L.or_uniform_index.restype = c_int


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a, t):
    return a.ctypes.data_as(t)


class CSR:
    """Symmetric affinity matrix in CSR form (0-based, ascending columns)."""

    def __init__(self, row_ptr, col, val):
        self.row_ptr, self.col, self.val = row_ptr, col, val

    @property
    def n(self):
        return len(self.row_ptr) - 1

    def dense(self):
        n = self.n
        out = np.zeros((n, n))
        for i in range(n):
            s, e = self.row_ptr[i], self.row_ptr[i + 1]
            out[i, self.col[s:e]] = self.val[s:e]
        return out


_ERRORS = {1: "KNN indices are out of range.",
           2: "KNN distances must be finite and non-negative.",
           3: "KNN graph produced no non-self t-SNE edges.",
           4: "t-SNE probability normalization failed.",
           5: "invalid argument", 6: "allocation failed"}


def row_probabilities(dist_row, perplexity):
    d = _d(dist_row)
    out = np.zeros_like(d)
    lib().or_row_probabilities(_ptr(d, _dp), len(d), float(perplexity), _ptr(out, _dp))
    return out


def build_p(indices, distances, perplexity, return_directed=False):
    """FP64 P: perplexity rows then symmetrised CSR (0-based indices)."""
    idx, dist = _i(indices), _d(distances)
    n, k = idx.shape
    row_ptr = np.zeros(n + 1, np.int32)
    col = np.zeros(2 * n * k, np.int32)
    val = np.zeros(2 * n * k)
    p_dir = np.zeros((n, k))
    nnz = lib().or_build_p(n, k, _ptr(idx, _ip), _ptr(dist, _dp), float(perplexity),
                           _ptr(row_ptr, _ip), _ptr(col, _ip), _ptr(val, _dp), _ptr(p_dir, _dp))
    if nnz < 0:
        raise ValueError(_ERRORS.get(-nnz, "oracle error"))
    csr = CSR(row_ptr, col[:nnz].copy(), val[:nnz].copy())
    return (csr, p_dir) if return_directed else csr


def attractive(csr, y, exaggeration=1.0):
    y = _d(y)
    out = np.zeros_like(y)
    lib().or_attractive(csr.n, _ptr(_i(csr.row_ptr), _ip), _ptr(_i(csr.col), _ip),
                        _ptr(_d(csr.val), _dp), _ptr(y, _dp), float(exaggeration), _ptr(out, _dp))
    return out


def exact_repulsive(y):
    y = _d(y)
    out = np.zeros_like(y)
    sq = lib().or_exact_repulsive(len(y), _ptr(y, _dp), _ptr(out, _dp))
    return out, sq


def sum_q(y):
    y = _d(y)
    return lib().or_sum_q(len(y), _ptr(y, _dp))


def kl(csr, y):
    y = _d(y)
    return lib().or_kl(csr.n, _ptr(_i(csr.row_ptr), _ip), _ptr(_i(csr.col), _ip),
                       _ptr(_d(csr.val), _dp), _ptr(y, _dp))


def fft_repulsive(y, grid_size, return_grids=False):
    """FP64 FFT-grid repulsion; returns (rep, Z[, grids(4,g,g), splat(3,g,g)])."""
    y = _d(y)
    n = len(y)
    g = int(grid_size)
    rep = np.zeros_like(y)
    z = np.zeros(1)
    grids = np.zeros((4, g, g))
    splat = np.zeros((3, g, g))
    rc = lib().or_fft_repulsive(n, _ptr(y, _dp), g, _ptr(rep, _dp), _ptr(z, _dp),
                                _ptr(grids, _dp), _ptr(splat, _dp))
    if rc != 0:
        raise ValueError("grid_size must be a power of two >= 4")
    if return_grids:
        return rep, float(z[0]), grids, splat
    return rep, float(z[0])


def grid_geometry(y, grid_size):
    y = _d(y)
    out = np.zeros(3)
    lib().or_grid_geometry(len(y), _ptr(y, _dp), int(grid_size), _ptr(out, _dp))
    return out


def fft2(re, im, inverse=False):
    re, im = _d(re).copy(), _d(im).copy()
    lib().or_fft2(_ptr(re, _dp), _ptr(im, _dp), re.shape[0], int(bool(inverse)))
    return re, im


def direct_convolution(src, spacing, power):
    src = _d(src)
    g = src.shape[0]
    out = np.zeros_like(src)
    lib().or_direct_convolution(g, float(spacing), _ptr(src, _dp), int(power), _ptr(out, _dp))
    return out


def update(y, upd, gains, grad, lr, momentum, min_gain, max_step):
    """In-place FP64 gains/momentum/clip update (arrays float64, n x 2)."""
    lib().or_update(len(y), _ptr(y, _dp), _ptr(upd, _dp), _ptr(gains, _dp), _ptr(_d(grad), _dp),
                    float(lr), float(momentum), float(min_gain), float(max_step))


def zero_mean(y):
    lib().or_zero_mean(len(y), _ptr(y, _dp))


def grid_size_cpu(n):
    return lib().or_grid_size_cpu(int(n))


def grid_size_gpu(n):
    return lib().or_grid_size_gpu(int(n))


def tsne_optimize(csr, y_init, grid_size, early_iter, n_iter, early_exag=12.0, exag=1.0,
                  lr_early=None, lr_normal=None, mom0=0.8, mom1=0.8, min_gain=0.01,
                  max_step=5.0):
    """FP64 optimisation from a fixed P; grid_size 0 selects exact repulsion."""
    y = _d(y_init).copy()
    n = len(y)
    lr_early = n / early_exag if lr_early is None else lr_early
    lr_normal = n / exag if lr_normal is None else lr_normal
    lib().or_tsne_optimize(n, _ptr(_i(csr.row_ptr), _ip), _ptr(_i(csr.col), _ip),
                           _ptr(_d(csr.val), _dp), _ptr(y, _dp), int(grid_size), int(early_iter),
                           int(n_iter), float(early_exag), float(exag), float(lr_early),
                           float(lr_normal), float(mom0), float(mom1), float(min_gain),
                           float(max_step))
    return y


def structure_metrics(high, low, k):
    """(trustworthiness, continuity, knn_preservation, rank_error) at k."""
    h, l_ = _d(high), _d(low)
    out = np.zeros(4)
    rc = lib().or_structure_metrics(len(h), h.shape[1], _ptr(h, _dp), l_.shape[1], _ptr(l_, _dp),
                                    int(k), _ptr(out, _dp))
    if rc != 0:
        raise ValueError("each k must satisfy 1 <= k < (2 * n - 1) / 3")
    return out


def exact_knn(data, k, queries=None):
    """Exact k-NN of the rows of ``data`` (n x p), self excluded.

    Returns (indices, distances), row-major len(queries) x k, 0-based, ties
    by the smaller index, distances (float32) sqrt of the FP64 squared
    distance.
    """
    x = np.asfortranarray(np.asarray(data, dtype=np.float32))
    n, p = x.shape
    q = None if queries is None else np.ascontiguousarray(queries, dtype=np.int32)
    nq = n if q is None else len(q)
    idx = np.zeros((nq, k), dtype=np.int32)
    dist = np.zeros((nq, k), dtype=np.float32)
    flat = x.ravel(order="F")
    rc = lib().or_exact_knn(_ptr(flat, _fp), n, p, int(k), None if q is None else _ptr(q, _ip), nq,
                            _ptr(idx, _ip), _ptr(dist, _fp))
    if rc != 0:
        raise ValueError("exact KNN needs n >= 2, p >= 1 and 1 <= k < n")
    return idx, dist


def exact_query_knn(data, query, k):
    """Exact k-NN of the rows of ``query`` among the rows of ``data`` (no row
    dropped): (indices, distances) row-major len(query) x k, 0-based."""
    x = np.asarray(data, dtype=np.float32)
    q = np.asarray(query, dtype=np.float32)
    n, p = x.shape
    m = q.shape[0]
    xf = np.asfortranarray(x).ravel(order="F")
    qf = np.asfortranarray(q).ravel(order="F")
    idx = np.zeros((m, k), dtype=np.int32)
    dist = np.zeros((m, k), dtype=np.float32)
    rc = lib().or_exact_query_knn(_ptr(xf, _fp), n, _ptr(qf, _fp), m, p, int(k), _ptr(idx, _ip),
                                  _ptr(dist, _fp))
    if rc != 0:
        raise ValueError("exact query KNN needs 1 <= k <= nrow(data)")
    return idx, dist


def knn_structure_score(layout, indices, keep, preserve_k, labels=None, n_levels=0):
    """The six totals of the KNN structure score (preservation, trust,
    continuity, label hits, label rows, scored rows), summed in row order.

    layout: n x dim; indices: index_rows x limit, 1-based (index_rows = n, or
    len(keep) for compact rows); keep: 1-based rows; labels: n factor codes.
    """
    y = np.asfortranarray(np.asarray(layout, dtype=np.float64))
    n, dim = y.shape
    ind = np.asfortranarray(np.asarray(indices, dtype=np.int32))
    index_rows, limit = ind.shape
    kp = np.ascontiguousarray(keep, dtype=np.int32)
    lab = None if labels is None else np.ascontiguousarray(labels, dtype=np.int32)
    out = np.zeros(6)
    yf, inf_ = y.ravel(order="F"), ind.ravel(order="F")
    rc = lib().or_knn_structure_score(
        _ptr(yf, _dp), n, dim, _ptr(inf_, _ip), index_rows, limit, _ptr(kp, _ip), len(kp),
        int(index_rows == len(kp)), int(preserve_k), None if lab is None else _ptr(lab, _ip),
        int(n_levels if lab is not None else 0), _ptr(out, _dp))
    if rc != 0:
        raise ValueError("invalid structure-score dimensions")
    return out


def port_tsne_iterate(row_ptr, col, val, y, iters, exag=12.0, learning_rate=None, momentum=0.8,
                      min_gain=0.01, max_step=5.0, grid=0, n_threads=None):
    """float32 reference-algorithm iterations from a given CSR P; returns
    (Y, seconds).  y is copied."""
    rp, c = _i(row_ptr), _i(col)
    v = np.ascontiguousarray(val, dtype=np.float32)
    yy = np.ascontiguousarray(y, dtype=np.float32).copy()
    n = len(rp) - 1
    lr = n / exag if learning_rate is None else learning_rate
    sec = np.zeros(1)
    rc = lib().port_tsne_iterate(n, _ptr(rp, _ip), _ptr(c, _ip), _ptr(v, _fp), _ptr(yy, _fp),
                                 int(iters), float(exag), float(lr), float(momentum),
                                 float(min_gain), float(max_step), int(grid),
                                 int(n_threads or os.cpu_count() or 1), _ptr(sec, _dp))
    if rc != 0:
        raise ValueError("port error %d" % rc)
    return yy, float(sec[0])


def port_tsne_run(indices, distances, y_init, perplexity=30.0, early_iter=250, n_iter=500,
                  early_exag=12.0, exag=1.0, learning_rate=None, mom0=0.8, mom1=0.8,
                  min_gain=0.01, max_step=5.0, grid=0, n_threads=None):
    """float32 reference-algorithm CPU run; returns (Y, affinity_s, optimise_s)."""
    idx = _i(indices)
    dist = np.ascontiguousarray(distances, dtype=np.float32)
    yi = np.ascontiguousarray(y_init, dtype=np.float32)
    n, k = idx.shape
    out = np.zeros((n, 2), np.float32)
    t = np.zeros(2)
    threads = n_threads or os.cpu_count() or 1
    rc = lib().port_tsne_run(n, k, _ptr(idx, _ip), _ptr(dist, _fp), _ptr(yi, _fp),
                             float(perplexity), int(early_iter), int(n_iter), float(early_exag),
                             float(exag), float(learning_rate or 0.0), float(mom0), float(mom1),
                             float(min_gain), float(max_step), int(grid), int(threads),
                             _ptr(out, _fp), _ptr(t, _dp))
    if rc != 0:
        raise ValueError(_ERRORS.get(rc, "port error %d" % rc))
    return out, float(t[0]), float(t[1])


# ------------------------------- UMAP -----------------------------------
# min_dist 0.01, spread 1 (find_ab_params, src/fast_knn_umap.cpp:437-497)
UMAP_A = 1.895605865596314
UMAP_B = 0.8006377738365004


def umap_fuzzy_graph(indices, distances, col_start=0, n_cols=None, index_offset=None):
    """The reference's default fuzzy graph (umap_graph_csr_float_cpp) of a KNN
    given as n x k arrays (rows = points).  index_offset None: detected as the
    reference does (1 when every used index lies in 1..n).  Returns a dict
    with offsets, neighbors, weights, epochs_per_sample, max_weight, sigmas,
    rhos."""
    idx = np.asarray(indices)
    n, k_total = idx.shape
    n_cols = k_total - col_start if n_cols is None else n_cols
    used = idx[:, col_start:col_start + n_cols]
    if index_offset is None:
        index_offset = 1 if (used.min() >= 1 and used.max() <= n) else 0
    ic = np.ascontiguousarray(idx.astype(np.int32).T)
    dc = np.ascontiguousarray(np.asarray(distances, np.float64).astype(np.float32).T)
    cap = n * n_cols * 2 + 1
    off = np.zeros(n + 1, np.int32)
    nb = np.zeros(cap, np.int32)
    w = np.zeros(cap, np.float32)
    eps = np.zeros(cap, np.float32)
    mw = np.zeros(1, np.float32)
    sig = np.zeros(n, np.float32)
    rho = np.zeros(n, np.float32)
    nnz = lib().or_umap_fuzzy_graph(_ptr(ic, _ip), _ptr(dc, _fp), n, k_total, col_start, n_cols,
                                    int(index_offset), _ptr(off, _ip), _ptr(nb, _ip), _ptr(w, _fp),
                                    _ptr(eps, _fp), _ptr(mw, _fp), _ptr(sig, _fp), _ptr(rho, _fp),
                                    cap)
    if nnz < 0:
        raise RuntimeError("or_umap_fuzzy_graph capacity")
    return {"offsets": off, "neighbors": nb[:nnz].copy(), "weights": w[:nnz].copy(),
            "epochs_per_sample": eps[:nnz].copy(), "max_weight": float(mw[0]), "sigmas": sig,
            "rhos": rho}


def umap_schedule(heads, tails, period, n, epoch, nsr, seed, max_neg):
    h, t = _i(heads), _i(tails)
    p = np.ascontiguousarray(period, dtype=np.float32)
    m = len(h)
    npos = np.zeros(m, np.int32)
    nneg = np.zeros(m, np.int32)
    ids = np.zeros((m, max(max_neg, 1)), np.int32)
    lib().or_umap_schedule(_ptr(h, _ip), _ptr(t, _ip), _ptr(p, _fp), m, int(n), int(epoch),
                           int(nsr), int(seed) & 0xFFFFFFFF, int(max_neg), _ptr(npos, _ip),
                           _ptr(nneg, _ip), _ptr(ids, _ip))
    return npos, nneg, ids[:, :max_neg]


def umap_optimize(heads, tails, period, y_init, n_epochs, nsr=5, lr=1.0, a=UMAP_A, b=UMAP_B,
                  gamma=1.0, seed=0, mode=1):
    h, t = _i(heads), _i(tails)
    p = np.ascontiguousarray(period, dtype=np.float32)
    y = np.ascontiguousarray(y_init, dtype=np.float32).copy()
    lib().or_umap_optimize(_ptr(h, _ip), _ptr(t, _ip), _ptr(p, _fp), len(h), len(y),
                           int(n_epochs), int(nsr), float(lr), float(a), float(b), float(gamma),
                           int(seed) & 0xFFFFFFFF, int(mode), _ptr(y, _fp))
    return y


def umap_epoch(heads, tails, period, y, epoch, n_epochs, nsr=5, lr=1.0, a=UMAP_A, b=UMAP_B,
               gamma=1.0, seed=0, mode=1):
    h, t = _i(heads), _i(tails)
    p = np.ascontiguousarray(period, dtype=np.float32)
    yy = np.ascontiguousarray(y, dtype=np.float32).copy()
    scratch = np.zeros(2 * len(yy))
    lib().or_umap_epoch(_ptr(h, _ip), _ptr(t, _ip), _ptr(p, _fp), len(h), _ptr(scratch, _dp),
                        len(yy), int(epoch), int(n_epochs), int(nsr), float(lr), float(a),
                        float(b), float(gamma), int(seed) & 0xFFFFFFFF, int(mode), _ptr(yy, _fp))
    return yy


# ------------------------- reference CPU UMAP (A16) ------------------------
def umap_cpu_active(offsets, neighbors, weights, n_epochs, index_offset=0):
    """The active edges of optimize_layout_csr: (rows, tails, epochs_per_sample)."""
    off, nb = _i(offsets), _i(neighbors)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    n = len(off) - 1
    m = len(nb)
    rows, tails = np.zeros(m, np.int32), np.zeros(m, np.int32)
    eps = np.zeros(m, np.float32)
    cnt = lib().or_umap_cpu_active(n, _ptr(off, _ip), _ptr(nb, _ip), _ptr(w, _fp),
                                   int(index_offset), int(n_epochs), _ptr(rows, _ip),
                                   _ptr(tails, _ip), _ptr(eps, _fp))
    if cnt < 0:
        raise ValueError("The CSR graph has no usable UMAP edges." if cnt == -1 else
                         "The CSR graph has no edges sampled by n_epochs.")
    return rows[:cnt], tails[:cnt], eps[:cnt]


def umap_cpu_optimize(rows, tails, eps, y_init, n_epochs, nsr=5, lr=1.0, a=UMAP_A, b=UMAP_B,
                      gamma=1.0, seed=1, n_threads=1, form=-1):
    """The reference CPU UMAP epochs; form -1 picks as the reference does
    (hogwild / delta / serial).  Returns (layout, form, seconds)."""
    r, t = _i(rows), _i(tails)
    e = np.ascontiguousarray(eps, dtype=np.float32)
    y = np.ascontiguousarray(y_init, dtype=np.float32).copy()
    sec = ctypes.c_double()
    used = lib().or_umap_cpu_optimize(len(r), _ptr(r, _ip), _ptr(t, _ip), _ptr(e, _fp), len(y),
                                      _ptr(y, _fp), int(n_epochs), int(nsr), float(a), float(b),
                                      float(gamma), float(lr), int(seed), int(n_threads),
                                      int(form), ctypes.byref(sec))
    if used < 0:
        raise ValueError("invalid UMAP CPU optimizer arguments")
    return y, int(used), sec.value


def port_row_probabilities_f(dist, perplexity):
    d = np.ascontiguousarray(dist, dtype=np.float32)
    out = np.zeros_like(d)
    lib().port_row_probabilities_f(_ptr(d, _fp), d.shape[0], d.shape[1], float(perplexity),
                                   _ptr(out, _fp))
    return out


# ----------------------------- transform ---------------------------------
def transform(reference_layout, indices, distances, Y_init=None, initialization="median",
              perplexity=5.0, n_iter=250, early_exaggeration_iter=0, learning_rate=0.1,
              early_exaggeration=4.0, exaggeration=1.5, initial_momentum=0.8, final_momentum=0.8,
              max_grad_norm=0.25, max_step_norm=np.inf, n_negatives=None,
              exact_repulsion_threshold=4096, seed=4):
    """FP64 fixed-reference transform (defaults of R/transform_tsne.R:58-67)."""
    ref = _d(reference_layout)
    idx, dist = _i(indices), _d(distances)
    nq, k = idx.shape
    nref = len(ref)
    if n_negatives is None:
        n_negatives = nref if nref <= exact_repulsion_threshold else min(256, nref)
    init = None if Y_init is None else _d(Y_init)
    out = np.zeros((nq, 2))
    mode = {"median": 0, "weighted": 1}[initialization] if init is None else 0
    rc = lib().or_transform(nref, _ptr(ref, _dp), nq, k, _ptr(idx, _ip), _ptr(dist, _dp),
                            _ptr(init, _dp) if init is not None else None, mode,
                            float(perplexity), int(n_iter), int(early_exaggeration_iter),
                            float(learning_rate), float(early_exaggeration), float(exaggeration),
                            float(initial_momentum), float(final_momentum), float(max_grad_norm),
                            float(max_step_norm), int(n_negatives), int(exact_repulsion_threshold),
                            int(seed), _ptr(out, _dp))
    if rc != 0:
        raise ValueError(_ERRORS.get(rc, "oracle error"))
    return out
