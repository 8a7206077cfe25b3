"""t-SNE on the HIP backend: the host side above the C-ABI.

``tsne_knn`` mirrors the reference's public ``tsne_knn()`` for
``backend = "hip"`` (R/fast_knn_tsne.R:216-261): KNN input in, a float32
n x 2 layout out, with the reference's parameter policy
(tsne_auto_parameters_cpp, src/tsne_neighbors.cpp:1820-1868;
R/tsne_optimizer_helpers.R:151-221) and a ``config`` dict in place of the R
attribute ``fastEmbedR_config``.  ``gradient_diagnostic`` is the HIP
counterpart of ``opentsne_force_diagnostic_cpp`` (tsne_neighbors.cpp:
1698-1817).  ``TsnePlan`` wraps the resident-plan entry points.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import _dp, _fp, _ip, _llp, check, ptr

__all__ = ["tsne_knn", "auto_parameters", "grid_size", "gradient_diagnostic",
           "force_diagnostic", "TsnePlan", "Embedding"]


class Embedding(np.ndarray):
    """float32 n x 2 layout carrying the resolved run configuration."""

    config: dict

    def __new__(cls, data, config):
        obj = np.asarray(data, dtype=np.float32).view(cls)
        obj.config = config
        return obj

    def __array_finalize__(self, obj):
        self.config = getattr(obj, "config", {})


def auto_parameters(n: int, k: int, perplexity: float | None = None) -> dict:
    """tsne_auto_parameters_cpp (src/tsne_neighbors.cpp:1820-1868)."""
    if n < 2:
        raise ValueError("`n` must be at least 2.")
    if k < 1:
        raise ValueError("`k` must be positive.")
    max_n = max(1, (n - 1) // 3)
    max_k = max(1, k // 3)
    missing = perplexity is None or not math.isfinite(perplexity) or perplexity <= 0
    p = float(min(30, max_n, max_k)) if missing else float(perplexity)
    cap = min(max_n, max_k) if missing else max_n
    p = max(1.0, min(p, float(cap)))
    early = 12.0
    return dict(perplexity=p, n_neighbors=max(1, min(n - 1, math.ceil(3.0 * p))),
                early_exaggeration=early, exaggeration=1.0, learning_rate=n / early,
                early_exaggeration_iter=250, n_iter=500)


def grid_size(n: int) -> int:
    """The backend's grid rule (resolve_cuda_tsne_fft_grid_size,
    src/embedding_cuda_kernels.cpp:5610-5623)."""
    return int(_lib.lib().fastembedr_hip_opentsne_fft_grid_size(int(n)))


def _knn_arrays(indices, distances):
    if isinstance(indices, dict):
        distances = indices["distances"] if distances is None else distances
        indices = indices["indices"]
    if distances is None:
        raise ValueError("KNN `distances` are required.")
    idx = np.asarray(indices)
    dist = np.asarray(distances)
    if idx.ndim != 2 or idx.shape != dist.shape:
        raise ValueError("KNN `indices` and `distances` must have the same dimensions.")
    if not np.issubdtype(idx.dtype, np.integer):
        raise ValueError("KNN `indices` must be integers.")
    return idx, dist


def _index_offset(idx) -> int:
    """knn_index_offset (src/embedding_cuda_impl.cpp:933-945): 1-based when all
    ids lie in [1, n]."""
    n = idx.shape[0]
    return 1 if idx.size and idx.min() >= 1 and idx.max() <= n else 0


def _column_major(idx, dist, dtype):
    return (np.asfortranarray(idx, dtype=np.int32), np.asfortranarray(dist, dtype=dtype))


def tsne_knn(indices, distances=None, n_neighbors=None, perplexity=None,
             affinity_support="standard", n_components=2, Y_init=None, seed=4,
             backend="hip", learning_rate="auto", early_exaggeration_iter=None,
             early_exaggeration="auto", n_iter=None, exaggeration=None, initial_momentum=0.8,
             final_momentum=0.8, max_step_norm="auto", min_gain=0.01,
             negative_gradient_method="auto", auto_config=True, grid_size=0):
    """Embed a KNN graph with the HIP FFT-grid t-SNE optimizer.

    ``indices``/``distances`` are n x k (0- or 1-based ids, self excluded), or a
    dict with those keys.  Returns an :class:`Embedding` (float32 n x 2).
    """
    if backend != "hip":
        raise ValueError("this package implements backend = \"hip\" only")
    # resolve_opentsne_gradient_method (R/tsne_optimizer_helpers.R:61-75) with
    # the reference's method names (tsne_repulsion_mode,
    # src/tsne_neighbors.cpp:365-400): "auto" is the FFT grid on a GPU backend
    # (as for backend = "cuda"); "exact" is the pair form every iteration
    method = str(negative_gradient_method).lower()
    if method in ("bh", "barnes_hut", "barnes-hut"):
        raise ValueError("Barnes-Hut openTSNE has been removed from fastEmbedR. Use "
                         "`negative_gradient_method = \"fft\"` or `\"exact\"`.")
    if method in ("exact", "pair", "pair_symmetric"):
        exact = True
    elif method in ("auto", "fft", "fitsne", "fit_sne", "interpolation"):
        exact = False
    else:
        raise ValueError(f"unknown `negative_gradient_method` \"{negative_gradient_method}\".")
    if affinity_support not in ("standard", "compact"):
        raise ValueError("`affinity_support` must be \"standard\" or \"compact\".")
    if n_components != 2:
        raise ValueError("HIP openTSNE FFT-grid currently supports exactly two output components.")
    idx, dist = _knn_arrays(indices, distances)
    n, k = idx.shape
    auto = auto_parameters(n, k, perplexity)
    perp = auto["perplexity"]
    required = math.ceil((3 if affinity_support == "standard" else 1) * perp)
    nn = k if n_neighbors is None else int(n_neighbors)
    if nn > k:
        raise ValueError("`n_neighbors` exceeds the supplied KNN width.")
    if nn < required:
        raise ValueError(f"`n_neighbors` is too small for `affinity_support = \"{affinity_support}\"`;"
                         f" need at least {required}.")
    idx, dist = idx[:, :nn], dist[:, :nn]
    if n - 1 < 3.0 * perp:
        raise ValueError("perplexity is too large for the number of samples.")
    early_iter = auto["early_exaggeration_iter"] if early_exaggeration_iter is None else int(early_exaggeration_iter)
    normal_iter = auto["n_iter"] if n_iter is None else int(n_iter)
    ee = auto["early_exaggeration"] if early_exaggeration == "auto" else float(early_exaggeration)
    ex = auto["exaggeration"] if exaggeration is None else float(exaggeration)
    if learning_rate == "auto":
        # auto_config = TRUE fixes n / early_exaggeration for both phases
        # (R/tsne_optimizer_helpers.R:200-205); otherwise n / phase exaggeration
        lr_auto = not auto_config
        lr = n / ee
    else:
        lr_auto, lr = False, float(learning_rate)
    step = 5.0 if max_step_norm == "auto" else (float("nan") if max_step_norm is None
                                                 else float(max_step_norm))
    has_init = Y_init is not None
    init = np.ascontiguousarray(Y_init, dtype=np.float32) if has_init else np.zeros((n, 2), np.float32)
    if has_init:
        if init.shape != (n, 2):
            raise ValueError("`Y_init` has the wrong shape.")
        if not np.all(np.isfinite(init)):
            raise ValueError("`Y_init` must contain only finite coordinates.")
    offset = _index_offset(idx)
    out = np.zeros((n, 2), np.float32)
    if exact:
        grid_size = -1
    L = _lib.lib()
    if dist.dtype == np.float32:
        ci, cd = _column_major(idx, dist, np.float32)
        rc = L.fastembedr_hip_opentsne_fft_from_knn_float(
            ptr(ci, _ip), ptr(cd, _fp), ptr(init, _fp), int(has_init), n, nn, 2, perp, early_iter,
            normal_iter, ee, ex, lr, int(lr_auto), initial_momentum, final_momentum, min_gain,
            step, seed & 0xFFFFFFFF, offset, int(grid_size), ptr(out, _fp))
    else:
        ci, cd = _column_major(idx, dist, np.float64)
        rc = L.fastembedr_hip_opentsne_fft_from_knn(
            ptr(ci, _ip), ptr(cd, _dp), ptr(init, _fp), int(has_init), n, nn, 2, perp, early_iter,
            normal_iter, ee, ex, lr, int(lr_auto), initial_momentum, final_momentum, min_gain,
            step, seed & 0xFFFFFFFF, offset, int(grid_size), ptr(out, _fp))
    check(rc, "HIP openTSNE FFT-grid")
    g = None if exact else (int(grid_size) or globals()["grid_size"](n))
    config = dict(backend="hip",
                  optimizer=("opentsne_exact_sparse_knn_hip" if exact
                             else "opentsne_fitsne_fft_grid_native_hip"),
                  repulsion="pair_symmetric_hip" if exact else "fft_grid_hip", fft_grid_size=g,
                  probabilities="symmetric_sparse_knn_hip_float32", precision="float32",
                  perplexity=perp, n_neighbors=nn, affinity_support=affinity_support,
                  early_exaggeration_iter=early_iter, n_iter=normal_iter,
                  early_exaggeration=ee, exaggeration=ex,
                  learning_rate_early=n / ee if lr_auto else lr,
                  learning_rate_normal=n / ex if lr_auto else lr,
                  initial_momentum=initial_momentum, final_momentum=final_momentum,
                  min_gain=min_gain, max_step_norm=step, seed=seed, index_offset=offset)
    return Embedding(out, config)


@dataclass
class Diagnostic:
    attractive_force: np.ndarray
    repulsive_force_fft: np.ndarray
    optimizer_gradient_fft: np.ndarray
    objective_gradient_fft: np.ndarray
    sum_q: float
    row_ptr: np.ndarray
    col: np.ndarray
    val: np.ndarray


def gradient_diagnostic(indices, distances, layout, perplexity, exaggeration=1.0, grid_size=64):
    """Forces from a fixed layout on the HIP path (host arrays, 0/1-based ids)."""
    idx, dist = _knn_arrays(indices, distances)
    n, k = idx.shape
    y = np.ascontiguousarray(layout, dtype=np.float32)
    if y.shape != (n, 2):
        raise ValueError("`layout` must have one row per KNN row and exactly two columns.")
    if not np.all(np.isfinite(y)):
        raise ValueError("`layout` must contain only finite coordinates.")
    ci, cd = _column_major(idx, dist, np.float32)
    att = np.zeros((n, 2), np.float32)
    rep = np.zeros((n, 2), np.float32)
    grad = np.zeros((n, 2), np.float32)
    sq = np.zeros(1)
    row_ptr = np.zeros(n + 1, np.int32)
    col = np.zeros(2 * n * k, np.int32)
    val = np.zeros(2 * n * k, np.float32)
    nnz = np.zeros(1, np.int64)
    rc = _lib.lib().fastembedr_hip_tsne_gradient_diagnostic(
        ptr(ci, _ip), ptr(cd, _fp), n, k, _index_offset(idx), float(perplexity),
        float(exaggeration), int(grid_size), ptr(y, _fp), ptr(att, _fp), ptr(rep, _fp),
        ptr(grad, _fp), ptr(sq, _dp), ptr(row_ptr, _ip), ptr(col, _ip), ptr(val, _fp),
        ptr(nnz, _llp))
    check(rc, "HIP t-SNE gradient diagnostic")
    m = int(nnz[0])
    return Diagnostic(att, rep, grad, 4.0 * grad, float(sq[0]), row_ptr, col[:m].copy(),
                      val[:m].copy())


def force_diagnostic(indices, distances, layout, perplexity, exaggeration=1.0, grid_size=64):
    """HIP counterpart of ``opentsne_force_diagnostic_cpp``
    (src/tsne_neighbors.cpp:1698-1817): the whole bundle from a fixed layout --
    attraction, exact and FFT-grid repulsion, both optimizer and objective
    gradients, the exact ``sum_q``, ``kl``, the norms and the affinities (with
    1-based columns, as ``affinity_col1``).  Returns a dict with the
    reference's field names."""
    idx, dist = _knn_arrays(indices, distances)
    n, k = idx.shape
    if not (0.0 < float(perplexity) <= k):
        raise ValueError("`perplexity` must be positive and no larger than the KNN width.")
    y = np.ascontiguousarray(layout, dtype=np.float32)
    if y.shape != (n, 2):
        raise ValueError("`layout` must have one row per KNN row and exactly two columns.")
    ci, cd = _column_major(idx, dist, np.float32)
    out = {key: np.zeros((n, 2), np.float32) for key in
           ("attractive_force", "repulsive_force_exact", "optimizer_gradient_exact",
            "repulsive_force_fft", "optimizer_gradient_fft")}
    sq, kl = np.zeros(1), np.zeros(1)
    row_ptr = np.zeros(n + 1, np.int32)
    col = np.zeros(2 * n * k, np.int32)
    val = np.zeros(2 * n * k, np.float32)
    nnz = np.zeros(1, np.int64)
    rc = _lib.lib().fastembedr_hip_opentsne_force_diagnostic(
        ptr(ci, _ip), ptr(cd, _fp), n, k, _index_offset(idx), float(perplexity),
        float(exaggeration), int(grid_size), ptr(y, _fp), ptr(out["attractive_force"], _fp),
        ptr(out["repulsive_force_exact"], _fp), ptr(out["optimizer_gradient_exact"], _fp),
        ptr(out["repulsive_force_fft"], _fp), ptr(out["optimizer_gradient_fft"], _fp),
        ptr(sq, _dp), ptr(kl, _dp), ptr(row_ptr, _ip), ptr(col, _ip), ptr(val, _fp),
        ptr(nnz, _llp))
    check(rc, "HIP t-SNE force diagnostic")
    m = int(nnz[0])
    out["objective_gradient_exact"] = 4.0 * out["optimizer_gradient_exact"]
    out["objective_gradient_fft"] = 4.0 * out["optimizer_gradient_fft"]
    out.update(sum_q=float(sq[0]), kl=float(kl[0]),
               attractive_norm=float(np.linalg.norm(out["attractive_force"].astype(np.float64))),
               repulsive_exact_norm=float(np.linalg.norm(
                   out["repulsive_force_exact"].astype(np.float64))),
               repulsive_fft_norm=float(np.linalg.norm(
                   out["repulsive_force_fft"].astype(np.float64))),
               affinity_row_ptr0=row_ptr, affinity_col1=col[:m] + 1,
               affinity_weight=val[:m].astype(np.float64), grid_size=int(grid_size),
               perplexity=float(perplexity), exaggeration=float(exaggeration),
               precision="float32", objective_gradient_scale=4.0)
    return out


def _device_inputs_ready(*arrays):
    """The library reads device inputs on its own stream: wait for the
    producer's work on torch tensors before handing their pointers over."""
    for a in arrays:
        if getattr(a, "is_cuda", False):
            import torch

            torch.cuda.current_stream(a.device).synchronize()


class TsnePlan:
    """A resident t-SNE job: P, layout, optimizer state and FFT work arrays on
    the device.  ``indices``/``distances`` are host n x k arrays or, with
    ``device=True``, objects exposing ``data_ptr()`` to column-major device
    memory (e.g. torch tensors created as ``t.T.contiguous()`` of a k x n).
    ``attraction`` (0 built in, 1 row groups, 2 segment-blocked, 3 every pair
    once) and ``pair_workgroups`` (0 built in) are the launch choices the
    autotuner picks per size bucket."""

    def __init__(self, indices, distances, perplexity=30.0, grid_size=0, Y_init=None, seed=4,
                 device=False, n=None, k=None, index_offset=None, attraction=0,
                 pair_workgroups=0):
        L = _lib.lib()
        self._h = _lib.ctypes.c_void_p()
        init = None
        if device:
            self.n, self.k = int(n), int(k)
            off = 0 if index_offset is None else int(index_offset)
            _device_inputs_ready(indices, distances)
            ip, dp = _lib.ctypes.c_void_p(indices.data_ptr()), _lib.ctypes.c_void_p(distances.data_ptr())
            self._keep = (indices, distances)
        else:
            idx, dist = _knn_arrays(indices, distances)
            self.n, self.k = idx.shape
            off = _index_offset(idx) if index_offset is None else int(index_offset)
            ci, cd = _column_major(idx, dist, np.float32)
            self._keep = (ci, cd)
            ip, dp = ci.ctypes.data_as(_lib.ctypes.c_void_p), cd.ctypes.data_as(_lib.ctypes.c_void_p)
        if Y_init is not None:
            init = np.ascontiguousarray(Y_init, dtype=np.float32)
            if init.shape != (self.n, 2):
                raise ValueError("`Y_init` has the wrong shape.")
        rc = L.fastembedr_hip_tsne_plan_create_shaped(
            ip, dp, int(bool(device)), self.n, self.k, off, float(perplexity), int(grid_size),
            int(attraction), int(pair_workgroups), ptr(init, _fp) if init is not None else None,
            seed & 0xFFFFFFFF, _lib.ctypes.byref(self._h))
        check(rc, "HIP t-SNE plan create")
        self._keep = None

    def run(self, n_iter, exaggeration, momentum, learning_rate, min_gain=0.01, max_step_norm=5.0):
        check(_lib.lib().fastembedr_hip_tsne_plan_run(self._h, int(n_iter), float(exaggeration),
                                                      float(momentum), float(learning_rate),
                                                      float(min_gain), float(max_step_norm)),
              "HIP t-SNE plan run")

    def prepare(self, n_iter, exaggeration, momentum, learning_rate, min_gain=0.01,
                max_step_norm=5.0, after=0):
        """Capture the iteration graphs a later ``run`` with these arguments
        replays (started after ``after`` further iterations) without running
        them."""
        check(_lib.lib().fastembedr_hip_tsne_plan_prepare(
            self._h, int(n_iter), float(exaggeration), float(momentum), float(learning_rate),
            float(min_gain), float(max_step_norm), int(after)), "HIP t-SNE plan prepare")

    def sync(self):
        check(_lib.lib().fastembedr_hip_tsne_plan_sync(self._h), "HIP t-SNE plan sync")

    def kl(self):
        """(KL divergence, exact Z) of the current layout (evaluate_kl_f)."""
        kl, sq = np.zeros(1), np.zeros(1)
        check(_lib.lib().fastembedr_hip_tsne_plan_kl(self._h, ptr(kl, _dp), ptr(sq, _dp)),
              "HIP t-SNE plan KL")
        return float(kl[0]), float(sq[0])

    def layout(self):
        out = np.zeros((self.n, 2), np.float32)
        check(_lib.lib().fastembedr_hip_tsne_plan_layout(self._h, ptr(out, _fp)),
              "HIP t-SNE plan layout")
        return out

    def info(self):
        v = np.zeros(8, np.int64)
        check(_lib.lib().fastembedr_hip_tsne_plan_info(self._h, ptr(v, _llp)), "plan info")
        keys = ("n", "k", "grid_size", "nnz", "lanes_per_row", "device_bytes", "iterations",
                "setup_us")
        return dict(zip(keys, (int(x) for x in v)))

    def attraction_info(self):
        """The attraction kernel the plan chose and the sizes of its tables."""
        v = np.zeros(6, np.int64)
        check(_lib.lib().fastembedr_hip_tsne_plan_attraction_info(self._h, ptr(v, _llp)),
              "plan attraction info")
        kinds = ("rows", "lanes", "staged", "pairs")
        return dict(kind=kinds[int(v[0])], entries=int(v[1]), rest_entries=int(v[2]),
                    workgroups=int(v[3]), chunks=int(v[4]), slots=int(v[5]))

    def affinities(self):
        """(row_ptr, col, val) of the symmetric P, 0-based CSR."""
        nnz = self.info()["nnz"]
        row_ptr = np.zeros(self.n + 1, np.int32)
        col = np.zeros(max(nnz, 1), np.int32)
        val = np.zeros(max(nnz, 1), np.float32)
        check(_lib.lib().fastembedr_hip_tsne_plan_affinities(self._h, ptr(row_ptr, _ip),
                                                             ptr(col, _ip), ptr(val, _fp)),
              "plan affinities")
        return row_ptr, col[:nnz], val[:nnz]

    def profile(self, n_iter, exaggeration, momentum, learning_rate, min_gain=0.01,
                max_step_norm=5.0):
        ms = np.zeros(7)
        check(_lib.lib().fastembedr_hip_tsne_plan_profile(
            self._h, int(n_iter), float(exaggeration), float(momentum), float(learning_rate),
            float(min_gain), float(max_step_norm), ptr(ms, _dp)), "plan profile")
        keys = ("splat", "conv_rows_fwd", "conv_cols", "conv_rows_inv", "attract", "update",
                "iteration")
        return dict(zip(keys, ms.tolist()))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.lib().fastembedr_hip_tsne_plan_destroy(self._h)
            self._h = _lib.ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
