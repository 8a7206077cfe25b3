"""UMAP SGD epochs on the HIP backend.

``optimize_coo`` mirrors the reference's CUDA COO optimizer entry
(fastembedr_cuda_umap_optimize_coo, src/embedding_cuda_kernels.cpp:5483-5608,
reached from umap_cuda_optimize_csr_cpp / umap_cuda_optimize_coo_impl,
src/embedding_cuda_bridge.cpp:345-378, src/embedding_cuda_impl.cpp:
2048-2107).  ``csr_to_coo`` does the bridge's CSR -> COO expansion with
epochs_per_sample = max_w / max(w, 1e-6) (:3019-3042; CPU
attach_csr_epoch_schedule, src/fast_knn_umap.cpp:1264-1282).  ``FuzzyGraph``
builds the reference's default fuzzy graph (umap_graph_csr_float_cpp,
src/fast_knn_umap.cpp:3995-4048) on the device and can feed the epochs
without a host round trip.
"""
from __future__ import annotations

import numpy as np

import ctypes

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


def _device_ptr(a):
    if hasattr(a, "data_ptr"):  # torch tensor on the GPU
        import torch

        if a.is_cuda:
            torch.cuda.current_stream(a.device).synchronize()
        return a.data_ptr()
    return int(a)


class FuzzyGraph:
    """The default UMAP fuzzy graph of a KNN (rows = points), kept on the GPU.

    ``indices`` / ``distances`` are n x k arrays (host), or with
    ``device=True`` column-major k x n device buffers (torch tensors or raw
    pointers) with ``n`` and ``k`` given.  ``index_offset`` None detects it
    as the reference does.  ``export()`` returns the reference's
    list(offsets, neighbors, weights, epochs_per_sample); ``optimize`` runs
    the epochs from the resident graph."""

    def __init__(self, indices, distances, col_start=0, n_cols=None, edge_budget=None,
                 index_offset=None, device=False, n=None, k=None):
        L = _lib.lib()
        if device:
            n, k_total = int(n), int(k)
            ip, dp = _device_ptr(indices), _device_ptr(distances)
            keep = ()
        else:
            idx = np.asarray(indices)
            dist = np.asarray(distances)
            if idx.ndim != 2 or idx.shape != dist.shape:
                raise ValueError("indices and distances must have the same dimensions")
            n, k_total = idx.shape
            ic = np.ascontiguousarray(idx.astype(np.int32).T)
            # the reference's float conversion of double distances
            dc = np.ascontiguousarray(np.asarray(dist, np.float64).astype(np.float32).T)
            ip, dp = ic.ctypes.data, dc.ctypes.data
            keep = (ic, dc)
        n_cols = k_total - col_start if n_cols is None else int(n_cols)
        edge_budget = n_cols if edge_budget is None else int(edge_budget)
        off = -1 if index_offset is None else int(index_offset)
        h = ctypes.c_void_p()
        check(L.fastembedr_hip_umap_graph_create(ip, dp, n, k_total, int(col_start), n_cols,
                                                 edge_budget, off, 1 if device else 0,
                                                 ctypes.byref(h)), "HIP UMAP fuzzy graph")
        del keep
        self._h = h
        self.n = n
        nnz = ctypes.c_longlong()
        mw = ctypes.c_float()
        check(L.fastembedr_hip_umap_graph_info(h, ctypes.byref(nnz), ctypes.byref(mw)),
              "HIP UMAP fuzzy graph info")
        self.nnz = int(nnz.value)
        self.max_weight = float(mw.value)

    def export(self):
        off = np.zeros(self.n + 1, np.int32)
        m = max(self.nnz, 1)
        nb = np.zeros(m, np.int32)
        w = np.zeros(m, np.float32)
        e = np.zeros(m, np.float32)
        check(_lib.lib().fastembedr_hip_umap_graph_export(self._h, ptr(off, _ip), ptr(nb, _ip),
                                                          ptr(w, _fp), ptr(e, _fp)),
              "HIP UMAP fuzzy graph export")
        k = self.nnz
        return {"offsets": off, "neighbors": nb[:k], "weights": w[:k],
                "epochs_per_sample": e[:k], "max_weight": self.max_weight, "n": self.n,
                "nnz": k}

    def optimize(self, init, n_epochs, negative_sample_rate=5, learning_rate=1.0,
                 a=A_MIN_DIST_001, b=B_MIN_DIST_001, repulsion_strength=1.0, seed=0,
                 optimizer_mode=1):
        """UMAP epochs on the resident graph; returns the float32 n x 2 layout."""
        y = np.ascontiguousarray(init, dtype=np.float32)
        out = np.zeros_like(y)
        rc = _lib.lib().fastembedr_hip_umap_graph_optimize(
            self._h, ptr(y, _fp), int(n_epochs), int(negative_sample_rate), float(learning_rate),
            float(a), float(b), float(repulsion_strength), int(seed) & 0xFFFFFFFF,
            int(optimizer_mode), ptr(out, _fp))
        check(rc, "HIP UMAP optimizer (resident graph)")
        return out

    def optimize_coo_device(self, init, n_epochs, negative_sample_rate=5, learning_rate=1.0,
                            a=A_MIN_DIST_001, b=B_MIN_DIST_001, repulsion_strength=1.0, seed=0,
                            optimizer_mode=1):
        """The same epochs through the general device-COO entry point (no use
        of the graph's symmetry)."""
        ptrs = [ctypes.c_void_p() for _ in range(4)]
        L = _lib.lib()
        check(L.fastembedr_hip_umap_graph_device(self._h, *[ctypes.byref(p) for p in ptrs]),
              "HIP UMAP fuzzy graph device")
        y = np.ascontiguousarray(init, dtype=np.float32)
        out = np.zeros_like(y)
        rc = L.fastembedr_hip_umap_optimize_coo_device(
            ptrs[0], ptrs[1], ptrs[2], ptrs[3], ptr(y, _fp), self.n, self.nnz, int(n_epochs),
            int(negative_sample_rate), float(learning_rate), float(a), float(b),
            float(repulsion_strength), int(seed) & 0xFFFFFFFF, int(optimizer_mode), ptr(out, _fp))
        check(rc, "HIP COO UMAP optimizer (device graph)")
        return out

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().fastembedr_hip_umap_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class UmapJobStruct(ctypes.Structure):
    """Mirror of fastembedr_hip_umap_job (include/fastembedr_hip.h)."""

    _fields_ = [("indices", ctypes.c_void_p), ("distances", ctypes.c_void_p),
                ("n", ctypes.c_int), ("k_total", ctypes.c_int), ("n_neighbors", ctypes.c_int),
                ("index_offset", ctypes.c_int), ("init", ctypes.c_void_p),
                ("n_epochs", ctypes.c_int), ("negative_sample_rate", ctypes.c_int),
                ("learning_rate", ctypes.c_float), ("a", ctypes.c_float), ("b", ctypes.c_float),
                ("repulsion_strength", ctypes.c_float), ("seed", ctypes.c_uint),
                ("optimizer_mode", ctypes.c_int)]


def umap_job_bytes(n, n_neighbors, optimizer_mode=1, setup_peak=True):
    return int(_lib.lib().fastembedr_hip_umap_job_bytes(int(n), int(n_neighbors),
                                                        int(optimizer_mode),
                                                        int(bool(setup_peak))))


class UmapNotAdmitted(Exception):
    """The budget or the slot table cannot take the job now."""


class UmapBatch:
    """Many UMAP jobs advancing together (``fastembedr_hip_umap_batch_*``):
    each job's fuzzy graph and layout stay resident, one set of job-table
    launches per epoch covers every job, and a job in optimizer_mode 1 gives
    the bits of ``FuzzyGraph.optimize`` on its own graph."""

    def __init__(self, max_jobs=64, hbm_budget_bytes=0):
        self.max_jobs = int(max_jobs)
        self._h = ctypes.c_void_p()
        check(_lib.lib().fastembedr_hip_umap_batch_create(self.max_jobs, int(hbm_budget_bytes),
                                                          ctypes.byref(self._h)),
              "UMAP batch create")
        self.n_of = {}

    def add(self, indices, distances, init, n_epochs, n_neighbors=None, negative_sample_rate=5,
            learning_rate=1.0, a=A_MIN_DIST_001, b=B_MIN_DIST_001, repulsion_strength=1.0,
            seed=0, optimizer_mode=1, index_offset=None):
        """Add a job from a host n x k KNN (the graph uses its first
        ``n_neighbors`` columns, all by default); returns its slot."""
        idx = np.asarray(indices)
        dist = np.asarray(distances)
        if idx.ndim != 2 or idx.shape != dist.shape:
            raise ValueError("indices and distances must have the same dimensions")
        n, k_total = idx.shape
        ic = np.asfortranarray(idx, dtype=np.int32)
        dc = np.asfortranarray(np.asarray(dist, np.float64).astype(np.float32))
        y = np.ascontiguousarray(init, dtype=np.float32)
        if y.shape != (n, 2):
            raise ValueError("`init` has the wrong shape.")
        nn = k_total if n_neighbors is None else int(n_neighbors)
        off = -1 if index_offset is None else int(index_offset)
        j = UmapJobStruct(ic.ctypes.data, dc.ctypes.data, n, k_total, nn, off, y.ctypes.data,
                          int(n_epochs), int(negative_sample_rate), float(learning_rate),
                          float(a), float(b), float(repulsion_strength), int(seed) & 0xFFFFFFFF,
                          int(optimizer_mode))
        slot = ctypes.c_int(-1)
        rc = _lib.lib().fastembedr_hip_umap_batch_add(self._h, ctypes.byref(j), 0,
                                                      ctypes.byref(slot))
        if rc == 2:
            raise UmapNotAdmitted(_lib.last_error())
        check(rc, "UMAP batch add")
        self.n_of[slot.value] = n
        return slot.value

    def step(self, n_epochs):
        check(_lib.lib().fastembedr_hip_umap_batch_step(self._h, int(n_epochs)), "UMAP batch step")

    def status(self):
        ep = np.zeros(self.max_jobs, np.int32)
        act = np.zeros(self.max_jobs, np.int32)
        check(_lib.lib().fastembedr_hip_umap_batch_status(self._h, ptr(ep, _ip), ptr(act, _ip)),
              "UMAP batch status")
        return ep, act

    def layout(self, slot):
        out = np.zeros((self.n_of[slot], 2), np.float32)
        check(_lib.lib().fastembedr_hip_umap_batch_layout(self._h, int(slot), ptr(out, _fp)),
              "UMAP batch layout")
        return out

    def remove(self, slot):
        check(_lib.lib().fastembedr_hip_umap_batch_remove(self._h, int(slot)), "UMAP batch remove")
        self.n_of.pop(slot, None)

    def run_to_completion(self, chunk=50):
        while True:
            _, act = self.status()
            if not act.any():
                return
            self.step(chunk)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.lib().fastembedr_hip_umap_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
