"""Batched t-SNE jobs on one GPU (job-table kernels, HBM-budget admission).

``TsneBatch`` wraps the ``fastembedr_hip_tsne_batch_*`` entry points: jobs are
added while the device-memory budget admits them, every ``step`` advances all
resident unfinished jobs by the same number of iterations (each job follows
its own schedule and stops at its own end), and finished jobs are read and
removed by the caller -- the continuous-batching primitive of ``serve.py``.
Each job's layout is bit-identical to the same job run alone through
``tsne_knn`` / ``TsnePlan``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import _fp, _ip, check, ptr
from .tsne import _column_major, _device_inputs_ready, _index_offset, _knn_arrays


class JobStruct(ctypes.Structure):
    """Mirror of fastembedr_hip_tsne_job (include/fastembedr_hip.h)."""

    _fields_ = [("indices", ctypes.c_void_p), ("distances", ctypes.c_void_p),
                ("n", ctypes.c_int), ("k", ctypes.c_int), ("index_offset", ctypes.c_int),
                ("grid_size", ctypes.c_int), ("perplexity", ctypes.c_float),
                ("init", ctypes.c_void_p), ("seed", ctypes.c_uint),
                ("early_exaggeration_iter", ctypes.c_int), ("n_iter", ctypes.c_int),
                ("early_exaggeration", ctypes.c_float), ("exaggeration", ctypes.c_float),
                ("learning_rate_early", ctypes.c_float), ("learning_rate_normal", ctypes.c_float),
                ("initial_momentum", ctypes.c_float), ("final_momentum", ctypes.c_float),
                ("min_gain", ctypes.c_float), ("max_step_norm", ctypes.c_float),
                ("attraction", ctypes.c_int), ("pair_workgroups", ctypes.c_int)]


def job_bytes(n, k, grid_size=0, setup_peak=True):
    return int(_lib.lib().fastembedr_hip_tsne_job_bytes(int(n), int(k), int(grid_size),
                                                        int(bool(setup_peak))))


class NotAdmitted(Exception):
    """The budget or the slot table cannot take the job now."""


class TsneBatch:
    def __init__(self, max_jobs=256, hbm_budget_bytes=0):
        self.max_jobs = int(max_jobs)
        self._h = ctypes.c_void_p()
        check(_lib.lib().fastembedr_hip_tsne_batch_create(self.max_jobs, int(hbm_budget_bytes),
                                                          ctypes.byref(self._h)), "batch create")
        self.n_of = {}

    def add(self, indices, distances, perplexity=30.0, Y_init=None, seed=4, grid_size=0,
            early_exaggeration_iter=250, n_iter=500, early_exaggeration=12.0, exaggeration=1.0,
            learning_rate=None, initial_momentum=0.8, final_momentum=0.8, min_gain=0.01,
            max_step_norm=5.0, device=False, n=None, k=None, index_offset=None, attraction=0,
            pair_workgroups=0):
        """Add a job; returns its slot.  ``indices``/``distances`` are host
        n x k arrays or, with ``device=True``, objects exposing ``data_ptr()``
        to column-major device memory (a k x n torch tensor), as for
        ``TsnePlan``.  The learning rate defaults to n / early_exaggeration for
        both phases (the reference's auto_config policy).  ``attraction`` and
        ``pair_workgroups`` are the launch choices of ``TsnePlan`` (0: built
        in; the autotuner's table gives them per size bucket)."""
        if device:
            n, k = int(n), int(k)
            off = 0 if index_offset is None else int(index_offset)
            _device_inputs_ready(indices, distances)
            ip, dp = int(indices.data_ptr()), int(distances.data_ptr())
        else:
            idx, dist = _knn_arrays(indices, distances)
            n, k = idx.shape
            ci, cd = _column_major(idx, dist, np.float32)
            off = _index_offset(idx) if index_offset is None else int(index_offset)
            ip, dp = ci.ctypes.data, cd.ctypes.data
        init = None if Y_init is None else np.ascontiguousarray(Y_init, dtype=np.float32)
        if init is not None and init.shape != (n, 2):
            raise ValueError("`Y_init` has the wrong shape.")
        lr = n / early_exaggeration if learning_rate is None else float(learning_rate)
        j = JobStruct(ip, dp, n, k, off, int(grid_size),
                      float(perplexity), init.ctypes.data if init is not None else None,
                      seed & 0xFFFFFFFF, int(early_exaggeration_iter), int(n_iter),
                      float(early_exaggeration), float(exaggeration), lr, lr,
                      float(initial_momentum), float(final_momentum), float(min_gain),
                      float(max_step_norm), int(attraction), int(pair_workgroups))
        slot = ctypes.c_int(-1)
        rc = _lib.lib().fastembedr_hip_tsne_batch_add(self._h, ctypes.byref(j), int(bool(device)),
                                                      ctypes.byref(slot))
        if rc == 2:
            raise NotAdmitted(_lib.last_error())
        check(rc, "batch add")
        self.n_of[slot.value] = n
        return slot.value

    def step(self, n_iter):
        check(_lib.lib().fastembedr_hip_tsne_batch_step(self._h, int(n_iter)), "batch step")

    def status(self):
        it = np.zeros(self.max_jobs, np.int32)
        act = np.zeros(self.max_jobs, np.int32)
        check(_lib.lib().fastembedr_hip_tsne_batch_status(self._h, ptr(it, _ip), ptr(act, _ip)),
              "batch status")
        return it, act

    def layout(self, slot):
        out = np.zeros((self.n_of[slot], 2), np.float32)
        check(_lib.lib().fastembedr_hip_tsne_batch_layout(self._h, int(slot), ptr(out, _fp)),
              "batch layout")
        return out

    def remove(self, slot):
        check(_lib.lib().fastembedr_hip_tsne_batch_remove(self._h, int(slot)), "batch remove")
        self.n_of.pop(slot, None)

    def run_to_completion(self, chunk=50):
        while True:
            it, act = self.status()
            if not act.any():
                return
            self.step(chunk)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.lib().fastembedr_hip_tsne_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
