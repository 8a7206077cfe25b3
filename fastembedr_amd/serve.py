"""Continuous-batching serving on one GPU, through the native session.

``Session`` wraps the ``fastembedr_hip_session_*`` C-ABI: a session keeps a
t-SNE job batch, a fixed reference layout and the transforms' workspace pools
resident on the device, and one engine thread inside the library serves the
requests:

* ``embed``: a full t-SNE job (KNN in, layout out).  It joins the job-table
  batch when a slot and the HBM budget allow; every engine tick advances all
  resident jobs by 25 iterations, so a job that arrives while others run
  starts at the next tick and its work shares their kernels.  Finished jobs
  are read back and removed, which frees their budget for waiting ones.
* ``umap``: a full UMAP job; its fuzzy graph joins the session's UMAP batch,
  which advances every resident job by 25 epochs per tick beside the t-SNE
  ticks (the two batches share the budget).
* ``transform``: placement of query points into the resident reference
  layout, enqueued on its own stream with workspace from the session's pool
  (its share of the budget), so it overlaps the tick on the device and never
  blocks the loop.

A request's result does not depend on what else is resident: the batch gives
every job the bits of its single-job plan and a transform is one
deterministic launch.  ``Server`` replays a request trace (Poisson arrivals)
against a session and reports latency from arrival to result.
"""
from __future__ import annotations

import ctypes
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import HipError, _fp, _ip, check, ptr
from .batch import JobStruct, job_bytes
from .transform import _INIT, _controls, _knn, _offset
from .tsne import _index_offset, _knn_arrays

QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3


class Session:
    """A native serving session (see the module docstring)."""

    def __init__(self, reference_layout=None, max_jobs=64, hbm_budget_bytes=0,
                 transform_pool_bytes=0):
        ref = None
        n_ref = 0
        if reference_layout is not None:
            ref = np.ascontiguousarray(reference_layout, dtype=np.float32)
            if ref.ndim != 2 or ref.shape[1] != 2 or len(ref) < 1:
                raise ValueError("`reference_layout` must be a non-empty n x 2 matrix.")
            n_ref = len(ref)
        self.n_ref = n_ref
        self._h = ctypes.c_void_p()
        check(_lib.lib().fastembedr_hip_session_create(
            ptr(ref, _fp) if ref is not None else None, n_ref, int(max_jobs),
            int(hbm_budget_bytes), int(transform_pool_bytes), ctypes.byref(self._h)),
            "session create")
        self._rows = {}

    def submit_embed(self, indices, distances, perplexity=30.0, Y_init=None, seed=4, grid_size=0,
                     early_exaggeration_iter=250, n_iter=500, early_exaggeration=12.0,
                     exaggeration=1.0, learning_rate=None, initial_momentum=0.8,
                     final_momentum=0.8, min_gain=0.01, max_step_norm=5.0, index_offset=None,
                     attraction=0, pair_workgroups=0):
        """Queue a t-SNE job (host n x k KNN; a Fortran-ordered array is used
        as is); returns the request id.  The learning rate defaults to
        n / early_exaggeration for both phases (the reference's auto_config
        policy)."""
        idx, dist = _knn_arrays(indices, distances)
        n, k = idx.shape
        ci = np.asfortranarray(idx, dtype=np.int32)
        cd = np.asfortranarray(dist, dtype=np.float32)
        off = _index_offset(idx) if index_offset is None else int(index_offset)
        init = None if Y_init is None else np.ascontiguousarray(Y_init, dtype=np.float32)
        if init is not None and init.shape != (n, 2):
            raise ValueError("`Y_init` has the wrong shape.")
        lr = n / early_exaggeration if learning_rate is None else float(learning_rate)
        j = JobStruct(ci.ctypes.data, cd.ctypes.data, n, k, off, int(grid_size), float(perplexity),
                      init.ctypes.data if init is not None else None, seed & 0xFFFFFFFF,
                      int(early_exaggeration_iter), int(n_iter), float(early_exaggeration),
                      float(exaggeration), lr, lr, float(initial_momentum), float(final_momentum),
                      float(min_gain), float(max_step_norm), int(attraction), int(pair_workgroups))
        rid = ctypes.c_longlong(0)
        check(_lib.lib().fastembedr_hip_session_submit_embed(self._h, ctypes.byref(j),
                                                             ctypes.byref(rid)), "session submit")
        self._rows[rid.value] = n
        return rid.value

    def submit_transform(self, indices, distances, perplexity=5.0, initialization="median",
                         Y_init=None, n_iter=250, early_exaggeration_iter=0, learning_rate=0.1,
                         early_exaggeration=4.0, exaggeration=1.5, initial_momentum=0.8,
                         final_momentum=0.8, max_grad_norm=0.25, max_step_norm=math.inf,
                         n_negatives=None, exact_repulsion_threshold=4096, seed=4):
        """Queue a transform against the session's reference (the arguments
        and defaults of ``TransformReference.transform``); returns the id."""
        if initialization not in _INIT:
            raise ValueError("`initialization` must be 'median', 'weighted', or 'random'.")
        idx, dist = _knn(None, indices, distances)
        nq, k = idx.shape
        nneg = _controls(self.n_ref, n_negatives, exact_repulsion_threshold)
        init = None if Y_init is None else np.ascontiguousarray(Y_init, dtype=np.float32)
        ci = np.asfortranarray(idx, dtype=np.int32)
        cd = np.asfortranarray(dist, dtype=np.float32)
        rid = ctypes.c_longlong(0)
        check(_lib.lib().fastembedr_hip_session_submit_transform(
            self._h, ptr(ci, _ip), ptr(cd, _fp), nq, k, _offset(idx, self.n_ref),
            ptr(init, _fp) if init is not None else None, _INIT[initialization],
            float(perplexity), int(n_iter), int(early_exaggeration_iter), float(learning_rate),
            float(early_exaggeration), float(exaggeration), float(initial_momentum),
            float(final_momentum), float(max_grad_norm), float(max_step_norm), nneg,
            int(exact_repulsion_threshold), int(seed), ctypes.byref(rid)), "session submit")
        self._rows[rid.value] = nq
        return rid.value

    def submit_umap(self, indices, distances, init, n_epochs, n_neighbors=None,
                    negative_sample_rate=5, learning_rate=1.0, a=None, b=None,
                    repulsion_strength=1.0, seed=0, optimizer_mode=1, index_offset=None):
        """Queue a UMAP job (host n x k KNN, its first ``n_neighbors``
        columns; the arguments of ``UmapBatch.add``); returns the id."""
        from .umap import A_MIN_DIST_001, B_MIN_DIST_001, UmapJobStruct

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
        j = UmapJobStruct(ic.ctypes.data, dc.ctypes.data, n, k_total,
                          k_total if n_neighbors is None else int(n_neighbors),
                          -1 if index_offset is None else int(index_offset), y.ctypes.data,
                          int(n_epochs), int(negative_sample_rate), float(learning_rate),
                          float(A_MIN_DIST_001 if a is None else a),
                          float(B_MIN_DIST_001 if b is None else b), float(repulsion_strength),
                          int(seed) & 0xFFFFFFFF, int(optimizer_mode))
        rid = ctypes.c_longlong(0)
        check(_lib.lib().fastembedr_hip_session_submit_umap(self._h, ctypes.byref(j),
                                                            ctypes.byref(rid)), "session submit")
        self._rows[rid.value] = n
        return rid.value

    def poll(self, rid, wait_ms=0):
        """(state, layout or None, latency_ms, error or None).  A done or
        failed request is released by this call."""
        state = ctypes.c_int(-1)
        lat = ctypes.c_double(-1.0)
        out = np.empty((self._rows.get(rid, 0), 2), np.float32)
        check(_lib.lib().fastembedr_hip_session_poll(self._h, int(rid), int(wait_ms),
                                                     ctypes.byref(state), ptr(out, _fp),
                                                     ctypes.byref(lat)), "session poll")
        s = state.value
        if s == DONE:
            self._rows.pop(rid, None)
            return s, out, lat.value, None
        if s == FAILED:
            self._rows.pop(rid, None)
            return s, None, lat.value, _lib.last_error()
        return s, None, None, None

    def wait(self, rid, timeout_s=600.0):
        """Blocks until the request ends; returns its layout or raises."""
        t_end = time.perf_counter() + timeout_s
        while True:
            s, out, _, err = self.poll(rid, wait_ms=100)
            if s == DONE:
                return out
            if s == FAILED:
                raise HipError(f"request {rid} failed: {err}")
            if time.perf_counter() > t_end:
                raise TimeoutError(f"request {rid} still in state {s}")

    def stats(self):
        s = np.zeros(12, np.int64)
        check(_lib.lib().fastembedr_hip_session_stats(self._h, s.ctypes.data_as(_lib._llp)),
              "session stats")
        keys = ("submitted", "completed", "failed", "queued_embed", "queued_transform",
                "running_embed", "running_transform", "pool_bytes_used", "ticks", "pool_bytes",
                "queued_umap", "running_umap")
        return dict(zip(keys, (int(v) for v in s)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.lib().fastembedr_hip_session_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Request:
    rid: int
    kind: str            # "embed", "umap" or "transform"
    arrival: float       # seconds on the loop's clock
    payload: dict
    started: float = -1.0
    finished: float = -1.0
    result: np.ndarray | None = None
    session_latency_ms: float | None = None
    error: str | None = None


@dataclass
class ServeStats:
    latencies: dict = field(default_factory=dict)
    completed: int = 0
    failed: int = 0
    wall: float = 0.0

    def summary(self):
        lat = np.array(sorted(self.latencies.values())) if self.latencies else np.zeros(1)
        return {"completed": self.completed, "failed": self.failed, "wall_s": round(self.wall, 4),
                "jobs_per_s": round(self.completed / max(self.wall, 1e-9), 3),
                "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
                "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3)}


class Server:
    """Replays a request trace against a native ``Session``."""

    def __init__(self, reference_layout=None, max_jobs=64, hbm_budget_bytes=0,
                 transform_pool_bytes=0):
        self.session = Session(reference_layout, max_jobs=max_jobs,
                               hbm_budget_bytes=hbm_budget_bytes,
                               transform_pool_bytes=transform_pool_bytes)

    def close(self):
        self.session.close()

    def _submit(self, r):
        if r.kind == "embed":
            return self.session.submit_embed(**r.payload)
        if r.kind == "umap":
            return self.session.submit_umap(**r.payload)
        return self.session.submit_transform(**r.payload)

    def run(self, requests, realtime=True):
        """Submit each request at its arrival time (all at once with
        realtime=False) and collect the results.  Latency is arrival (or
        submission, when later) to the moment the loop sees the result."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        outstanding = {}
        stats = ServeStats()
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        while pending or outstanding:
            while pending and (not realtime or pending[0].arrival <= now()):
                r = pending.popleft()
                r.started = now()
                try:
                    outstanding[self._submit(r)] = r
                except HipError as exc:
                    r.error = str(exc)
                    r.finished = now()
                    stats.failed += 1
            seen = False
            for sid in list(outstanding):
                state, out, lat, err = self.session.poll(sid)
                if state not in (DONE, FAILED):
                    continue
                seen = True
                r = outstanding.pop(sid)
                r.finished = now()
                r.session_latency_ms = lat
                if state == DONE:
                    r.result = out
                    stats.latencies[r.rid] = r.finished - (max(r.arrival, 0.0) if realtime
                                                           else r.started)
                    stats.completed += 1
                else:
                    r.error = err
                    stats.failed += 1
            if not seen:
                nxt = pending[0].arrival - now() if (pending and realtime) else 0.0005
                time.sleep(max(0.0, min(nxt, 0.0005)))
        stats.wall = now()
        return stats


def poisson_trace(rate_per_s, n_requests, seed, make_embed, make_transform, embed_fraction=0.2):
    """Poisson arrivals; each request is an embed job with probability
    `embed_fraction`, else a transform; the factories build the payloads."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n_requests):
        t += rng.exponential(1.0 / rate_per_s)
        if rng.random() < embed_fraction:
            out.append(Request(rid, "embed", t, make_embed(rng)))
        else:
            out.append(Request(rid, "transform", t, make_transform(rng)))
    return out


__all__ = ["Request", "Server", "ServeStats", "Session", "poisson_trace", "job_bytes"]
