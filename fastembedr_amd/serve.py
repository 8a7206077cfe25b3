"""Continuous-batching request loop on one GPU.

Two kinds of request share the device:

* ``embed``: a full t-SNE job (KNN in, layout out).  Admitted into the
  job-table batch (``batch.TsneBatch``) when a slot and the HBM budget allow;
  every engine tick advances all resident jobs by ``tick`` iterations, so a
  job that arrives while others run starts at the next tick and its work
  overlaps theirs inside the same kernels.  Finished jobs are read back and
  removed, which frees their budget for waiting requests.
* ``transform``: placement of query points into a fixed reference embedding
  that stays resident (``transform.TransformReference``).  Each request is one
  fused launch; transforms run between engine ticks.

The loop reports per-request latency (arrival to result) and throughput.  A
request's result does not depend on what else is resident: the batch gives
every job the block count of its single-job launch and all reductions have
fixed orders, so results are bit-identical alone and in the mix.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from .batch import NotAdmitted, TsneBatch, job_bytes
from .transform import TransformReference


@dataclass
class Request:
    rid: int
    kind: str            # "embed" or "transform"
    arrival: float       # seconds on the loop's clock
    payload: dict
    started: float = -1.0
    finished: float = -1.0
    result: np.ndarray | None = None
    slot: int = -1
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
    """Single-GPU continuous-batching loop over a request trace."""

    def __init__(self, reference_layout=None, max_jobs=64, hbm_budget_bytes=0, tick=25):
        self.batch = TsneBatch(max_jobs=max_jobs, hbm_budget_bytes=hbm_budget_bytes)
        self.ref = TransformReference(reference_layout) if reference_layout is not None else None
        self.tick = int(tick)
        self.budget = int(hbm_budget_bytes)

    def close(self):
        self.batch.close()
        if self.ref is not None:
            self.ref.close()

    def run(self, requests, realtime=True):
        """Serve `requests` (sorted by arrival).  With realtime=True a request
        is only visible once the wall clock passes its arrival time."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting_embed = deque()
        waiting_tf = deque()
        running = {}
        stats = ServeStats()
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        while pending or waiting_embed or waiting_tf or running:
            t = now()
            while pending and (not realtime or pending[0].arrival <= t):
                r = pending.popleft()
                (waiting_embed if r.kind == "embed" else waiting_tf).append(r)
            # admit embed jobs while slots and budget allow (FIFO, no overtaking)
            while waiting_embed:
                r = waiting_embed[0]
                p = r.payload
                try:
                    r.slot = self.batch.add(**p)
                except NotAdmitted as exc:
                    if running:
                        break  # retried when a running job frees its slot and budget
                    # nothing holds the device: this job never fits the budget
                    waiting_embed.popleft()
                    r.error = f"not admissible: {exc}"
                    r.finished = now()
                    stats.failed += 1
                    continue
                r.started = now()
                running[r.slot] = r
                waiting_embed.popleft()
            # enqueue the tick (asynchronous), then run the waiting transforms on
            # the reference's own stream so they overlap the tick on the device
            if running:
                self.batch.step(self.tick)
            while waiting_tf:
                r = waiting_tf.popleft()
                r.started = now()
                r.result = self.ref.transform(**r.payload)
                r.finished = now()
                stats.latencies[r.rid] = r.finished - max(r.arrival, 0.0)
                stats.completed += 1
            if running:
                it, act = self.batch.status()
                for slot in [s for s in running if act[s] == 0]:
                    r = running.pop(slot)
                    r.result = self.batch.layout(slot)
                    self.batch.remove(slot)
                    r.finished = now()
                    stats.latencies[r.rid] = r.finished - max(r.arrival, 0.0)
                    stats.completed += 1
            elif pending and realtime:
                time.sleep(max(0.0, min(pending[0].arrival - now(), 0.002)))
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


__all__ = ["Request", "Server", "ServeStats", "poisson_trace", "job_bytes"]
