"""The reference's own C++ for the hot-path functions that need only the
standard library, compiled into ``oracle/_ref/libref.so`` by
``oracle/build_ref.sh`` from line ranges of the reference sources.

TEST INFRASTRUCTURE ONLY: the tests use it to pin the oracle and the CPU
ports, and ``bench.py``'s ``cpu_baseline`` leg times it.  ``lib()`` returns
None when the library was not built (no reference sources on this machine);
callers skip.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libref.so")
_lib = None

_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_u32p = ctypes.POINTER(ctypes.c_uint32)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            return None
        L = ctypes.CDLL(LIB_PATH)
        ci, cf, cd, cu, cll = (ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_uint32,
                               ctypes.c_longlong)
        L.ref_row_probabilities_float.argtypes = [_fp, ci, ci, cd, _fp]
        L.ref_mix_uint32.argtypes = [cu]
        L.ref_mix_uint32.restype = cu
        L.ref_xorshift_stream.argtypes = [cu, ci, ci, _u32p, _ip]
        L.ref_fft_grid_size.argtypes = [ci]
        L.ref_tsne_gradient.argtypes = [ci, _ip, _ip, _fp, _fp, cf, ci, ci, _fp]
        L.ref_tsne_sum_q.argtypes = [_fp, ci, ci]
        L.ref_tsne_sum_q.restype = cd
        L.ref_tsne_kl.argtypes = [ci, _ip, _ip, _fp, _fp, ci]
        L.ref_tsne_kl.restype = cd
        L.ref_tsne_iterate.argtypes = [ci, _ip, _ip, _fp, _fp, _fp, _fp, ci, cf, cf, cf, cf, cf,
                                       ci, ci, _dp]
        L.ref_umap_powf_fast.argtypes = [cf, cf]
        L.ref_umap_powf_fast.restype = cf
        L.ref_umap_pow.argtypes = [cd, cd]
        L.ref_umap_pow.restype = cd
        L.ref_clip4f.argtypes = [cf]
        L.ref_clip4f.restype = cf
        L.ref_mix_seed.argtypes = [cu]
        L.ref_mix_seed.restype = cu
        L.ref_deterministic_vertex.argtypes = [ci, ci, ci, cll, ci]
        L.ref_deterministic_vertex_cuda_style.argtypes = [ci, ci, ci, ci, cll, ci]
        L.ref_positive_samples.argtypes = [cf, ci]
        L.ref_negative_samples.argtypes = [cf, ci, ci]
        L.ref_tau_vertices.argtypes = [ci, ci, cll, ci, ci, ci, _ip]
        L.ref_find_ab.argtypes = [cd, cd, _dp, _dp]
        L.ref_effective_cpu_threads.argtypes = [ci, ci]
        L.ref_umap_hogwild.argtypes = [_ip, _ip, _fp, cll, ci, _fp, ci, ci, cd, cd, cd, cd, ci,
                                       ci, _dp]
        _lib = L
    return _lib


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _p(a, t):
    return a.ctypes.data_as(t)


def row_probabilities_float(dist, perplexity):
    """compute_row_probabilities_float of every row (n x k float32)."""
    d = _c(dist, np.float32)
    out = np.zeros_like(d)
    lib().ref_row_probabilities_float(_p(d, _fp), d.shape[0], d.shape[1], float(perplexity),
                                      _p(out, _fp))
    return out


def xorshift_stream(state, count, n):
    raw = np.zeros(count, np.uint32)
    idx = np.zeros(count, np.int32)
    lib().ref_xorshift_stream(int(state) & 0xFFFFFFFF, count, n, _p(raw, _u32p), _p(idx, _ip))
    return raw, idx


def tsne_gradient(row_ptr, col, val, y, exaggeration=1.0, grid=0, threads=1):
    """The reference CPU optimizer gradient: FFT grid (grid > 0, or the CPU
    rule at 0) or exact pairs (grid < 0) plus the attraction."""
    rp, cl, vl = _c(row_ptr, np.int32), _c(col, np.int32), _c(val, np.float32)
    yy = _c(y, np.float32)
    out = np.zeros_like(yy)
    rc = lib().ref_tsne_gradient(len(rp) - 1, _p(rp, _ip), _p(cl, _ip), _p(vl, _fp), _p(yy, _fp),
                                 float(exaggeration), int(grid), int(threads), _p(out, _fp))
    assert rc == 0
    return out


def tsne_kl(row_ptr, col, val, y, threads=1):
    rp, cl, vl = _c(row_ptr, np.int32), _c(col, np.int32), _c(val, np.float32)
    yy = _c(y, np.float32)
    return lib().ref_tsne_kl(len(rp) - 1, _p(rp, _ip), _p(cl, _ip), _p(vl, _fp), _p(yy, _fp),
                             int(threads))


def tsne_iterate(row_ptr, col, val, y, iters, exaggeration, learning_rate, momentum,
                 min_gain=0.01, max_step_norm=5.0, grid=0, threads=1, update=None, gains=None):
    """`iters` iterations of one phase of the reference CPU optimizer.
    Returns (layout, update, gains, seconds)."""
    rp, cl, vl = _c(row_ptr, np.int32), _c(col, np.int32), _c(val, np.float32)
    yy = _c(y, np.float32).copy()
    up = np.zeros_like(yy) if update is None else _c(update, np.float32).copy()
    ga = np.ones_like(yy) if gains is None else _c(gains, np.float32).copy()
    sec = ctypes.c_double()
    rc = lib().ref_tsne_iterate(len(rp) - 1, _p(rp, _ip), _p(cl, _ip), _p(vl, _fp), _p(yy, _fp),
                                _p(up, _fp), _p(ga, _fp), int(iters), float(exaggeration),
                                float(learning_rate), float(momentum), float(min_gain),
                                float(max_step_norm), int(grid), int(threads), ctypes.byref(sec))
    assert rc == 0
    return yy, up, ga, sec.value


def tau_vertices(seed, epoch, window_end, thread_id, n, count):
    out = np.zeros(count, np.int32)
    lib().ref_tau_vertices(seed, epoch, window_end, thread_id, n, count, _p(out, _ip))
    return out


def find_ab(spread, min_dist):
    a, b = ctypes.c_double(), ctypes.c_double()
    lib().ref_find_ab(float(spread), float(min_dist), ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def umap_hogwild(rows, tails, eps, y, n_epochs, nsr=5, a=1.895605865596314,
                 b=0.8006377738365004, gamma=1.0, lr=1.0, seed=1, threads=1):
    """The reference CPU hogwild epochs (its own edge loop) from the active
    edges.  Returns (layout, seconds)."""
    r, t, e = _c(rows, np.int32), _c(tails, np.int32), _c(eps, np.float32)
    yy = _c(y, np.float32).copy()
    sec = ctypes.c_double()
    rc = lib().ref_umap_hogwild(_p(r, _ip), _p(t, _ip), _p(e, _fp), len(r), yy.shape[0],
                                _p(yy, _fp), int(n_epochs), int(nsr), float(a), float(b),
                                float(gamma), float(lr), int(seed), int(threads),
                                ctypes.byref(sec))
    assert rc == 0
    return yy, sec.value
