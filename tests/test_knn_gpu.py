"""Exact k-NN on the GPU against the FP64 brute-force oracle: the same
indices and the same float32 distances, bit for bit, over list lengths,
column counts across the 64-column chunks, duplicates (rows the error bound
cannot settle go to the full scan) and device buffers."""
import os

import numpy as np
import pytest

import oracle
from tests.helpers import DeviceArray, gaussian_mixture

pytestmark = pytest.mark.gpu


def _check(x, k, index_base=0):
    from fastembedr_amd import knn

    print(f"[knn] n={x.shape[0]} p={x.shape[1]} k={k} base={index_base}", flush=True)
    gi, gd = knn.exact_knn(x, k, index_base=index_base)
    oi, od = oracle.exact_knn(x, k)
    assert gi.shape == (k, x.shape[0])
    assert np.array_equal(gi.T - index_base, oi)
    assert np.array_equal(gd.T, od)
    return gi, gd


def test_knn_matches_oracle_across_list_lengths():
    rng = np.random.default_rng(901)
    x = rng.normal(size=(1500, 50)).astype(np.float32)
    for k in (1, 15, 47, 90, 111, 200):  # lists of 64, 128, 256 and 512 keys
        _check(x, k)
    _check(x, 90, index_base=1)


def test_knn_matches_oracle_across_shapes():
    rng = np.random.default_rng(902)
    for n, p, k in ((777, 1, 10), (1000, 3, 30), (650, 13, 7), (900, 64, 20), (700, 65, 25),
                    (500, 130, 12), (257, 9, 256), (2, 4, 1), (65, 2, 64)):
        x = (rng.normal(size=(n, p)) * rng.uniform(0.1, 10.0)).astype(np.float32)
        _check(x, k)


def test_knn_clustered_and_offset_data():
    x, _ = gaussian_mixture(3000, 50, 12, seed=903)
    _check(x.astype(np.float32), 90)
    # far from the origin: the keys are centred first
    y = np.random.default_rng(904).normal(size=(800, 6)).astype(np.float32) + 1000.0
    _check(y, 16)


def test_knn_duplicates_and_full_scan():
    from fastembedr_amd import _lib, knn

    rng = np.random.default_rng(905)
    x = rng.normal(size=(1200, 8)).astype(np.float32)
    x[100:400] = x[100]          # 300 copies: more than any list holds
    x[500:520] = x[7]            # 21 copies
    x[900:] = np.round(x[900:])  # many equal distances
    _check(x, 40)
    import ctypes

    col = np.asfortranarray(x).ravel(order="F")
    idx = np.zeros((40, 1200), np.int32)
    dist = np.zeros((40, 1200), np.float32)
    scanned = ctypes.c_int(-1)
    _lib.check(_lib.lib().fastembedr_hip_exact_knn(
        col.ctypes.data, 1200, 8, 40, 0, 0, idx.ctypes.data, dist.ctypes.data, 0,
        ctypes.byref(scanned)), "exact KNN")
    assert scanned.value >= 300  # the copies of row 100 cannot be settled by the bound
    # every row through the full scan gives the same answer
    os.environ["FASTEMBEDR_HIP_KNN_SCAN_ALL"] = "1"
    try:
        si, sd = knn.exact_knn(x, 40)
    finally:
        del os.environ["FASTEMBEDR_HIP_KNN_SCAN_ALL"]
    assert np.array_equal(si, idx) and np.array_equal(sd, dist)


def test_knn_bound_settles_tie_free_rows():
    import ctypes

    from fastembedr_amd import _lib

    x = np.random.default_rng(906).normal(size=(4000, 50)).astype(np.float32)
    col = np.asfortranarray(x).ravel(order="F")
    idx = np.zeros((90, 4000), np.int32)
    dist = np.zeros((90, 4000), np.float32)
    scanned = ctypes.c_int(-1)
    _lib.check(_lib.lib().fastembedr_hip_exact_knn(
        col.ctypes.data, 4000, 50, 90, 0, 0, idx.ctypes.data, dist.ctypes.data, 0,
        ctypes.byref(scanned)), "exact KNN")
    assert scanned.value == 0
    oi, od = oracle.exact_knn(x, 90, queries=np.arange(0, 4000, 37))
    assert np.array_equal(idx.T[::37], oi) and np.array_equal(dist.T[::37], od)


def test_knn_device_buffers():
    import ctypes

    from fastembedr_amd import _lib

    x = np.random.default_rng(907).normal(size=(1000, 20)).astype(np.float32)
    k = 30
    hi, hd = _check(x, k, index_base=1)
    dx = DeviceArray(np.asfortranarray(x).ravel(order="F"))
    di = DeviceArray(np.zeros(k * 1000, np.int32))
    dd = DeviceArray(np.zeros(k * 1000, np.float32))
    try:
        scanned = ctypes.c_int(-1)
        _lib.check(_lib.lib().fastembedr_hip_exact_knn(
            dx.ptr, 1000, 20, k, 1, 1, di.ptr, dd.ptr, 1, ctypes.byref(scanned)), "exact KNN")
        gi = di.to_host(np.int32, (k, 1000))
        gd = dd.to_host(np.float32, (k, 1000))
    finally:
        dx.free()
        di.free()
        dd.free()
    assert np.array_equal(gi, hi) and np.array_equal(gd, hd)


def test_query_knn_matches_oracle():
    import ctypes

    from fastembedr_amd import _lib, knn

    rng = np.random.default_rng(908)
    for n, m, p, k in ((2000, 700, 50, 90), (500, 333, 3, 15), (300, 40, 70, 300 - 45),
                       (64, 10, 8, 64), (1, 3, 5, 1)):
        data = (rng.normal(size=(n, p)) * 3.0).astype(np.float32)
        query = (rng.normal(size=(m, p)) * 3.0).astype(np.float32)
        print(f"[query knn] n={n} m={m} p={p} k={k}", flush=True)
        gi, gd = knn.exact_query_knn(data, query, k, index_base=1)
        oi, od = oracle.exact_query_knn(data, query, k)
        assert np.array_equal(gi.T - 1, oi) and np.array_equal(gd.T, od)
    # queries equal to data rows (distance 0) and duplicated data rows
    data = rng.normal(size=(900, 6)).astype(np.float32)
    data[200:500] = data[200]
    query = np.concatenate([data[:100], data[200:210], rng.normal(size=(50, 6))]).astype(np.float32)
    gi, gd = knn.exact_query_knn(data, query, 40)
    oi, od = oracle.exact_query_knn(data, query, 40)
    assert np.array_equal(gi.T, oi) and np.array_equal(gd.T, od)
    col = np.asfortranarray(data).ravel(order="F")
    qcol = np.asfortranarray(query).ravel(order="F")
    idx = np.zeros((40, len(query)), np.int32)
    dist = np.zeros((40, len(query)), np.float32)
    scanned = ctypes.c_int(-1)
    _lib.check(_lib.lib().fastembedr_hip_exact_query_knn(
        col.ctypes.data, 900, qcol.ctypes.data, len(query), 6, 40, 0, 0, idx.ctypes.data,
        dist.ctypes.data, 0, ctypes.byref(scanned)), "exact query KNN")
    assert scanned.value >= 10 and np.array_equal(idx, gi) and np.array_equal(dist, gd)
