"""Pins the UMAP and transform parts of the oracle (CPU only) with independent
Python restatements, closed forms and invariants."""
import struct

import numpy as np
import pytest

import oracle
from tests.helpers import exact_knn

M32 = 0xFFFFFFFF


def py_mix_uint(x):
    # embedding_cuda_kernels.cpp:253-260 with Python integers
    x ^= x >> 16
    x = (x * 0x7FEB352D) & M32
    x ^= x >> 15
    x = (x * 0x846CA68B) & M32
    x ^= x >> 16
    return x


def py_negative_vertex(n, seed, epoch, head, tail, s):
    x = seed & M32
    x ^= (epoch * 0x9E3779B9) & M32
    x ^= ((head + 1) * 0x85EBCA6B) & M32
    x ^= ((tail + 1) * 0xC2B2AE35) & M32
    x ^= ((s + 1) * 0x27D4EB2D) & M32
    return py_mix_uint(x) % n


def test_hash_and_negative_vertex_match_integer_restatement():
    rng = np.random.default_rng(1)
    L = oracle.lib()
    for _ in range(200):
        x = int(rng.integers(0, 2 ** 32))
        assert L.or_umap_mix_uint(x) == py_mix_uint(x)
        n = int(rng.integers(2, 10 ** 6))
        args = [int(v) for v in rng.integers(0, 2 ** 20, size=4)]
        seed = int(rng.integers(0, 2 ** 32))
        assert L.or_umap_negative_vertex(n, seed, *args) == py_negative_vertex(n, seed, *args)


def test_positive_schedule_telescopes():
    # sum over epochs 1..E of n_pos = floor(E / P) in float32 division
    for period in np.float32([1.0, 1.5, 2.0, 3.3333333, 7.25, 50.0, 499.9]):
        total = sum(oracle.lib().or_umap_n_positive(e, float(period)) for e in range(0, 501))
        assert total == int(np.floor(np.float32(500) / period))
        assert oracle.lib().or_umap_n_positive(0, float(period)) == 0
    assert oracle.lib().or_umap_n_positive(5, 0.0) == 0


def test_negative_schedule_matches_float32_restatement():
    def n_neg(epoch, period, nsr):
        f = np.float32
        if epoch == 0:
            return 0
        now, prev = f(epoch), f(epoch - 1)
        cur, pre = int(np.floor(now / period)), int(np.floor(prev / period))
        if cur <= pre:
            return 0
        npd = f(period / f(nsr))
        ct = int(np.floor((now - npd) / npd))
        pt = 0
        if cur > 1:
            pa = np.ceil(f(cur - 1) * period)
            pt = int(np.floor((pa - npd) / npd))
        return max(ct - pt, 0)

    for period in np.float32([1.0, 1.7, 2.5, 9.0, 33.3]):
        for e in range(0, 120):
            assert oracle.lib().or_umap_n_negative(e, float(period), 5) == n_neg(e, period, 5)
    # every-epoch edge, nsr 5: floor((1 - 0.2) / 0.2) = 4 at the first active
    # epoch, then 5 per epoch
    assert [oracle.lib().or_umap_n_negative(e, 1.0, 5) for e in range(1, 6)] == [4, 5, 5, 5, 5]


def test_fast_pow_restatement():
    def py_fast_pow(x, b):
        if x <= 0:
            return 0.0
        xb = struct.unpack("<I", struct.pack("<f", np.float32(x)))[0]
        bias = np.float32(1064866805.0)
        whole = int(b)
        frac = np.float32(np.float32(b) - np.float32(whole))
        ib = int(np.float32(np.float32(frac * np.float32(np.float32(xb) - bias)) + bias))
        fp = struct.unpack("<f", struct.pack("<I", ib))[0]
        return np.float32(np.float32(x) ** whole) * np.float32(fp)

    for x in (0.01, 0.5, 1.0, 2.0, 17.5, 1e4):
        for b in (0.8006377738365004, 1.0, 1.5):
            got = oracle.lib().or_umap_fast_pow(x, b)
            assert abs(got - py_fast_pow(x, b)) <= 1e-6 * abs(got)
            # within 7% of the true power (Schraudolph-style approximation)
            assert abs(got - x ** b) <= 0.07 * x ** b


def _graph(n, k, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 5))
    idx, dist = exact_knn(x, k)
    heads = np.repeat(np.arange(n), k).astype(np.int32)
    tails = idx.ravel().astype(np.int32)
    w = np.exp(-(dist - dist[:, :1]) / (dist.std() + 1e-6)).ravel().astype(np.float32)
    per = (w.max() / np.maximum(w, 1e-6)).astype(np.float32)
    return heads, tails, per, x


def test_umap_delta_epoch_is_order_free_and_attracts():
    heads, tails, per, x = _graph(200, 8, 3)
    y0 = np.random.default_rng(4).uniform(-10, 10, size=(200, 2)).astype(np.float32)
    a = oracle.umap_epoch(heads, tails, per, y0, 3, 50, mode=1)
    perm = np.random.default_rng(5).permutation(len(heads))
    b = oracle.umap_epoch(heads[perm], tails[perm], per[perm], y0, 3, 50, mode=1)
    assert np.abs(a - b).max() < 1e-5
    assert np.all(np.abs(a) <= 16)
    # with no negatives, one epoch pulls connected pairs together
    c = oracle.umap_epoch(heads, tails, per, y0, 1, 50, nsr=0, mode=1)
    d0 = np.linalg.norm(y0[heads] - y0[tails], axis=1).mean()
    d1 = np.linalg.norm(c[heads] - c[tails], axis=1).mean()
    assert d1 < d0


def test_transform_rng_and_sampled_indices():
    L = oracle.lib()
    for v in (0, 1, 12345, 2 ** 31):
        got = L.or_mix_uint32(v)
        exp = py_mix_uint(v)
        assert got == (exp if exp != 0 else 0x6D2B79F5)
    import ctypes
    s = ctypes.c_uint32(0x12345678)
    t = 0x12345678
    for _ in range(50):
        idx = L.or_uniform_index(ctypes.byref(s), 1000)
        t ^= (t << 13) & M32
        t ^= t >> 17
        t ^= (t << 5) & M32
        assert idx == (t * 1000) >> 32


def test_transform_exact_small_problem_converges_near_neighbours():
    rng = np.random.default_rng(8)
    ref = rng.normal(size=(400, 2)) * 3
    q = ref[:30] + rng.normal(scale=0.05, size=(30, 2))
    d = np.sqrt(((q[:, None] - ref[None]) ** 2).sum(-1))
    idx = np.argsort(d, 1)[:, :15]
    dist = np.take_along_axis(d, idx, 1)
    out = oracle.transform(ref, idx, dist)
    # each query lands near its own source point
    assert np.median(np.linalg.norm(out - ref[:30], axis=1)) < 1.0
    # median start is the coordinate-wise median of the neighbours
    one = oracle.transform(ref, idx, dist, n_iter=1, learning_rate=1e-12)
    med = np.median(ref[idx], axis=1)
    assert np.abs(one - med).max() < 1e-9
    wgt = oracle.transform(ref, idx, dist, initialization="weighted", n_iter=1,
                           learning_rate=1e-12)
    w = 1.0 / (dist + 1e-6)
    exp = (w[:, :, None] * ref[idx]).sum(1) / w.sum(1)[:, None]
    assert np.abs(wgt - exp).max() < 1e-9


def test_transform_rejects_bad_knn():
    ref = np.zeros((10, 2))
    with pytest.raises(ValueError):
        oracle.transform(ref, np.array([[11]]), np.array([[1.0]]))


def _dense_fuzzy(idx, dist, sig, rho, offset=0):
    """Dense restatement of the reference's memberships and union
    (fast_knn_umap.cpp:1608-1634, 1702-1768) from the oracle's sigmas/rhos,
    float32 arithmetic, exp of the float argument correctly rounded via
    float64."""
    n = idx.shape[0]
    D = np.zeros((n, n), np.float32)
    d32 = np.asarray(dist, np.float64).astype(np.float32)
    for i in range(n):
        for j in range(idx.shape[1]):
            nb = int(idx[i, j]) - offset
            d = d32[i, j]
            if nb < 0 or nb >= n or nb == i or not np.isfinite(d):
                continue
            if d <= rho[i]:
                w = np.float32(1.0)
            else:
                x = np.float32(-(d - rho[i])) / sig[i]
                w = np.float32(np.exp(np.float64(np.float32(x))))
            if w > 0:
                D[i, nb] = max(D[i, nb], w)
    s = (D + D.T).astype(np.float32)
    p = (D * D.T).astype(np.float32)
    return (s - p).astype(np.float32)


def test_fuzzy_graph_oracle_pinned():
    rng = np.random.default_rng(41)
    x = np.vstack([rng.normal(size=(120, 6)), rng.normal(loc=4.0, size=(80, 6))])
    idx, dist = exact_knn(x, 15)
    g = oracle.umap_fuzzy_graph(idx, dist, index_offset=0)
    n, k = idx.shape
    sig, rho = g["sigmas"], g["rhos"]
    # smooth_knn_dist: rho is the smallest positive distance and the search
    # meets sum_j exp(-max(d - rho, 0)/sigma) = log2(k) to its tolerance
    d32 = dist.astype(np.float32)
    assert np.array_equal(rho, d32.min(1))
    psum = np.exp(-np.maximum(d32.astype(np.float64) - rho[:, None], 0) / sig[:, None]).sum(1)
    assert np.abs(psum - np.log2(k)).max() < 1e-4
    # the union against a dense restatement
    W = _dense_fuzzy(idx, dist, sig, rho)
    assert np.array_equal(W, W.T)
    rows, cols = np.nonzero(W > np.float32(1e-6))
    assert np.array_equal(np.diff(g["offsets"]), np.bincount(rows, minlength=n))
    assert np.array_equal(g["neighbors"], cols.astype(np.int32))
    ref_w = W[rows, cols]
    assert np.abs(g["weights"] - ref_w).max() <= 2 * np.spacing(np.float32(1.0))
    # epochs_per_sample = max_w / max(w, 1e-6) (fast_knn_umap.cpp:1264-1282)
    mw = np.float32(g["weights"].max())
    assert g["max_weight"] == mw
    assert np.array_equal(g["epochs_per_sample"],
                          (mw / np.maximum(g["weights"], np.float32(1e-6))).astype(np.float32))


def test_fuzzy_graph_oracle_known_answers():
    # k = 2: the target log2(2) = 1 is met by the nearest neighbour alone, so
    # every directed edge but the nearest vanishes; the union of a mutual
    # nearest pair is 1 + 1 - 1 = 1
    x = np.array([[0.0], [1.0], [3.0], [10.0]])
    idx, dist = exact_knn(x, 2)
    g = oracle.umap_fuzzy_graph(idx, dist, index_offset=0)
    W = np.zeros((4, 4), np.float32)
    for i in range(4):
        for p in range(g["offsets"][i], g["offsets"][i + 1]):
            W[i, g["neighbors"][p]] = g["weights"][p]
    # nearest: 0->1, 1->0, 2->1, 3->2
    assert W[0, 1] == 1.0 and W[1, 0] == 1.0 and W[1, 2] == 1.0 and W[2, 3] == 1.0
    assert (W == W.T).all()
    assert W[0, 2] == 0.0 and W[0, 3] == 0.0
    # 1-based indices are detected like the reference (all in 1..n)
    g1 = oracle.umap_fuzzy_graph(idx + 1, dist)
    assert np.array_equal(g1["neighbors"], g["neighbors"])
    # duplicate points: rho falls back to 0 for an all-zero row and sigma to
    # its floor 1e-3 * global mean
    y = np.array([[0.0], [0.0], [0.0], [5.0]])
    idy, dsy = exact_knn(y, 2)
    gd = oracle.umap_fuzzy_graph(idy, dsy, index_offset=0)
    assert gd["rhos"][0] == 0.0
    gmean = dsy[np.isfinite(dsy)].astype(np.float32).astype(np.float64).mean()
    assert abs(gd["sigmas"][0] - 1e-3 * gmean) / (1e-3 * gmean) < 1e-6 or gd["sigmas"][0] >= 1e-3 * gmean
