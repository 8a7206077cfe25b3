"""UMAP epoch and fixed-reference transform on the GPU against the oracle."""
import numpy as np
import pytest

import oracle
from tests.helpers import exact_knn, gaussian_mixture, relative_l2

pytestmark = pytest.mark.gpu


def _graph(n, k, seed, d=5):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, d))
    idx, dist = exact_knn(x, k)
    heads = np.repeat(np.arange(n), k).astype(np.int32)
    tails = idx.ravel().astype(np.int32)
    w = np.exp(-(dist - dist[:, :1]) / (dist.std() + 1e-6)).ravel().astype(np.float32)
    per = (w.max() / np.maximum(w, 1e-6)).astype(np.float32)
    return heads, tails, w, per, x


def test_umap_schedule_is_bit_exact():
    from fastembedr_amd import umap

    heads, tails, w, per, _ = _graph(500, 10, 11)
    for epoch in (0, 1, 2, 7, 50, 199):
        g = umap.schedule(heads, tails, per, 500, epoch, 5, seed=1234, max_neg=12)
        o = oracle.umap_schedule(heads, tails, per, 500, epoch, 5, 1234, 12)
        for a, b in zip(g, o):
            assert np.array_equal(a, b)


def test_umap_schedule_negative_ids_at_large_n():
    # the negative vertex is the hash modulo n, taken on the device by a
    # multiplication (an exact remainder for every 32-bit hash): the ids must
    # equal the oracle's % for vertex counts up to 2^31 - 1
    from fastembedr_amd import umap

    rng = np.random.default_rng(15)
    for n in (7, 65537, 1_000_003, 2_097_152, 2**31 - 1):
        m = 2000
        heads = rng.integers(0, min(n, 2**31 - 1), size=m).astype(np.int32)
        tails = ((heads.astype(np.int64) + 1) % n).astype(np.int32)
        per = rng.uniform(1.0, 3.0, size=m).astype(np.float32)
        for epoch in (1, 3, 77):
            g = umap.schedule(heads, tails, per, n, epoch, 5, seed=99, max_neg=12)
            o = oracle.umap_schedule(heads, tails, per, n, epoch, 5, 99, 12)
            for a, b in zip(g, o):
                assert np.array_equal(a, b), (n, epoch)


def test_umap_delta_mode_matches_oracle_and_is_deterministic():
    from fastembedr_amd import umap

    heads, tails, w, per, _ = _graph(800, 10, 12)
    y0 = np.random.default_rng(13).uniform(-10, 10, size=(800, 2)).astype(np.float32)
    for epochs, tol in ((1, 1e-6), (5, 1e-5)):
        g = umap.optimize_coo(heads, tails, w, per, y0, epochs, seed=7, optimizer_mode=1)
        o = oracle.umap_optimize(heads, tails, per, y0, epochs, seed=7, mode=1)
        assert relative_l2(g, o) < tol
    a = umap.optimize_coo(heads, tails, w, per, y0, 40, seed=7, optimizer_mode=1)
    b = umap.optimize_coo(heads, tails, w, per, y0, 40, seed=7, optimizer_mode=1)
    assert np.array_equal(a, b)


def test_umap_delta_mode_independent_of_edge_order():
    # the delta is an integer sum of per-edge updates: any edge order (heads
    # sorted or not, so any grouping of the head sums) gives the same bits
    from fastembedr_amd import umap

    heads, tails, w, per, _ = _graph(900, 12, 14)
    y0 = np.random.default_rng(15).uniform(-10, 10, size=(900, 2)).astype(np.float32)
    a = umap.optimize_coo(heads, tails, w, per, y0, 25, seed=9, optimizer_mode=1)
    perm = np.random.default_rng(16).permutation(len(heads))
    b = umap.optimize_coo(heads[perm], tails[perm], w[perm], per[perm], y0, 25, seed=9,
                          optimizer_mode=1)
    assert np.array_equal(a, b)


def test_umap_end_to_end_quality():
    from fastembedr_amd import umap

    x, _ = gaussian_mixture(2000, 20, 8, seed=31)
    idx, dist = exact_knn(x, 15)
    heads = np.repeat(np.arange(2000), 15).astype(np.int32)
    tails = idx.ravel().astype(np.int32)
    sig = dist.mean(1, keepdims=True) + 1e-6
    w = np.exp(-(dist - dist[:, :1]) / sig).ravel().astype(np.float32)
    per = (w.max() / np.maximum(w, 1e-6)).astype(np.float32)
    y0 = np.random.default_rng(32).uniform(-10, 10, size=(2000, 2)).astype(np.float32)
    g1 = umap.optimize_coo(heads, tails, w, per, y0, 200, seed=3, optimizer_mode=1)
    g0 = umap.optimize_coo(heads, tails, w, per, y0, 200, seed=3, optimizer_mode=0)
    o = oracle.umap_optimize(heads, tails, per, y0, 200, seed=3, mode=1)
    mo = oracle.structure_metrics(x, o, 15)
    # 200 SGD epochs from different float rounding end in different local
    # layouts; trustworthiness agrees to 0.01, kNN preservation (about 0.2 on
    # this graph) to 0.05
    for g in (g1, g0):
        assert np.all(np.isfinite(g)) and np.abs(g).max() <= 16
        mg = oracle.structure_metrics(x, g, 15)
        assert abs(mg[0] - mo[0]) <= 0.01 and abs(mg[2] - mo[2]) <= 0.05, (mg, mo)


def _ref_case(n_ref, n_q, k, seed):
    rng = np.random.default_rng(seed)
    ref = (rng.normal(size=(n_ref, 2)) * 5).astype(np.float32)
    q = ref[rng.choice(n_ref, n_q)] + rng.normal(scale=0.2, size=(n_q, 2))
    d = np.sqrt(((q[:, None] - ref[None].astype(np.float64)) ** 2).sum(-1))
    idx = np.argsort(d, 1)[:, :k].astype(np.int32)
    dist = np.take_along_axis(d, idx, 1)
    return ref, idx, dist


@pytest.mark.parametrize("n_ref", [600, 20000])
def test_transform_matches_fp64_oracle(n_ref):
    from fastembedr_amd.transform import transform_tsne

    ref, idx, dist = _ref_case(n_ref, 64, 30, seed=n_ref)
    for init in ("median", "weighted"):
        g = transform_tsne(ref, indices=idx, distances=dist, initialization=init)
        o = oracle.transform(ref.astype(np.float64), idx, dist, initialization=init)
        assert relative_l2(g, o) < 1e-4, init


def test_transform_resident_reference_and_random_init():
    from fastembedr_amd.transform import TransformReference, transform_tsne

    ref, idx, dist = _ref_case(9000, 40, 20, seed=5)
    R = TransformReference(ref)
    a = R.transform(idx, dist.astype(np.float32), n_iter=100)
    b = transform_tsne(ref, indices=idx, distances=dist.astype(np.float32), n_iter=100)
    assert relative_l2(a, b) < 1e-5
    c = R.transform(idx, dist.astype(np.float32), n_iter=100)
    assert np.array_equal(a, c)
    r = R.transform(idx, dist.astype(np.float32), initialization="random", n_iter=50)
    assert np.all(np.isfinite(r))
    R.close()


def test_transform_invalid_inputs():
    import fastembedr_amd as fe
    from fastembedr_amd.transform import transform_tsne

    ref, idx, dist = _ref_case(100, 5, 5, seed=1)
    bad = idx.copy()
    bad[0, 0] = 1000
    with pytest.raises(fe.HipError, match="out of range"):
        transform_tsne(ref, indices=bad, distances=dist)
    d2 = dist.copy()
    d2[0, 0] = np.nan
    with pytest.raises(fe.HipError, match="finite"):
        transform_tsne(ref, indices=idx, distances=d2)
