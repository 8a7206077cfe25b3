"""The exact k-NN oracle (oracle/oracle_knn.c) pinned by independent
references: scipy's k-d tree and scikit-learn's brute-force search on
tie-free data, and known answers where ties decide the order."""
import numpy as np
import pytest

import oracle


def test_knn_oracle_matches_kdtree_and_sklearn():
    from scipy.spatial import cKDTree
    from sklearn.neighbors import NearestNeighbors

    for n, p, k, seed in ((500, 3, 10, 1), (800, 20, 30, 2), (300, 1, 5, 3), (400, 50, 90, 4)):
        x = np.random.default_rng(seed).normal(size=(n, p)).astype(np.float32)
        idx, dist = oracle.exact_knn(x, k)
        xd = x.astype(np.float64)
        d_tree, i_tree = cKDTree(xd).query(xd, k=k + 1)
        assert np.array_equal(idx, i_tree[:, 1:]), (n, p, k)
        assert np.allclose(dist, d_tree[:, 1:], rtol=1e-6, atol=0), (n, p, k)
        d_sk, i_sk = NearestNeighbors(n_neighbors=k, algorithm="brute").fit(xd).kneighbors()
        assert np.array_equal(idx, i_sk)
        assert np.allclose(dist, d_sk, rtol=1e-5, atol=1e-6)


def test_knn_oracle_known_answers():
    # points 0, 1, ..., n-1 on a line: row i's neighbours at distance 1 are
    # i - 1 and i + 1, the smaller index first
    n = 50
    x = np.arange(n, dtype=np.float32)[:, None]
    idx, dist = oracle.exact_knn(x, 4)
    i = 20
    assert list(idx[i]) == [19, 21, 18, 22] and list(dist[i]) == [1, 1, 2, 2]
    assert list(idx[0]) == [1, 2, 3, 4] and list(dist[0]) == [1, 2, 3, 4]
    assert list(idx[n - 1]) == [48, 47, 46, 45]
    # duplicates: the other copies first, in index order, at distance 0
    y = np.random.default_rng(5).normal(size=(40, 3)).astype(np.float32)
    y[[3, 11, 17, 30]] = y[3]
    idx, dist = oracle.exact_knn(y, 5)
    assert list(idx[17][:3]) == [3, 11, 30] and np.all(dist[17][:3] == 0)
    assert dist[17][3] > 0
    # a subset of query rows gives those rows of the full answer
    full_i, full_d = oracle.exact_knn(y, 5)
    sub_i, sub_d = oracle.exact_knn(y, 5, queries=[17, 0, 39])
    assert np.array_equal(sub_i, full_i[[17, 0, 39]]) and np.array_equal(sub_d, full_d[[17, 0, 39]])


def test_knn_oracle_rejects_bad_sizes():
    with pytest.raises(ValueError):
        oracle.exact_knn(np.zeros((5, 2), np.float32), 5)


def test_query_knn_oracle_matches_kdtree():
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(21)
    data = rng.normal(size=(700, 9)).astype(np.float32)
    query = rng.normal(size=(120, 9)).astype(np.float32)
    idx, dist = oracle.exact_query_knn(data, query, 20)
    d_tree, i_tree = cKDTree(data.astype(np.float64)).query(query.astype(np.float64), k=20)
    assert np.array_equal(idx, i_tree) and np.allclose(dist, d_tree, rtol=1e-6, atol=0)
    # queries that are data rows find themselves first, at distance 0; the
    # rest is the self-excluded answer
    qi, qd = oracle.exact_query_knn(data, data[:50], 11)
    si, sd = oracle.exact_knn(data, 10, queries=np.arange(50))
    assert np.array_equal(qi[:, 0], np.arange(50)) and np.all(qd[:, 0] == 0)
    assert np.array_equal(qi[:, 1:], si) and np.array_equal(qd[:, 1:], sd)
    # k = nrow(data)
    ai, _ = oracle.exact_query_knn(data[:30], query[:5], 30)
    assert sorted(ai[0]) == list(range(30))
