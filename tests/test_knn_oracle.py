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
