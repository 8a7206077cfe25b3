"""The KNN structure-score oracle (oracle/oracle_knn.c) pinned by an
independent route -- full lexicographic sorts of every row's layout
distances, set intersections and bincount majorities -- and by known
answers (a layout that is its own high-dimensional space scores 1)."""
import numpy as np

import oracle


def _numpy_score(layout, indices, keep, k, labels=None, n_levels=0):
    y = np.asarray(layout, np.float64)
    n = len(y)
    compact = len(indices) == len(keep)
    limit = indices.shape[1]
    tot = np.zeros(6)
    for kk, q1 in enumerate(keep):
        q = q1 - 1
        row = indices[kk if compact else q] - 1
        d2 = ((y - y[q]) ** 2).sum(1)
        others = np.array([c for c in range(n) if c != q])
        order = others[np.lexsort((others, d2[others]))]  # by (d2, index)
        low = order[:k]
        if len(low) < k:
            continue
        pos = {c: r for r, c in enumerate(row)}
        ranks = np.array([pos[c] + 1 if c in pos else limit + 1 for c in low])
        shared = int((ranks <= k).sum())
        trust_pen = float(np.maximum(ranks - k, 0).sum())
        low_rank = {c: r + 1 for r, c in enumerate(order)}
        targets = [c for c in row[:k] if 0 <= c < n]
        cont_pen = float(sum(max(low_rank[c] - k, 0) for c in targets))
        tot[0] += shared / k
        tot[1] += min(1.0, max(0.0, 1.0 - trust_pen / (k * max(1, limit + 1 - k))))
        tot[2] += min(1.0, max(0.0, 1.0 - cont_pen / (k * max(1, n - k))))
        if labels is not None and 1 <= labels[q] <= n_levels:
            counts = np.bincount(labels[low], minlength=n_levels + 1)[1:n_levels + 1]
            if counts.max() > 0:
                tot[3] += float(np.argmax(counts) + 1 == labels[q])
                tot[4] += 1.0
        tot[5] += 1.0
    return tot


def test_knn_score_oracle_matches_independent_route():
    rng = np.random.default_rng(11)
    x = rng.normal(size=(240, 6))
    y = x[:, :2] + rng.normal(scale=0.4, size=(240, 2))
    idx, _ = oracle.exact_knn(x.astype(np.float32), 25)
    ind = idx + 1
    labels = rng.integers(1, 5, size=240).astype(np.int32)
    keep = np.sort(rng.choice(240, 60, replace=False)) + 1
    for k in (1, 7, 25):
        got = oracle.knn_structure_score(y, ind, keep, k, labels, 4)
        ref = _numpy_score(y, ind, keep, k, labels, 4)
        assert np.allclose(got, ref, rtol=0, atol=1e-12), (k, got, ref)
    # compact rows: the indices of the kept rows only
    got = oracle.knn_structure_score(y, ind[keep - 1], keep, 10)
    ref = _numpy_score(y, ind[keep - 1], keep, 10)
    assert np.allclose(got, ref, rtol=0, atol=1e-12)
    # a grid layout: many equal distances, order by index
    g = rng.integers(0, 5, size=(150, 2)).astype(np.float64)
    gi, _ = oracle.exact_knn(rng.normal(size=(150, 3)).astype(np.float32), 12)
    got = oracle.knn_structure_score(g, gi + 1, np.arange(1, 151), 8)
    assert np.allclose(got, _numpy_score(g, gi + 1, np.arange(1, 151), 8), rtol=0, atol=1e-12)


def test_knn_score_oracle_known_answers():
    rng = np.random.default_rng(12)
    y = rng.normal(size=(200, 2))
    idx, _ = oracle.exact_knn(y.astype(np.float32), 15)
    # the layout's own exact neighbours: every kept row scores 1, 1, 1
    keep = np.arange(1, 201, 3)
    got = oracle.knn_structure_score(y.astype(np.float32).astype(np.float64), idx + 1, keep, 15)
    assert got[5] == len(keep)
    assert np.array_equal(got[:3], [len(keep)] * 3)
    # invalid kept rows are not scored; fewer candidates than k: not scored
    got = oracle.knn_structure_score(y, idx + 1, np.array([0, 201, 5]), 15)
    assert got[5] == 1
    tiny = rng.normal(size=(5, 2))
    ti = np.tile(np.arange(1, 6), (5, 1)).astype(np.int32)
    assert oracle.knn_structure_score(tiny, ti, np.arange(1, 6), 5)[5] == 0
