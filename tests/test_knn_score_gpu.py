"""KNN structure score on the GPU against the oracle's restatement of
knn_structure_score (src/embedding_utils.cpp:1386-1580; the GPU form
src/embedding_cuda_kernels.cpp:1827-1944): the six totals equal, bit for
bit, over neighbour counts, compact index rows, labels, ties and the limits."""
import numpy as np
import pytest

import oracle
from tests.helpers import gaussian_mixture

pytestmark = pytest.mark.gpu


def _check(layout, indices, keep, k, labels=None, levels=0):
    from fastembedr_amd import quality

    print(f"[score] n={len(layout)} rows={len(indices)} keep={len(keep)} k={k}", flush=True)
    got = quality.knn_structure_totals(layout, indices, keep, k, labels, levels)
    ref = oracle.knn_structure_score(layout, indices, keep, k, labels, levels)
    assert np.array_equal(got, ref), (k, got, ref)
    return got


def test_structure_score_matches_oracle():
    x, lab = gaussian_mixture(1500, 10, 5, seed=811)
    rng = np.random.default_rng(812)
    layout = x[:, :2] * 2.0 + rng.normal(scale=0.8, size=(1500, 2))
    idx, _ = oracle.exact_knn(x.astype(np.float32), 30)
    ind = idx + 1
    labels = (np.asarray(lab) + 1).astype(np.int32)
    labels[rng.choice(1500, 100, replace=False)] = 0  # unlabelled rows
    keep = np.sort(rng.choice(1500, 300, replace=False)) + 1
    for k in (1, 10, 30):
        tot = _check(layout, ind, keep, k, labels, 5)
        assert tot[5] == 300 and 0 < tot[0] <= 300
    # compact index rows
    _check(layout, ind[keep - 1], keep, 12, labels, 5)
    # the Python surface: the reference's five values
    from fastembedr_amd import quality

    s = quality.knn_structure_score(layout, ind, keep, 10, labels, 5)
    assert abs(s["structure_score"] - (s["knn_preservation"] + s["local_trustworthiness"] +
                                       s["local_continuity"]) / 3) < 1e-15


def test_structure_score_ties_invalid_rows_and_limits():
    rng = np.random.default_rng(813)
    # integer grid: many equal layout distances
    g = rng.integers(0, 6, size=(800, 2)).astype(np.float64)
    gi, _ = oracle.exact_knn(rng.normal(size=(800, 4)).astype(np.float32), 20)
    keep = np.concatenate([[0, 801], np.arange(1, 801, 5)]).astype(np.int32)  # two invalid
    tot = _check(g, gi + 1, keep, 20)
    assert tot[5] == 160
    # a neighbour listed twice and the row itself in its own list
    ind = gi + 1
    ind[:, 3] = ind[:, 2]
    ind[:, 5] = np.arange(1, 801)
    _check(g, ind, np.arange(1, 801, 7), 8)
    # k = n - 1 (every other row) and the largest k
    y = rng.normal(size=(40, 2))
    yi, _ = oracle.exact_knn(rng.normal(size=(40, 3)).astype(np.float32), 39)
    _check(y, yi + 1, np.arange(1, 41), 39)
    z = rng.normal(size=(3000, 2))
    zi, _ = oracle.exact_knn(rng.normal(size=(3000, 5)).astype(np.float32), 256)
    _check(z, zi + 1, np.arange(1, 3001, 97), 256)
    # the layout's own neighbours score 1, 1, 1
    own, _ = oracle.exact_knn(z.astype(np.float32), 15)
    zf = z.astype(np.float32).astype(np.float64)
    tot = _check(zf, own + 1, np.arange(1, 3001, 11), 15)
    assert tot[0] == tot[1] == tot[2] == tot[5]


def test_structure_score_large_layout():
    x, lab = gaussian_mixture(100_000, 8, 20, seed=814)
    rng = np.random.default_rng(815)
    layout = x[:, :2] + rng.normal(scale=0.3, size=(100_000, 2))
    keep = np.sort(rng.choice(100_000, 40, replace=False)) + 1
    from fastembedr_amd import knn

    idx, _ = knn.exact_knn(x.astype(np.float32), 15, index_base=1)
    _check(layout, idx.T.copy(), keep, 15, (np.asarray(lab) + 1).astype(np.int32), 20)
