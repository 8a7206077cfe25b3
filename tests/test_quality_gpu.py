"""Exact structure metrics on the GPU against the oracle's restatement of
exact_structure_metrics_cpp (src/embedding_utils.cpp:1582-1752)."""
import numpy as np
import pytest

import oracle
from tests.helpers import gaussian_mixture

pytestmark = pytest.mark.gpu


def _check(high, low, ks):
    from fastembedr_amd import quality

    got = quality.exact_structure_metrics(high, low, ks)
    assert got.shape == (len(ks), 4)
    for q, k in enumerate(ks):
        ref = oracle.structure_metrics(np.asarray(high, np.float64), np.asarray(low, np.float64), k)
        # integer rank sums: the same bits as the reference's formula
        assert np.array_equal(got[q], np.asarray(ref)), (k, got[q], ref)
    return got


def test_structure_metrics_match_oracle():
    x, _ = gaussian_mixture(600, 12, 4, seed=801)
    rng = np.random.default_rng(802)
    low = x[:, :2] + rng.normal(scale=0.5, size=(600, 2))
    got = _check(x, low, [1, 5, 15, 60])
    assert (got[:, :3] > 0.0).all() and (got[:, :3] <= 1.0).all()


def test_structure_metrics_identity_and_ties():
    from fastembedr_amd import quality

    x = np.random.default_rng(803).normal(size=(300, 5))
    m = quality.exact_structure_metrics(x, x, [10, 50])
    assert np.array_equal(m[:, :3], np.ones((2, 3))) and np.array_equal(m[:, 3], [0.0, 0.0])
    # integer grids: many equal distances, ranks broken by index
    g = np.random.default_rng(804).integers(0, 4, size=(250, 3)).astype(np.float64)
    lowg = np.random.default_rng(805).integers(0, 3, size=(250, 2)).astype(np.float64)
    _check(g, lowg, [3, 20])


def test_structure_metrics_smallest_and_largest():
    rng = np.random.default_rng(806)
    h = rng.normal(size=(3, 4))
    _check(h, rng.normal(size=(3, 2)), [1])
    # the reference's limits: n = 5000 and k just below (2n - 1) / 3
    x, _ = gaussian_mixture(5000, 8, 6, seed=807)
    low = x[:, :2] * 3.0 + rng.normal(size=(5000, 2))
    _check(x, low, [15])
    _check(x[:400], low[:400], [266])
