"""Pins the CPU oracle against the reference tests' own known answers and
against what the mathematics fixes (dense FP64 restatement of
helper-tsne-reference.R, finite differences, textbook FFT, direct
convolution, closed forms).  CPU only."""
import numpy as np
import pytest

import oracle
from tests.helpers import (exact_knn, relative_l2, tsne_reference_forces,
                           tsne_reference_probabilities)


def _case(seed, n, d, k, sd):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, d))
    idx, dist = exact_knn(x, k)
    layout = rng.normal(scale=sd, size=(n, 2))
    return idx, dist, layout


def test_affinities_and_forces_match_dense_fp64_reference():
    # test-tsne-numerics.R:1-34 (n=48, d=6, k=15, perplexity 5)
    idx, dist, y = _case(701, 48, 6, 15, 0.3)
    dense = tsne_reference_probabilities(idx, dist, 5)
    csr = oracle.build_p(idx, dist, 5)
    assert np.abs(csr.dense() - dense).max() < 1e-12
    ref = tsne_reference_forces(dense, y)
    att = oracle.attractive(csr, y)
    rep, sq = oracle.exact_repulsive(y)
    assert relative_l2(att, ref["attractive_force"]) < 1e-12
    assert relative_l2(rep, ref["repulsive_force"]) < 1e-12
    assert abs(sq - ref["sum_q"]) / ref["sum_q"] < 1e-12
    # the reference's KL carries 1e-9 guards (tsne_neighbors.cpp:1360); its
    # test accepts 5e-6 against the unguarded dense value
    kl = oracle.kl(csr, y)
    assert abs(kl - ref["kl"]) / ref["kl"] < 5e-6
    pd = csr.dense()
    num = 1.0 / (1.0 + ((y[:, None] - y[None]) ** 2).sum(-1))
    np.fill_diagonal(num, 0)
    qn = num / num.sum()
    m = pd > 0
    guarded = (pd[m] * np.log((pd[m] + 1e-9) / (qn[m] + 1e-9))).sum()
    assert abs(kl - guarded) < 1e-12
    # rows of the CSR are ascending and symmetric
    for i in range(csr.n):
        cols = csr.col[csr.row_ptr[i]:csr.row_ptr[i + 1]]
        assert np.all(np.diff(cols) > 0)
    dm = csr.dense()
    assert np.allclose(dm, dm.T, atol=0, rtol=0)
    assert abs(dm.sum() - 1.0) < 1e-12


def test_objective_gradient_passes_finite_difference():
    # test-tsne-numerics.R:36-66 (tolerance 2e-7 for the FP64 reference)
    idx, dist, y = _case(702, 18, 5, 9, 0.2)
    csr = oracle.build_p(idx, dist, 3)
    rep, _ = oracle.exact_repulsive(y)
    grad = 4.0 * (oracle.attractive(csr, y) + rep)
    p = csr.dense()
    eps = 1e-6
    num = np.zeros_like(y)
    for t in range(y.size):
        up, lo = y.copy(), y.copy()
        up.flat[t] += eps
        lo.flat[t] -= eps
        num.flat[t] = (tsne_reference_forces(p, up)["kl"] - tsne_reference_forces(p, lo)["kl"]) / (2 * eps)
    assert relative_l2(grad, num) < 2e-7


def test_fft_repulsion_converges_to_exact():
    # test-tsne-numerics.R:68-95: limits 1e-2 / 3e-3 / 1e-3 at 32 / 64 / 128
    idx, dist, y = _case(703, 48, 6, 15, 0.3)
    exact, _ = oracle.exact_repulsive(y)
    errors = []
    for g, limit in ((32, 1e-2), (64, 3e-3), (128, 1e-3)):
        rep, z = oracle.fft_repulsive(y, g)
        err = relative_l2(rep, exact)
        errors.append(err)
        assert err < limit
        assert np.corrcoef(rep.ravel(), exact.ravel())[0, 1] > 0.999
    assert all(np.diff(errors) < 0)


def test_fft2_matches_numpy():
    rng = np.random.default_rng(3)
    a = rng.normal(size=(32, 32)) + 1j * rng.normal(size=(32, 32))
    re, im = oracle.fft2(a.real, a.imag)
    ref = np.fft.fft2(a)
    assert np.abs(re + 1j * im - ref).max() < 1e-10
    re2, im2 = oracle.fft2(re, im, inverse=True)
    assert np.abs(re2 + 1j * im2 - a).max() < 1e-12


def test_fft_convolution_matches_direct_convolution():
    rng = np.random.default_rng(4)
    y = rng.normal(scale=0.5, size=(40, 2))
    g = 8
    rep, z, grids, splat = oracle.fft_repulsive(y, g, return_grids=True)
    h = oracle.grid_geometry(y, g)[2]
    for out, ch, power in ((0, 0, 1), (1, 0, 2), (2, 1, 2), (3, 2, 2)):
        direct = oracle.direct_convolution(splat[ch], h, power)
        assert np.abs(grids[out] - direct).max() < 1e-10 * max(1, np.abs(direct).max())
    # splat conserves mass and first moments exactly
    assert abs(splat[0].sum() - len(y)) < 1e-10
    assert abs(splat[1].sum() - y[:, 0].sum()) < 1e-10
    assert abs(splat[2].sum() - y[:, 1].sum()) < 1e-10


def test_fft_grid_z_close_to_exact_sum_q():
    idx, dist, y = _case(705, 64, 6, 15, 0.3)
    _, z = oracle.fft_repulsive(y, 128)
    assert abs(z - oracle.sum_q(y)) / oracle.sum_q(y) < 1e-3


@pytest.mark.parametrize("perplexity", [2, 4, 8])
@pytest.mark.parametrize("mult", [1, 3, 4])
def test_row_calibration_hits_target_entropy(perplexity, mult):
    # test-tsne-numerics.R:97-118 grid of perplexity x support
    rng = np.random.default_rng(704)
    x = rng.normal(size=(40, 6))
    k = perplexity * mult
    idx, dist = exact_knn(x, k)
    for i in range(0, 40, 7):
        p = oracle.row_probabilities(dist[i], perplexity)
        assert abs(p.sum() - 1.0) < 1e-12
        if mult > 1:
            nz = p[p > 0]
            entropy = -(nz * np.log(nz)).sum()
            assert abs(entropy - np.log(perplexity)) < 1e-4
    dense = tsne_reference_probabilities(idx, dist, perplexity)
    assert np.abs(oracle.build_p(idx, dist, perplexity).dense() - dense).max() < 1e-12


def test_degenerate_rows():
    # uniform distances -> 1/k (tsne_neighbors.cpp:467-471)
    assert np.allclose(oracle.row_probabilities(np.ones(7), 3), 1.0 / 7)
    # one overwhelming tie group -> mass shared by the tied nearest entries
    p = oracle.row_probabilities(np.array([0.0, 0.0, 1e6, 1e6]), 1.5)
    assert abs(p.sum() - 1) < 1e-12 and p[0] == p[1]


def test_invalid_knn_inputs_rejected():
    idx, dist, _ = _case(706, 20, 4, 5, 0.1)
    bad = idx.copy()
    bad[0, 0] = 99
    with pytest.raises(ValueError, match="out of range"):
        oracle.build_p(bad, dist, 2)
    d2 = dist.copy()
    d2[1, 1] = np.inf
    with pytest.raises(ValueError, match="finite"):
        oracle.build_p(idx, d2, 2)


def test_update_closed_form_sign_cases():
    # tsne_neighbors.cpp:1297-1336: sign(upd) != sign(grad) -> gain + 0.2,
    # else 0.8 gain + min_gain; zero sign counts as its own state
    y = np.zeros((3, 2))
    upd = np.array([[1.0, -1.0], [0.0, 2.0], [0.0, 0.0]])
    gains = np.ones((3, 2))
    grad = np.array([[1.0, -1.0], [0.0, -1.0], [1.0, 0.0]])
    oracle.update(y, upd, gains, grad, lr=0.5, momentum=0.8, min_gain=0.01, max_step=np.inf)
    exp_g = np.array([[0.81, 0.81], [0.81, 1.2], [1.2, 0.81]])
    assert np.allclose(gains, exp_g)
    exp_u = np.array([[0.8 - 0.5 * 0.81, -0.8 + 0.5 * 0.81], [0.0, 1.6 + 0.5 * 1.2],
                      [-0.5 * 1.2, 0.0]])
    assert np.allclose(upd, exp_u)
    assert np.allclose(y, exp_u)
    # clipping to max_step
    y = np.zeros((1, 2))
    upd = np.zeros((1, 2))
    gains = np.ones((1, 2))
    oracle.update(y, upd, gains, np.array([[-30.0, -40.0]]), 1.0, 0.0, 0.01, 5.0)
    assert np.allclose(np.linalg.norm(upd), 5.0) and np.allclose(upd, [[3.0, 4.0]])


def test_grid_rules():
    # CPU floor 128 (test-tsne-numerics.R:358-361), CUDA 256/512/1024
    assert oracle.grid_size_cpu(64) == 128 and oracle.grid_size_cpu(50000) == 256
    assert oracle.grid_size_gpu(2000) == 256 and oracle.grid_size_gpu(20000) == 512
    assert oracle.grid_size_gpu(1_000_000) == 1024


def test_structure_metrics():
    rng = np.random.default_rng(9)
    x = rng.normal(size=(60, 5))
    m = oracle.structure_metrics(x, x, 10)
    assert np.allclose(m[:3], 1.0) and m[3] == 0.0
    low = rng.normal(size=(60, 2))
    m = oracle.structure_metrics(x, low, 10)
    # brute force kNN preservation
    _, hi = exact_knn(x, 10)
    hi_idx, _ = exact_knn(x, 10)
    lo_idx, _ = exact_knn(low, 10)
    shared = sum(len(set(a) & set(b)) for a, b in zip(hi_idx, lo_idx))
    assert abs(m[2] - shared / 600.0) < 1e-12


def test_cpu_port_first_step_matches_fp64_oracle():
    # float32 port vs FP64 oracle, first optimiser step (grid 64, exact)
    rng = np.random.default_rng(706)
    x = rng.normal(size=(64, 6))
    idx, dist = exact_knn(x, 15)
    y0 = rng.normal(scale=1e-4, size=(64, 2))
    csr = oracle.build_p(idx, dist, 5)
    for grid, tol in ((64, 2e-3), (-1, 5e-4)):
        ref = oracle.tsne_optimize(csr, y0, max(grid, 0), 0, 1, 1.0, 1.0, 1.0, 1.0, 0.0, 0.0,
                                   0.01, 1e6)
        got, _, _ = oracle.port_tsne_run(idx, dist, y0, perplexity=5, early_iter=0, n_iter=1,
                                         early_exag=1, exag=1, learning_rate=1.0, mom0=0,
                                         mom1=0, max_step=1e6, grid=grid, n_threads=2)
        assert relative_l2(got, ref) < tol


def _lex_ranks(d2):
    """1-based ranks of every column in each row, self excluded, ties by index
    (np.lexsort: the last key is primary)."""
    n = d2.shape[0]
    ranks = np.zeros((n, n), np.int64)
    cols = np.arange(n)
    for i in range(n):
        order = np.lexsort((cols, d2[i]))
        order = order[order != i]
        ranks[i, order] = np.arange(1, n)
    return ranks, None


def test_structure_metrics_against_sklearn_and_brute_force():
    # trustworthiness is scikit-learn's (the same definition, tie-free data);
    # continuity is trustworthiness with the two spaces swapped; the mean
    # neighbour rank error, |rank in high space of the r-th low neighbour - r|
    # averaged over n k, is recounted from full lexicographic rank tables
    from sklearn.manifold import trustworthiness

    rng = np.random.default_rng(31)
    for n, k in ((120, 5), (300, 10), (300, 15), (61, 20)):
        high = rng.normal(size=(n, 7))
        low = high[:, :2] + rng.normal(scale=0.7, size=(n, 2))
        m = oracle.structure_metrics(high, low, k)
        assert abs(m[0] - trustworthiness(high, low, n_neighbors=k)) < 1e-12
        assert abs(m[1] - trustworthiness(low, high, n_neighbors=k)) < 1e-12
        d_hi = ((high[:, None] - high[None]) ** 2).sum(-1)
        d_lo = ((low[:, None] - low[None]) ** 2).sum(-1)
        r_hi, _ = _lex_ranks(d_hi)
        r_lo, _ = _lex_ranks(d_lo)
        err = 0
        for i in range(n):
            low_nb = np.argsort(r_lo[i] + (np.arange(n) == i) * n)[:k]  # ranks 1..k in order
            err += np.abs(r_hi[i, low_nb] - np.arange(1, k + 1)).sum()
        assert abs(m[3] - err / (n * k)) < 1e-12
    # a wrong normaliser or a dropped term would move these values
    assert not np.isclose(trustworthiness(high, low, n_neighbors=k), 1.0)
