"""Parity of the HIP t-SNE path with the CPU oracle (GPU tests).

Tolerances are the reference suite's (tests/testthat/test-tsne-numerics.R):
P 2e-6, attraction 5e-5 relative L2, objective gradient 1e-4, FFT first step
vs CPU 2e-3; the interpolated gradient from a fixed Y must match the FP64
evaluation of the same interpolated gradient to 1e-4 relative L2.
"""
import numpy as np
import pytest

import oracle
from tests.helpers import exact_knn, gaussian_mixture, relative_l2

pytestmark = pytest.mark.gpu


def fe():
    import fastembedr_amd

    assert fastembedr_amd.available(), "no gfx950 device visible"
    return fastembedr_amd


def _case(seed, n, d, k, sd):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, d))
    idx, dist = exact_knn(x, k)
    y = rng.normal(scale=sd, size=(n, 2))
    return idx, dist, y


def test_device_is_gfx950(hip):
    assert hip.fastembedr_hip_available() == 1


@pytest.mark.parametrize("grid", [32, 64, 128])
def test_diagnostic_matches_oracle_small(grid):
    # test-tsne-numerics.R:1-34 shapes (n=48, k=15, perplexity 5)
    idx, dist, y = _case(701, 48, 6, 15, 0.3)
    dg = fe().gradient_diagnostic(idx, dist.astype(np.float32), y, 5, 1.0, grid)
    csr = oracle.build_p(idx, dist.astype(np.float32).astype(np.float64), 5)
    assert np.array_equal(dg.row_ptr, csr.row_ptr)
    assert np.array_equal(dg.col, csr.col)
    assert np.abs(dg.val - csr.val).max() <= 2e-6 * csr.val.max()
    yf = y.astype(np.float32).astype(np.float64)
    att = oracle.attractive(csr, yf)
    assert relative_l2(dg.attractive_force, att) < 5e-5
    rep, z = oracle.fft_repulsive(yf, grid)
    assert relative_l2(dg.repulsive_force_fft, rep) < 1e-4
    assert relative_l2(dg.optimizer_gradient_fft, att + rep) < 1e-4
    assert abs(dg.sum_q - z) / z < 1e-5


def test_fft_repulsion_converges_to_exact_on_gpu():
    # test-tsne-numerics.R:68-95
    idx, dist, y = _case(703, 48, 6, 15, 0.3)
    exact, _ = oracle.exact_repulsive(y.astype(np.float32).astype(np.float64))
    errs = []
    for g, lim in ((32, 1e-2), (64, 3e-3), (128, 1e-3)):
        dg = fe().gradient_diagnostic(idx, dist, y, 5, 1.0, g)
        e = relative_l2(dg.repulsive_force_fft, exact)
        errs.append(e)
        assert e < lim
        assert np.corrcoef(dg.repulsive_force_fft.ravel(), exact.ravel())[0, 1] > 0.999
    assert all(np.diff(errs) < 0)


@pytest.mark.parametrize("n,grid,exag", [(3000, 256, 12.0), (20000, 512, 1.0), (60000, 1024, 4.0)])
def test_interpolated_gradient_matches_fp64_oracle(n, grid, exag):
    x, _ = gaussian_mixture(n, 20, 10, seed=801 + n)
    idx, dist = exact_knn_clusters(x, 30)
    rng = np.random.default_rng(n)
    y = rng.normal(scale=5.0, size=(n, 2)).astype(np.float32)
    dg = fe().gradient_diagnostic(idx, dist, y, 10, exag, grid)
    csr = oracle.build_p(idx, dist.astype(np.float64), 10)
    yd = y.astype(np.float64)
    ref = oracle.attractive(csr, yd, exag)
    rep, z = oracle.fft_repulsive(yd, grid)
    assert relative_l2(dg.attractive_force, ref) < 5e-5
    assert relative_l2(dg.repulsive_force_fft, rep) < 1e-4
    assert relative_l2(dg.optimizer_gradient_fft, ref + rep) < 1e-4
    assert abs(dg.sum_q - z) / z < 1e-5


def exact_knn_clusters(x, k):
    """Exact KNN in chunks (float64)."""
    n = len(x)
    idx = np.zeros((n, k), np.int32)
    dist = np.zeros((n, k), np.float32)
    xd = x.astype(np.float64)
    sq = (xd * xd).sum(1)
    for s in range(0, n, 2048):
        e = min(n, s + 2048)
        d2 = sq[s:e, None] + sq[None, :] - 2.0 * xd[s:e] @ xd.T
        d2[np.arange(e - s), np.arange(s, e)] = np.inf
        part = np.argpartition(d2, k, axis=1)[:, :k]
        pd = np.take_along_axis(d2, part, 1)
        order = np.argsort(pd, axis=1, kind="stable")
        idx[s:e] = np.take_along_axis(part, order, 1)
        dist[s:e] = np.sqrt(np.maximum(np.take_along_axis(pd, order, 1), 0))
    return idx, dist


def test_pathological_inputs_stay_finite():
    # test-tsne-numerics.R:120-171
    rng = np.random.default_rng(705)
    x = np.vstack([np.zeros((8, 4)), np.ones((8, 4)), rng.normal(scale=1e-8, size=(8, 4))])
    idx, dist = exact_knn(x, 6)
    dx = np.vstack([rng.normal(-20, 0.01, size=(12, 4)), rng.normal(20, 0.01, size=(12, 4))])
    didx, ddist = exact_knn(dx, 6)
    cases = [(idx, dist, rng.normal(scale=0.1, size=(24, 2))),
             (idx, dist, np.zeros((24, 2))),
             (idx, dist, np.tile([-1e4, 1e4], (24, 1))),
             (didx, ddist, rng.normal(scale=0.1, size=(24, 2)))]
    for i, d, y in cases:
        dg = fe().gradient_diagnostic(i, d, y, 2, 1.0, 64)
        assert np.all(np.isfinite(dg.objective_gradient_fft))
        assert dg.sum_q > 0
        out = fe().tsne_knn(i, d, n_neighbors=6, perplexity=2, Y_init=y, early_exaggeration_iter=0,
                            n_iter=1, early_exaggeration=1, exaggeration=1, learning_rate=1,
                            initial_momentum=0, final_momentum=0, max_step_norm=1e6,
                            auto_config=False)
        assert np.all(np.isfinite(out))
        assert out.config["backend"] == "hip"
    bad = cases[0][2].copy()
    bad[0, 0] = np.inf
    with pytest.raises(ValueError, match="finite"):
        fe().gradient_diagnostic(idx, dist, bad, 2, 1.0, 64)
    with pytest.raises(ValueError, match="finite"):
        fe().tsne_knn(idx, dist, perplexity=2, Y_init=bad, n_iter=1, early_exaggeration_iter=0)


def test_invalid_knn_is_rejected():
    idx, dist, y = _case(709, 40, 4, 6, 0.1)
    bad = idx.copy()
    bad[3, 2] = 400
    with pytest.raises(fe().HipError, match="out of range"):
        fe().gradient_diagnostic(bad, dist, y, 2, 1.0, 64)
    dneg = dist.copy()
    dneg[1, 1] = -1.0
    with pytest.raises(fe().HipError, match="finite and non-negative"):
        fe().gradient_diagnostic(idx, dneg, y, 2, 1.0, 64)


def _step_args():
    return dict(perplexity=5, early_exaggeration_iter=0, n_iter=1, early_exaggeration=1,
                exaggeration=1, learning_rate=1, initial_momentum=0, final_momentum=0,
                max_step_norm=1e6, auto_config=False)


def test_first_step_matches_cpu():
    # test-tsne-numerics.R:239-316 (grid 64, 2e-3 against the CPU FFT step)
    rng = np.random.default_rng(706)
    x = rng.normal(size=(64, 6))
    idx, dist = exact_knn(x, 15)
    y0 = rng.normal(scale=1e-4, size=(64, 2))
    hip = fe().tsne_knn(idx, dist.astype(np.float32), n_neighbors=15, Y_init=y0, grid_size=64,
                        **_step_args())
    port, _, _ = oracle.port_tsne_run(idx, dist, y0, perplexity=5, early_iter=0, n_iter=1,
                                      early_exag=1, exag=1, learning_rate=1.0, mom0=0, mom1=0,
                                      max_step=1e6, grid=64, n_threads=2)
    csr = oracle.build_p(idx, dist.astype(np.float32).astype(np.float64), 5)
    ref = oracle.tsne_optimize(csr, y0.astype(np.float32).astype(np.float64), 64, 0, 1, 1, 1,
                               1.0, 1.0, 0.0, 0.0, 0.01, 1e6)
    assert relative_l2(hip, port) < 2e-3
    assert relative_l2(hip, ref) < 1e-4


def test_short_trajectory_tracks_fp64_oracle():
    x, _ = gaussian_mixture(1500, 10, 5, seed=77)
    idx, dist = exact_knn(x, 30)
    rng = np.random.default_rng(78)
    y0 = rng.normal(scale=1e-4, size=(1500, 2)).astype(np.float32)
    hip = fe().tsne_knn(idx, dist.astype(np.float32), perplexity=10, Y_init=y0, grid_size=128,
                        early_exaggeration_iter=20, n_iter=0)
    csr = oracle.build_p(idx, dist.astype(np.float32).astype(np.float64), 10)
    ref = oracle.tsne_optimize(csr, y0.astype(np.float64), 128, 20, 0, 12.0, 1.0, 1500 / 12.0,
                               1500 / 12.0, 0.8, 0.8, 0.01, 5.0)
    assert relative_l2(hip, ref) < 1e-3


def test_runs_are_bitwise_deterministic_and_graph_equals_eager(monkeypatch):
    x, _ = gaussian_mixture(4000, 10, 8, seed=91)
    idx, dist = exact_knn(x, 30)
    kw = dict(perplexity=10, grid_size=128, early_exaggeration_iter=60, n_iter=60, seed=5)
    a = fe().tsne_knn(idx, dist.astype(np.float32), **kw)
    b = fe().tsne_knn(idx, dist.astype(np.float32), **kw)
    assert np.array_equal(a, b)
    monkeypatch.setenv("FASTEMBEDR_HIP_GRAPH_CAPTURE", "0")
    c = fe().tsne_knn(idx, dist.astype(np.float32), **kw)
    assert np.array_equal(a, c)


def test_plan_phases_equal_one_shot():
    x, _ = gaussian_mixture(3000, 10, 6, seed=92)
    idx, dist = exact_knn(x, 30)
    y0 = np.random.default_rng(3).normal(scale=1e-4, size=(3000, 2)).astype(np.float32)
    one = fe().tsne_knn(idx, dist.astype(np.float32), perplexity=10, Y_init=y0, grid_size=256,
                        early_exaggeration_iter=100, n_iter=50)
    plan = fe().TsnePlan(idx, dist.astype(np.float32), perplexity=10, grid_size=256, Y_init=y0)
    lr = 3000 / 12.0
    plan.run(100, 12.0, 0.8, lr)
    plan.run(50, 1.0, 0.8, lr)
    out = plan.layout()
    info = plan.info()
    plan.close()
    assert np.array_equal(out, one)
    assert info["iterations"] == 150 and info["grid_size"] == 256


def test_end_to_end_quality_matches_cpu_reference_algorithm():
    # config 1 shape: 2000 points, 10-cluster mixture in 50 dims, k=90, 1000 iterations
    x, labels = gaussian_mixture(2000, 50, 10, seed=101)
    idx, dist = exact_knn(x, 90)
    y0 = np.random.default_rng(102).normal(scale=1e-4, size=(2000, 2)).astype(np.float32)
    y0 -= y0.mean(0)
    hip = fe().tsne_knn(idx, dist.astype(np.float32), perplexity=30, Y_init=y0, grid_size=128,
                        early_exaggeration_iter=250, n_iter=750)
    cpu, _, _ = oracle.port_tsne_run(idx, dist, y0, perplexity=30, early_iter=250, n_iter=750,
                                     learning_rate=2000 / 12.0, grid=128, n_threads=8)
    mh = oracle.structure_metrics(x, hip, 15)
    mc = oracle.structure_metrics(x, cpu, 15)
    assert abs(mh[0] - mc[0]) <= 0.01, (mh, mc)
    assert abs(mh[2] - mc[2]) <= 0.01, (mh, mc)
    assert abs(hip.mean(0)).max() < 1e-3


def test_attraction_kernels_agree_on_hub_rows(monkeypatch):
    # The centre of a 20-d Gaussian is every point's nearest neighbour, so its
    # row of the symmetric P holds all n points: rows of every length, one
    # spanning many 64-entry rounds.
    rng = np.random.default_rng(711)
    x = np.vstack([np.zeros((1, 20)), rng.normal(size=(4999, 20))])
    idx, dist = exact_knn(x, 30)
    y = rng.normal(scale=3.0, size=(5000, 2)).astype(np.float32)
    csr = oracle.build_p(idx, dist.astype(np.float32).astype(np.float64), 10)
    assert np.diff(csr.row_ptr).max() >= 500
    ref = oracle.attractive(csr, y.astype(np.float64), 4.0)
    got = {}
    for kind in ("lanes", "staged", "rows", "pairs"):
        monkeypatch.setenv("FASTEMBEDR_HIP_ATTRACTION", kind)
        got[kind] = fe().gradient_diagnostic(idx, dist.astype(np.float32), y, 10, 4.0, 128)
        assert relative_l2(got[kind].attractive_force, ref) < 5e-5, kind
    assert relative_l2(got["pairs"].attractive_force, got["lanes"].attractive_force) < 1e-5
    lanes = got["lanes"]
    # the symmetric P itself: its rows are sorted and merged by one wave up to
    # 512 entries and by one block above (the hub row)
    assert np.array_equal(lanes.row_ptr, csr.row_ptr)
    assert np.array_equal(lanes.col, csr.col)
    assert np.abs(lanes.val - csr.val).max() <= 2e-6 * csr.val.max()
    # the lane-padded kernels sum in one order; the staged window changes
    # only where the points are read from
    assert np.array_equal(got["staged"].attractive_force, lanes.attractive_force)
    assert relative_l2(lanes.attractive_force, got["rows"].attractive_force) < 1e-5
    assert np.array_equal(lanes.repulsive_force_fft, got["rows"].repulsive_force_fft)


def test_collapsed_layout_splits_heavy_tiles():
    # 90% of the points inside one grid cell: their tile is cut into several
    # work items whose sums meet in the 64-bit grid
    n = 20000
    x, _ = gaussian_mixture(n, 10, 5, seed=712)
    idx, dist = exact_knn_clusters(x, 30)
    rng = np.random.default_rng(713)
    y = np.where(rng.random((n, 1)) < 0.9, rng.normal(scale=1e-3, size=(n, 2)),
                 rng.normal(scale=10.0, size=(n, 2))).astype(np.float32)
    dg = fe().gradient_diagnostic(idx, dist, y, 10, 1.0, 256)
    yd = y.astype(np.float64)
    rep, z = oracle.fft_repulsive(yd, 256)
    assert relative_l2(dg.repulsive_force_fft, rep) < 1e-4
    assert abs(dg.sum_q - z) / z < 1e-5
    again = fe().gradient_diagnostic(idx, dist, y, 10, 1.0, 256)
    assert np.array_equal(dg.repulsive_force_fft, again.repulsive_force_fft)


@pytest.mark.parametrize("dense_min", [None, "1000000000"])
def test_pair_attraction_staged_and_rest_streams(monkeypatch, dense_min):
    # every pair once: tiles of 4096-point blocks staged in LDS, integer
    # sums.  With the staging threshold raised only the diagonal tiles are
    # staged and the pairs between blocks go through the rest stream (both
    # directions, global columns).  Sums are integers: repeatable bit for bit.
    n = 30000
    x, _ = gaussian_mixture(n, 12, 8, seed=721)
    idx, dist = exact_knn_clusters(x, 30)
    rng = np.random.default_rng(722)
    y = rng.normal(scale=8.0, size=(n, 2)).astype(np.float32)
    monkeypatch.setenv("FASTEMBEDR_HIP_ATTRACTION", "pairs")
    if dense_min:
        monkeypatch.setenv("FASTEMBEDR_HIP_PAIR_DENSE_MIN", dense_min)
    plan = fe().TsnePlan(idx, dist, perplexity=10, grid_size=256, Y_init=y)
    info = plan.attraction_info()
    plan.close()
    assert info["kind"] == "pairs"
    assert (info["rest_entries"] > 0) == bool(dense_min)
    a = fe().gradient_diagnostic(idx, dist, y, 10, 12.0, 256)
    b = fe().gradient_diagnostic(idx, dist, y, 10, 12.0, 256)
    csr = oracle.build_p(idx, dist.astype(np.float64), 10)
    ref = oracle.attractive(csr, y.astype(np.float64), 12.0)
    assert relative_l2(a.attractive_force, ref) < 5e-5
    assert np.array_equal(a.attractive_force, b.attractive_force)


@pytest.mark.parametrize("dense_min", [None, "1000000000"])
def test_pair_attraction_same_bits_for_any_workgroup_count(monkeypatch, dense_min):
    # The streams are dealt to workgroups as contiguous ranges of equal
    # estimated cost, and a chunk is cut where a range ends; every pair must
    # be summed exactly once whatever the cuts.  The sums are integers, so one
    # workgroup, a few (chunks cut at the range ends) and the default (as
    # many workgroups as chunks at this size) give the same bits.
    n = 30000
    x, _ = gaussian_mixture(n, 12, 8, seed=721)
    idx, dist = exact_knn_clusters(x, 30)
    y0 = np.random.default_rng(742).normal(scale=1e-4, size=(n, 2)).astype(np.float32)
    if dense_min:
        monkeypatch.setenv("FASTEMBEDR_HIP_PAIR_DENSE_MIN", dense_min)
    outs, shapes = [], []
    for wg in (1, 3, 7, 0):
        plan = fe().TsnePlan(idx, dist, perplexity=10, grid_size=256, Y_init=y0, attraction=3,
                             pair_workgroups=wg)
        info = plan.attraction_info()
        assert info["kind"] == "pairs"
        assert (info["rest_entries"] > 0) == bool(dense_min)
        plan.run(30, 12.0, 0.5, n / 12.0)
        plan.run(20, 1.0, 0.8, n / 12.0)
        outs.append(plan.layout())
        shapes.append((info["workgroups"], info["chunks"]))
        plan.close()
    assert [w for w, _ in shapes[:3]] == [1, 3, 7], shapes
    # the cuts add chunks
    assert shapes[2][1] > shapes[0][1], shapes
    assert np.all(np.isfinite(outs[0]))
    for o, s in zip(outs[1:], shapes[1:]):
        assert np.array_equal(o, outs[0]), s


def _hub_knn(n=30000, k=15):
    """A KNN-shaped input with hub points: row i's neighbours are i+1 .. i+k
    (mod n) at increasing distances; the last neighbour of 20000 rows is
    replaced by point 0 and of 9000 rows by point 1, and the next-to-last
    neighbour of 8200 rows by point 2, so the symmetric rows of the hubs
    exceed the 8192 entries sorted in LDS (point 0's needs two merge passes)."""
    i = np.arange(n)[:, None]
    idx = ((i + np.arange(1, k + 1)[None, :]) % n).astype(np.int32)
    dist = ((np.arange(k)[None, :] + 1.0) * (1.0 + 0.01 * (i % 7))).astype(np.float32)
    rows = np.arange(100, n - 100)
    np.random.default_rng(731).shuffle(rows)
    for hub, col, take in ((0, k - 1, rows[:20000]), (1, k - 1, rows[20000:29000]),
                           (2, k - 2, rows[:8200])):
        idx[take, col] = hub
    return idx, dist


def test_hub_rows_beyond_lds_sort_cap():
    idx, dist = _hub_knn()
    csr = oracle.build_p(idx, dist.astype(np.float64), 5.0)
    lengths = np.diff(csr.row_ptr)
    assert lengths[0] > 16384 and lengths[1] > 8192 and lengths[2] > 8192
    y = np.random.default_rng(732).normal(scale=5.0, size=(len(idx), 2)).astype(np.float32)
    print("[tsne] hub rows beyond the LDS sort", flush=True)
    dg = fe().gradient_diagnostic(idx, dist, y, 5.0, 1.0, 128)
    assert np.array_equal(dg.row_ptr, csr.row_ptr)
    assert np.array_equal(dg.col, csr.col)
    assert np.abs(dg.val - csr.val).max() <= 2e-6 * csr.val.max()
    again = fe().gradient_diagnostic(idx, dist, y, 5.0, 1.0, 128)
    assert np.array_equal(dg.val, again.val)
    assert relative_l2(dg.attractive_force, oracle.attractive(csr, y.astype(np.float64))) < 5e-5
