"""The measured configurations, on the GPU, at their real shapes.

* single-job t-SNE (1M points, 100-cluster 50-d Dirichlet mixture, exact
  k = 90, perplexity 30, grid 1024): P and the interpolated gradient of a
  100k-point draw of the same generator against the FP64 oracle (P identical
  up to 2e-6 of the largest value, gradient 1e-4 relative L2); on the full
  1M points a finite, bitwise repeatable run whose captured graphs give the
  eager launches' bits;
* UMAP (the same data, k = 15 fuzzy graph, U[-10, 10] start, 500 epochs,
  5 negatives): the graph of a 100k draw identical to the oracle's, 5 epochs of
  the deterministic mode against the oracle, the reference-default atomic mode
  on the quality gate against the reference CPU optimizer (its own hogwild
  edge loop, 4 threads), and 500 epochs of both modes at 1M points;
* batched t-SNE: 24 jobs of the batched workload's generator (2k-100k points,
  both grid sizes, k = 90, perplexity 30, 750 iterations), each bit-identical to
  the job run alone;
* the served stream: a 500k-point reference embedding, transforms and embed
  jobs in one trace, each result bit-identical alone and in the mix.
"""
import numpy as np
import pytest

import oracle
from oracle import ref as oref
from tests.helpers import relative_l2

pytestmark = pytest.mark.gpu


def fe():
    import fastembedr_amd

    assert fastembedr_amd.available(), "no gfx950 device visible"
    return fastembedr_amd


def _data(n, seed=201):
    from fastembedr_amd import workloads

    x, labels, _ = workloads.gaussian_mixture(n, 50, 100, seed=seed, dirichlet=True)
    return x, labels


def _knn(x, k):
    """Exact KNN from the device producer as host n x k arrays."""
    from fastembedr_amd import knn

    ki, kd = knn.exact_knn(x, k)
    return np.ascontiguousarray(ki.T), np.ascontiguousarray(kd.T)


_CACHE = {}


def _big(n=1_000_000, k=90):
    if (n, k) not in _CACHE:
        print(f"[configs] data and exact KNN, n={n} k={k}", flush=True)
        x, labels = _data(n)
        _CACHE[(n, k)] = (x, labels) + _knn(x, k)
    return _CACHE[(n, k)]


def test_config2_slice_p_and_gradient_match_fp64_oracle():
    print("[configs] single job: 100k draw, P and gradient at grid 1024", flush=True)
    x, _ = _data(100_000)
    idx, dist = _knn(x, 90)
    csr = oracle.build_p(idx, dist.astype(np.float64), 30.0)
    for y in ((x[:, :2] * 3.0).astype(np.float32),
              np.random.default_rng(202).normal(scale=1e-4, size=(100_000, 2)).astype(np.float32)):
        dg = fe().gradient_diagnostic(idx, dist, y, 30.0, 12.0, 1024)
        assert np.array_equal(dg.row_ptr, csr.row_ptr)
        assert np.array_equal(dg.col, csr.col)
        assert np.abs(dg.val - csr.val).max() <= 2e-6 * csr.val.max()
        yd = y.astype(np.float64)
        att = oracle.attractive(csr, yd, 12.0)
        rep, _ = oracle.fft_repulsive(yd, 1024)
        assert relative_l2(dg.attractive_force, att) < 5e-5
        assert relative_l2(dg.optimizer_gradient_fft, att + rep) < 1e-4


def test_config2_1m_run_is_finite_repeatable_and_graph_equals_eager(monkeypatch):
    x, _, idx, dist = _big()
    y0 = np.random.default_rng(202).normal(scale=1e-4, size=(len(x), 2)).astype(np.float32)
    lr = len(x) / 12.0

    def run(capture):
        monkeypatch.setenv("FASTEMBEDR_HIP_GRAPH_CAPTURE", "1" if capture else "0")
        p = fe().TsnePlan(idx, dist, perplexity=30.0, Y_init=y0)
        assert p.info()["grid_size"] == 1024
        p.run(12, 12.0, 0.8, lr)
        p.run(6, 1.0, 0.8, lr)
        out = p.layout()
        p.close()
        return out

    print("[configs] single job: 1M points, graph / graph / eager", flush=True)
    a, b, c = run(True), run(True), run(False)
    assert np.isfinite(a).all()
    assert np.array_equal(a, b) and np.array_equal(a, c)


def _fuzzy(idx, dist):
    from fastembedr_amd import umap

    return umap.FuzzyGraph(idx, dist, index_offset=0)


def test_config3_slice_graph_and_epochs_match_oracle():
    print("[configs] UMAP: 100k draw, fuzzy graph and 5 epochs", flush=True)
    x, _ = _data(100_000)
    idx, dist = _knn(x, 15)
    g = _fuzzy(idx, dist)
    got = g.export()
    want = oracle.umap_fuzzy_graph(idx, dist.astype(np.float64), index_offset=0)
    assert np.array_equal(got["offsets"], want["offsets"])
    assert np.array_equal(got["neighbors"], want["neighbors"])
    assert np.abs(got["weights"] - want["weights"]).max() <= 2.4e-7
    y0 = np.random.default_rng(301).uniform(-10, 10, size=(100_000, 2)).astype(np.float32)
    dev = g.optimize(y0, 5, negative_sample_rate=5, seed=302, optimizer_mode=1)
    # the epochs from the same graph (weights within 2 ulp can move a period
    # across an integer and change one edge's sample count)
    off = got["offsets"]
    heads = np.repeat(np.arange(100_000, dtype=np.int32), np.diff(off))
    ref = oracle.umap_optimize(heads, got["neighbors"], got["epochs_per_sample"], y0, 5, nsr=5,
                               seed=302, mode=1)
    assert relative_l2(dev, ref) < 1e-5
    g.close()


@pytest.mark.skipif(oref.lib() is None, reason="oracle/_ref not built")
def test_config3_atomic_mode_quality_matches_reference_cpu():
    # the reference's R layer runs CUDA UMAP in mode 0 (R/umap_core_helpers.R:
    # 306-319); the CPU reference is optimize_layout_csr at its 4-thread cap
    n = 50_000
    print(f"[configs] UMAP: {n}-point draw, 500 epochs, GPU mode 0 vs reference CPU", flush=True)
    x, _ = _data(n, seed=311)
    idx, dist = _knn(x, 15)
    g = _fuzzy(idx, dist)
    gr = g.export()
    y0 = np.random.default_rng(301).uniform(-10, 10, size=(n, 2)).astype(np.float32)
    dev = g.optimize(y0, 500, negative_sample_rate=5, seed=302, optimizer_mode=0)
    g.close()
    rows, tails, eps = oracle.umap_cpu_active(gr["offsets"], gr["neighbors"], gr["weights"], 500)
    cpu, _ = oref.umap_hogwild(rows, tails, eps, y0, 500, nsr=5, seed=302, threads=4)
    sample = np.random.default_rng(7).choice(n, 2000, replace=False)
    hx = x[sample].astype(np.float64)
    mg = oracle.structure_metrics(hx, dev[sample].astype(np.float64), 15)
    mc = oracle.structure_metrics(hx, cpu[sample].astype(np.float64), 15)
    print(f"[configs] trust {mg[0]:.4f} vs {mc[0]:.4f}, knn {mg[2]:.4f} vs {mc[2]:.4f}", flush=True)
    assert np.isfinite(dev).all() and np.abs(dev).max() <= 16.0
    assert abs(mg[0] - mc[0]) <= 0.01
    assert abs(mg[2] - mc[2]) <= 0.01


def test_config3_1m_epochs_both_modes():
    x, _, idx, dist = _big()
    print("[configs] UMAP: 1M points, 500 epochs, both modes", flush=True)
    g = _fuzzy(np.ascontiguousarray(idx[:, :15]), np.ascontiguousarray(dist[:, :15]))
    y0 = np.random.default_rng(301).uniform(-10, 10, size=(len(x), 2)).astype(np.float32)
    a = g.optimize(y0, 500, negative_sample_rate=5, seed=302, optimizer_mode=1)
    b = g.optimize(y0, 500, negative_sample_rate=5, seed=302, optimizer_mode=1)
    c = g.optimize(y0, 500, negative_sample_rate=5, seed=302, optimizer_mode=0)
    g.close()
    assert np.array_equal(a, b)
    for y in (a, c):
        assert np.isfinite(y).all() and np.abs(y).max() <= 16.0
        assert y.std() > 0.5


def test_config4_batched_jobs_equal_alone():
    from fastembedr_amd import workloads
    from fastembedr_amd.batch import TsneBatch

    specs = workloads.batch_job_specs(24, seed=301)
    grids = {fe().grid_size(s["n"]) for s in specs}
    assert grids == {256, 512}
    print(f"[configs] batched: 24 jobs, {sum(s['n'] for s in specs)} points", flush=True)
    jobs = []
    for s in specs:
        x, _, _ = workloads.gaussian_mixture(s["n"], s["d"], s["clusters"], seed=s["seed"],
                                             dirichlet=True)
        idx, dist = _knn(x, 90)
        jobs.append((idx, dist, workloads.centred_normal_init(s["n"], seed=s["seed"] + 1)))
    b = TsneBatch(max_jobs=24)
    slots = [b.add(i, d, perplexity=30.0, Y_init=y, early_exaggeration_iter=250, n_iter=500)
             for i, d, y in jobs]
    b.run_to_completion(chunk=50)
    outs = [b.layout(s) for s in slots]
    b.close()
    for (i, d, y), got in zip(jobs, outs):
        alone = fe().tsne_knn(i, d, perplexity=30.0, Y_init=y)
        assert np.array_equal(got, alone)


def test_config5_served_results_equal_alone():
    from fastembedr_amd import workloads
    from fastembedr_amd.serve import Request, Server

    n_ref = 500_000
    print("[configs] served: 500k reference", flush=True)
    xr, lr_, centres = workloads.gaussian_mixture(n_ref, 50, 50, seed=402, dirichlet=True)
    ri, rd = _knn(xr, 90)
    ref_y = fe().tsne_knn(ri, rd, perplexity=30.0,
                          Y_init=workloads.centred_normal_init(n_ref, 403))
    weights = np.bincount(lr_, minlength=len(centres)) / n_ref
    rng = np.random.default_rng(501)
    reqs = []
    for rid, (kind, size) in enumerate([("transform", 1000), ("embed", 12000),
                                        ("transform", 20000), ("transform", 3000),
                                        ("embed", 30000), ("transform", 50000)]):
        if kind == "embed":
            x, _, _ = workloads.gaussian_mixture(size, 40, 20, seed=600 + rid, dirichlet=True)
            idx, dist = _knn(x, 90)
            payload = dict(indices=idx, distances=dist, perplexity=30.0,
                           Y_init=workloads.centred_normal_init(size, 700 + rid))
        else:
            xq, _ = workloads.sample_mixture(size, centres, weights, rng)
            qi, qd = workloads.knn_to_reference(xq, xr, 90)
            payload = dict(indices=qi, distances=qd)
        reqs.append(Request(rid, kind, 0.0, payload))
    srv = Server(reference_layout=ref_y, max_jobs=8)
    stats = srv.run(reqs, realtime=False)
    assert stats.completed == len(reqs)
    mixed = {r.rid: r.result for r in reqs}
    for r in reqs:
        alone = Request(r.rid, r.kind, 0.0, r.payload)
        assert srv.run([alone], realtime=False).completed == 1
        assert np.array_equal(alone.result, mixed[r.rid]), (r.rid, r.kind)
    srv.close()
