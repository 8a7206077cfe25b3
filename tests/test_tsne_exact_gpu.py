"""The exact route and the full force diagnostic on the GPU against the oracle.

Tolerances (tests/testthat/test-tsne-numerics.R and the reference's own
FP32/FP64 gaps): attraction and exact repulsion 5e-5 relative L2, FFT-grid
repulsion 1e-4, the exact Z 1e-6 relative (a float), KL 5e-6 relative, the
first exact step 1e-4 and 20 exact steps 1e-3 against the FP64 oracle.
Config 1 (2000 points, k = 90, perplexity 30, exact route, 1000 iterations,
the reference CPU's default for n <= 3000) is compared on the quality gate:
trustworthiness and kNN preservation at k = 15 within 0.01 of the reference
CPU algorithm run on the same P and start.
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


def _case(n, d, k, seed, sd):
    x, _ = gaussian_mixture(n, d, 6, seed=seed)
    idx, dist = exact_knn(x, k)
    y = np.random.default_rng(seed + 1).normal(scale=sd, size=(n, 2)).astype(np.float32)
    return x, idx, dist.astype(np.float32), y


@pytest.mark.parametrize("n,k,perp,sd,grid", [(300, 15, 5.0, 0.3, 64), (1500, 30, 10.0, 4.0, 128),
                                              (4000, 30, 10.0, 20.0, 256)])
def test_force_diagnostic_bundle_matches_oracle(n, k, perp, sd, grid):
    print(f"[exact] diagnostic n={n} k={k} grid={grid}", flush=True)
    _, idx, dist, y = _case(n, 8, k, 900 + n, sd)
    d = fe().force_diagnostic(idx, dist, y, perp, 2.0, grid)
    csr = oracle.build_p(idx, dist.astype(np.float64), perp)
    assert np.array_equal(d["affinity_row_ptr0"], csr.row_ptr)
    assert np.array_equal(d["affinity_col1"] - 1, csr.col)
    yd = y.astype(np.float64)
    att = oracle.attractive(csr, yd, 2.0)
    rep, _ = oracle.exact_repulsive(yd)
    frep, _ = oracle.fft_repulsive(yd, grid)
    assert relative_l2(d["attractive_force"], att) < 5e-5
    assert relative_l2(d["repulsive_force_exact"], rep) < 5e-5
    assert relative_l2(d["optimizer_gradient_exact"], att + rep) < 5e-5
    assert relative_l2(d["repulsive_force_fft"], frep) < 1e-4
    assert np.allclose(d["objective_gradient_exact"], 4.0 * d["optimizer_gradient_exact"])
    z = oracle.sum_q(yd)
    assert abs(d["sum_q"] - z) <= 1e-6 * z
    kl = oracle.kl(csr, yd)
    assert abs(d["kl"] - kl) <= 5e-6 * abs(kl)


def test_exact_route_steps_match_fp64_oracle():
    _, idx, dist, _ = _case(800, 6, 20, 950, 1.0)
    y0 = np.random.default_rng(951).normal(scale=1e-4, size=(800, 2)).astype(np.float32)
    csr = oracle.build_p(idx, dist.astype(np.float64), 6.0)
    for steps, tol in ((1, 1e-4), (20, 1e-3)):
        out = fe().tsne_knn(idx, dist, perplexity=6.0, Y_init=y0, early_exaggeration_iter=steps,
                            n_iter=0, negative_gradient_method="exact", auto_config=True)
        ref = oracle.tsne_optimize(csr, y0, 0, steps, 0, 12.0, 1.0, 800 / 12.0, 800 / 12.0)
        assert relative_l2(out, ref) < tol, steps
        assert out.config["repulsion"] == "pair_symmetric_hip"


def test_exact_plan_graph_equals_eager_and_kl(monkeypatch):
    _, idx, dist, y = _case(1200, 6, 20, 960, 1.0)
    lr = 1200 / 12.0

    def run(capture):
        monkeypatch.setenv("FASTEMBEDR_HIP_GRAPH_CAPTURE", "1" if capture else "0")
        p = fe().TsnePlan(idx, dist, perplexity=6.0, grid_size=-1, Y_init=y)
        p.run(30, 12.0, 0.5, lr)
        p.run(7, 1.0, 0.8, lr)
        out = p.layout()
        kl = p.kl()
        info = p.info()
        p.close()
        return out, kl, info

    a, (kla, za), info = run(True)
    b, (klb, zb), _ = run(False)
    assert info["grid_size"] == -1
    assert np.array_equal(a, b) and kla == klb and za == zb
    csr = oracle.build_p(idx, dist.astype(np.float64), 6.0)
    kl = oracle.kl(csr, a.astype(np.float64))
    assert abs(kla - kl) <= 5e-6 * abs(kl)
    assert abs(za - oracle.sum_q(a.astype(np.float64))) <= 1e-6 * za


def test_config1_exact_route_quality_matches_reference_cpu():
    # 2000 points, 10 clusters of 200 in 50 dims (centres N(0, 5^2 I)), k = 90,
    # perplexity 30, 250 early (12) + 750 normal, explicit centred start
    x, _ = gaussian_mixture(2000, 50, 10, seed=101)
    idx, dist = exact_knn(x, 90)
    dist = dist.astype(np.float32)
    y0 = np.random.default_rng(102).normal(scale=1e-4, size=(2000, 2)).astype(np.float32)
    y0 -= y0.mean(0)
    print("[exact] config 1 on the GPU", flush=True)
    g = fe().tsne_knn(idx, dist, perplexity=30.0, Y_init=y0, early_exaggeration_iter=250,
                      n_iter=750, negative_gradient_method="exact")
    print("[exact] config 1 on the CPU (reference algorithm)", flush=True)
    c, _, _ = oracle.port_tsne_run(idx, dist, y0, perplexity=30.0, early_iter=250, n_iter=750,
                                   early_exag=12.0, exag=1.0, learning_rate=2000 / 12.0,
                                   mom0=0.8, mom1=0.8, max_step=5.0, grid=-1, n_threads=16)
    mg = oracle.structure_metrics(x.astype(np.float64), g.astype(np.float64), 15)
    mc = oracle.structure_metrics(x.astype(np.float64), c.astype(np.float64), 15)
    print(f"[exact] trust {mg[0]:.4f} vs {mc[0]:.4f}, knn {mg[2]:.4f} vs {mc[2]:.4f}", flush=True)
    assert abs(mg[0] - mc[0]) <= 0.01
    assert abs(mg[2] - mc[2]) <= 0.01
