"""Autotuned batched jobs: the tuner's choices per bucket, tuned jobs in a
batch equal the same tuned job alone bit for bit, and a tuned job's first
step stays within the oracle tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tuned(tmp_path_factory):
    from fastembedr_amd.tune import Autotuner

    path = str(tmp_path_factory.mktemp("tune") / "table.json")
    t = Autotuner(k=30, perplexity=10.0, cache_path=path, quality_sample=1500, batch_copies=4,
                  time_iters=25)
    print("tune: buckets of 3000 and 7000 points", flush=True)
    t.tune(sizes=[3000, 7000])
    return t


def test_tuner_table(tuned):
    from fastembedr_amd.tune import Autotuner, cpu_rule_grid, cuda_rule_grid

    for b in (0, 1):
        e = tuned.table[b]
        n = e["n"]
        assert cpu_rule_grid(n) <= e["grid_size"] <= cuda_rule_grid(n)
        assert e["attraction"] in (2, 3) and e["pair_workgroups"] >= 0
        # the grid kept satisfies the tuner's own quality rule
        ref = e["grid_trials"][0]
        kept = [g for g in e["grid_trials"] if g["grid_size"] == e["grid_size"]][0]
        assert kept["trustworthiness"] >= ref["trustworthiness"] - tuned.quality_tol
        assert kept["knn_preservation"] >= ref["knn_preservation"] - tuned.quality_tol
    # the cache gives the same table back
    t2 = Autotuner(k=30, perplexity=10.0, cache_path=tuned.cache_path)
    assert t2.shape_for(3500) == tuned.shape_for(3500)


def test_tuned_batch_equals_tuned_alone(tuned):
    from fastembedr_amd import workloads
    from fastembedr_amd.batch import TsneBatch
    from fastembedr_amd.tune import run_alone
    from tests.helpers import exact_knn

    jobs = []
    for n, seed in ((2600, 31), (4200, 32), (6100, 33), (8800, 34)):
        x, _, _ = workloads.gaussian_mixture(n, 20, 8, seed=seed, dirichlet=True)
        idx, dist = exact_knn(x, 30)
        jobs.append((idx, dist.astype(np.float32), workloads.centred_normal_init(n, seed)))
    b = TsneBatch(max_jobs=4)
    slots = []
    for idx, dist, y0 in jobs:
        sh = tuned.shape_for(len(idx))
        slots.append(b.add(idx, dist, perplexity=10.0, Y_init=y0, early_exaggeration_iter=60,
                           n_iter=90, **sh))
    b.run_to_completion(chunk=25)
    for (idx, dist, y0), s in zip(jobs, slots):
        got = b.layout(s)
        alone = run_alone(idx, dist, y0, tuned.shape_for(len(idx)), perplexity=10.0, early=60,
                          normal=90)
        assert np.array_equal(got, alone), len(idx)
    b.close()


def test_tuned_first_step_matches_oracle(tuned):
    import oracle
    from fastembedr_amd import workloads
    from fastembedr_amd.tsne import TsnePlan
    from tests.helpers import exact_knn

    n = 3000
    x, _, _ = workloads.gaussian_mixture(n, 20, 8, seed=35, dirichlet=True)
    idx, dist = exact_knn(x, 30)
    y = np.random.default_rng(36).normal(size=(n, 2)).astype(np.float32)
    sh = tuned.shape_for(n)
    g = sh["grid_size"]
    plan = TsnePlan(idx, dist.astype(np.float32), perplexity=10.0, grid_size=g, Y_init=y,
                    attraction=sh["attraction"], pair_workgroups=sh["pair_workgroups"])
    plan.run(1, exaggeration=1.0, momentum=0.0, learning_rate=1.0, max_step_norm=1e6)
    got = plan.layout()
    plan.close()
    csr = oracle.build_p(idx, dist.astype(np.float64), 10.0)
    yd = y.astype(np.float64)
    ref = oracle.tsne_optimize(csr, yd, g, 0, 1, 1.0, 1.0, 1.0, 1.0, 0.0, 0.0, 0.01, 1e6)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 1e-4, err
