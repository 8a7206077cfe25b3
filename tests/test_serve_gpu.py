"""The request loop: every result equals the same request served alone."""
import numpy as np
import pytest

from tests.helpers import exact_knn, gaussian_mixture

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone():
    from fastembedr_amd.serve import Request, Server

    rng = np.random.default_rng(0)
    ref_x, _ = gaussian_mixture(3000, 8, 5, seed=40)
    ref_y = rng.normal(size=(3000, 2)).astype(np.float32) * 5
    reqs = []
    for rid in range(8):
        if rid % 3 == 0:
            n = int(rng.integers(800, 2000))
            x, _ = gaussian_mixture(n, 8, 4, seed=100 + rid)
            idx, dist = exact_knn(x, 30)
            payload = dict(indices=idx, distances=dist.astype(np.float32), perplexity=10,
                           grid_size=128, early_exaggeration_iter=30, n_iter=30,
                           Y_init=rng.normal(scale=1e-4, size=(n, 2)).astype(np.float32))
            reqs.append(Request(rid, "embed", 0.001 * rid, payload))
        else:
            q = ref_x[rng.choice(3000, 200)] + rng.normal(scale=0.1, size=(200, 8)).astype(np.float32)
            idx, dist = exact_knn(ref_x, 20, query=q)
            reqs.append(Request(rid, "transform", 0.001 * rid,
                                dict(indices=idx, distances=dist.astype(np.float32), n_iter=60)))
    srv = Server(reference_layout=ref_y, max_jobs=4, tick=7)
    stats = srv.run(reqs, realtime=False)
    mixed = {r.rid: r.result for r in reqs}
    assert stats.completed == len(reqs)
    summary = stats.summary()
    assert summary["p99_ms"] >= summary["p50_ms"] > 0
    for r in reqs:
        alone = Request(r.rid, r.kind, 0.0, r.payload)
        s2 = srv.run([alone], realtime=False)
        assert s2.completed == 1
        assert np.array_equal(alone.result, mixed[r.rid]), r.rid
    srv.close()
