"""Batched UMAP jobs: in optimizer_mode 1 each job's layout equals its
resident graph optimised alone, bit for bit, also when jobs join and leave
while others run; mode 0 jobs run beside them; admission by budget."""
import numpy as np
import pytest

from tests.helpers import exact_knn, gaussian_mixture

pytestmark = pytest.mark.gpu


def _job(n, seed, k=15):
    x, _ = gaussian_mixture(n, 8, 5, seed=seed)
    idx, dist = exact_knn(x, k)
    y0 = np.random.default_rng(seed + 1).uniform(-10, 10, size=(n, 2)).astype(np.float32)
    return idx, dist.astype(np.float32), y0


def _alone(idx, dist, y0, epochs, seed, mode=1):
    from fastembedr_amd.umap import FuzzyGraph

    g = FuzzyGraph(idx, dist)
    out = g.optimize(y0, epochs, seed=seed, optimizer_mode=mode)
    g.close()
    return out


@pytest.mark.parametrize("persistent", ["1", "0"])
def test_umap_batch_equals_alone_bitwise(monkeypatch, persistent):
    # with FASTEMBEDR_HIP_UMAP_PERSISTENT=1 jobs of up to 6144 points run in
    # the persistent kernel (layout in LDS), larger ones and hogwild jobs in
    # the job-table launches; with 0 (the default) every job takes the launches
    from fastembedr_amd.umap import UmapBatch

    monkeypatch.setenv("FASTEMBEDR_HIP_UMAP_PERSISTENT", persistent)
    specs = [(1200, 1, 60), (3100, 2, 45), (700, 3, 80), (2400, 4, 33), (7000, 5, 52),
             (6144, 6, 30)]
    jobs = [_job(n, s) for n, s, _ in specs]
    b = UmapBatch(max_jobs=8)
    slots = [b.add(i, d, y, n_epochs=e, seed=100 + s) for (n, s, e), (i, d, y) in zip(specs, jobs)]
    # one hogwild job beside them
    hi, hd, hy = _job(1500, 9)
    hs = b.add(hi, hd, hy, n_epochs=40, seed=7, optimizer_mode=0)
    b.run_to_completion(chunk=26)
    ep, act = b.status()
    for (n, s, e), slot, (i, d, y) in zip(specs, slots, jobs):
        assert ep[slot] == e and act[slot] == 0
        assert np.array_equal(b.layout(slot), _alone(i, d, y, e, 100 + s)), n
    assert ep[hs] == 40 and np.isfinite(b.layout(hs)).all()
    b.close()


def test_umap_jobs_join_and_leave_mid_run():
    from fastembedr_amd.umap import UmapBatch

    a, c = _job(2000, 11), _job(1300, 12)
    b = UmapBatch(max_jobs=2)
    sa = b.add(*a, n_epochs=70, seed=1)
    b.step(30)
    sc = b.add(*c, n_epochs=30, seed=2)
    b.step(30)
    _, act = b.status()
    assert act[sa] == 1 and act[sc] == 0
    lc = b.layout(sc)
    b.remove(sc)
    b.run_to_completion(chunk=10)
    assert np.array_equal(b.layout(sa), _alone(*a, 70, 1))
    assert np.array_equal(lc, _alone(*c, 30, 2))
    b.close()


def test_umap_batch_budget():
    from fastembedr_amd.umap import UmapBatch, UmapNotAdmitted, umap_job_bytes

    j = _job(3000, 21)
    need = umap_job_bytes(3000, 15, 1, True)
    # the first job holds what it measured after its build (below the
    # estimate); the second job's estimate with its build peak no longer fits
    b = UmapBatch(max_jobs=4, hbm_budget_bytes=need + need // 8)
    s0 = b.add(*j, n_epochs=10)
    with pytest.raises(UmapNotAdmitted):
        b.add(*j, n_epochs=10)
    b.remove(s0)
    b.add(*j, n_epochs=10)
    b.close()
