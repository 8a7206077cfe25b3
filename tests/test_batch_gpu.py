"""Batched jobs: each job's layout equals the same job run alone, bit for
bit, also when jobs join and leave while others run (continuous batching)."""
import numpy as np
import pytest

from tests.helpers import exact_knn, gaussian_mixture

pytestmark = pytest.mark.gpu


def _job(n, seed, k=30):
    x, _ = gaussian_mixture(n, 10, 6, seed=seed)
    idx, dist = exact_knn(x, k)
    y0 = np.random.default_rng(seed + 1).normal(scale=1e-4, size=(n, 2)).astype(np.float32)
    return idx, dist.astype(np.float32), y0


def _alone(idx, dist, y0, grid, early, normal):
    import fastembedr_amd as fe

    return fe.tsne_knn(idx, dist, perplexity=10, Y_init=y0, grid_size=grid,
                       early_exaggeration_iter=early, n_iter=normal)


def _add(b, job, grid, early, normal):
    idx, dist, y0 = job
    return b.add(idx, dist, perplexity=10, Y_init=y0, grid_size=grid,
                 early_exaggeration_iter=early, n_iter=normal)


def test_batch_equals_alone_bitwise():
    from fastembedr_amd.batch import TsneBatch

    specs = [(1500, 11, 128, 30, 40), (2600, 12, 256, 50, 25), (900, 13, 64, 20, 20),
             (3100, 14, 256, 60, 60)]
    jobs = [_job(n, s) for n, s, *_ in specs]
    b = TsneBatch(max_jobs=8)
    slots = []
    for (n, s, g, e, m), (idx, dist, y0) in zip(specs, jobs):
        slots.append(b.add(idx, dist, perplexity=10, Y_init=y0, grid_size=g,
                           early_exaggeration_iter=e, n_iter=m))
    b.run_to_completion(chunk=17)
    it, act = b.status()
    for (n, s, g, e, m), slot, (idx, dist, y0) in zip(specs, slots, jobs):
        assert it[slot] == e + m and act[slot] == 0
        got = b.layout(slot)
        ref = _alone(idx, dist, y0, g, e, m)
        assert np.array_equal(got, ref), (n, g)
    b.close()


def test_jobs_join_and_leave_mid_run():
    from fastembedr_amd.batch import TsneBatch

    a, c = _job(2000, 21), _job(1200, 22)
    b = TsneBatch(max_jobs=2)
    sa = _add(b, a, 128, 40, 40)
    b.step(30)
    sc = _add(b, c, 128, 30, 10)
    b.run_to_completion(chunk=10)
    ya, yc = b.layout(sa), b.layout(sc)
    b.remove(sa)
    # the freed slot takes a new job
    d = _job(1000, 23)
    sd = _add(b, d, 128, 10, 10)
    b.run_to_completion(chunk=10)
    yd = b.layout(sd)
    b.close()
    assert np.array_equal(ya, _alone(*a, 128, 40, 40))
    assert np.array_equal(yc, _alone(*c, 128, 30, 10))
    assert np.array_equal(yd, _alone(*d, 128, 10, 10))


def test_budget_admission():
    from fastembedr_amd.batch import NotAdmitted, TsneBatch, job_bytes

    a = _job(2000, 31)
    need = job_bytes(2000, 30, 128)
    b = TsneBatch(max_jobs=4, hbm_budget_bytes=need + need // 4)
    _add(b, a, 128, 5, 5)
    with pytest.raises(NotAdmitted):
        _add(b, a, 128, 5, 5)
    b.close()
    # the admission estimate bounds every device byte the job holds (P, the
    # attraction tables, state, grids), at the estimate's own grid and at the
    # rule's
    from fastembedr_amd import TsnePlan

    idx, dist, y0 = a
    for g in (128, 0):
        plan = TsnePlan(idx, dist, perplexity=10, grid_size=g, Y_init=y0)
        held = plan.info()["device_bytes"]
        plan.close()
        assert 0 < held <= job_bytes(2000, 30, g, setup_peak=False), (g, held)
