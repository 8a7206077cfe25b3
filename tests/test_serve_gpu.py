"""The native serving session: every result equals the same request served
alone and the plain single-request path, bit for bit; admission and errors."""
import numpy as np
import pytest

from tests.helpers import exact_knn, gaussian_mixture

pytestmark = pytest.mark.gpu


def _trace(rng, ref_x, n_req=10):
    from fastembedr_amd.serve import Request

    reqs = []
    for rid in range(n_req):
        if rid % 4 == 2:
            n = int(rng.integers(600, 1800))
            x, _ = gaussian_mixture(n, 8, 4, seed=200 + rid)
            idx, dist = exact_knn(x, 15)
            payload = dict(indices=idx, distances=dist.astype(np.float32),
                           init=rng.uniform(-10, 10, size=(n, 2)).astype(np.float32),
                           n_epochs=40 + 9 * rid, seed=rid)
            reqs.append(Request(rid, "umap", 0.001 * rid, payload))
        elif rid % 3 == 0:
            n = int(rng.integers(800, 2000))
            x, _ = gaussian_mixture(n, 8, 4, seed=100 + rid)
            idx, dist = exact_knn(x, 30)
            payload = dict(indices=idx, distances=dist.astype(np.float32), perplexity=10,
                           grid_size=128, early_exaggeration_iter=30, n_iter=30 + 7 * rid,
                           Y_init=rng.normal(scale=1e-4, size=(n, 2)).astype(np.float32))
            reqs.append(Request(rid, "embed", 0.001 * rid, payload))
        else:
            m = int(rng.integers(50, 400))
            q = ref_x[rng.choice(len(ref_x), m)] + rng.normal(scale=0.1, size=(m, 8)).astype(np.float32)
            idx, dist = exact_knn(ref_x, 20, query=q)
            init = ("median", "weighted", "random")[rid % 3]
            reqs.append(Request(rid, "transform", 0.001 * rid,
                                dict(indices=idx, distances=dist.astype(np.float32), n_iter=60,
                                     initialization=init, n_negatives=64,
                                     exact_repulsion_threshold=1000, seed=rid)))
    return reqs


def test_session_results_equal_alone_and_plain_path():
    import fastembedr_amd as fe
    from fastembedr_amd.serve import Request, Server
    from fastembedr_amd.transform import TransformReference
    from fastembedr_amd.umap import FuzzyGraph

    rng = np.random.default_rng(0)
    ref_x, _ = gaussian_mixture(3000, 8, 5, seed=40)
    ref_y = rng.normal(size=(3000, 2)).astype(np.float32) * 5
    reqs = _trace(rng, ref_x, n_req=12)
    srv = Server(reference_layout=ref_y, max_jobs=4, transform_pool_bytes=64 << 20)
    print("session: mixed trace of", len(reqs), "requests", flush=True)
    stats = srv.run(reqs, realtime=False)
    assert stats.completed == len(reqs) and stats.failed == 0
    summary = stats.summary()
    assert summary["p99_ms"] >= summary["p50_ms"] > 0
    st = srv.session.stats()
    assert st["completed"] == len(reqs) and st["ticks"] > 0
    assert st["queued_embed"] == st["running_embed"] == 0 and st["pool_bytes_used"] == 0
    assert st["queued_umap"] == st["running_umap"] == 0
    mixed = {r.rid: r.result for r in reqs}
    tref = TransformReference(ref_y)
    for r in reqs:
        alone = Request(r.rid, r.kind, 0.0, r.payload)
        s2 = srv.run([alone], realtime=False)
        assert s2.completed == 1
        assert np.array_equal(alone.result, mixed[r.rid]), r.rid
        p = r.payload
        if r.kind == "umap":
            g = FuzzyGraph(p["indices"], p["distances"])
            plain = g.optimize(p["init"], p["n_epochs"], seed=p["seed"])
            g.close()
        elif r.kind == "embed":
            plain = fe.tsne_knn(p["indices"], p["distances"], perplexity=p["perplexity"],
                                Y_init=p["Y_init"], grid_size=p["grid_size"],
                                early_exaggeration_iter=p["early_exaggeration_iter"],
                                n_iter=p["n_iter"])
        else:
            plain = tref.transform(**p)
        assert np.array_equal(plain, mixed[r.rid]), ("plain", r.rid)
    tref.close()
    srv.close()


def test_session_admission_and_errors():
    from fastembedr_amd._lib import HipError
    from fastembedr_amd.serve import DONE, FAILED, Session

    rng = np.random.default_rng(1)
    ref_x, _ = gaussian_mixture(2000, 8, 5, seed=41)
    ref_y = rng.normal(size=(2000, 2)).astype(np.float32)
    # a 256 KiB pool: a 2000-query transform at k = 20 needs 640 KB of
    # workspace and is refused at submit; a small one runs
    s = Session(ref_y, max_jobs=2, hbm_budget_bytes=64 << 20, transform_pool_bytes=256 << 10)
    q = ref_x[:2000]
    idx, dist = exact_knn(ref_x, 20, query=q)
    with pytest.raises(HipError, match="transform pool"):
        s.submit_transform(idx, dist.astype(np.float32))
    small = s.submit_transform(idx[:50], dist[:50].astype(np.float32), n_iter=20)
    out = s.wait(small)
    assert out.shape == (50, 2) and np.isfinite(out).all()
    # invalid KNN ids are refused at submit with the transform's message
    bad = idx[:10].copy()
    bad[3, 4] = 5000
    with pytest.raises(HipError, match="out of range"):
        s.submit_transform(bad, dist[:10].astype(np.float32))
    # a job larger than the whole 63.75 MiB left for jobs never fits: it fails
    # instead of waiting for ever
    x, _ = gaussian_mixture(30000, 8, 4, seed=42)
    jidx = rng.integers(0, 30000, size=(30000, 90)).astype(np.int32)
    jdist = rng.random(size=(30000, 90)).astype(np.float32) + 0.1
    big = s.submit_embed(jidx, jdist, early_exaggeration_iter=1, n_iter=1)
    state, res, _, err = s.poll(big, wait_ms=30000)
    assert state == FAILED and "not admissible" in err
    # a small job runs to completion beside it
    xs, _ = gaussian_mixture(600, 8, 4, seed=43)
    sidx, sdist = exact_knn(xs, 15)
    ok = s.submit_embed(sidx, sdist.astype(np.float32), perplexity=5, grid_size=64,
                        early_exaggeration_iter=10, n_iter=10)
    state, res, lat, err = s.poll(ok, wait_ms=60000)
    assert state == DONE and res.shape == (600, 2) and np.isfinite(res).all() and lat > 0
    with pytest.raises(HipError, match="unknown"):
        s.poll(ok)
    st = s.stats()
    assert st["failed"] == 1 and st["completed"] == 2
    s.close()


def test_session_concurrent_submitters():
    """Requests submitted and polled from several threads at once: every one
    completes and equals its request served alone."""
    import threading

    from fastembedr_amd.serve import DONE, Session

    rng = np.random.default_rng(2)
    ref_x, _ = gaussian_mixture(2500, 8, 5, seed=44)
    ref_y = rng.normal(size=(2500, 2)).astype(np.float32) * 3
    reqs = []
    for t in range(12):
        m = 100 + 37 * t
        q = ref_x[rng.choice(2500, m)] + rng.normal(scale=0.1, size=(m, 8)).astype(np.float32)
        idx, dist = exact_knn(ref_x, 15, query=q)
        reqs.append(dict(indices=idx, distances=dist.astype(np.float32), n_iter=40, seed=t))
    s = Session(ref_y, max_jobs=2)
    out = [None] * len(reqs)

    def worker(ids):
        for i in ids:
            rid = s.submit_transform(**reqs[i])
            out[i] = s.wait(rid)

    threads = [threading.Thread(target=worker, args=(list(range(w, len(reqs), 4)),))
               for w in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    for i, p in enumerate(reqs):
        rid = s.submit_transform(**p)
        state, alone, _, _ = s.poll(rid, wait_ms=60000)
        assert state == DONE and np.array_equal(alone, out[i]), i
    assert s.stats()["completed"] == 2 * len(reqs)
    s.close()
