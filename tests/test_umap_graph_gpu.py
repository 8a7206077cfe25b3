"""The UMAP fuzzy graph built on the GPU against the oracle's restatement of
umap_graph_csr_float_cpp (src/fast_knn_umap.cpp:3995-4048)."""
import numpy as np
import pytest

import oracle
from tests.helpers import DeviceArray, exact_knn, gaussian_mixture

pytestmark = pytest.mark.gpu


def _compare(g, ref):
    assert np.array_equal(g["offsets"], ref["offsets"])
    assert np.array_equal(g["neighbors"], ref["neighbors"])
    # float memberships: exp of a float argument may differ in the last place
    # from the C library's
    assert np.abs(g["weights"] - ref["weights"]).max() <= 2 * np.spacing(np.float32(1.0))
    assert abs(g["max_weight"] - ref["max_weight"]) <= 2 * np.spacing(np.float32(1.0))
    assert np.allclose(g["epochs_per_sample"], ref["epochs_per_sample"], rtol=1e-6, atol=0)


def test_fuzzy_graph_matches_oracle():
    from fastembedr_amd import umap

    x, _ = gaussian_mixture(3000, 10, 6, seed=901)
    idx, dist = exact_knn(x, 15)
    fg = umap.FuzzyGraph(idx, dist)
    g = fg.export()
    _compare(g, oracle.umap_fuzzy_graph(idx, dist, index_offset=0))
    # symmetric weights, no self edges
    rows = np.repeat(np.arange(3000), np.diff(g["offsets"]))
    pairs = dict(zip(zip(rows, g["neighbors"]), g["weights"]))
    assert all(pairs[(j, i)] == w for (i, j), w in pairs.items())
    assert not np.any(rows == g["neighbors"])
    fg.close()


def test_fuzzy_graph_hub_rows():
    """Points in many KNN lists: union rows in every length class of the
    sort (at most 64, 128, 256 and 512 entries by one wave, longer by one
    block), duplicate entries of one neighbour in a list, and the symmetric
    epochs on such a graph."""
    from fastembedr_amd import umap

    x, _ = gaussian_mixture(2000, 6, 4, seed=904)
    idx, dist = exact_knn(x, 15)
    for hub, rows in ((0, range(1000, 2000)), (1, range(10, 60)), (2, range(100, 250)),
                      (3, range(300, 600)), (4, range(700, 740))):
        for r in rows:
            idx[r, 14] = hub  # keeps the distance of the entry it replaces
    idx[700, 13] = idx[700, 12]  # a duplicated neighbour
    print("[umap graph] hub rows n=2000 k=15", flush=True)
    fg = umap.FuzzyGraph(idx, dist)
    g = fg.export()
    lengths = np.diff(g["offsets"])
    assert lengths.max() > 512 and (lengths <= 64).sum() > 1000
    for lo, hi in ((64, 128), (128, 256), (256, 512)):
        assert ((lengths > lo) & (lengths <= hi)).any()
    _compare(g, oracle.umap_fuzzy_graph(idx, dist, index_offset=0))
    y0 = np.random.default_rng(905).uniform(-10, 10, size=(2000, 2)).astype(np.float32)
    r1 = fg.optimize(y0, 20, seed=6, optimizer_mode=1)
    r2 = fg.optimize_coo_device(y0, 20, seed=6, optimizer_mode=1)
    assert np.array_equal(r1, r2) and np.isfinite(r1).all()
    fg.close()


def test_fuzzy_graph_columns_offsets_and_duplicates():
    from fastembedr_amd import umap

    rng = np.random.default_rng(902)
    x = rng.normal(size=(1500, 5))
    x[100:140] = x[100]  # duplicates: zero distances, rho = 0 rows
    idx, dist = exact_knn(x, 20)
    # self in column 0 (as KNN tools return it), 1-based, skipped by col_start
    idx1 = np.hstack([np.arange(1500)[:, None], idx]) + 1
    dist1 = np.hstack([np.zeros((1500, 1)), dist])
    g = umap.FuzzyGraph(idx1, dist1, col_start=1, n_cols=12).export()
    _compare(g, oracle.umap_fuzzy_graph(idx1, dist1, col_start=1, n_cols=12))
    with pytest.raises(Exception, match="default edge budget"):
        umap.FuzzyGraph(idx, dist, edge_budget=5)


def test_fuzzy_graph_device_inputs_and_resident_epochs():
    from fastembedr_amd import umap

    x, _ = gaussian_mixture(2500, 8, 5, seed=903)
    idx, dist = exact_knn(x, 15)
    host = umap.FuzzyGraph(idx, dist)
    ti = DeviceArray(idx.T.astype(np.int32))
    td = DeviceArray(dist.T.astype(np.float32))
    dev = umap.FuzzyGraph(ti.ptr.value, td.ptr.value, device=True, n=2500, k=15, index_offset=0)
    ti.free()
    td.free()
    a, b = host.export(), dev.export()
    for key in ("offsets", "neighbors", "weights", "epochs_per_sample"):
        assert np.array_equal(a[key], b[key])
    # the epochs from the resident graph equal the epochs from the exported one
    y0 = np.random.default_rng(904).uniform(-10, 10, size=(2500, 2)).astype(np.float32)
    r1 = dev.optimize(y0, 30, seed=5, optimizer_mode=1)
    rows = np.repeat(np.arange(2500, dtype=np.int32), np.diff(b["offsets"]))
    r2 = umap.optimize_coo(rows, b["neighbors"], b["weights"], b["epochs_per_sample"], y0, 30,
                           seed=5, optimizer_mode=1)
    assert np.array_equal(r1, r2)
    assert np.isfinite(r1).all()
    # the symmetric form (attraction through the heads) against the general
    # device-COO path on the same resident graph
    r3 = dev.optimize_coo_device(y0, 30, seed=5, optimizer_mode=1)
    assert np.array_equal(r1, r3)


def test_fuzzy_graph_rows_beyond_lds_sort_cap():
    """Union rows of more than 8192 raw entries (a point in over 8192 KNN
    lists) take the global sort: the graph is still the oracle's."""
    from fastembedr_amd import umap
    from tests.test_tsne_gpu import _hub_knn

    idx, dist = _hub_knn()
    print("[umap graph] rows beyond the LDS sort", flush=True)
    fg = umap.FuzzyGraph(idx, dist, index_offset=0)
    g = fg.export()
    lengths = np.diff(g["offsets"])
    assert lengths[0] > 16384 and lengths[1] > 8192
    _compare(g, oracle.umap_fuzzy_graph(idx, dist.astype(np.float64), index_offset=0))
    fg.close()
