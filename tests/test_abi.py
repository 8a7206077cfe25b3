"""CPU-side checks of the C-ABI library: it exists, loads, and exports every
symbol include/fastembedr_hip.h declares (no device calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fastembedr_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fastembedr_hip_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "fastembedr_hip_opentsne_fft_from_knn_float" in names
    assert "fastembedr_hip_tsne_gradient_diagnostic" in names
    assert len(names) >= 12


def test_library_loads_and_exports_every_declared_symbol():
    from fastembedr_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libfastembedr_hip.so not built; run __graft_entry__.build()")
    raw = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(raw, n)]
    assert not missing, missing
    # the ctypes signature table covers exactly the header
    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_grid_rule_needs_no_device():
    from fastembedr_amd import grid_size
    import oracle

    for n in (100, 19999, 20000, 99999, 100000, 1_000_000):
        assert grid_size(n) == oracle.grid_size_gpu(n)


def test_auto_parameters_known_answers():
    # test-tsne.R:31-61: n=150,k=30 -> perplexity 10, 30 neighbours, lr 150/12;
    # n=70000, k=90 -> perplexity 30, 90 neighbours
    from fastembedr_amd import auto_parameters

    a = auto_parameters(150, 30)
    assert a["perplexity"] == 10 and a["n_neighbors"] == 30
    assert abs(a["learning_rate"] - 150 / 12) < 1e-12
    b = auto_parameters(70000, 90)
    assert b["perplexity"] == 30 and b["n_neighbors"] == 90
    assert b["early_exaggeration_iter"] == 250 and b["n_iter"] == 500


def test_structure_metrics_argument_errors_need_no_device():
    # embedding_utils.cpp:1587-1605: the checks run before any device work
    import numpy as np

    from fastembedr_amd import HipError, quality

    x = np.zeros((10, 3))
    cases = [
        (np.zeros((2, 3)), np.zeros((2, 2)), [1], "at least three observations are required"),
        (np.zeros((5001, 1)), np.zeros((5001, 1)), [5], "at most 5000 sampled observations"),
        (x, np.zeros((10, 2)), [], "requested_k must not be empty"),
        (x, np.zeros((10, 2)), [0], "each k must satisfy"),
        (x, np.zeros((10, 2)), [7], "each k must satisfy"),
        (x, np.zeros((10, 2)), [3, 10], "each k must satisfy"),
    ]
    for high, low, ks, msg in cases:
        with pytest.raises(HipError, match=msg):
            quality.exact_structure_metrics(high, low, ks)
    with pytest.raises(ValueError, match="same row count"):
        quality.exact_structure_metrics(x, np.zeros((9, 2)), [2])


def test_exact_knn_argument_errors_need_no_device():
    # native_knn_cuda_impl's checks (src/native_knn_cuda_impl.cpp:536-543)
    # run before any device work
    import numpy as np

    from fastembedr_amd import HipError, knn

    x = np.zeros((10, 3), np.float32)
    with pytest.raises(HipError, match="at least two rows and one column"):
        knn.exact_knn(np.zeros((1, 3), np.float32), 1)
    with pytest.raises(HipError, match="at least two rows and one column"):
        knn.exact_knn(np.zeros((10, 0), np.float32), 1)
    for k in (0, 10, 11):
        with pytest.raises(HipError, match=r"k in \[1, min\(n - 1, 256\)\]"):
            knn.exact_knn(x, k)
    with pytest.raises(HipError, match=r"k in \[1, min\(n - 1, 256\)\]"):
        knn.exact_knn(np.zeros((300, 2), np.float32), 257)
    with pytest.raises(HipError, match="index_base must be 0 or 1"):
        knn.exact_knn(x, 3, index_base=2)


def test_structure_score_argument_errors_need_no_device():
    # fastembedr_cuda_knn_structure_score's checks run before any device work
    import ctypes

    import numpy as np

    from fastembedr_amd import HipError, _lib, quality

    y = np.zeros((10, 2))
    ind = np.ones((10, 5), np.int32)
    for k in (0, 6):
        with pytest.raises(HipError, match="invalid structure-score dimensions"):
            quality.knn_structure_totals(y, ind, [1, 2], k)
    big = np.ones((10, 300), np.int32)
    with pytest.raises(HipError, match="invalid structure-score dimensions"):
        quality.knn_structure_totals(y, big, [1], 257)
    yc = np.zeros(20)
    ic = np.ones(50, np.int32)
    kp = np.ones(1, np.int32)
    tot = np.zeros(6)
    rc = _lib.lib().fastembedr_hip_knn_structure_score(
        yc.ctypes.data_as(_lib._dp), ic.ctypes.data_as(_lib._ip), kp.ctypes.data_as(_lib._ip),
        None, 10, 10, 5, 3, 1, 0, 4, tot.ctypes.data_as(_lib._dp))
    assert rc == 1 and "labels pointer is null" in _lib.last_error()
    with pytest.raises(ValueError, match="two-dimensional"):
        quality.knn_structure_score(np.zeros((10, 3)), ind, [1], 2)
    assert np.isnan(quality.knn_structure_score(y, ind, [], 2)["structure_score"])
    del ctypes


def test_query_knn_argument_errors_need_no_device():
    # native_cuda_query_knn_impl's checks (src/native_knn_cuda_impl.cpp:894-906)
    import numpy as np

    from fastembedr_amd import HipError, knn

    x = np.zeros((10, 3), np.float32)
    with pytest.raises(ValueError, match="matching columns"):
        knn.exact_query_knn(x, np.zeros((4, 2), np.float32), 2)
    for k in (0, 11):
        with pytest.raises(HipError, match=r"k in \[1, min\(nrow\(data\), 256\)\]"):
            knn.exact_query_knn(x, x[:2], k)
    with pytest.raises(HipError, match="at least one row"):
        knn.exact_query_knn(x, np.zeros((0, 3), np.float32), 1)


def test_gradient_method_and_grid_arguments_need_no_device():
    # tsne_repulsion_mode (src/tsne_neighbors.cpp:365-400): Barnes-Hut is
    # refused; the C-ABI takes grid_size -1 (exact), 0 (rule) or a power of two
    import numpy as np

    from fastembedr_amd import _lib, tsne_knn
    from fastembedr_amd._lib import _fp, _ip, ptr

    idx = np.tile(np.arange(1, 4, dtype=np.int32), (40, 1)) % 40
    dist = np.ones((40, 3), np.float32)
    with pytest.raises(ValueError, match="Barnes-Hut"):
        tsne_knn(idx, dist, negative_gradient_method="bh")
    with pytest.raises(ValueError, match="unknown"):
        tsne_knn(idx, dist, negative_gradient_method="spectral")
    out = np.zeros((40, 2), np.float32)
    ci, cd = np.asfortranarray(idx), np.asfortranarray(dist)
    rc = _lib.lib().fastembedr_hip_opentsne_fft_from_knn_float(
        ptr(ci, _ip), ptr(cd, _fp), None, 0, 40, 3, 2, 1.0, 0, 1, 12.0, 1.0, 1.0, 1, 0.5, 0.8,
        0.01, 5.0, 1, 0, -2, ptr(out, _fp))
    assert rc == 1 and "grid_size must be -1" in _lib.last_error()
