"""Pins of the oracle and the CPU ports against the reference's own C++.

``oracle/_ref/libref.so`` is compiled by ``oracle/build_ref.sh`` from line
ranges of the reference sources (the functions on the hot path that need only
the standard library).  Each restated function must equal it bit for bit on
seeded random inputs chosen to reach every branch (edge periods below one,
integral and huge periods, epoch 0, seeds and indices near 2^31, rows that
trigger the uniform and the tied fallbacks).  The tests skip where the
reference sources were not present to build ``_ref``.
"""
import numpy as np
import pytest

import oracle
from oracle import ref

pytestmark = pytest.mark.skipif(ref.lib() is None, reason="oracle/_ref not built "
                                "(reference sources absent)")


def _periods(rng, m):
    p = np.concatenate([rng.uniform(1.0, 3.0, m // 4), rng.uniform(1.0, 600.0, m // 4),
                        rng.integers(1, 40, m // 4).astype(np.float64),
                        np.exp(rng.uniform(0.0, 12.0, m - 3 * (m // 4)))])
    return p.astype(np.float32)


def test_fast_pow_matches_reference():
    L = ref.lib()
    rng = np.random.default_rng(1)
    xs = np.concatenate([np.exp(rng.uniform(-30, 30, 3000)), [0.0, -1.0, 1e-38, 1.0, 2.0]])
    for b in (0.8006377738365004, 0.5, 1.0, 1.5, 2.3, 3.0):
        bf = np.float32(b)
        for x in xs.astype(np.float32):
            assert (np.float32(oracle.lib().or_umap_fast_pow(float(x), float(bf))).tobytes() ==
                    np.float32(L.ref_umap_powf_fast(float(x), float(bf))).tobytes())
            xd = float(x)
            assert oracle.lib().or_umap_cpu_pow(xd, b) == L.ref_umap_pow(xd, b)


def test_umap_hashes_match_reference():
    L = ref.lib()
    O = oracle.lib()
    rng = np.random.default_rng(2)
    for _ in range(4000):
        x = int(rng.integers(0, 2**32))
        assert O.or_umap_mix_uint(x) == L.ref_mix_seed(x)
        n = int(rng.choice([2, 3, 17, 1000, 1_000_000, 2**31 - 1]))
        seed = int(rng.integers(0, 2**31))
        epoch = int(rng.integers(0, 5000))
        head = int(rng.integers(0, n))
        tail = int(rng.integers(0, n))
        s = int(rng.integers(0, 64))
        # the GPU schedule's negative id: CUDA's deterministic_vertex(n, seed,
        # epoch, head, tail, s) is the reference CPU file's cuda_style form
        # with edge = tail (embedding_cuda_kernels.cpp:262-274, 3138-3145)
        assert (O.or_umap_negative_vertex(n, seed, epoch, head, tail, s) ==
                L.ref_deterministic_vertex_cuda_style(n, seed, epoch, head, tail, s))
        edge = int(rng.integers(0, 2**40))
        assert (O.or_umap_cpu_det_vertex(n, seed, epoch, edge, s) ==
                L.ref_deterministic_vertex(n, seed, epoch, edge, s))


def test_umap_schedules_match_reference():
    L = ref.lib()
    O = oracle.lib()
    rng = np.random.default_rng(3)
    periods = np.concatenate([_periods(rng, 400), np.float32([1.0, 2.0, 0.5, 1e-3, 0.0, -1.0,
                                                              np.inf, np.nan])])
    for p in periods:
        for e in list(range(0, 60)) + list(rng.integers(60, 5000, 40)):
            assert O.or_umap_n_positive(int(e), float(p)) == L.ref_positive_samples(float(p), int(e))
            for nsr in (0, 1, 5, 7):
                assert (O.or_umap_n_negative(int(e), float(p), nsr) ==
                        L.ref_negative_samples(float(p), nsr, int(e)))


def test_tau_prng_matches_reference():
    rng = np.random.default_rng(4)
    for _ in range(200):
        args = (int(rng.integers(0, 2**31)), int(rng.integers(0, 600)),
                int(rng.integers(0, 2**36)), int(rng.integers(0, 4)),
                int(rng.choice([7, 1000, 1_000_000])), 64)
        got = np.zeros(64, np.int32)
        oracle.lib().or_umap_cpu_tau_vertices(*args, got.ctypes.data_as(oracle._ip))
        assert np.array_equal(got, ref.tau_vertices(*args))


def test_umap_ab_constants_are_the_references():
    assert ref.find_ab(1.0, 0.01) == (oracle.UMAP_A, oracle.UMAP_B)


def test_transform_rng_matches_reference():
    L = ref.lib()
    O = oracle.lib()
    rng = np.random.default_rng(5)
    for v in list(rng.integers(0, 2**32, 3000)) + [0, 1, 2**32 - 1]:
        assert O.or_mix_uint32(int(v)) == L.ref_mix_uint32(int(v))
    for _ in range(50):
        state = int(rng.integers(1, 2**32))
        n = int(rng.choice([1, 2, 255, 4096, 500_000, 2**31 - 1]))
        raw, idx = ref.xorshift_stream(state, 300, n)
        s = oracle.ctypes.c_uint32(state)
        got = [O.or_uniform_index(oracle.ctypes.byref(s), n) for _ in range(300)]
        assert np.array_equal(np.array(got, np.int32), idx)
        assert all(0 <= i < n for i in got)


def test_float_perplexity_rows_match_reference():
    rng = np.random.default_rng(6)
    d = np.sort(rng.gamma(2.0, 1.0, size=(400, 90)), axis=1).astype(np.float32)
    d[0] = 1.5                       # uniform row
    d[1, :3] = 0.0                   # ties at the minimum
    d[2] = np.linspace(0, 1e4, 90)   # underflowing tail
    d[3, :10] = 1e-20
    for perp in (5.0, 30.0, 29.0):
        assert np.array_equal(oracle.port_row_probabilities_f(d, perp),
                              ref.row_probabilities_float(d, perp))


def _problem(n, k, seed):
    from tests.helpers import exact_knn, gaussian_mixture

    x, _ = gaussian_mixture(n, 10, 5, seed=seed)
    idx, dist = exact_knn(x, k)
    csr = oracle.build_p(idx, dist, 10.0)
    y = np.random.default_rng(seed + 1).normal(scale=2.0, size=(n, 2))
    return csr, y


def test_cpu_port_iteration_matches_reference_bitwise():
    # the float32 CPU port (the timed baseline when _ref is absent) against
    # the reference's own iteration code: same P, layout and state
    csr, y = _problem(1500, 20, 11)
    val = csr.val.astype(np.float32)
    yf = y.astype(np.float32)
    for grid, threads in ((64, 1), (128, 3), (0, 4)):
        got, _ = oracle.port_tsne_iterate(csr.row_ptr, csr.col, val, yf, 3, exag=12.0,
                                          learning_rate=125.0, momentum=0.5, grid=grid,
                                          n_threads=threads)
        want, _, _, _ = ref.tsne_iterate(csr.row_ptr, csr.col, val, yf, 3, 12.0, 125.0, 0.5,
                                         grid=grid if grid else ref.lib().ref_fft_grid_size(1500),
                                         threads=threads)
        assert np.array_equal(got, want), (grid, threads)


def test_oracle_gradients_and_kl_against_reference():
    # FP64 oracle against the reference's float32 functions: exact pair
    # repulsion, FFT-grid repulsion at the same grid, KL
    csr, y = _problem(800, 15, 12)
    val = csr.val.astype(np.float32)
    yf = y.astype(np.float32)
    yd = yf.astype(np.float64)
    exact = oracle.attractive(csr, yd, 1.0) + oracle.exact_repulsive(yd)[0]
    got = ref.tsne_gradient(csr.row_ptr, csr.col, val, yf, 1.0, grid=-1, threads=2)
    assert np.linalg.norm(got - exact) / np.linalg.norm(exact) < 2e-5
    rep, _ = oracle.fft_repulsive(yd, 128)
    fft = oracle.attractive(csr, yd, 1.0) + rep
    got = ref.tsne_gradient(csr.row_ptr, csr.col, val, yf, 1.0, grid=128, threads=2)
    assert np.linalg.norm(got - fft) / np.linalg.norm(fft) < 1e-4
    kl = oracle.kl(csr, yd)
    assert abs(ref.tsne_kl(csr.row_ptr, csr.col, val, yf, 2) - kl) <= 5e-6 * max(1.0, abs(kl))


def _umap_graph(n, seed):
    from tests.helpers import exact_knn

    x = np.random.default_rng(seed).normal(size=(n, 5))
    idx, dist = exact_knn(x, 8)
    g = oracle.umap_fuzzy_graph(idx, dist, index_offset=0)
    y0 = np.random.default_rng(seed + 1).uniform(-10, 10, size=(n, 2)).astype(np.float32)
    return g, y0


def test_cpu_umap_hogwild_matches_reference_edge_loop():
    # one thread: the reference's own hogwild edge loop is deterministic
    g, y0 = _umap_graph(2000, 21)
    rows, tails, eps = oracle.umap_cpu_active(g["offsets"], g["neighbors"], g["weights"], 60)
    got, form, _ = oracle.umap_cpu_optimize(rows, tails, eps, y0, 60, seed=9, n_threads=1, form=2)
    assert form == 2
    want, _ = ref.umap_hogwild(rows, tails, eps, y0, 60, seed=9, threads=1)
    assert np.array_equal(got, want)


def test_cpu_umap_forms_and_threads():
    # the reference's choice of form; several threads agree in distribution
    # (the hogwild races are nondeterministic in the reference too)
    L = ref.lib()
    for nt, items in ((1, 10), (8, 10**6), (3, 2), (16, 1)):
        assert L.ref_effective_cpu_threads(nt, items) == max(1, min(nt, 4, items))
    g, y0 = _umap_graph(1200, 22)
    rows, tails, eps = oracle.umap_cpu_active(g["offsets"], g["neighbors"], g["weights"], 40)
    _, form, _ = oracle.umap_cpu_optimize(rows, tails, eps, y0, 5, n_threads=1)
    assert form == 0
    _, form, _ = oracle.umap_cpu_optimize(rows, tails, eps, y0, 5, n_threads=4)
    assert form == 1
    a, _, _ = oracle.umap_cpu_optimize(rows, tails, eps, y0, 40, seed=3, n_threads=4, form=2)
    b, _ = ref.umap_hogwild(rows, tails, eps, y0, 40, seed=3, threads=4)
    assert np.isfinite(a).all() and np.isfinite(b).all()
    assert abs(np.std(a) - np.std(b)) < 0.1 * np.std(b)
