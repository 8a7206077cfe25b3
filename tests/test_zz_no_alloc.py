"""The allocation-free UMAP entry point against the allocating one.

fastembedr_hip_umap_optimize_coo_device_noalloc keeps optimizer_mode 1's
fixed-point delta in a file-scope device array and takes device pointers
only; its result must equal fastembedr_hip_umap_optimize_coo_device's bit for
bit, on two calls in a row (the second smaller, so a stale array would show),
and it must refuse inputs beyond its stated maximum.
"""
import numpy as np
import pytest

from tests.helpers import DeviceArray, exact_knn

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 21


def _graph(n, k, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 6))
    idx, dist = exact_knn(x, k)
    heads = np.repeat(np.arange(n), k).astype(np.int32)
    tails = idx.ravel().astype(np.int32)
    w = np.exp(-(dist - dist[:, :1]) / (dist.std() + 1e-6)).ravel().astype(np.float32)
    per = (w.max() / np.maximum(w, 1e-6)).astype(np.float32)
    y0 = rng.uniform(-10, 10, size=(n, 2)).astype(np.float32)
    return heads, tails, w, per, y0


def _noalloc(hip, heads, tails, w, per, y0, epochs, seed, mode):
    dev = [DeviceArray(a) for a in (heads, tails, w, per)]
    lay = DeviceArray(y0)
    rc = hip.fastembedr_hip_umap_optimize_coo_device_noalloc(
        dev[0].ptr, dev[1].ptr, dev[2].ptr, dev[3].ptr, lay.ptr, len(y0), len(heads), epochs, 5,
        1.0, 1.895605865596314, 0.8006377738365004, 1.0, seed, mode, None)
    assert rc == 0, hip.fastembedr_hip_last_error()
    assert lay._hip.hipDeviceSynchronize() == 0
    out = lay.to_host(np.float32, y0.shape)
    for d in dev + [lay]:
        d.free()
    return out


def _alloc(hip, heads, tails, w, per, y0, epochs, seed, mode):
    from fastembedr_amd import _lib

    dev = [DeviceArray(a) for a in (heads, tails, w, per)]
    out = np.zeros_like(y0)
    rc = hip.fastembedr_hip_umap_optimize_coo_device(
        dev[0].ptr, dev[1].ptr, dev[2].ptr, dev[3].ptr, y0.ctypes.data_as(_lib._fp), len(y0),
        len(heads), epochs, 5, 1.0, 1.895605865596314, 0.8006377738365004, 1.0, seed, mode,
        out.ctypes.data_as(_lib._fp))
    assert rc == 0, _lib.last_error()
    for d in dev:
        d.free()
    return out


def test_noalloc_matches_allocating_entry_point_twice(hip):
    for n, k, epochs, seed in ((3000, 12, 30, 5), (700, 8, 20, 6)):
        print(f"[no-alloc] n={n} k={k} epochs={epochs} seed={seed}", flush=True)
        g = _graph(n, k, seed)
        ref = _alloc(hip, *g, epochs, seed, 1)
        got = _noalloc(hip, *g, epochs, seed, 1)
        assert np.array_equal(got, ref)


def test_noalloc_atomic_mode_runs(hip):
    g = _graph(1500, 10, 8)
    got = _noalloc(hip, *g, 15, 3, 0)
    assert np.isfinite(got).all() and np.abs(got).max() <= 16.0


def test_noalloc_rejects_oversize_and_bad_arguments(hip):
    # both calls are refused on the host before any device work
    small = DeviceArray(np.zeros(16, np.int32))
    lay = DeviceArray(np.zeros(16, np.float32))
    p = small.ptr
    rc = hip.fastembedr_hip_umap_optimize_coo_device_noalloc(
        p, p, lay.ptr, lay.ptr, lay.ptr, MAX_POINTS + 1, 4, 1, 5, 1.0, 1.9, 0.8, 1.0, 1, 1, None)
    assert rc == 1
    msg = hip.fastembedr_hip_last_error().decode()
    assert "at most 2097152 points" in msg
    rc = hip.fastembedr_hip_umap_optimize_coo_device_noalloc(
        None, p, lay.ptr, lay.ptr, lay.ptr, 8, 4, 1, 5, 1.0, 1.9, 0.8, 1.0, 1, 1, None)
    assert rc == 1
    assert hip.fastembedr_hip_last_error()
    small.free()
    lay.free()
