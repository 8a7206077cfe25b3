"""ctypes binding of libfastembedr_hip.so (the C-ABI in include/fastembedr_hip.h).

The library is built in-tree by ``python -m fastembedr_amd.build``.  Loading
fails loudly when it is missing: there is no CPU fallback for the optimizer.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfastembedr_hip.so")

c_int, c_uint, c_float, c_double, c_ll = (ctypes.c_int, ctypes.c_uint, ctypes.c_float,
                                          ctypes.c_double, ctypes.c_longlong)
_ip = ctypes.POINTER(ctypes.c_int)
_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_llp = ctypes.POINTER(ctypes.c_longlong)
_vp = ctypes.c_void_p
_vpp = ctypes.POINTER(ctypes.c_void_p)

# every symbol include/fastembedr_hip.h declares, with (restype, argtypes)
SIGNATURES = {
    "fastembedr_hip_available": (c_int, []),
    "fastembedr_hip_last_error": (ctypes.c_char_p, []),
    "fastembedr_hip_opentsne_fft_grid_size": (c_int, [c_int]),
    "fastembedr_hip_opentsne_fft_from_knn": (
        c_int, [_ip, _dp, _fp, c_int, c_int, c_int, c_int, c_float, c_int, c_int, c_float,
                c_float, c_float, c_int, c_float, c_float, c_float, c_float, c_uint, c_int, c_int,
                _fp]),
    "fastembedr_hip_opentsne_fft_from_knn_float": (
        c_int, [_ip, _fp, _fp, c_int, c_int, c_int, c_int, c_float, c_int, c_int, c_float,
                c_float, c_float, c_int, c_float, c_float, c_float, c_float, c_uint, c_int, c_int,
                _fp]),
    "fastembedr_hip_opentsne_fft_from_device_knn_float": (
        c_int, [_vp, _vp, _fp, c_int, c_int, c_int, c_int, c_float, c_int, c_int, c_float,
                c_float, c_float, c_int, c_float, c_float, c_float, c_float, c_uint, c_int, c_int,
                _fp]),
    "fastembedr_hip_tsne_gradient_diagnostic": (
        c_int, [_ip, _fp, c_int, c_int, c_int, c_float, c_float, c_int, _fp, _fp, _fp, _fp, _dp,
                _ip, _ip, _fp, _llp]),
    "fastembedr_hip_opentsne_force_diagnostic": (
        c_int, [_ip, _fp, c_int, c_int, c_int, c_float, c_float, c_int, _fp, _fp, _fp, _fp, _fp,
                _fp, _dp, _dp, _ip, _ip, _fp, _llp]),
    "fastembedr_hip_tsne_plan_kl": (c_int, [_vp, _dp, _dp]),
    "fastembedr_hip_tsne_plan_create": (
        c_int, [_vp, _vp, c_int, c_int, c_int, c_int, c_float, c_int, _fp, c_uint, _vpp]),
    "fastembedr_hip_tsne_plan_create_shaped": (
        c_int, [_vp, _vp, c_int, c_int, c_int, c_int, c_float, c_int, c_int, c_int, _fp, c_uint,
                _vpp]),
    "fastembedr_hip_tsne_plan_run": (
        c_int, [_vp, c_int, c_float, c_float, c_float, c_float, c_float]),
    "fastembedr_hip_tsne_plan_prepare": (
        c_int, [_vp, c_int, c_float, c_float, c_float, c_float, c_float, c_ll]),
    "fastembedr_hip_tsne_plan_sync": (c_int, [_vp]),
    "fastembedr_hip_tsne_plan_layout": (c_int, [_vp, _fp]),
    "fastembedr_hip_tsne_plan_info": (c_int, [_vp, _llp]),
    "fastembedr_hip_tsne_plan_affinities": (c_int, [_vp, _ip, _ip, _fp]),
    "fastembedr_hip_tsne_plan_attraction_info": (c_int, [_vp, _llp]),
    "fastembedr_hip_tsne_plan_profile": (
        c_int, [_vp, c_int, c_float, c_float, c_float, c_float, c_float, _dp]),
    "fastembedr_hip_tsne_plan_destroy": (None, [_vp]),
    "fastembedr_hip_tsne_batch_create": (c_int, [c_int, c_ll, _vpp]),
    "fastembedr_hip_tsne_job_bytes": (c_ll, [c_int, c_int, c_int, c_int]),
    "fastembedr_hip_tsne_batch_add": (c_int, [_vp, _vp, c_int, _ip]),
    "fastembedr_hip_tsne_batch_step": (c_int, [_vp, c_int]),
    "fastembedr_hip_tsne_batch_status": (c_int, [_vp, _ip, _ip]),
    "fastembedr_hip_tsne_batch_layout": (c_int, [_vp, c_int, _fp]),
    "fastembedr_hip_tsne_batch_remove": (c_int, [_vp, c_int]),
    "fastembedr_hip_tsne_batch_destroy": (None, [_vp]),
    "fastembedr_hip_session_create": (c_int, [_fp, c_int, c_int, c_ll, c_ll, _vpp]),
    "fastembedr_hip_session_submit_embed": (c_int, [_vp, _vp, _llp]),
    "fastembedr_hip_session_submit_transform": (
        c_int, [_vp, _ip, _fp, c_int, c_int, c_int, _fp, c_int, c_double, c_int, c_int, c_double,
                c_double, c_double, c_double, c_double, c_double, c_double, c_int, c_int, c_int,
                _llp]),
    "fastembedr_hip_session_submit_umap": (c_int, [_vp, _vp, _llp]),
    "fastembedr_hip_session_poll": (c_int, [_vp, c_ll, c_int, _ip, _fp, _dp]),
    "fastembedr_hip_session_stats": (c_int, [_vp, _llp]),
    "fastembedr_hip_session_destroy": (None, [_vp]),
    "fastembedr_hip_umap_optimize_coo": (
        c_int, [_ip, _ip, _fp, _fp, _fp, c_int, c_int, c_int, c_int, c_float, c_float, c_float,
                c_float, c_uint, c_int, _fp]),
    "fastembedr_hip_umap_optimize_coo_device": (
        c_int, [_vp, _vp, _vp, _vp, _fp, c_int, c_int, c_int, c_int, c_float, c_float, c_float,
                c_float, c_uint, c_int, _fp]),
    "fastembedr_hip_umap_optimize_coo_device_noalloc": (
        c_int, [_vp, _vp, _vp, _vp, _vp, c_int, c_int, c_int, c_int, c_float, c_float, c_float,
                c_float, c_uint, c_int, _vp]),
    "fastembedr_hip_umap_schedule": (
        c_int, [_ip, _ip, _fp, c_int, c_int, c_int, c_int, c_uint, c_int, _ip, _ip, _ip]),
    "fastembedr_hip_umap_batch_create": (c_int, [c_int, c_ll, _vpp]),
    "fastembedr_hip_umap_job_bytes": (c_ll, [c_int, c_int, c_int, c_int]),
    "fastembedr_hip_umap_batch_add": (c_int, [_vp, _vp, c_int, _ip]),
    "fastembedr_hip_umap_batch_step": (c_int, [_vp, c_int]),
    "fastembedr_hip_umap_batch_status": (c_int, [_vp, _ip, _ip]),
    "fastembedr_hip_umap_batch_layout": (c_int, [_vp, c_int, _fp]),
    "fastembedr_hip_umap_batch_remove": (c_int, [_vp, c_int]),
    "fastembedr_hip_umap_batch_destroy": (None, [_vp]),
    "fastembedr_hip_tsne_transform": (
        c_int, [_fp, c_int, _ip, _dp, c_int, c_int, c_int, _fp, c_int, c_double, c_int, c_int,
                c_double, c_double, c_double, c_double, c_double, c_double, c_double, c_int,
                c_int, c_int, _fp]),
    "fastembedr_hip_transform_reference_create": (c_int, [_fp, c_int, _vpp]),
    "fastembedr_hip_transform_reference_run": (
        c_int, [_vp, _ip, _fp, c_int, c_int, c_int, _fp, c_int, c_double, c_int, c_int, c_double,
                c_double, c_double, c_double, c_double, c_double, c_double, c_int, c_int, c_int,
                _fp]),
    "fastembedr_hip_transform_reference_destroy": (None, [_vp]),
    "fastembedr_hip_umap_graph_create": (
        c_int, [_vp, _vp, c_int, c_int, c_int, c_int, c_int, c_int, c_int, _vpp]),
    "fastembedr_hip_umap_graph_info": (c_int, [_vp, _llp, _fp]),
    "fastembedr_hip_umap_graph_export": (c_int, [_vp, _ip, _ip, _fp, _fp]),
    "fastembedr_hip_umap_graph_device": (c_int, [_vp, _vpp, _vpp, _vpp, _vpp]),
    "fastembedr_hip_umap_graph_destroy": (None, [_vp]),
    "fastembedr_hip_umap_graph_optimize": (
        c_int, [_vp, _fp, c_int, c_int, c_float, c_float, c_float, c_float, c_uint, c_int, _fp]),
    "fastembedr_hip_exact_structure_metrics": (
        c_int, [_dp, c_int, c_int, _dp, c_int, _ip, c_int, _dp]),
    "fastembedr_hip_knn_structure_score": (
        c_int, [_dp, _ip, _ip, _ip, c_int, c_int, c_int, c_int, c_int, c_int, c_int, _dp]),
    "fastembedr_hip_exact_knn": (
        c_int, [_vp, c_int, c_int, c_int, c_int, c_int, _vp, _vp, c_int, _ip]),
    "fastembedr_hip_exact_query_knn": (
        c_int, [_vp, c_int, _vp, c_int, c_int, c_int, c_int, c_int, _vp, _vp, c_int, _ip]),
}

_lib = None


class HipError(RuntimeError):
    """A C-ABI call returned failure; the message is the library's last error."""


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libfastembedr_hip.so is not built (run `python -m fastembedr_amd.build`); "
                "the HIP backend has no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    msg = lib().fastembedr_hip_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str = "HIP backend call") -> None:
    if rc != 0:
        raise HipError(f"{what} failed: {last_error()}")


def ptr(a, t):
    return a.ctypes.data_as(t)


def available() -> bool:
    return bool(lib().fastembedr_hip_available())
