"""Loader and ctypes bindings of libdress_hip.so.

Mirrors the symbol binding of the reference's CUDA selector
(python/src/dress/cuda/__init__.py:142-193): the core API
(dress_init_graph / dress_free_graph) and the *_hip entry points are bound
from the one shared library.  There is no fallback: if the library is
missing or fails to load, every call raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdress_hip.so")

c_int_p = ctypes.POINTER(ctypes.c_int)
c_dbl_p = ctypes.POINTER(ctypes.c_double)
c_i64_p = ctypes.POINTER(ctypes.c_int64)


class DressGraph(ctypes.Structure):
    """dress_graph_t, field for field (include/dress/dress.h)."""
    _fields_ = [
        ("variant", ctypes.c_int), ("N", ctypes.c_int), ("E", ctypes.c_int),
        ("U", c_int_p), ("V", c_int_p),
        ("adj_offset", c_int_p), ("adj_target", c_int_p), ("adj_edge_idx", c_int_p),
        ("max_degree", ctypes.c_int),
        ("W", c_dbl_p), ("edge_weight", c_dbl_p), ("edge_dress", c_dbl_p),
        ("edge_dress_next", c_dbl_p),
        ("NW", c_dbl_p), ("vertex_dress", c_dbl_p),
        ("precompute_intercepts", ctypes.c_int),
        ("intercept_offset", c_int_p), ("intercept_edge_ux", c_int_p),
        ("intercept_edge_vx", c_int_p),
    ]


class HistPair(ctypes.Structure):
    _fields_ = [("value", ctypes.c_double), ("count", ctypes.c_int64)]


class FitStats(ctypes.Structure):
    _fields_ = [("iterations_executed", ctypes.c_int), ("fp32_iterations", ctypes.c_int),
                ("fp64_reruns", ctypes.c_int), ("device_ms", ctypes.c_double),
                ("intersections", ctypes.c_int64), ("half_edges", ctypes.c_int64)]


class LiftedStats(ctypes.Structure):
    _fields_ = [("subgraphs", ctypes.c_int64), ("values", ctypes.c_int64),
                ("fp64_reruns", ctypes.c_int), ("device_ms", ctypes.c_double)]


class LiftedBatch(ctypes.Structure):
    _fields_ = [("n_graphs", ctypes.c_int),
                ("hist_offset", ctypes.POINTER(ctypes.c_int64)),
                ("hist_value", ctypes.POINTER(ctypes.c_double)),
                ("hist_count", ctypes.POINTER(ctypes.c_int64)),
                ("qhist_offset", ctypes.POINTER(ctypes.c_int64)),
                ("qhist_key", ctypes.POINTER(ctypes.c_int64)),
                ("qhist_count", ctypes.POINTER(ctypes.c_int64)),
                ("qbound_offset", ctypes.POINTER(ctypes.c_int64)),
                ("qbound_key", ctypes.POINTER(ctypes.c_int64)),
                ("qbound_count", ctypes.POINTER(ctypes.c_int64)),
                ("subgraphs", ctypes.c_int64), ("values", ctypes.c_int64),
                ("fp64_reruns", ctypes.c_int), ("device_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("work", ctypes.c_int64 * 4)]


class FitBatch(ctypes.Structure):
    _fields_ = [("n_graphs", ctypes.c_int), ("iterations", ctypes.POINTER(ctypes.c_int)),
                ("delta", ctypes.POINTER(ctypes.c_double)), ("edge_dress", ctypes.POINTER(ctypes.c_double)),
                ("vertex_offset", ctypes.POINTER(ctypes.c_int64)),
                ("vertex_dress", ctypes.POINTER(ctypes.c_double)), ("fp64_reruns", ctypes.c_int),
                ("device_ms", ctypes.c_double), ("total_ms", ctypes.c_double)]


class ServeRequest(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("k", ctypes.c_int), ("N", ctypes.c_int), ("E", ctypes.c_int),
                ("U", ctypes.POINTER(ctypes.c_int)), ("V", ctypes.POINTER(ctypes.c_int)),
                ("W", ctypes.POINTER(ctypes.c_double)), ("variant", ctypes.c_int),
                ("max_iterations", ctypes.c_int), ("epsilon", ctypes.c_double),
                ("precision", ctypes.c_int), ("arrival", ctypes.c_double)]


class ServeResponse(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int64), ("status", ctypes.c_int), ("kind", ctypes.c_int),
                ("N", ctypes.c_int), ("E", ctypes.c_int), ("iterations", ctypes.c_int),
                ("delta", ctypes.c_double), ("edge_dress", ctypes.POINTER(ctypes.c_double)),
                ("vertex_dress", ctypes.POINTER(ctypes.c_double)), ("hist", ctypes.POINTER(HistPair)),
                ("hist_size", ctypes.c_int), ("subgraphs", ctypes.c_int64), ("batch_size", ctypes.c_int),
                ("footprint", ctypes.c_int64), ("arrival", ctypes.c_double), ("admitted", ctypes.c_double),
                ("finished", ctypes.c_double), ("error", ctypes.c_char * 192)]


class ServeStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "submitted", "completed", "failed", "rejected", "queued", "running", "max_queued",
        "budget_bytes", "reserved_bytes", "peak_reserved_bytes", "lifted_launches", "lifted_requests",
        "rejected_host", "chunked_requests", "device_failed")]


PG = ctypes.POINTER(DressGraph)

_lib = None
_libc = ctypes.CDLL(None)
_libc.malloc.restype = ctypes.c_void_p
_libc.malloc.argtypes = [ctypes.c_size_t]
_libc.free.argtypes = [ctypes.c_void_p]


class HipError(RuntimeError):
    pass


def _lifted_args():
    return [PG, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_uint,
            c_int_p, ctypes.c_int, ctypes.POINTER(c_dbl_p), c_i64_p]


def _bind(L):
    L.dress_init_graph.restype = PG
    L.dress_init_graph.argtypes = [ctypes.c_int, ctypes.c_int, c_int_p, c_int_p, c_dbl_p, c_dbl_p,
                                   ctypes.c_int, ctypes.c_int]
    L.dress_free_graph.restype = None
    L.dress_free_graph.argtypes = [PG]
    for name in ("dress_fit", "dress_fit_hip"):
        f = getattr(L, name)
        f.restype = None
        f.argtypes = [PG, ctypes.c_int, ctypes.c_double, c_int_p, c_dbl_p]
    L.dress_fit_hip_ex.restype = ctypes.c_int
    L.dress_fit_hip_ex.argtypes = [PG, ctypes.c_int, ctypes.c_double, ctypes.c_int, c_int_p, c_dbl_p]
    L.dress_hip_last_error.restype = ctypes.c_char_p
    L.dress_hip_last_error.argtypes = []
    L.dress_hip_device_count.restype = ctypes.c_int
    L.dress_hip_set_precision.argtypes = [ctypes.c_int]
    L.dress_hip_get_precision.restype = ctypes.c_int
    L.dress_hip_last_fit_stats.argtypes = [ctypes.POINTER(FitStats)]
    L.dress_hip_graph_info.restype = ctypes.c_int
    L.dress_hip_graph_info.argtypes = [PG, c_i64_p, ctypes.c_int]
    L.dress_hip_resident_reset.restype = ctypes.c_int
    L.dress_hip_resident_reset.argtypes = [PG]
    L.dress_hip_fit_resident.restype = ctypes.c_int
    L.dress_hip_fit_resident.argtypes = [PG, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                         c_int_p, c_dbl_p]
    L.dress_hip_resident_fill.restype = ctypes.c_int
    L.dress_hip_resident_fill.argtypes = [PG, ctypes.c_double]
    L.dress_hip_time_phases.restype = ctypes.c_int
    L.dress_hip_time_phases.argtypes = [PG, ctypes.c_int, c_dbl_p]
    L.dress_hip_copy_probe.restype = ctypes.c_int
    L.dress_hip_copy_probe.argtypes = [ctypes.c_int64, ctypes.c_int, c_dbl_p]
    L.dress_hip_gather_probe.restype = ctypes.c_int
    L.dress_hip_gather_probe.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int, c_dbl_p]
    L.dress_hip_lds_probe.restype = ctypes.c_int
    L.dress_hip_lds_probe.argtypes = [ctypes.c_int, c_dbl_p, c_dbl_p]
    L.dress_hip_resident_download.restype = ctypes.c_int
    L.dress_hip_resident_download.argtypes = [PG]
    L.dress_hip_set_lifted_mode.argtypes = [ctypes.c_int]
    L.dress_hip_enumerate_subgraphs.restype = ctypes.c_int64
    L.dress_hip_enumerate_subgraphs.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_uint, ctypes.c_int, ctypes.c_int, c_i64_p,
                                                c_int_p, ctypes.c_int64]
    L.dress_hip_last_lifted_stats.argtypes = [ctypes.POINTER(LiftedStats)]
    L.dress_hip_lifted_batch.restype = ctypes.c_int
    L.dress_hip_lifted_batch.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, c_int_p, c_i64_p,
                                         c_int_p, c_int_p, c_dbl_p, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                         ctypes.POINTER(LiftedBatch)]
    L.dress_hip_lifted_batch_free.argtypes = [ctypes.POINTER(LiftedBatch)]
    L.dress_delta_binom.restype = ctypes.c_int64
    L.dress_delta_binom.argtypes = [ctypes.c_int, ctypes.c_int]
    L.dress_nabla_perm_count.restype = ctypes.c_int64
    L.dress_nabla_perm_count.argtypes = [ctypes.c_int, ctypes.c_int]
    L.dress_get.restype = ctypes.c_double
    L.dress_get.argtypes = [PG, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                            ctypes.c_double]
    L.dress_hip_get_batch.restype = ctypes.c_int
    L.dress_hip_get_batch.argtypes = [PG, ctypes.c_int, c_int_p, c_int_p, c_dbl_p, ctypes.c_int,
                                      ctypes.c_double, c_dbl_p]
    L.dress_hip_get_batch_resident.restype = ctypes.c_int
    L.dress_hip_get_batch_resident.argtypes = [PG, ctypes.c_int, c_int_p, c_int_p, c_dbl_p, ctypes.c_int,
                                               ctypes.c_double, c_dbl_p]
    L.dress_hip_get_batch_resident_noalloc.restype = ctypes.c_int
    L.dress_hip_get_batch_resident_noalloc.argtypes = [PG, ctypes.c_int, c_int_p, c_int_p, c_dbl_p, ctypes.c_int,
                                                       ctypes.c_double, c_dbl_p]
    L.dress_hip_fit_batch_eligible.restype = ctypes.c_int
    L.dress_hip_fit_batch_eligible.argtypes = [ctypes.c_int, ctypes.c_int]
    L.dress_hip_fit_batch.restype = ctypes.c_int
    L.dress_hip_fit_batch.argtypes = [ctypes.c_int, c_int_p, c_i64_p, c_int_p, c_int_p, c_dbl_p, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.POINTER(FitBatch)]
    L.dress_hip_fit_batch_free.restype = None
    L.dress_hip_fit_batch_free.argtypes = [ctypes.POINTER(FitBatch)]
    L.dress_hip_clock.restype = ctypes.c_double
    L.dress_hip_clock.argtypes = []
    L.dress_hip_server_create.restype = ctypes.c_void_p
    L.dress_hip_server_create.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int]
    L.dress_hip_server_submit.restype = ctypes.c_int64
    L.dress_hip_server_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(ServeRequest)]
    L.dress_hip_server_next.restype = ctypes.c_int
    L.dress_hip_server_next.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(ServeResponse)]
    L.dress_hip_response_free.restype = None
    L.dress_hip_response_free.argtypes = [ctypes.POINTER(ServeResponse)]
    L.dress_hip_server_stats.restype = None
    L.dress_hip_server_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ServeStats)]
    L.dress_hip_server_inject_error.restype = ctypes.c_int
    L.dress_hip_server_inject_error.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.dress_hip_server_destroy.restype = None
    L.dress_hip_server_destroy.argtypes = [ctypes.c_void_p]
    strided = _lifted_args() + [ctypes.c_int, ctypes.c_int]
    for name, args, res in (
            ("dress_delta_fit_hip", _lifted_args(), ctypes.POINTER(HistPair)),
            ("dress_delta_fit_hip_strided", strided, ctypes.POINTER(HistPair)),
            ("dress_delta_fit_hip_flat", _lifted_args(), c_i64_p),
            ("dress_delta_fit_hip_strided_flat", strided, c_i64_p),
            ("dress_nabla_fit_hip", _lifted_args(), ctypes.POINTER(HistPair)),
            ("dress_nabla_fit_hip_strided", strided, ctypes.POINTER(HistPair)),
            ("dress_nabla_fit_hip_flat", _lifted_args(), c_i64_p),
            ("dress_nabla_fit_hip_strided_flat", strided, c_i64_p),
            ("dress_delta_fit", _lifted_args(), ctypes.POINTER(HistPair)),
            ("dress_delta_fit_flat", _lifted_args(), c_i64_p),
            ("dress_delta_fit_strided", strided, ctypes.POINTER(HistPair)),
            ("dress_delta_fit_strided_flat", strided, c_i64_p),
            ("dress_nabla_fit", _lifted_args(), ctypes.POINTER(HistPair)),
            ("dress_nabla_fit_flat", _lifted_args(), c_i64_p),
            ("dress_nabla_fit_strided_flat", strided, c_i64_p)):
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def lib():
    """The loaded libdress_hip.so (raises if it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            raise HipError(f"{LIB_PATH} not found: build it with __graft_entry__.build() "
                           "(make -C dress-graph_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        _bind(L)
        _lib = L
    return _lib


def last_error():
    return lib().dress_hip_last_error().decode()


def check(ok, what):
    if not ok:
        raise HipError(f"{what} failed: {last_error()}")


def malloc_copy(arr, ctype):
    """malloc'd copy of a numpy array; the C API takes ownership."""
    np_t = {ctypes.c_int: np.int32, ctypes.c_double: np.float64}[ctype]
    a = np.ascontiguousarray(arr, dtype=np_t)
    p = _libc.malloc(max(a.nbytes, 8))
    if not p:
        raise MemoryError("malloc failed")
    ctypes.memmove(p, a.ctypes.data, a.nbytes)
    return ctypes.cast(p, ctypes.POINTER(ctype))


def c_free(ptr):
    if ptr:
        _libc.free(ctypes.cast(ptr, ctypes.c_void_p))


def as_np(ptr, n, dtype):
    if n <= 0:
        return np.empty(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)
