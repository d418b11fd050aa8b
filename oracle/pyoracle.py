"""ctypes access to the CPU oracle (liboracle.so) and, when it was built, the
unchanged reference library (_ref/libdress_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product (dress-graph_amd/) never imports it.

Both libraries share the dress_graph_t layout; ``Lib("oracle")`` binds the
``ora_*`` restatement and ``Lib("ref")`` binds libdress's own ``dress_*``
symbols.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libdress_ref.so")

c_int_p = ctypes.POINTER(ctypes.c_int)
c_dbl_p = ctypes.POINTER(ctypes.c_double)
c_i64_p = ctypes.POINTER(ctypes.c_int64)


class DressGraph(ctypes.Structure):
    """dress_graph_t (libdress/include/dress/dress.h:43-75)."""
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


PG = ctypes.POINTER(DressGraph)

_libc = ctypes.CDLL(None)
_libc.malloc.restype = ctypes.c_void_p
_libc.malloc.argtypes = [ctypes.c_size_t]
_libc.free.argtypes = [ctypes.c_void_p]


def malloc_copy(arr, ctype):
    """malloc'd copy of a numpy array (the C API takes ownership)."""
    np_t = {ctypes.c_int: np.int32, ctypes.c_double: np.float64}[ctype]
    a = np.ascontiguousarray(arr, dtype=np_t)
    p = _libc.malloc(max(a.nbytes, 8))
    if not p:
        raise MemoryError
    ctypes.memmove(p, a.ctypes.data, a.nbytes)
    return ctypes.cast(p, ctypes.POINTER(ctype))


def as_np(ptr, n, dtype):
    if n == 0:
        return np.empty(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def flat_to_pairs(flat):
    """[bits, count, ...] int64 -> (values float64, counts int64)."""
    flat = np.asarray(flat, dtype=np.int64)
    return flat[0::2].view(np.float64).copy(), flat[1::2].copy()


def available(which):
    return os.path.isfile(ORACLE_SO if which == "oracle" else REF_SO)


class Lib:
    def __init__(self, which="oracle"):
        path = ORACLE_SO if which == "oracle" else REF_SO
        if not os.path.isfile(path):
            raise RuntimeError(f"{path} not built (make -C oracle {'ref' if which == 'ref' else ''})")
        self.which = which
        self.lib = ctypes.CDLL(path)
        p = "ora_" if which == "oracle" else "dress_"
        L = self.lib
        self.init = getattr(L, p + "init_graph")
        self.init.restype = PG
        self.init.argtypes = [ctypes.c_int, ctypes.c_int, c_int_p, c_int_p, c_dbl_p, c_dbl_p,
                              ctypes.c_int, ctypes.c_int]
        self.free_graph = getattr(L, p + "free_graph")
        self.free_graph.argtypes = [PG]
        fit_args = [PG, ctypes.c_int, ctypes.c_double, c_int_p, c_dbl_p]
        self.fit_seq = getattr(L, p + "fit")
        self.fit_seq.argtypes = fit_args
        self.fit_par = getattr(L, p + "fit_omp")
        self.fit_par.argtypes = fit_args
        self.get = getattr(L, p + "get")
        self.get.restype = ctypes.c_double
        self.get.argtypes = [PG, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                             ctypes.c_double]
        lifted = [PG, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_uint,
                  c_int_p, ctypes.c_int, ctypes.POINTER(c_dbl_p), c_i64_p,
                  ctypes.c_int, ctypes.c_int]
        if which == "oracle":
            self.delta = L.ora_delta_fit_flat
            self.delta.argtypes = lifted + [ctypes.c_int]
            self.nabla = L.ora_nabla_fit_flat
            self.nabla.argtypes = lifted + [ctypes.c_int]
            L.ora_graph_from_parts.restype = PG
            L.ora_graph_from_parts.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               c_int_p, c_int_p, c_dbl_p, c_dbl_p,
                                               c_int_p, c_int_p, c_int_p, c_dbl_p,
                                               c_int_p, c_int_p, c_int_p]
            L.ora_sample_picks.restype = ctypes.c_int
            L.ora_sample_picks.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_uint,
                                           ctypes.c_int, c_i64_p]
            L.ora_kbn_sorted_sum.restype = ctypes.c_double
            L.ora_kbn_sorted_sum.argtypes = [c_dbl_p, ctypes.c_int]
        else:
            self.delta = L.dress_delta_fit_strided_flat
            self.delta.argtypes = lifted
            self.nabla = L.dress_nabla_fit_strided_flat
            self.nabla.argtypes = lifted
            self.delta_omp = L.dress_delta_fit_omp_strided_flat
            self.delta_omp.argtypes = lifted
            self.delta_omp.restype = c_i64_p
        for f in (self.delta, self.nabla):
            f.restype = c_i64_p
        if which == "ref":
            for nm in ("dress_delta_binom", "dress_nabla_perm_count"):
                getattr(L, nm).restype = ctypes.c_int64
                getattr(L, nm).argtypes = [ctypes.c_int, ctypes.c_int]

    # -- construction -------------------------------------------------
    def graph(self, N, U, V, W=None, NW=None, variant=0, precompute=0):
        U = np.asarray(U)
        E = int(U.size)
        pW = malloc_copy(W, ctypes.c_double) if W is not None else c_dbl_p()
        pNW = malloc_copy(NW, ctypes.c_double) if NW is not None else c_dbl_p()
        g = self.init(int(N), E, malloc_copy(U, ctypes.c_int), malloc_copy(V, ctypes.c_int),
                      pW, pNW, int(variant), int(precompute))
        return g

    @staticmethod
    def csr(g):
        c = g.contents
        S = c.adj_offset[c.N]
        out = {
            "adj_offset": as_np(c.adj_offset, c.N + 1, np.int32),
            "adj_target": as_np(c.adj_target, S, np.int32),
            "adj_edge_idx": as_np(c.adj_edge_idx, S, np.int32),
            "edge_weight": as_np(c.edge_weight, c.E, np.float64),
            "max_degree": c.max_degree,
        }
        if c.precompute_intercepts:
            T = c.intercept_offset[c.E]
            out["intercept_offset"] = as_np(c.intercept_offset, c.E + 1, np.int32)
            out["intercept_edge_ux"] = as_np(c.intercept_edge_ux, T, np.int32)
            out["intercept_edge_vx"] = as_np(c.intercept_edge_vx, T, np.int32)
        return out

    # -- fit ------------------------------------------------------------
    def fit_graph(self, g, max_iterations=100, epsilon=1e-6, parallel=False):
        it = ctypes.c_int(-1)
        dl = ctypes.c_double(-1.0)
        (self.fit_par if parallel else self.fit_seq)(g, max_iterations, epsilon,
                                                      ctypes.byref(it), ctypes.byref(dl))
        return it.value, dl.value

    def fit(self, N, U, V, W=None, NW=None, variant=0, precompute=0,
            max_iterations=100, epsilon=1e-6, parallel=False):
        """Returns dict(edge_dress, vertex_dress, edge_weight, iterations, delta)."""
        g = self.graph(N, U, V, W, NW, variant, precompute)
        try:
            it, dl = self.fit_graph(g, max_iterations, epsilon, parallel)
            c = g.contents
            return {
                "edge_dress": as_np(c.edge_dress, c.E, np.float64),
                "vertex_dress": as_np(c.vertex_dress, c.N, np.float64),
                "edge_weight": as_np(c.edge_weight, c.E, np.float64),
                "iterations": it, "delta": dl,
            }
        finally:
            self.free_graph(g)

    # -- lifted ----------------------------------------------------------
    def lifted(self, kind, N, U, V, W=None, NW=None, variant=0, k=1,
               max_iterations=100, epsilon=1e-6, n_samples=0, seed=0,
               keep_multisets=False, compute_histogram=True, offset=0, stride=1,
               precompute=0, use_omp=False):
        """delta^k / nabla^k -> (values, counts, multisets or None, count)."""
        g = self.graph(N, U, V, W, NW, variant, precompute)
        try:
            hs = ctypes.c_int(0)
            ns = ctypes.c_int64(0)
            ms = c_dbl_p()
            args = [g, int(k), int(max_iterations), float(epsilon), int(n_samples),
                    ctypes.c_uint(seed),
                    ctypes.byref(hs) if compute_histogram else None,
                    1 if keep_multisets else 0, ctypes.byref(ms), ctypes.byref(ns),
                    int(offset), int(stride)]
            if self.which == "oracle":
                f = self.delta if kind == "delta" else self.nabla
                flat = f(*args, 1 if use_omp else 0)
            else:
                f = self.delta if kind == "delta" else self.nabla
                if kind == "delta" and use_omp:
                    f = self.delta_omp
                flat = f(*args)
            n = hs.value
            vals, cnts = (np.empty(0), np.empty(0, dtype=np.int64))
            if flat and n > 0:
                arr = np.ctypeslib.as_array(flat, shape=(n,)).copy()
                vals, cnts = flat_to_pairs(arr)
            if flat:
                _libc.free(ctypes.cast(flat, ctypes.c_void_p))
            mset = None
            if keep_multisets and ms:
                E = int(np.asarray(U).size)
                mset = np.ctypeslib.as_array(ms, shape=(ns.value * E,)).copy().reshape(ns.value, E)
                _libc.free(ctypes.cast(ms, ctypes.c_void_p))
            return vals, cnts, mset, ns.value
        finally:
            self.free_graph(g)

    def sample_picks(self, count, n_samples, seed, offset=0):
        out = np.empty(max(int(n_samples), 1), dtype=np.int64)
        n = self.lib.ora_sample_picks(int(count), int(n_samples), ctypes.c_uint(seed), int(offset),
                                      out.ctypes.data_as(c_i64_p))
        return out[:n]


def ref_graph_from_csr(lib, N, U, V, csr, variant=0):
    """A dress_graph_t for the compiled reference filled from construction
    arrays built elsewhere (CSR, variant weights, intercept index), exactly the
    fields libdress's dress_init_graph(..., precompute=1) sets
    (dress.c:378-427); every array is malloc'd so the reference's
    dress_free_graph releases it."""
    E = int(np.asarray(U).size)
    p = _libc.malloc(ctypes.sizeof(DressGraph))
    if not p:
        raise MemoryError
    g = ctypes.cast(p, PG)
    c = g.contents
    c.variant = int(variant)
    c.N = int(N)
    c.E = E
    c.U = malloc_copy(U, ctypes.c_int)
    c.V = malloc_copy(V, ctypes.c_int)
    c.adj_offset = malloc_copy(csr["adj_offset"], ctypes.c_int)
    c.adj_target = malloc_copy(csr["adj_target"], ctypes.c_int)
    c.adj_edge_idx = malloc_copy(csr["adj_edge_idx"], ctypes.c_int)
    c.max_degree = int(csr["max_degree"])
    c.W = c_dbl_p()
    c.edge_weight = malloc_copy(csr["edge_weight"], ctypes.c_double)
    c.edge_dress = malloc_copy(np.ones(E), ctypes.c_double)
    c.edge_dress_next = malloc_copy(np.ones(E), ctypes.c_double)
    c.NW = malloc_copy(np.ones(N), ctypes.c_double)
    c.vertex_dress = malloc_copy(np.zeros(N), ctypes.c_double)
    c.precompute_intercepts = 1
    c.intercept_offset = malloc_copy(csr["intercept_offset"], ctypes.c_int)
    c.intercept_edge_ux = malloc_copy(csr["intercept_edge_ux"], ctypes.c_int)
    c.intercept_edge_vx = malloc_copy(csr["intercept_edge_vx"], ctypes.c_int)
    return g
