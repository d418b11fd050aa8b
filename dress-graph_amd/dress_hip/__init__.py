"""dress_hip -- the MI355X (HIP) backend of DRESS, selected by import.

Mirrors the reference's ``dress.cuda`` module
(python/src/dress/cuda/__init__.py:1-643)::

    from dress_hip import fit, delta_fit, nabla_fit, DRESS

    r = fit(N, sources, targets)                 # runs on the GPU
    h = delta_fit(N, sources, targets, k=1)      # one batched GPU launch

Same parameters, defaults and result fields as ``dress.fit`` /
``dress.delta_fit`` / ``dress.nabla_fit``.  Every call goes through
libdress_hip.so's C ABI; there is no CPU path.  ``is_available()`` probes
for the library and a GPU.
"""
import ctypes
from collections.abc import Sequence
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from . import _lib
from ._lib import HipError, c_dbl_p, c_int_p

UNDIRECTED = 0
DIRECTED = 1
FORWARD = 2
BACKWARD = 3

FP64 = 0
FP32_FP64 = 1


@dataclass
class DRESSResult:
    """Fields of the reference's DRESSResult (python/src/dress/core.py)."""
    sources: List[int]
    targets: List[int]
    edge_dress: np.ndarray
    edge_weight: np.ndarray
    vertex_dress: np.ndarray
    iterations: int
    delta: float


@dataclass
class DeltaDRESSResult:
    histogram: List[Tuple[float, int]]
    multisets: Optional[np.ndarray]
    num_subgraphs: int


@dataclass
class NablaDRESSResult:
    histogram: List[Tuple[float, int]]
    multisets: Optional[np.ndarray]
    num_tuples: int


@dataclass
class FitStats:
    iterations_executed: int
    fp32_iterations: int
    fp64_reruns: int
    device_ms: float
    intersections: int
    half_edges: int


def is_available():
    try:
        return _lib.lib().dress_hip_device_count() > 0
    except (HipError, OSError):
        return False


def set_precision(precision):
    _lib.lib().dress_hip_set_precision(int(precision))


def get_precision():
    return _lib.lib().dress_hip_get_precision()


def last_fit_stats():
    s = _lib.FitStats()
    _lib.lib().dress_hip_last_fit_stats(ctypes.byref(s))
    return FitStats(s.iterations_executed, s.fp32_iterations, s.fp64_reruns, s.device_ms,
                    s.intersections, s.half_edges)


@dataclass
class LiftedBatchResult:
    """Per-graph outputs of :func:`lifted_batch`."""
    histograms: Sequence        # per graph: exact (values, counts)
    quantised: Sequence         # per graph: floor(d/eps) keys, counts
    boundary: Sequence          # per graph: keys, counts of values near a bin edge
    subgraphs: int
    values: int
    fp64_reruns: int
    device_ms: float
    total_ms: float
    work: Tuple[int, int, int, int] = (0, 0, 0, 0)   # DRESS_HIP_LIFT_WORK=1: iterations x entries,
                                                     # half-edges, edges, vertices


def copy_probe(n_bytes=1 << 30, reps=5):
    """Mean ms of a plain device-to-device copy kernel over n_bytes."""
    ms = ctypes.c_double()
    _lib.check(_lib.lib().dress_hip_copy_probe(int(n_bytes), int(reps), ctypes.byref(ms)) == 0,
               "dress_hip_copy_probe")
    return ms.value


def gather_probe(src_bytes=1 << 30, n_gathers=1 << 26, reps=5):
    """Mean ms of n_gathers random 8-byte device reads from src_bytes."""
    ms = ctypes.c_double()
    _lib.check(_lib.lib().dress_hip_gather_probe(int(src_bytes), int(n_gathers), int(reps), ctypes.byref(ms)) == 0,
               "dress_hip_gather_probe")
    return ms.value


def lds_probe(rounds=4096):
    """(ms, bytes) of one LDS read kernel over the whole device."""
    ms = ctypes.c_double()
    nb = ctypes.c_double()
    _lib.check(_lib.lib().dress_hip_lds_probe(int(rounds), ctypes.byref(ms), ctypes.byref(nb)) == 0,
               "dress_hip_lds_probe")
    return ms.value, nb.value


def set_lifted_mode(mode):
    """0: batched persistent kernel when the graph fits its LDS state, else
    fits on the graph's device copy, many subgraphs per pass (default);
    1: one device fit per rebuilt subgraph; 2: always one-at-a-time fits on
    the device copy; 3: always multi-subgraph passes where they apply."""
    _lib.lib().dress_hip_set_lifted_mode(int(mode))


def enumerate_subgraphs(kind, n, k, n_samples=0, seed=0, offset=0, stride=1):
    """Host-side: (indices, vertex tuples) a lifted call would process."""
    L = _lib.lib()
    kd = 0 if kind == "delta" else 1
    cnt = L.dress_hip_enumerate_subgraphs(kd, n, k, n_samples, ctypes.c_uint(seed), offset, stride,
                                          None, None, 0)
    _lib.check(cnt >= 0, "dress_hip_enumerate_subgraphs")
    sel = np.zeros(max(cnt, 1), dtype=np.int64)
    verts = np.zeros(max(cnt, 1) * k, dtype=np.int32)
    L.dress_hip_enumerate_subgraphs(kd, n, k, n_samples, ctypes.c_uint(seed), offset, stride,
                                    sel.ctypes.data_as(_lib.c_i64_p), verts.ctypes.data_as(c_int_p), cnt)
    return sel[:cnt], verts[:cnt * k].reshape(cnt, k)


def last_lifted_stats():
    s = _lib.LiftedStats()
    _lib.lib().dress_hip_last_lifted_stats(ctypes.byref(s))
    return {"subgraphs": s.subgraphs, "values": s.values, "fp64_reruns": s.fp64_reruns,
            "device_ms": s.device_ms}


@dataclass(frozen=True)
class GraphBatch:
    """Graphs of one batch in the flat arrays the batch entry points take:
    vertex counts ``N``, edge ``offsets`` (graph i owns edges
    ``offsets[i]:offsets[i+1]``), endpoints ``U``/``V`` and optional weights
    ``W``.  Made by :func:`pack_graphs`; pass it to :func:`lifted_batch` or
    :func:`fit_batch` in place of the list of graphs."""
    N: np.ndarray
    offsets: np.ndarray
    U: np.ndarray
    V: np.ndarray
    W: Optional[np.ndarray] = None

    def __len__(self):
        return len(self.N)


def pack_graphs(graphs, weights=None):
    """Packs a list of ``(n, U, V)`` (and optional per-graph weight arrays)
    into a :class:`GraphBatch`."""
    if isinstance(graphs, GraphBatch):
        if weights is not None:
            raise ValueError("a GraphBatch carries its own weights")
        return graphs
    n = len(graphs)
    Ns = np.fromiter((g[0] for g in graphs), dtype=np.int32, count=n)
    offs = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.fromiter((len(g[1]) for g in graphs), dtype=np.int64, count=n), out=offs[1:])
    if n:
        U = np.ascontiguousarray(np.concatenate([g[1] for g in graphs]), dtype=np.int32)
        V = np.ascontiguousarray(np.concatenate([g[2] for g in graphs]), dtype=np.int32)
    else:
        U = V = np.zeros(0, np.int32)
    if U.size != offs[-1] or V.size != offs[-1]:
        raise ValueError("every graph needs as many targets as sources")
    W = None
    if weights is not None:
        W = np.ascontiguousarray(np.concatenate([np.asarray(w, dtype=np.float64) for w in weights]),
                                 dtype=np.float64)
        if W.size != offs[-1]:
            raise ValueError("every graph needs one weight per edge")
    return GraphBatch(Ns, offs, U, V, W)


def lifted_batch(graphs, k=1, kind="delta", variant=UNDIRECTED, max_iterations=100, epsilon=1e-6,
                 precision=FP64, qtol=None, weights=None):
    """delta^k / nabla^k of many graphs in one batched launch.

    ``graphs`` is a list of ``(n, U, V)``; ``weights`` an optional list of
    per-graph weight arrays.  ``qtol`` (e.g. 1e-9) asks for the
    floor(d/epsilon) tables as well, with values within ``qtol`` of a bin
    edge counted apart; ``None`` leaves them empty.  ``graphs`` may also be
    a :class:`GraphBatch` (then ``weights`` are its own).  Returns a
    :class:`LiftedBatchResult`."""
    L = _lib.lib()
    b = pack_graphs(graphs, weights)
    n, Ns, offs, U, V, W = len(b), b.N, b.offsets, b.U, b.V, b.W
    r = _lib.LiftedBatch()
    rc = L.dress_hip_lifted_batch(0 if kind == "delta" else 1, int(k), n,
                                  Ns.ctypes.data_as(c_int_p), offs.ctypes.data_as(_lib.c_i64_p),
                                  U.ctypes.data_as(c_int_p), V.ctypes.data_as(c_int_p),
                                  W.ctypes.data_as(c_dbl_p) if W is not None else c_dbl_p(),
                                  int(variant), int(max_iterations), float(epsilon), int(precision),
                                  -1.0 if qtol is None else float(qtol), ctypes.byref(r))
    _lib.check(rc == 0, "dress_hip_lifted_batch")
    # the result arrays stay where the library wrote them: numpy views own
    # the block and release it (dress_hip_lifted_batch_free) with the last view
    owner = _BatchOwner(r)

    def split(off_p, a_p, b_p, typestr_a):
        off = np.ctypeslib.as_array(off_p, shape=(n + 1,)).copy()
        tot = int(off[-1])
        return _PerGraph(off, _owned_view(a_p, tot, typestr_a, owner), _owned_view(b_p, tot, "<i8", owner))
    hists = split(r.hist_offset, r.hist_value, r.hist_count, "<f8")
    quant = split(r.qhist_offset, r.qhist_key, r.qhist_count, "<i8")
    bound = split(r.qbound_offset, r.qbound_key, r.qbound_count, "<i8")
    return LiftedBatchResult(hists, quant, bound, r.subgraphs, r.values, r.fp64_reruns, r.device_ms, r.total_ms,
                             tuple(int(x) for x in r.work))


class _PerGraph(Sequence):
    """Per-graph (a, b) views of two concatenated result arrays, made on
    access: item i is (a[off[i]:off[i+1]], b[off[i]:off[i+1]])."""

    def __init__(self, off, a, b):
        self._off, self._a, self._b = off, a, b

    def __len__(self):
        return len(self._off) - 1

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[j] for j in range(*i.indices(len(self)))]
        n = len(self)
        if i < 0:
            i += n
        if not 0 <= i < n:
            raise IndexError(i)
        lo, hi = int(self._off[i]), int(self._off[i + 1])
        return self._a[lo:hi], self._b[lo:hi]


class _BatchOwner:
    """Frees a dress_hip_lifted_batch_t when the last array viewing it goes."""

    def __init__(self, r):
        self.r = r

    def __del__(self):
        try:
            _lib.lib().dress_hip_lifted_batch_free(ctypes.byref(self.r))
        except Exception:
            pass


class _CFree:
    """Frees one malloc'd block with libc free() when the last view goes."""

    def __init__(self, ptr):
        self.addr = ctypes.cast(ptr, ctypes.c_void_p).value

    def __del__(self):
        try:
            _lib.c_free(ctypes.c_void_p(self.addr))
        except Exception:
            pass


class _OwnedBuffer:
    def __init__(self, addr, count, typestr, owner):
        self.__array_interface__ = {"data": (addr, False), "shape": (count,), "typestr": typestr, "version": 3}
        self._owner = owner


def _owned_view(ptr, count, typestr, owner):
    if count == 0 or not ptr:
        return np.zeros(0, dtype=np.dtype(typestr))
    return np.asarray(_OwnedBuffer(ctypes.cast(ptr, ctypes.c_void_p).value, count, typestr, owner))


@dataclass
class FitBatchResult:
    iterations: np.ndarray
    delta: np.ndarray
    edge_dress: List[np.ndarray]
    vertex_dress: List[np.ndarray]
    fp64_reruns: int
    device_ms: float
    total_ms: float


def fit_batch_eligible(n_vertices, n_edges):
    """True when a graph of this size runs in the batched (LDS) kernel."""
    return bool(_lib.lib().dress_hip_fit_batch_eligible(int(n_vertices), int(n_edges)))


def fit_batch(graphs, variant=UNDIRECTED, max_iterations=100, epsilon=1e-6, precision=FP64, weights=None):
    """Whole-graph fits of many small graphs in one launch; ``graphs`` is a
    list of ``(n, U, V)`` or a :class:`GraphBatch`.  Per graph the results
    equal a single fit."""
    L = _lib.lib()
    b = pack_graphs(graphs, weights)
    n, Ns, offs, U, V, W = len(b), b.N, b.offsets, b.U, b.V, b.W
    r = _lib.FitBatch()
    rc = L.dress_hip_fit_batch(n, Ns.ctypes.data_as(c_int_p), offs.ctypes.data_as(_lib.c_i64_p),
                               U.ctypes.data_as(c_int_p), V.ctypes.data_as(c_int_p),
                               W.ctypes.data_as(c_dbl_p) if W is not None else c_dbl_p(),
                               int(variant), int(max_iterations), float(epsilon), int(precision), ctypes.byref(r))
    _lib.check(rc == 0, "dress_hip_fit_batch")
    try:
        totE, totN = int(offs[-1]), int(Ns.sum())
        it = np.ctypeslib.as_array(r.iterations, shape=(max(n, 1),))[:n].copy()
        dl = np.ctypeslib.as_array(r.delta, shape=(max(n, 1),))[:n].copy()
        ed = np.ctypeslib.as_array(r.edge_dress, shape=(max(totE, 1),))[:totE].copy()
        vo = np.ctypeslib.as_array(r.vertex_offset, shape=(n + 1,)).copy()
        vd = np.ctypeslib.as_array(r.vertex_dress, shape=(max(totN, 1),))[:totN].copy()
        return FitBatchResult(it, dl, [ed[offs[i]:offs[i + 1]] for i in range(n)],
                              [vd[vo[i]:vo[i + 1]] for i in range(n)], r.fp64_reruns, r.device_ms, r.total_ms)
    finally:
        L.dress_hip_fit_batch_free(ctypes.byref(r))


class Graph:
    """Owns one dress_graph_t built by libdress_hip.so (device-resident)."""

    def __init__(self, n_vertices, sources, targets, weights=None, vertex_weights=None,
                 variant=UNDIRECTED, precompute_intercepts=False):
        L = _lib.lib()
        U = np.ascontiguousarray(sources, dtype=np.int32)
        V = np.ascontiguousarray(targets, dtype=np.int32)
        if U.shape != V.shape:
            raise ValueError("sources and targets differ in length")
        self.N = int(n_vertices)
        self.E = int(U.size)
        pW = _lib.malloc_copy(weights, ctypes.c_double) if weights is not None else c_dbl_p()
        pNW = _lib.malloc_copy(vertex_weights, ctypes.c_double) if vertex_weights is not None else c_dbl_p()
        g = L.dress_init_graph(self.N, self.E, _lib.malloc_copy(U, ctypes.c_int),
                               _lib.malloc_copy(V, ctypes.c_int), pW, pNW, int(variant),
                               1 if precompute_intercepts else 0)
        if not g:
            raise HipError(f"dress_init_graph failed: {_lib.last_error()}")
        self.g = g

    def close(self):
        if getattr(self, "g", None):
            _lib.lib().dress_free_graph(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- views of the host struct --------------------------------------
    @property
    def edge_dress(self):
        return _lib.as_np(self.g.contents.edge_dress, self.E, np.float64)

    @property
    def vertex_dress(self):
        return _lib.as_np(self.g.contents.vertex_dress, self.N, np.float64)

    @property
    def edge_weight(self):
        return _lib.as_np(self.g.contents.edge_weight, self.E, np.float64)

    def csr(self):
        c = self.g.contents
        S = c.adj_offset[c.N] if c.N > 0 else 0
        out = {
            "adj_offset": _lib.as_np(c.adj_offset, c.N + 1, np.int32),
            "adj_target": _lib.as_np(c.adj_target, S, np.int32),
            "adj_edge_idx": _lib.as_np(c.adj_edge_idx, S, np.int32),
            "edge_weight": _lib.as_np(c.edge_weight, c.E, np.float64),
            "max_degree": c.max_degree,
        }
        if c.precompute_intercepts:
            T = c.intercept_offset[c.E]
            out["intercept_offset"] = _lib.as_np(c.intercept_offset, c.E + 1, np.int32)
            out["intercept_edge_ux"] = _lib.as_np(c.intercept_edge_ux, T, np.int32)
            out["intercept_edge_vx"] = _lib.as_np(c.intercept_edge_vx, T, np.int32)
        return out

    def set_edge_dress(self, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        ctypes.memmove(self.g.contents.edge_dress, a.ctypes.data, 8 * self.E)

    # -- fits -------------------------------------------------------------
    def fit(self, max_iterations=100, epsilon=1e-6, precision=None):
        it = ctypes.c_int(0)
        dl = ctypes.c_double(0.0)
        L = _lib.lib()
        if precision is None:
            L.dress_fit_hip(self.g, int(max_iterations), float(epsilon), ctypes.byref(it),
                            ctypes.byref(dl))
            if it.value == -1 and _lib.last_error():
                raise HipError(f"dress_fit_hip failed: {_lib.last_error()}")
        else:
            r = L.dress_fit_hip_ex(self.g, int(max_iterations), float(epsilon), int(precision),
                                   ctypes.byref(it), ctypes.byref(dl))
            _lib.check(r == 0, "dress_fit_hip_ex")
        return it.value, dl.value

    def info(self):
        """Device graph shape: N, E, S, T, max degree, common weight, per-bin
        edge/vertex counts and intersection entries."""
        v = np.zeros(21, dtype=np.int64)
        r = _lib.lib().dress_hip_graph_info(self.g, v.ctypes.data_as(_lib.c_i64_p), 21)
        _lib.check(r == 0, "dress_hip_graph_info")
        return {"N": int(v[0]), "E": int(v[1]), "S": int(v[2]), "T": int(v[3]),
                "max_degree": int(v[4]), "packed_index": bool(v[5]),
                "edge_bins": v[6:11].tolist(), "vertex_bins": v[11:16].tolist(),
                "bin_terms": v[16:21].tolist()}

    def resident_reset(self):
        _lib.check(_lib.lib().dress_hip_resident_reset(self.g) == 0, "dress_hip_resident_reset")

    def fit_resident(self, max_iterations=100, epsilon=1e-6, precision=FP64):
        it = ctypes.c_int(0)
        dl = ctypes.c_double(0.0)
        r = _lib.lib().dress_hip_fit_resident(self.g, int(max_iterations), float(epsilon),
                                              int(precision), ctypes.byref(it), ctypes.byref(dl))
        _lib.check(r == 0, "dress_hip_fit_resident")
        return it.value, dl.value

    def resident_fill(self, value=1.0):
        _lib.check(_lib.lib().dress_hip_resident_fill(self.g, float(value)) == 0, "dress_hip_resident_fill")

    def time_phases(self, reps=3):
        """Average ms per kernel of one FP64 iteration: (vertex[5], edge[5])
        with every bin launched alone, and (vertex, edge) as fits launch
        them (the vertex phase one kernel, the edge bins back to back)."""
        ms = np.zeros(12, dtype=np.float64)
        r = _lib.lib().dress_hip_time_phases(self.g, int(reps), ms.ctypes.data_as(c_dbl_p))
        _lib.check(r == 0, "dress_hip_time_phases")
        return ms[:5], ms[5:10], ms[10:12]

    def resident_download(self):
        _lib.check(_lib.lib().dress_hip_resident_download(self.g) == 0, "dress_hip_resident_download")

    def get(self, u, v, max_iterations=100, epsilon=1e-6, edge_weight=1.0):
        L = _lib.lib()
        val = L.dress_get(self.g, int(u), int(v), int(max_iterations), float(epsilon),
                          float(edge_weight))
        err = _lib.last_error()
        if err:
            raise HipError(f"dress_get failed: {err}")
        return val

    def get_batch(self, u, v, max_iterations=100, epsilon=1e-6, edge_weight=1.0, resident=False):
        """dress_get for arrays of pairs in one device launch.  With
        ``resident``, against the state the last fit left on the device
        (dress_hip_get_batch_resident) instead of the graph's host arrays."""
        u = np.ascontiguousarray(u, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.int32)
        w = np.ascontiguousarray(np.broadcast_to(np.asarray(edge_weight, dtype=np.float64), u.shape))
        out = np.empty(u.size, dtype=np.float64)
        L = _lib.lib()
        f = L.dress_hip_get_batch_resident if resident else L.dress_hip_get_batch
        r = f(self.g, int(u.size), u.ctypes.data_as(c_int_p), v.ctypes.data_as(c_int_p),
              w.ctypes.data_as(c_dbl_p), int(max_iterations), float(epsilon), out.ctypes.data_as(c_dbl_p))
        _lib.check(r == 0, "dress_hip_get_batch_resident" if resident else "dress_hip_get_batch")
        return out

    def lifted(self, kind, k, max_iterations=100, epsilon=1e-6, n_samples=0, seed=0,
               keep_multisets=False, compute_histogram=True, offset=0, stride=1):
        """delta^k / nabla^k through dress_{delta,nabla}_fit_hip_strided_flat.
        Returns (values, counts, multisets or None, count)."""
        L = _lib.lib()
        f = L.dress_delta_fit_hip_strided_flat if kind == "delta" else L.dress_nabla_fit_hip_strided_flat
        hs = ctypes.c_int(0)
        ns = ctypes.c_int64(0)
        ms = c_dbl_p()
        flat = f(self.g, int(k), int(max_iterations), float(epsilon), int(n_samples),
                 ctypes.c_uint(seed), ctypes.byref(hs) if compute_histogram else None,
                 1 if keep_multisets else 0, ctypes.byref(ms), ctypes.byref(ns),
                 int(offset), int(stride))
        err = _lib.last_error()
        if err:
            _lib.c_free(flat)
            _lib.c_free(ms)
            raise HipError(f"{kind} fit failed: {err}")
        n = hs.value
        vals = np.empty(0)
        cnts = np.empty(0, dtype=np.int64)
        if flat and n > 0:
            # zero-copy views of the malloc'd (value bits, count) pairs; the
            # block is freed when the last view goes
            arr = _owned_view(flat, n, "<i8", _CFree(flat))
            vals = arr.view(np.float64)[0::2]
            cnts = arr[1::2]
        else:
            _lib.c_free(flat)
        mset = None
        if keep_multisets and ms:
            if ns.value * self.E > 0:
                mset = _owned_view(ms, ns.value * self.E, "<f8", _CFree(ms)).reshape(ns.value, self.E)
            else:
                _lib.c_free(ms)
                mset = np.zeros((ns.value, self.E))
        elif ms:
            _lib.c_free(ms)
        return vals, cnts, mset, ns.value


# ---------------------------------------------------------------------------
#  Unified API -- same signatures as dress.fit / dress.delta_fit / nabla_fit
# ---------------------------------------------------------------------------

def fit(n_vertices, sources, targets, weights=None, vertex_weights=None, variant=UNDIRECTED,
        max_iterations=100, epsilon=1e-6, precompute_intercepts=False, precision=None):
    """GPU DRESS fit -- drop-in for ``dress.fit`` (cuda/__init__.py:451-494)."""
    src = list(sources) if not isinstance(sources, np.ndarray) else sources
    tgt = list(targets) if not isinstance(targets, np.ndarray) else targets
    with Graph(n_vertices, src, tgt, weights, vertex_weights, variant,
               precompute_intercepts) as g:
        it, dl = g.fit(max_iterations, epsilon, precision)
        return DRESSResult(sources=list(np.asarray(src).tolist()), targets=list(np.asarray(tgt).tolist()),
                           edge_dress=g.edge_dress, edge_weight=g.edge_weight,
                           vertex_dress=g.vertex_dress, iterations=it, delta=dl)


def _hist_list(vals, cnts):
    return [(float(v), int(c)) for v, c in zip(vals, cnts)]


def delta_fit(n_vertices, sources, targets, weights=None, vertex_weights=None, k=0,
              variant=UNDIRECTED, max_iterations=100, epsilon=1e-6, precompute=False,
              keep_multisets=False, offset=0, stride=1, n_samples=0, seed=0,
              compute_histogram=True):
    """GPU delta^k-DRESS -- drop-in for ``dress.delta_fit`` (cuda/__init__.py:497-574)."""
    with Graph(n_vertices, sources, targets, weights, vertex_weights, variant, precompute) as g:
        vals, cnts, ms, ns = g.lifted("delta", k, max_iterations, epsilon, n_samples, seed,
                                      keep_multisets, compute_histogram, offset, stride)
    return DeltaDRESSResult(histogram=_hist_list(vals, cnts), multisets=ms, num_subgraphs=ns)


def nabla_fit(n_vertices, sources, targets, weights=None, vertex_weights=None, k=0,
              variant=UNDIRECTED, max_iterations=100, epsilon=1e-6, precompute=False,
              keep_multisets=False, n_samples=0, seed=0, compute_histogram=True):
    """GPU nabla^k-DRESS -- drop-in for ``dress.nabla_fit`` (cuda/__init__.py:306-375)."""
    with Graph(n_vertices, sources, targets, weights, vertex_weights, variant, precompute) as g:
        vals, cnts, ms, ns = g.lifted("nabla", k, max_iterations, epsilon, n_samples, seed,
                                      keep_multisets, compute_histogram, 0, 1)
    return NablaDRESSResult(histogram=_hist_list(vals, cnts), multisets=ms, num_tuples=ns)


class DRESS:
    """Object API of ``dress.cuda.DRESS`` (cuda/__init__.py:586-643):
    ``fit`` / ``delta_fit`` / ``nabla_fit`` / ``get`` on one graph."""

    def __init__(self, n_vertices, sources, targets, weights=None, vertex_weights=None,
                 variant=UNDIRECTED, precompute_intercepts=False):
        self._g = Graph(n_vertices, sources, targets, weights, vertex_weights, variant,
                        precompute_intercepts)
        self.n_vertices = int(n_vertices)
        self.n_edges = self._g.E

    def fit(self, max_iterations=100, epsilon=1e-6):
        it, dl = self._g.fit(max_iterations, epsilon)
        return it, dl

    def delta_fit(self, k=0, max_iterations=100, epsilon=1e-6, n_samples=0, seed=0,
                  keep_multisets=False, compute_histogram=True, offset=0, stride=1):
        vals, cnts, ms, ns = self._g.lifted("delta", k, max_iterations, epsilon, n_samples, seed,
                                            keep_multisets, compute_histogram, offset, stride)
        return DeltaDRESSResult(histogram=_hist_list(vals, cnts), multisets=ms, num_subgraphs=ns)

    def nabla_fit(self, k=0, max_iterations=100, epsilon=1e-6, n_samples=0, seed=0,
                  keep_multisets=False, compute_histogram=True):
        vals, cnts, ms, ns = self._g.lifted("nabla", k, max_iterations, epsilon, n_samples, seed,
                                            keep_multisets, compute_histogram)
        return NablaDRESSResult(histogram=_hist_list(vals, cnts), multisets=ms, num_tuples=ns)

    def get(self, u, v, max_iterations=100, epsilon=1e-6, edge_weight=1.0):
        return self._g.get(u, v, max_iterations, epsilon, edge_weight)

    @property
    def edge_dress(self):
        return self._g.edge_dress

    @property
    def vertex_dress(self):
        return self._g.vertex_dress

    def close(self):
        self._g.close()

    def __repr__(self):
        return f"DRESS(n_vertices={self.n_vertices}, n_edges={self.n_edges}, backend=hip)"
