"""Deterministic synthetic graphs for the DRESS workloads.

Every generator returns ``(N, U, V)`` or ``(N, U, V, W)`` with ``U, V`` int32
arrays of edge records (and ``W`` float64 weights or ``None``).  The same
parameters and seed always produce the same edge list.

Workloads (sizes from the DRESS benchmark configurations):

* ``er_both_directions``  Erdos-Renyi G(2000, 8/1999), seed 0, every edge
  listed in both directions (16,040 records).
* ``chung_lu``            power-law graph, 4,000,000 vertices, exponent 2.1,
  expected degrees capped at 20,000, mean 14, seed 1, 28,000,000 undirected
  edges listed once.
* ``random_regular``      configuration model with whole-matching rejection
  (uniform over simple d-regular graphs); 512 graphs of degree 6 with
  n ~ U{16..128} for the delta^1 batch.
* ``regular_switch``      degree-12, n=256 graph by a switch chain (whole
  matching rejection would need ~1e15 attempts at d=12).
* ``gnp``                 G(n, p) requests of the serving stream.
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_gen = None


def _genlib():
    global _gen
    if _gen is None:
        path = os.path.join(_HERE, "libdress_gen.so")
        if not os.path.isfile(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        i32p = ctypes.POINTER(ctypes.c_int)
        lib.dress_gen_random_regular.restype = ctypes.c_int
        lib.dress_gen_random_regular.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                                 ctypes.c_int, i32p, i32p]
        lib.dress_gen_regular_switch.restype = ctypes.c_int
        lib.dress_gen_regular_switch.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                                 ctypes.c_int, i32p, i32p]
        lib.dress_gen_gnp.restype = ctypes.c_int64
        lib.dress_gen_gnp.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_uint64, ctypes.c_int64,
                                      i32p, i32p, ctypes.POINTER(ctypes.c_double)]
        _gen = lib
    return _gen


def _i32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def er_edges(n, p, seed):
    """Upper-triangle Bernoulli draws in ``np.triu_indices(n, 1)`` order."""
    rng = np.random.default_rng(seed)
    keep = rng.random(n * (n - 1) // 2) < p
    iu, ju = np.triu_indices(n, 1)
    return iu[keep].astype(np.int32), ju[keep].astype(np.int32)


def er_both_directions(n=2000, p=8 / 1999, seed=0):
    """ER graph with every edge listed as (i, j) records then (j, i) records."""
    a, b = er_edges(n, p, seed)
    return n, np.concatenate([a, b]), np.concatenate([b, a])


def er_single(n=2000, p=8 / 1999, seed=0):
    a, b = er_edges(n, p, seed)
    return n, a, b


def chung_lu(n=4_000_000, gamma=2.1, max_deg=20_000, mean_deg=14.0,
             n_edges=28_000_000, seed=1):
    """Chung-Lu graph with truncated power-law expected degrees.

    Expected degrees follow a Pareto law with density exponent ``gamma``
    (x_min scaled so the capped mean is ``mean_deg``), capped at ``max_deg``.
    Endpoints are drawn proportionally to the expected degree; self-loops and
    repeated pairs are dropped and draws continue until exactly ``n_edges``
    distinct undirected edges exist.  Edges are returned once each (u < v),
    sorted by (u, v).
    """
    rng = np.random.default_rng(seed)
    raw = (1.0 - rng.random(n)) ** (-1.0 / (gamma - 1.0))
    scale = mean_deg / raw.mean()
    for _ in range(60):                      # fixed point for the capped mean
        w = np.minimum(raw * scale, max_deg)
        m = w.mean()
        if abs(m - mean_deg) < 1e-9 * mean_deg:
            break
        scale *= mean_deg / m
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    keys = np.empty(0, dtype=np.int64)
    want = n_edges
    while keys.size < n_edges:
        draw = int((want - keys.size) * 1.08) + 1024
        a = np.searchsorted(cdf, rng.random(draw), side="right").astype(np.int64)
        b = np.searchsorted(cdf, rng.random(draw), side="right").astype(np.int64)
        np.minimum(a, n - 1, out=a)
        np.minimum(b, n - 1, out=b)
        ok = a != b
        lo = np.minimum(a[ok], b[ok])
        hi = np.maximum(a[ok], b[ok])
        keys = np.unique(np.concatenate([keys, lo * n + hi]))
    if keys.size > n_edges:
        drop = rng.choice(keys.size, keys.size - n_edges, replace=False)
        mask = np.ones(keys.size, dtype=bool)
        mask[drop] = False
        keys = keys[mask]
    return n, (keys // n).astype(np.int32), (keys % n).astype(np.int32)


def random_regular(n, d, seed, max_attempts=10_000_000):
    """Uniform simple d-regular graph (configuration model, whole rejection)."""
    m = n * d // 2
    U = np.empty(m, dtype=np.int32)
    V = np.empty(m, dtype=np.int32)
    r = _genlib().dress_gen_random_regular(n, d, seed, max_attempts, _i32p(U), _i32p(V))
    if r < 0:
        raise RuntimeError(f"no simple {d}-regular graph on {n} vertices after {max_attempts} draws")
    return n, U, V


def regular_switch(n=256, d=12, seed=7, swaps_per_edge=100):
    """d-regular graph from a switch chain started at the circulant graph."""
    m = n * d // 2
    U = np.empty(m, dtype=np.int32)
    V = np.empty(m, dtype=np.int32)
    if _genlib().dress_gen_regular_switch(n, d, seed, swaps_per_edge, _i32p(U), _i32p(V)) != 0:
        raise RuntimeError("bad regular_switch parameters")
    return n, U, V


def delta1_batch_graphs(count=512, d=6, n_lo=16, n_hi=128):
    """The delta^1 batch: graph i has n_i = default_rng(i).integers(n_lo, n_hi+1)
    vertices and is the uniform d-regular graph drawn with seed i."""
    out = []
    for i in range(count):
        n = int(np.random.default_rng(i).integers(n_lo, n_hi + 1))
        out.append(random_regular(n, d, i))
    return out


def gnp(n, mean_deg, seed, weighted=False):
    """G(n, p) with p = mean_deg / (n - 1), drawn in C by geometric skipping
    over the vertex pairs (dress_gen_gnp).  Weighted requests carry float32
    weights uniform in (0.1, 1], passed on as float64."""
    p = min(1.0, mean_deg / max(n - 1, 1))
    pairs = n * (n - 1) // 2
    mu = pairs * p
    cap = int(mu + 12.0 * math.sqrt(mu + 1.0)) + 64
    L = _genlib()
    while True:
        U = np.empty(cap, dtype=np.int32)
        V = np.empty(cap, dtype=np.int32)
        W = np.empty(cap, dtype=np.float64) if weighted else None
        m = L.dress_gen_gnp(int(n), float(p), int(seed) & ((1 << 64) - 1), cap, _i32p(U), _i32p(V),
                            W.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if weighted else None)
        if m >= 0:
            return n, U[:m].copy(), V[:m].copy(), (W[:m].copy() if weighted else None)
        cap *= 2
