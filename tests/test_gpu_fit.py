"""dress_fit_hip through the C ABI against the oracle and the golden fixtures.

FP64 results are compared bit for bit (the device sums are exactly rounded,
as the reference's sorted KBN sums are); the FP32->FP64 precision is checked
to the 1e-9 per-edge tolerance of the sorted fingerprint.
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

TOL_MIXED = 1e-9   # FP32->FP64: max |edge value - FP64 value|


def _fit(gpu, N, U, V, W=None, NW=None, variant=0, pre=0, max_it=100, eps=1e-6, precision=None):
    with gpu.Graph(N, U, V, W, NW, variant, pre) as g:
        it, dl = g.fit(max_it, eps, precision)
        return {"iterations": it, "delta": dl, "edge_dress": g.edge_dress, "vertex_dress": g.vertex_dress,
                "csr": g.csr(), "stats": gpu.last_fit_stats()}


def test_er2000_bitwise_golden(gpu):
    z = np.load(os.path.join(GOLDEN, "er2000_both.npz"))
    r = _fit(gpu, 2000, z["U"], z["V"])
    assert r["iterations"] == int(z["iterations"]) == 27
    assert r["delta"] == float(z["delta"])
    assert np.array_equal(r["edge_dress"], z["edge_dress"])
    assert np.array_equal(r["vertex_dress"], z["vertex_dress"])


def test_er2000_mixed_precision(gpu):
    z = np.load(os.path.join(GOLDEN, "er2000_both.npz"))
    r = _fit(gpu, 2000, z["U"], z["V"], precision=gpu.FP32_FP64)
    assert r["iterations"] == int(z["iterations"])
    err = np.max(np.abs(np.sort(r["edge_dress"]) - np.sort(z["edge_dress"])))
    assert err <= TOL_MIXED
    assert r["stats"].fp32_iterations > 0 or r["stats"].fp64_reruns == 1


def test_variants_golden(gpu):
    z = np.load(os.path.join(GOLDEN, "variants.npz"))
    for var in range(4):
        for pre in (0, 1):
            key = f"v{var}_p{pre}"
            r = _fit(gpu, 40, z[key + "_U"], z[key + "_V"], z[key + "_W"], z[key + "_NW"], var, pre, 100, 1e-8)
            assert r["iterations"] == int(z[key + "_it"][0]), key
            assert r["delta"] == float(z[key + "_delta"][0]), key
            assert np.array_equal(r["edge_dress"], z[key + "_edge_dress"]), key
            assert np.array_equal(r["vertex_dress"], z[key + "_vertex_dress"]), key
            for f in ("adj_offset", "adj_target", "adj_edge_idx", "edge_weight"):
                assert np.array_equal(r["csr"][f], z[key + "_" + f]), (key, f)
            if pre:
                for f in ("intercept_offset", "intercept_edge_ux", "intercept_edge_vx"):
                    assert np.array_equal(r["csr"][f], z[key + "_" + f]), (key, f)


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_random_multigraphs_bitwise(gpu, ora, variant):
    rng = np.random.default_rng(40 + variant)
    for t in range(3):
        n, m = int(rng.integers(5, 120)), int(rng.integers(1, 600))
        U = rng.integers(0, n, m).astype(np.int32)
        V = rng.integers(0, n, m).astype(np.int32)
        W = rng.random(m) + 0.05 if t else None
        NW = rng.random(n) + 0.5 if t == 2 else None
        a = _fit(gpu, n, U, V, W, NW, variant)
        b = ora.fit(n, U, V, W, NW, variant, 0, 100, 1e-6)
        assert a["iterations"] == b["iterations"] and a["delta"] == b["delta"]
        assert np.array_equal(a["edge_dress"], b["edge_dress"])
        assert np.array_equal(a["vertex_dress"], b["vertex_dress"])


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_hub_bitmaps_bitwise(gpu, ora, variant, monkeypatch):
    """Every row of two or more neighbours indexed through its hub bitmap
    (rows with a repeated neighbour keep the binary search): fits equal the
    oracle's bit for bit, on simple graphs and on multigraphs."""
    monkeypatch.setenv("DRESS_HIP_HUB_DEGREE", "2")
    rng = np.random.default_rng(70 + variant)
    for t in range(4):
        n, m = int(rng.integers(5, 150)), int(rng.integers(1, 900))
        U = rng.integers(0, n, m).astype(np.int32)
        V = rng.integers(0, n, m).astype(np.int32)
        if t < 2:
            # simple graph: no loops, each unordered pair once
            keep = U != V
            U, V = U[keep], V[keep]
            _, first = np.unique(np.stack([np.minimum(U, V), np.maximum(U, V)]), axis=1, return_index=True)
            first = np.sort(first)
            U, V = U[first], V[first]
        W = rng.random(U.size) + 0.05 if t % 2 else None
        a = _fit(gpu, n, U, V, W, None, variant)
        b = ora.fit(n, U, V, W, None, variant, 0, 100, 1e-6)
        assert a["iterations"] == b["iterations"] and a["delta"] == b["delta"], (variant, t)
        assert np.array_equal(a["edge_dress"], b["edge_dress"]), (variant, t)
        assert np.array_equal(a["vertex_dress"], b["vertex_dress"]), (variant, t)


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_hub_bitmaps_intercepts(gpu, ora, variant, monkeypatch):
    """The reference-layout intercept arrays (precompute_intercepts) built
    through hub bitmaps of the reference rows equal the oracle's."""
    monkeypatch.setenv("DRESS_HIP_HUB_DEGREE", "2")
    rng = np.random.default_rng(90 + variant)
    for t in range(3):
        n, m = int(rng.integers(5, 120)), int(rng.integers(1, 700))
        U = rng.integers(0, n, m).astype(np.int32)
        V = rng.integers(0, n, m).astype(np.int32)
        if t == 0:
            keep = U != V
            U, V = U[keep], V[keep]
        with gpu.Graph(n, U, V, None, None, variant, 1) as g:
            a = g.csr()
        og = ora.graph(n, U, V, None, None, variant, 1)
        try:
            b = ora.csr(og)
        finally:
            ora.free_graph(og)
        for f in ("intercept_offset", "intercept_edge_ux", "intercept_edge_vx"):
            assert np.array_equal(a[f], b[f]), (variant, t, f)


def test_hub_bitmaps_power_law(gpu, ora, gen, monkeypatch):
    """Rows of 64 or more neighbours of a 150,000-vertex power-law graph as hub
    bitmaps, and the intersection count in row order (graphs of 1M edges and
    more): the fit equals the oracle's bit for bit."""
    monkeypatch.setenv("DRESS_HIP_HUB_DEGREE", "64")
    n, U, V = gen.chung_lu(n=150_000, n_edges=1_100_000, seed=5)
    assert U.size >= 1 << 20
    a = _fit(gpu, n, U, V)
    b = ora.fit(n, U, V, parallel=True)
    assert a["iterations"] == b["iterations"] and a["delta"] == b["delta"]
    assert np.array_equal(a["edge_dress"], b["edge_dress"])
    assert np.array_equal(a["vertex_dress"], b["vertex_dress"])


def test_iteration_limits(gpu, ora, gen):
    n, U, V = gen.er_single(300, 0.03, 2)
    for max_it, eps in ((1, 1e-6), (2, 1e-6), (5, 0.0), (0, 1e-6)):
        a = _fit(gpu, n, U, V, max_it=max_it, eps=eps)
        b = ora.fit(n, U, V, max_iterations=max_it, epsilon=eps)
        assert a["iterations"] == b["iterations"] == (max_it if (max_it <= 2 or eps == 0) else a["iterations"])
        assert np.array_equal(a["edge_dress"], b["edge_dress"])
        if max_it > 0:
            assert a["delta"] == b["delta"]


def test_fit_continues_from_current_values(gpu, ora, gen):
    n, U, V = gen.er_single(200, 0.05, 4)
    with gpu.Graph(n, U, V) as g:
        g.fit(3, 1e-6)
        it, dl = g.fit(100, 1e-6)
        d = g.edge_dress
    og = ora.graph(n, U, V, None, None, 0, 0)
    ora.fit_graph(og, 3, 1e-6)
    oit, odl = ora.fit_graph(og, 100, 1e-6)
    od = np.ctypeslib.as_array(og.contents.edge_dress, shape=(U.size,)).copy()
    ora.free_graph(og)
    assert (it, dl) == (oit, odl)
    assert np.array_equal(d, od)


def test_empty_and_isolated(gpu, ora):
    a = _fit(gpu, 5, np.zeros(0, np.int32), np.zeros(0, np.int32))
    b = ora.fit(5, np.zeros(0, np.int32), np.zeros(0, np.int32))
    assert a["iterations"] == b["iterations"] == 0
    assert np.array_equal(a["vertex_dress"], b["vertex_dress"])
    U = np.array([0, 1], np.int32)
    V = np.array([1, 2], np.int32)
    a = _fit(gpu, 10, U, V)
    b = ora.fit(10, U, V)
    assert np.array_equal(a["edge_dress"], b["edge_dress"]) and np.array_equal(a["vertex_dress"], b["vertex_dress"])


def test_relabel_invariance_on_device(gpu, gen):
    n, U, V = gen.er_single(500, 0.02, 8)
    rng = np.random.default_rng(1)
    base = np.sort(_fit(gpu, n, U, V)["edge_dress"])
    perm = rng.permutation(n).astype(np.int32)
    order = rng.permutation(U.size)
    other = np.sort(_fit(gpu, n, perm[V][order], perm[U][order])["edge_dress"])
    assert np.array_equal(base, other)


def test_power_law_graph_vs_oracle(gpu, ora, gen):
    n, U, V = gen.chung_lu(n=60_000, n_edges=420_000, seed=5)
    a = _fit(gpu, n, U, V, pre=1)
    b = ora.fit(n, U, V, precompute=1, parallel=True)
    assert a["iterations"] == b["iterations"]
    assert np.array_equal(a["edge_dress"], b["edge_dress"])
    assert np.array_equal(a["csr"]["intercept_offset"], ora.csr(g := ora.graph(n, U, V, None, None, 0, 1))["intercept_offset"])
    ora.free_graph(g)
    m = _fit(gpu, n, U, V, precision=gpu.FP32_FP64)
    assert m["iterations"] == b["iterations"]
    assert np.max(np.abs(m["edge_dress"] - b["edge_dress"])) <= TOL_MIXED


def _edge_residual(U, V, csr, d, sample):
    """|F(d) - d| on sampled edges, F one UNDIRECTED unweighted update."""
    off, tgt, eid = csr["adj_offset"], csr["adj_target"], csr["adj_edge_idx"]
    def norm(u):
        return math.sqrt(math.fsum([4.0] + [2.0 * d[e] for e in eid[off[u]:off[u + 1]]]))
    out = []
    for e in sample:
        u, v = int(U[e]), int(V[e])
        ru = dict(zip(tgt[off[u]:off[u + 1]], eid[off[u]:off[u + 1]]))
        terms = [2.0 * d[ru[x]] + 2.0 * d[f] for x, f in zip(tgt[off[v]:off[v + 1]], eid[off[v]:off[v + 1]])
                 if x in ru]
        terms.append(8.0 + 4.0 * d[e])
        out.append(abs(math.fsum(terms) / (norm(u) * norm(v)) - d[e]))
    return np.array(out)


def test_power_law_4m_full_size_properties(gpu, gen):
    """4M-vertex power-law graph: FP64 and FP32->FP64 agree to 1e-9 with the
    same stopping step; values lie in (0, 2]; sampled edges are fixed points
    to within the stopping tolerance."""
    n, U, V = gen.chung_lu()
    with gpu.Graph(n, U, V) as g:
        g.resident_reset()
        it64, dl64 = g.fit_resident(100, 1e-6, gpu.FP64)
        g.resident_download()
        d64 = g.edge_dress
        csr = g.csr()
        g.set_edge_dress(np.ones(U.size))       # start again from d = 1
        g.resident_reset()
        it32, dl32 = g.fit_resident(100, 1e-6, gpu.FP32_FP64)
        g.resident_download()
        d32 = g.edge_dress
    assert it64 == it32 and 20 <= it64 <= 40
    assert dl64 < 1e-6
    assert np.max(np.abs(d64 - d32)) <= TOL_MIXED
    assert np.all(d64 > 0) and np.all(d64 <= 2.0)
    sample = np.random.default_rng(0).choice(U.size, 300, replace=False)
    assert np.max(_edge_residual(U, V, csr, d64, sample)) < 2e-6


def test_precision_setting(gpu):
    assert gpu.get_precision() == gpu.FP64
    gpu.set_precision(gpu.FP32_FP64)
    try:
        assert gpu.get_precision() == gpu.FP32_FP64
    finally:
        gpu.set_precision(gpu.FP64)


def test_graph_captured_iterations(gpu, ora, gen, monkeypatch):
    """DRESS_HIP_GRAPH=1 replays one captured iteration per step: repeated
    fits (eager work and synchronisations between the replays) stay bitwise
    equal to the oracle, in FP64 and through a precision change."""
    monkeypatch.setenv("DRESS_HIP_GRAPH", "1")
    z = np.load(os.path.join(GOLDEN, "er2000_both.npz"))
    with gpu.Graph(2000, z["U"], z["V"]) as g:
        for _ in range(3):
            g.set_edge_dress(np.ones(z["U"].size))
            it, dl = g.fit(100, 1e-6)
            assert it == int(z["iterations"]) and dl == float(z["delta"])
            assert np.array_equal(g.edge_dress, z["edge_dress"])
        g.set_edge_dress(np.ones(z["U"].size))
        g.fit(100, 1e-6, gpu.FP32_FP64)
        assert np.max(np.abs(g.edge_dress - z["edge_dress"])) <= TOL_MIXED
        g.set_edge_dress(np.ones(z["U"].size))
        g.fit(100, 1e-6)
        assert np.array_equal(g.edge_dress, z["edge_dress"])
    n, U, V = gen.er_single(400, 0.02, 12)
    with gpu.Graph(n, U, V) as g:
        for eps in (1e-6, 1e-8):
            g.set_edge_dress(np.ones(U.size))
            it, dl = g.fit(100, eps)
            b = ora.fit(n, U, V, max_iterations=100, epsilon=eps)
            assert (it, dl) == (b["iterations"], b["delta"])
            assert np.array_equal(g.edge_dress, b["edge_dress"])


def _oracle_from(ora, n, U, V, W, NW, variant, value, max_it, eps):
    """oracle fit of the graph whose every edge value starts at `value`"""
    og = ora.graph(n, U, V, W, NW, variant, 0)
    try:
        np.ctypeslib.as_array(og.contents.edge_dress, shape=(U.size,))[:] = value
        it, dl = ora.fit_graph(og, max_it, eps)
        d = np.ctypeslib.as_array(og.contents.edge_dress, shape=(U.size,)).copy()
        vd = np.ctypeslib.as_array(og.contents.vertex_dress, shape=(n,)).copy()
    finally:
        ora.free_graph(og)
    return it, dl, d, vd


@pytest.mark.parametrize("case", ["uniform", "weighted", "vertex_weights", "directed"])
def test_resident_fill_then_fit_equals_oracle(gpu, ora, gen, case):
    """dress_hip_resident_fill + dress_hip_fit_resident: a fill of a uniform
    graph prepares the half-edge copies itself (the next FP64 fit skips
    their set-up); every case equals the oracle's fit from the same start,
    also after an earlier fit moved the values, and a mixed fit after a fill
    stays within the mixed tolerance."""
    n, U, V = gen.er_single(300, 0.04, 9)
    rng = np.random.default_rng(5)
    W = rng.uniform(0.1, 1.0, U.size) if case == "weighted" else None
    NW = rng.uniform(0.5, 2.0, n) if case == "vertex_weights" else None
    variant = 1 if case == "directed" else 0
    with gpu.Graph(n, U, V, W, NW, variant) as g:
        g.resident_reset()
        g.fit_resident(4, 1e-6, gpu.FP64)            # values away from the start
        for value in (1.0, 0.5, 1.0):
            g.resident_fill(value)
            it, dl = g.fit_resident(100, 1e-6, gpu.FP64)
            g.resident_download()
            oit, odl, od, ovd = _oracle_from(ora, n, U, V, W, NW, variant, value, 100, 1e-6)
            assert (it, dl) == (oit, odl), (case, value)
            assert np.array_equal(g.edge_dress, od), (case, value)
            assert np.array_equal(g.vertex_dress, ovd), (case, value)
        g.resident_fill(1.0)
        it, dl = g.fit_resident(100, 1e-6, gpu.FP32_FP64)
        g.resident_download()
        oit, odl, od, _ = _oracle_from(ora, n, U, V, W, NW, variant, 1.0, 100, 1e-6)
        assert it == oit
        assert np.max(np.abs(g.edge_dress - od)) <= TOL_MIXED
