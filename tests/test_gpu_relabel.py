"""Internal vertex labels on the device.

Graphs of at least 2^18 half-edges are fitted over vertices relabelled by
decreasing degree; smaller ones keep their input labels.  DRESS_HIP_RELABEL
forces either path, so the bitwise checks of the other GPU tests are
repeated here with relabelling forced on small graphs and forced off on a
multi-bin one: CSR, intercept arrays, values, norms and dress_get answers
must not depend on the internal labels.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from test_gpu_fit import TOL_MIXED, _fit

pytestmark = pytest.mark.gpu


@pytest.fixture
def relabel_always(monkeypatch):
    monkeypatch.setenv("DRESS_HIP_RELABEL", "always")


@pytest.fixture
def relabel_never(monkeypatch):
    monkeypatch.setenv("DRESS_HIP_RELABEL", "never")


def test_variants_golden_relabelled(gpu, relabel_always):
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
def test_multigraphs_relabelled_vs_oracle(gpu, ora, relabel_always, variant):
    rng = np.random.default_rng(140 + variant)
    for t in range(3):
        n, m = int(rng.integers(5, 150)), int(rng.integers(1, 900))
        U = rng.integers(0, n, m).astype(np.int32)
        V = rng.integers(0, n, m).astype(np.int32)
        W = rng.random(m) + 0.05 if t else None
        NW = rng.random(n) + 0.5 if t == 2 else None
        a = _fit(gpu, n, U, V, W, NW, variant, pre=t % 2)
        b = ora.fit(n, U, V, W, NW, variant, t % 2, 100, 1e-6)
        assert a["iterations"] == b["iterations"] and a["delta"] == b["delta"]
        assert np.array_equal(a["edge_dress"], b["edge_dress"])
        assert np.array_equal(a["vertex_dress"], b["vertex_dress"])
        og = ora.graph(n, U, V, W, NW, variant, t % 2)
        oc = ora.csr(og)
        ora.free_graph(og)
        fields = ["adj_offset", "adj_target", "adj_edge_idx", "edge_weight"]
        if t % 2:
            fields += ["intercept_offset", "intercept_edge_ux", "intercept_edge_vx"]
        for f in fields:
            assert np.array_equal(a["csr"][f], oc[f]), f


def test_get_relabelled(gpu, ora, relabel_always):
    z = np.load(os.path.join(GOLDEN, "get_queries.npz"))
    with gpu.Graph(80, z["U"], z["V"]) as g:
        g.fit(100, 1e-6)
        assert np.array_equal(g.get_batch(z["qu"], z["qv"], 100, 1e-6, z["qw"]), z["ans"])
    rng = np.random.default_rng(77)
    n, m = 40, 160
    U = rng.integers(0, n, m).astype(np.int32)
    V = rng.integers(0, n, m).astype(np.int32)
    W = rng.random(m) + 0.2
    qu = rng.integers(0, n, 80).astype(np.int32)
    qv = rng.integers(0, n, 80).astype(np.int32)
    with gpu.Graph(n, U, V, W, None, 1) as g:
        g.fit(100, 1e-7)
        got = g.get_batch(qu, qv, 50, 1e-7, 0.4)
    og = ora.graph(n, U, V, W, None, 1, 0)
    ora.fit_graph(og, 100, 1e-7)
    want = np.array([ora.get(og, int(a), int(b), 50, 1e-7, 0.4) for a, b in zip(qu, qv)])
    ora.free_graph(og)
    assert np.array_equal(got, want)


def test_power_law_input_labels_vs_oracle(gpu, ora, gen, relabel_never):
    n, U, V = gen.chung_lu(n=60_000, n_edges=420_000, seed=5)
    a = _fit(gpu, n, U, V)
    b = ora.fit(n, U, V, parallel=True)
    assert a["iterations"] == b["iterations"] and a["delta"] == b["delta"]
    assert np.array_equal(a["edge_dress"], b["edge_dress"])
    assert np.array_equal(a["vertex_dress"], b["vertex_dress"])
    m = _fit(gpu, n, U, V, precision=gpu.FP32_FP64)
    assert np.max(np.abs(m["edge_dress"] - b["edge_dress"])) <= TOL_MIXED


def test_both_label_paths_agree(gpu, gen, monkeypatch):
    n, U, V = gen.chung_lu(n=30_000, n_edges=200_000, seed=9)
    W = np.random.default_rng(3).random(U.size) + 0.1
    out = {}
    for mode in ("never", "always"):
        monkeypatch.setenv("DRESS_HIP_RELABEL", mode)
        out[mode] = _fit(gpu, n, U, V, W)
    assert out["never"]["iterations"] == out["always"]["iterations"]
    assert np.array_equal(out["never"]["edge_dress"], out["always"]["edge_dress"])
    assert np.array_equal(out["never"]["vertex_dress"], out["always"]["vertex_dress"])
