"""dress_get (virtual-edge queries) on the device against the golden answers
of the compiled reference and the oracle."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def test_get_golden(gpu):
    z = np.load(os.path.join(GOLDEN, "get_queries.npz"))
    with gpu.Graph(80, z["U"], z["V"]) as g:
        g.fit(100, 1e-6)
        batch = g.get_batch(z["qu"], z["qv"], 100, 1e-6, z["qw"])
        single = np.array([g.get(int(a), int(b), 100, 1e-6, float(w))
                           for a, b, w in zip(z["qu"][:20], z["qv"][:20], z["qw"][:20])])
    assert np.array_equal(batch, z["ans"])
    assert np.array_equal(single, z["ans"][:20])


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_get_variants_vs_oracle(gpu, ora, variant):
    rng = np.random.default_rng(70 + variant)
    n, m = 50, 200
    U = rng.integers(0, n, m).astype(np.int32)
    V = rng.integers(0, n, m).astype(np.int32)
    W = rng.random(m) + 0.2
    NW = rng.random(n) + 0.5
    qu = rng.integers(0, n, 100).astype(np.int32)
    qv = rng.integers(0, n, 100).astype(np.int32)
    with gpu.Graph(n, U, V, W, NW, variant) as g:
        g.fit(100, 1e-7)
        got = g.get_batch(qu, qv, 50, 1e-7, 0.7)
    og = ora.graph(n, U, V, W, NW, variant, 0)
    ora.fit_graph(og, 100, 1e-7)
    want = np.array([ora.get(og, int(a), int(b), 50, 1e-7, 0.7) for a, b in zip(qu, qv)])
    ora.free_graph(og)
    assert np.array_equal(got, want)


def test_get_out_of_range_is_an_error(gpu):
    with gpu.Graph(3, [0, 1], [1, 2]) as g:
        g.fit()
        with pytest.raises(gpu.HipError):
            g.get(0, 7)


@pytest.mark.parametrize("relabel", ["0", "1"])
@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_get_resident_vs_oracle(gpu, ora, variant, relabel, monkeypatch):
    """Queries against the state the fit left on the device (internal edge
    order and, with relabelling, internal vertex labels) equal the host-array
    path and the oracle bit for bit."""
    monkeypatch.setenv("DRESS_HIP_RELABEL", relabel)
    rng = np.random.default_rng(170 + variant)
    n, m = 60, 260
    U = rng.integers(0, n, m).astype(np.int32)
    V = rng.integers(0, n, m).astype(np.int32)
    W = rng.random(m) + 0.2
    NW = rng.random(n) + 0.5
    qu = rng.integers(0, n, 300).astype(np.int32)
    qv = rng.integers(0, n, 300).astype(np.int32)
    with gpu.Graph(n, U, V, W, NW, variant) as g:
        g.fit(100, 1e-7)
        res = g.get_batch(qu, qv, 50, 1e-7, 0.7, resident=True)
        host = g.get_batch(qu, qv, 50, 1e-7, 0.7)
    og = ora.graph(n, U, V, W, NW, variant, 0)
    ora.fit_graph(og, 100, 1e-7)
    want = np.array([ora.get(og, int(a), int(b), 50, 1e-7, 0.7) for a, b in zip(qu, qv)])
    ora.free_graph(og)
    assert np.array_equal(res, host)
    assert np.array_equal(res, want)


def test_get_resident_state(gpu, gen):
    """The resident state exists from a completed fit until other fits run on
    the graph's device copy; host-array queries leave it alone."""
    n, U, V = gen.random_regular(40, 6, 3)
    qu, qv = np.arange(20, dtype=np.int32), (np.arange(20, dtype=np.int32) * 7 + 3) % 40
    with gpu.Graph(n, U, V) as g:
        with pytest.raises(gpu.HipError):
            g.get_batch(qu, qv, resident=True)
        g.fit(100, 1e-6)
        a = g.get_batch(qu, qv, resident=True)
        b = g.get_batch(qu, qv)
        assert np.array_equal(a, b)
        assert np.array_equal(g.get_batch(qu, qv, resident=True), a)
        gpu.set_lifted_mode(2)                 # subgraph fits on this graph's device copy
        try:
            g.lifted("delta", 1, n_samples=3, seed=1)
        finally:
            gpu.set_lifted_mode(0)
        with pytest.raises(gpu.HipError):
            g.get_batch(qu, qv, resident=True)
        g.fit(100, 1e-6)                       # continues from the first fit's values
        assert np.array_equal(g.get_batch(qu, qv, resident=True), g.get_batch(qu, qv))


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_get_few_queries_vs_oracle(gpu, ora, variant):
    """Few queries on a larger graph upload only the state entries they read
    (the rows of their vertices); single gets, a small batch and a batch that
    uploads the whole state answer alike, bit for bit with the oracle."""
    rng = np.random.default_rng(270 + variant)
    n, m = 3000, 15000
    U = rng.integers(0, n, m).astype(np.int32)
    V = rng.integers(0, n, m).astype(np.int32)
    W = rng.random(m) + 0.2
    NW = rng.random(n) + 0.5
    qu = rng.integers(0, n, 40).astype(np.int32)
    qv = rng.integers(0, n, 40).astype(np.int32)
    big_u = rng.integers(0, n, 4000).astype(np.int32)
    big_v = rng.integers(0, n, 4000).astype(np.int32)
    with gpu.Graph(n, U, V, W, NW, variant) as g:
        g.fit(100, 1e-7)
        single = np.array([g.get(int(a), int(b), 50, 1e-7, 0.7) for a, b in zip(qu[:10], qv[:10])])
        few = g.get_batch(qu, qv, 50, 1e-7, 0.7)
        whole = g.get_batch(np.concatenate([qu, big_u]), np.concatenate([qv, big_v]), 50, 1e-7, 0.7)[:qu.size]
    og = ora.graph(n, U, V, W, NW, variant, 0)
    ora.fit_graph(og, 100, 1e-7)
    want = np.array([ora.get(og, int(a), int(b), 50, 1e-7, 0.7) for a, b in zip(qu, qv)])
    ora.free_graph(og)
    assert np.array_equal(few, want)
    assert np.array_equal(whole, want)
    assert np.array_equal(single, want[:10])
