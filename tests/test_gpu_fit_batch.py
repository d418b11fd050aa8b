"""dress_hip_fit_batch: many small whole-graph fits in one launch equal the
single-graph fit and the oracle bit for bit (iterations, delta, edge values,
vertex norms), for every variant, weighted or not, and for FP32->FP64."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _graphs(seed, count, weighted):
    rng = np.random.default_rng(seed)
    gs, ws = [], []
    for i in range(count):
        n = int(rng.integers(1, 90))
        m = int(rng.integers(0, 4 * n + 1))
        U = rng.integers(0, n, m).astype(np.int32)
        V = rng.integers(0, n, m).astype(np.int32)
        gs.append((n, U, V))
        ws.append(rng.random(m) + 0.05 if weighted else np.ones(m))
    return gs, (ws if weighted else None)


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("weighted", [False, True])
def test_fit_batch_equals_single_fits(gpu, ora, variant, weighted):
    gs, ws = _graphs(600 + variant + 10 * weighted, 24, weighted)
    r = gpu.fit_batch(gs, variant=variant, max_iterations=100, epsilon=1e-7, weights=ws)
    for i, (n, U, V) in enumerate(gs):
        W = ws[i] if ws else None
        with gpu.Graph(n, U, V, W, None, variant) as g:
            it, dl = g.fit(100, 1e-7)
            assert (int(r.iterations[i]), float(r.delta[i])) == (it, dl), i
            assert np.array_equal(r.edge_dress[i], g.edge_dress), i
            assert np.array_equal(r.vertex_dress[i], g.vertex_dress), i
        if i % 6 == 0:
            o = ora.fit(n, U, V, W, None, variant, 0, 100, 1e-7)
            assert o["iterations"] == it and np.array_equal(o["edge_dress"], r.edge_dress[i])
            assert np.array_equal(o["vertex_dress"], r.vertex_dress[i])


def test_fit_batch_mixed_precision(gpu, gen):
    """FP32->FP64 in the batched kernel: deterministic, the same stopping
    step as FP64, every value within 1e-9 of the FP64 fit."""
    gs = [gen.gnp(int(n), 6.0, 40 + i)[:3] for i, n in enumerate(np.linspace(30, 1200, 16).astype(int))]
    r = gpu.fit_batch(gs, precision=gpu.FP32_FP64)
    r2 = gpu.fit_batch(gs[::-1], precision=gpu.FP32_FP64)
    f = gpu.fit_batch(gs)
    for i in range(len(gs)):
        j = len(gs) - 1 - i
        assert np.array_equal(r.edge_dress[i], r2.edge_dress[j])
        assert int(r.iterations[i]) == int(f.iterations[i])
        assert np.max(np.abs(r.edge_dress[i] - f.edge_dress[i]), initial=0.0) <= 1e-9
        assert np.max(np.abs(r.vertex_dress[i] - f.vertex_dress[i]), initial=0.0) <= 1e-9


def test_fit_batch_iteration_limits_and_eligibility(gpu, ora):
    gs, _ = _graphs(5, 6, False)
    for max_it, eps in ((1, 1e-6), (3, 0.0), (0, 1e-6)):
        r = gpu.fit_batch(gs, max_iterations=max_it, epsilon=eps)
        for i, (n, U, V) in enumerate(gs):
            o = ora.fit(n, U, V, max_iterations=max_it, epsilon=eps)
            assert int(r.iterations[i]) == o["iterations"]
            assert np.array_equal(r.edge_dress[i], o["edge_dress"])
            if max_it > 0:
                assert float(r.delta[i]) == o["delta"]
                assert np.array_equal(r.vertex_dress[i], o["vertex_dress"])
    assert gpu.fit_batch_eligible(100, 2000) and not gpu.fit_batch_eligible(100, 70000)
    with pytest.raises(gpu.HipError):
        gpu.fit_batch([(200, np.zeros(20000, np.int32), np.ones(20000, np.int32))])


def test_batches_of_mixed_sizes_are_split_by_lds(gpu, gen):
    """A dense graph with few vertices beside a sparse one with many: each
    fits the batched kernel alone, both together would not; the launch is
    split by LDS and every result still equals the single-graph path."""
    dense = gen.gnp(140, 125.0, 3)[:3]          # ~8.7k edges, 140 vertices
    sparse = gen.gnp(3500, 3.0, 4)[:3]          # ~5.2k edges, 3500 vertices
    small = [gen.gnp(60, 5.0, 10 + i)[:3] for i in range(5)]
    graphs = [small[0], dense, small[1], sparse] + small[2:]
    for n, U, V in graphs:
        assert gpu.fit_batch_eligible(n, U.size)
    r = gpu.fit_batch(graphs)
    for i, (n, U, V) in enumerate(graphs):
        with gpu.Graph(n, U, V) as g:
            it, dl = g.fit(100, 1e-6)
            assert (int(r.iterations[i]), float(r.delta[i])) == (it, dl)
            assert np.array_equal(r.edge_dress[i], g.edge_dress)
            assert np.array_equal(r.vertex_dress[i], g.vertex_dress)
    lb = gpu.lifted_batch([dense, sparse[:1] + (sparse[1][:600], sparse[2][:600])], k=1)
    for (vals, cnts), (n, U, V) in zip(lb.histograms, [dense, (3500, sparse[1][:600], sparse[2][:600])]):
        with gpu.Graph(n, U, V) as g:
            sv, sc, _, _ = g.lifted("delta", 1)
        assert np.array_equal(vals, sv) and np.array_equal(cnts, sc)
