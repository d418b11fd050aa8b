"""Host-side parts of the serving stream and the batch API (no GPU): the
arrival process, the G(n, p) request generator and batch packing."""
import math

import numpy as np
import pytest


def test_request_stream_is_seeded_and_in_range(dh):
    from dress_hip import serve
    a = serve.request_stream(5.0, 400.0, seed=3)
    b = serve.request_stream(5.0, 400.0, seed=3)
    c = serve.request_stream(5.0, 400.0, seed=4)
    assert a == b and a != c
    t = np.array([x[0] for x in a])
    assert np.all(np.diff(t) > 0) and t[0] >= 0 and t[-1] < 5.0
    # Poisson count: mean 2000, sd ~45
    assert 1800 < len(a) < 2200
    kinds = np.array([x[1] for x in a])
    assert abs(np.mean(kinds == serve.FIT) - 0.7) < 0.04
    n = np.array([x[2] for x in a])
    d = np.array([x[3] for x in a])
    assert n.min() >= 20 and n.max() <= 20000 and d.min() >= 3 and d.max() <= 30
    # log-uniform n: median near sqrt(20 * 20000) = 632
    assert 450 < np.median(n) < 880
    assert abs(np.mean([x[4] for x in a]) - 0.5) < 0.05
    assert len({x[5] for x in a}) == len(a)


def test_gnp_generator_properties(gen):
    for n, d, seed in ((20, 3.0, 1), (500, 12.0, 2), (5000, 30.0, 3)):
        N, U, V, W = gen.gnp(n, d, seed, weighted=True)
        assert N == n and U.dtype == np.int32 and U.size == V.size == W.size
        assert np.all(U < V) and U.min(initial=0) >= 0 and V.max(initial=0) < n
        keys = U.astype(np.int64) * n + V
        assert np.unique(keys).size == keys.size
        mu = n * d / 2.0
        assert abs(U.size - mu) < 6 * math.sqrt(mu) + 5
        assert np.all(W > 0.1) and np.all(W <= 1.0)
        assert np.array_equal(W.astype(np.float32).astype(np.float64), W)
        N2, U2, V2, W2 = gen.gnp(n, d, seed, weighted=True)
        assert np.array_equal(U, U2) and np.array_equal(V, V2) and np.array_equal(W, W2)
    _, U, V, W = gen.gnp(30, 29.0, 5)
    assert U.size == 30 * 29 // 2 and W is None


def test_gnp_pair_frequencies(gen):
    # every pair equally likely: counts over many seeds of a 12-vertex graph
    n, p = 12, 0.3
    cnt = np.zeros((n, n))
    reps = 3000
    for s in range(reps):
        _, U, V, _ = gen.gnp(n, p * (n - 1), 1000 + s)
        cnt[U, V] += 1
    iu = np.triu_indices(n, 1)
    freq = cnt[iu] / reps
    assert abs(freq.mean() - p) < 0.01
    assert np.all(np.abs(freq - p) < 5 * math.sqrt(p * (1 - p) / reps))


def test_pack_graphs_flattens_a_batch(dh, gen):
    """pack_graphs gives the flat arrays the batch entry points take (vertex
    counts, edge offsets, int32 endpoints, float64 weights) and checks the
    per-graph lengths; a GraphBatch passes through unchanged."""
    graphs = gen.delta1_batch_graphs(7, n_lo=8, n_hi=20)
    graphs[3] = (5, np.zeros(0, np.int64), np.zeros(0, np.int64))        # no edges, other dtype
    w = [np.linspace(0.5, 1.0, len(U)) for _, U, _ in graphs]
    b = dh.pack_graphs(graphs, w)
    assert b.N.tolist() == [n for n, _, _ in graphs] and len(b) == 7
    assert b.offsets.tolist() == np.concatenate([[0], np.cumsum([len(U) for _, U, _ in graphs])]).tolist()
    assert b.U.dtype == np.int32 and b.V.dtype == np.int32 and b.W.dtype == np.float64
    for i, (n, U, V) in enumerate(graphs):
        lo, hi = b.offsets[i], b.offsets[i + 1]
        assert np.array_equal(b.U[lo:hi], U) and np.array_equal(b.V[lo:hi], V) and np.array_equal(b.W[lo:hi], w[i])
    assert dh.pack_graphs(b) is b
    with pytest.raises(ValueError):
        dh.pack_graphs(b, w)
    with pytest.raises(ValueError):
        dh.pack_graphs([(3, np.array([0, 1], np.int32), np.array([1], np.int32))])
    with pytest.raises(ValueError):
        dh.pack_graphs(graphs, [x[:-1] for x in w])
    empty = dh.pack_graphs([])
    assert len(empty) == 0 and empty.offsets.tolist() == [0]
