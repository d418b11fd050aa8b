"""Host-side parts of the serving stream (no GPU): the arrival process and
the G(n, p) request generator."""
import math

import numpy as np


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
