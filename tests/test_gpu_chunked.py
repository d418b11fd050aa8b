"""delta^k / nabla^k requests larger than the device budget of one call run
in chunks of subgraphs; each chunk's exact histogram is merged on the host.
The chunk limit is forced small here (DRESS_HIP_LIFT_CHUNK_VALUES), so every
path -- the batched kernel, masked fits with DIRECTED weight patches, and
multi-subgraph passes -- runs as many chunks, and the histograms and
multisets still equal the oracle's bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(gpu, mode, kind, n, U, V, W=None, NW=None, variant=0, **kw):
    gpu.set_lifted_mode(mode)
    try:
        with gpu.Graph(n, U, V, W, NW, variant) as g:
            return g.lifted(kind, **kw)
    finally:
        gpu.set_lifted_mode(0)


def _same(a, b):
    assert a[3] == b[3]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert (a[2] is None) == (b[2] is None)
    if a[2] is not None:
        assert np.array_equal(a[2], b[2], equal_nan=True)


def _directed_with_reciprocals():
    rng = np.random.default_rng(4)
    n = 12
    U, V = [], []
    for a, b in ((0, 5), (2, 7), (3, 9), (6, 1), (10, 4)):
        U.append(a); V.append(b)
        for _ in range(int(rng.integers(2, 5))):
            U.append(b); V.append(a)
    for _ in range(40):
        U.append(int(rng.integers(0, n))); V.append(int(rng.integers(0, n)))
    U = np.array(U, np.int32)
    V = np.array(V, np.int32)
    W = np.round(rng.random(U.size) * 4 + 0.25, 3)
    return n, U, V, W


@pytest.mark.parametrize("mode", [0, 2, 3])
@pytest.mark.parametrize("chunk", [1, 64, 1000])
def test_chunked_small_graphs_match_oracle(gpu, ora, monkeypatch, mode, chunk):
    monkeypatch.setenv("DRESS_HIP_LIFT_CHUNK_VALUES", str(chunk))
    rng = np.random.default_rng(31)
    n, m = 11, 30
    U = rng.integers(0, n, m).astype(np.int32)
    V = rng.integers(0, n, m).astype(np.int32)
    W = rng.random(m) + 0.1
    for kind in ("delta", "nabla"):
        for variant in (0, 2):
            for k, ns in ((1, 0), (2, 0), (3, 25)):
                a = _run(gpu, mode, kind, n, U, V, W, None, variant, k=k, n_samples=ns, seed=3,
                         keep_multisets=True)
                b = ora.lifted(kind, n, U, V, W, variant=variant, k=k, n_samples=ns, seed=3, keep_multisets=True)
                _same(a, b)
    # DIRECTED with weight patches that move between subgraphs: each
    # chunk carries its own items' patches
    n, U, V, W = _directed_with_reciprocals()
    for k in (1, 2):
        a = _run(gpu, mode, "delta", n, U, V, W, None, 1, k=k, keep_multisets=True)
        b = ora.lifted("delta", n, U, V, W, None, variant=1, k=k, keep_multisets=True)
        _same(a, b)


def test_chunked_large_parent_matches_unchunked(gpu, ora, gen, monkeypatch):
    """A parent beyond the LDS state (multi-subgraph passes): 200 sampled
    delta^1 subgraphs (190 distinct picks) of ~10k edges in chunks of 300k
    values equal the oracle and the call without chunks."""
    n, U, V, _ = gen.gnp(2000, 10.0, 23)
    whole = _run(gpu, 0, "delta", n, U, V, k=1, n_samples=200, seed=8)
    monkeypatch.setenv("DRESS_HIP_LIFT_CHUNK_VALUES", "300000")
    parts = _run(gpu, 0, "delta", n, U, V, k=1, n_samples=200, seed=8)
    st = gpu.last_lifted_stats()
    _same(whole, parts)
    b = ora.lifted("delta", n, U, V, k=1, n_samples=200, seed=8, use_omp=True)
    _same(parts, b)
    # the sampled picks are deduplicated (delta_dress_impl.c:234-254)
    sel, _ = gpu.enumerate_subgraphs("delta", n, 1, 200, 8)
    assert st["subgraphs"] == sel.size and st["values"] == int(b[1].sum())
