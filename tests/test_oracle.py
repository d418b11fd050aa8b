"""The CPU oracle, pinned before it is trusted as the checker.

* bit-for-bit against libdress v0.8.3 compiled unchanged (oracle/_ref), over
  every variant, weights, vertex weights, repeated edges and self-loops, the
  intercept and merge paths, the OpenMP paths and the delta/nabla drivers
  with sampling, striding and multisets;
* against the committed golden fixtures (written by tests/golden/make_golden.py
  from the compiled reference);
* against what the mathematics fixes independently of any implementation:
  closed-form fixed points of edge-transitive graphs, delta^k of complete
  graphs (= 2), exact rounding of the sort + KBN sum (math.fsum), relabelling
  invariance, one iteration restated over a dict adjacency with exact sums,
  and histogram totals.
"""
import itertools
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN
from graphs import named_graphs, both_directions, complete, complete_bipartite, petersen, rook4


def _rand_multigraph(rng, n, m):
    U = rng.integers(0, n, m).astype(np.int32)
    V = rng.integers(0, n, m).astype(np.int32)
    return U, V, rng.random(m) + 0.1, rng.random(n) + 0.5


# ---------------------------------------------------------------- vs reference

@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("precompute", [0, 1])
@pytest.mark.parametrize("weighted", [False, True])
def test_fit_bitwise_vs_reference(ora, ref, variant, precompute, weighted):
    rng = np.random.default_rng(100 + variant * 7 + precompute)
    U, V, W, NW = _rand_multigraph(rng, 50, 220)
    if not weighted:
        W, NW = None, None
    for par in (False, True):
        a = ora.fit(50, U, V, W, NW, variant, precompute, 100, 1e-9, parallel=par)
        b = ref.fit(50, U, V, W, NW, variant, precompute, 100, 1e-9, parallel=par)
        assert a["iterations"] == b["iterations"]
        assert a["delta"] == b["delta"]
        assert np.array_equal(a["edge_dress"], b["edge_dress"])
        assert np.array_equal(a["vertex_dress"], b["vertex_dress"])
        assert np.array_equal(a["edge_weight"], b["edge_weight"])


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_construction_bitwise_vs_reference(ora, ref, variant):
    rng = np.random.default_rng(5 + variant)
    U, V, W, NW = _rand_multigraph(rng, 30, 120)
    ga = ora.graph(30, U, V, W, NW, variant, 1)
    gb = ref.graph(30, U, V, W, NW, variant, 1)
    ca, cb = ora.csr(ga), ref.csr(gb)
    ora.free_graph(ga)
    ref.free_graph(gb)
    for k in cb:
        assert np.array_equal(ca[k], cb[k]), k


def test_fit_er2000_vs_reference(ora, ref, gen):
    N, U, V = gen.er_both_directions()
    a = ora.fit(N, U, V, parallel=True)
    b = ref.fit(N, U, V, parallel=True)
    assert (a["iterations"], a["delta"]) == (b["iterations"], b["delta"])
    assert np.array_equal(a["edge_dress"], b["edge_dress"])


def test_max_iterations_edges(ora, ref):
    n, U, V = petersen()
    for it in (0, 1, 2):
        a = ora.fit(n, U, V, max_iterations=it, epsilon=1e-12)
        b = ref.fit(n, U, V, max_iterations=it, epsilon=1e-12)
        assert a["iterations"] == b["iterations"] == it
        assert np.array_equal(a["edge_dress"], b["edge_dress"])


@pytest.mark.parametrize("kind", ["delta", "nabla"])
@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_lifted_bitwise_vs_reference(ora, ref, kind, variant):
    rng = np.random.default_rng(900 + variant)
    U, V, W, _ = _rand_multigraph(rng, 9, 22)
    for k in (0, 1, 2, 3, 9):
        for ns, seed in ((0, 0), (4, 11)):
            for off, st in ((0, 1), (2, 3)):
                a = ora.lifted(kind, 9, U, V, W, variant=variant, k=k, n_samples=ns, seed=seed,
                               keep_multisets=True, offset=off, stride=st)
                b = ref.lifted(kind, 9, U, V, W, variant=variant, k=k, n_samples=ns, seed=seed,
                               keep_multisets=True, offset=off, stride=st)
                assert a[3] == b[3]
                assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
                if a[2] is None or b[2] is None:
                    assert a[2] is None and b[2] is None
                else:
                    assert np.array_equal(a[2], b[2], equal_nan=True)


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("weighted", [False, True])
def test_get_bitwise_vs_reference(ora, ref, variant, weighted):
    """dress_get (dress.c:651-739) in every variant, with edge and vertex
    weights: existing edges (the binary search's copy of a repeated edge)
    and virtual edges (common neighbours of both rows, self term, cross
    factor and the 1-D fixed point), on a fitted multigraph with loops."""
    rng = np.random.default_rng(1200 + variant + 10 * weighted)
    n = 14
    U, V, W, _ = _rand_multigraph(rng, n, 40)
    NW = rng.random(n) + 0.5 if weighted else None
    Wg = W if weighted else None
    ga = ora.graph(n, U, V, Wg, NW, variant, 0)
    gb = ref.graph(n, U, V, Wg, NW, variant, 0)
    try:
        ora.fit_graph(ga, 100, 1e-8)
        ref.fit_graph(gb, 100, 1e-8)
        qu = rng.integers(0, n, 300)
        qv = rng.integers(0, n, 300)
        qw = rng.random(300) + 0.25
        for a, b, w in zip(qu.tolist(), qv.tolist(), qw.tolist()):
            for it, eps in ((100, 1e-8), (3, 0.0), (0, 1e-6)):
                x = ora.get(ga, a, b, it, eps, w)
                y = ref.get(gb, a, b, it, eps, w)
                assert (x == y) or (x != x and y != y), (variant, a, b, w, it, x, y)
    finally:
        ora.free_graph(ga)
        ref.free_graph(gb)


def test_lifted_omp_vs_reference(ora, ref, gen):
    n, U, V = gen.random_regular(36, 6, 2)
    a = ora.lifted("delta", n, U, V, k=1, use_omp=True)
    b = ref.lifted("delta", n, U, V, k=1, use_omp=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# ---------------------------------------------------------------- vs golden

def test_golden_er2000(ora):
    z = np.load(os.path.join(GOLDEN, "er2000_both.npz"))
    r = ora.fit(2000, z["U"], z["V"], max_iterations=100, epsilon=1e-6)
    assert r["iterations"] == int(z["iterations"]) == 27
    assert r["delta"] == float(z["delta"])
    assert np.array_equal(r["edge_dress"], z["edge_dress"])
    assert np.array_equal(r["vertex_dress"], z["vertex_dress"])


def test_golden_generator_er2000(gen):
    z = np.load(os.path.join(GOLDEN, "er2000_both.npz"))
    N, U, V = gen.er_both_directions()
    assert N == 2000 and U.size == 16040
    assert np.array_equal(U, z["U"]) and np.array_equal(V, z["V"])


def test_golden_variants(ora):
    z = np.load(os.path.join(GOLDEN, "variants.npz"))
    for var in range(4):
        for pre in (0, 1):
            key = f"v{var}_p{pre}"
            r = ora.fit(40, z[key + "_U"], z[key + "_V"], z[key + "_W"], z[key + "_NW"], var, pre,
                        100, 1e-8)
            assert r["iterations"] == int(z[key + "_it"][0])
            assert np.array_equal(r["edge_dress"], z[key + "_edge_dress"])
            assert np.array_equal(r["vertex_dress"], z[key + "_vertex_dress"])
            g = ora.graph(40, z[key + "_U"], z[key + "_V"], z[key + "_W"], z[key + "_NW"], var, 1)
            c = ora.csr(g)
            ora.free_graph(g)
            for f in ("adj_offset", "adj_target", "adj_edge_idx", "intercept_offset",
                      "intercept_edge_ux", "intercept_edge_vx"):
                assert np.array_equal(c[f], z[key + "_" + f]), f


def test_golden_named_graphs(ora):
    with open(os.path.join(GOLDEN, "named_graphs.json")) as f:
        named = json.load(f)
    for name, rec in named.items():
        U = np.array(rec["U"], np.int32)
        V = np.array(rec["V"], np.int32)
        for key in ("delta0", "delta1", "nabla1"):
            kind, k = key[:-1], int(key[-1])
            vals, cnts, _, num = ora.lifted(kind, rec["n"], U, V, k=k)
            assert num == rec[key]["items"]
            assert [float(x).hex() for x in vals] == rec[key]["values"], (name, key)
            assert cnts.tolist() == rec[key]["counts"], (name, key)


def test_golden_rook_shrikhande_published(ora):
    # values printed in notebooks/delta_dress_rook_shrikhande.ipynb (6 decimals)
    expect = {"rook4": [(0.991596, 96), (1.209662, 288), (1.350236, 288)],
              "shrikhande": [(1.081855, 96), (1.182154, 96), (1.211156, 192), (1.323048, 192),
                             (1.333430, 96)]}
    g = named_graphs()
    for name, ex in expect.items():
        n, U, V = g[name]
        vals, cnts, _, num = ora.lifted("delta", n, U, V, k=1)
        assert num == 16
        assert len(vals) == len(ex)
        for (v, c), (ev, ec) in zip(zip(vals, cnts), ex):
            assert abs(v - ev) < 1.5e-6 and c == ec


def test_golden_get_queries(ora):
    z = np.load(os.path.join(GOLDEN, "get_queries.npz"))
    g = ora.graph(80, z["U"], z["V"], None, None, 0, 0)
    ora.fit_graph(g, 100, 1e-6)
    got = np.array([ora.get(g, int(a), int(b), 100, 1e-6, float(w))
                    for a, b, w in zip(z["qu"], z["qv"], z["qw"])])
    ora.free_graph(g)
    assert np.array_equal(got, z["ans"])


# ---------------------------------------------------------------- mathematics

def _closed_single(r, lam):
    return (lam + math.sqrt(lam * lam + 4 * r)) / r


def _closed_both(r, lam):
    return (lam + math.sqrt(lam * lam + 2 * r)) / r


@pytest.mark.parametrize("g,r,lam,both", [
    (complete(3), 2, 1, False), (complete(4), 3, 2, False), (complete(5), 4, 3, False),
    (complete_bipartite(3, 3), 3, 0, False), (petersen(), 3, 0, False), (rook4(), 6, 2, False),
    (complete(3), 2, 1, True), (complete_bipartite(3, 3), 3, 0, True),
])
def test_closed_form_fixed_points(ora, g, r, lam, both):
    n, U, V = both_directions(g) if both else g
    res = ora.fit(n, U, V, max_iterations=200, epsilon=1e-15)
    x = _closed_both(r, lam) if both else _closed_single(r, lam)
    assert np.all(res["edge_dress"] == res["edge_dress"][0])   # edge-transitive: one value
    assert abs(res["edge_dress"][0] - x) < 1e-12


def test_complete_graph_lifted_is_two(ora):
    for n, k in ((3, 1), (4, 1), (4, 2), (5, 2)):
        _, U, V = complete(n)
        vals, cnts, _, num = ora.lifted("delta", n, U, V, k=k, epsilon=1e-12)
        assert num == math.comb(n, k)
        assert np.all(np.abs(vals - 2.0) < 1e-9)
        m = n - k
        assert cnts.sum() == math.comb(n, k) * m * (m - 1) // 2


def test_kbn_sum_is_correctly_rounded(ora):
    import ctypes
    rng = np.random.default_rng(4)
    for n in (1, 2, 3, 7, 30, 300, 3000):
        for _ in range(40):
            x = rng.random(n) * rng.choice([1e-3, 1.0, 8.0], n)
            buf = np.ascontiguousarray(x.copy())
            s = ora.lib.ora_kbn_sorted_sum(buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n)
            assert s == math.fsum(x)


def _simple(U, V, W):
    """drop self-loops and repeated pairs (parallel edges of different weights
    pair up in edge-index order, which relabelling changes)"""
    a, b = np.minimum(U, V), np.maximum(U, V)
    _, first = np.unique(a.astype(np.int64) * 100000 + b, return_index=True)
    first = np.sort(first[a[first] != b[first]])
    return U[first], V[first], (W[first] if W is not None else None)


def test_relabel_invariance_bitwise(ora):
    rng = np.random.default_rng(12)
    n = 60
    U, V, W, _ = _rand_multigraph(rng, n, 300)
    U, V, W = _simple(U, V, W)
    base = np.sort(ora.fit(n, U, V, W, None, 0, 0, 100, 1e-10)["edge_dress"])
    for t in range(3):
        perm = rng.permutation(n).astype(np.int32)
        order = rng.permutation(U.size)
        r = ora.fit(n, perm[U][order], perm[V][order], W[order], None, 0, 0, 100, 1e-10)
        assert np.array_equal(np.sort(r["edge_dress"]), base)


def _one_iteration_dict(n, U, V, W):
    """One UNDIRECTED iteration from d = 1 over a dict adjacency, exact sums."""
    adj = {u: [] for u in range(n)}
    for e, (u, v) in enumerate(zip(U, V)):
        w = 2.0 * (W[e] if W is not None else 1.0)
        adj[u].append((v, e, w))
        adj[v].append((u, e, w))
    norm = [math.sqrt(math.fsum([4.0] + [w * 1.0 for _, _, w in adj[u]])) for u in range(n)]
    out = []
    for e, (u, v) in enumerate(zip(U, V)):
        nu, nv = {}, {}
        for x, f, w in adj[u]:
            nu.setdefault(x, []).append((f, w))
        for x, f, w in adj[v]:
            nv.setdefault(x, []).append((f, w))
        terms = []
        for x in nu:
            if x in nv:
                for (fa, wa), (fb, wb) in zip(sorted(nu[x]), sorted(nv[x])):
                    terms.append(wa * 1.0 + wb * 1.0)
        w = 2.0 * (W[e] if W is not None else 1.0)
        terms.append(4.0 + 4.0 + 2.0 * (w * 1.0))
        den = norm[u] * norm[v]
        out.append(math.fsum(terms) / den if den > 0 else 0.0)
    return np.array(out)


def test_one_iteration_independent_restatement(ora):
    rng = np.random.default_rng(31)
    for t in range(5):
        n = 25
        U, V, W, _ = _rand_multigraph(rng, n, 90)
        keep = U != V
        U, V, W = U[keep], V[keep], W[keep]
        got = ora.fit(n, U, V, W, None, 0, 0, max_iterations=1, epsilon=0.0)["edge_dress"]
        assert np.array_equal(got, _one_iteration_dict(n, U, V, W))


def test_histogram_totals(ora, gen):
    n, U, V = gen.random_regular(30, 6, 9)
    vals, cnts, _, num = ora.lifted("delta", n, U, V, k=2)
    live = 0
    for a, b in itertools.combinations(range(n), 2):
        live += int(np.sum((U != a) & (U != b) & (V != a) & (V != b)))
    assert num == math.comb(n, 2) and int(cnts.sum()) == live
    vals, cnts, _, num = ora.lifted("nabla", n, U, V, k=1)
    assert num == n and int(cnts.sum()) == n * U.size


def test_sample_picks_match_reference_multisets(ora, ref, gen):
    # the rows a sampled run fills are exactly the picks, for both libraries
    n, U, V = gen.random_regular(12, 4, 1)
    total = math.comb(n, 2)
    for seed in (0, 5, 77):
        b = ref.lifted("delta", n, U, V, k=2, n_samples=9, seed=seed, keep_multisets=True)
        rows = np.where(~np.all(np.isnan(b[2]), axis=1))[0]
        picks = ora.sample_picks(total, 9, seed)
        assert np.array_equal(rows, picks)
