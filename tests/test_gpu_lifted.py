"""delta^k / nabla^k on the GPU against the oracle and the golden fixtures.

FP64: exact (value, count) histograms and multisets bit for bit.
FP32->FP64: values within 1e-9, counts equal, and the quantised histogram
floor(d/eps) bit-exact once values within the tolerance of a bin edge are
set apart on both sides.
"""
import hashlib
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _lifted(gpu, kind, n, U, V, W=None, variant=0, **kw):
    with gpu.Graph(n, U, V, W, None, variant) as g:
        return g.lifted(kind, **kw)


def _digest(vals, cnts):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(vals, dtype=np.float64).tobytes())
    h.update(np.ascontiguousarray(cnts, dtype=np.int64).tobytes())
    return h.hexdigest()


def test_named_graphs_golden(gpu):
    with open(os.path.join(GOLDEN, "named_graphs.json")) as f:
        named = json.load(f)
    for name, rec in named.items():
        U = np.array(rec["U"], np.int32)
        V = np.array(rec["V"], np.int32)
        for key in ("delta0", "delta1", "nabla1"):
            kind, k = key[:-1], int(key[-1])
            vals, cnts, _, num = _lifted(gpu, kind, rec["n"], U, V, k=k)
            assert num == rec[key]["items"]
            assert [float(x).hex() for x in vals] == rec[key]["values"], (name, key)
            assert cnts.tolist() == rec[key]["counts"], (name, key)


@pytest.mark.parametrize("kind", ["delta", "nabla"])
@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_random_multigraphs_bitwise(gpu, ora, kind, variant):
    rng = np.random.default_rng(300 + variant)
    n, m = 10, 26
    U = rng.integers(0, n, m).astype(np.int32)
    V = rng.integers(0, n, m).astype(np.int32)
    W = rng.random(m) + 0.1
    for k in (0, 1, 2, 3, 10):
        for ns, seed in ((0, 0), (6, 9)):
            for off, st in ((0, 1), (1, 2)):
                a = _lifted(gpu, kind, n, U, V, W, variant, k=k, n_samples=ns, seed=seed, keep_multisets=True,
                            offset=off, stride=st)
                b = ora.lifted(kind, n, U, V, W, variant=variant, k=k, n_samples=ns, seed=seed,
                               keep_multisets=True, offset=off, stride=st)
                assert a[3] == b[3]
                assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (k, ns, off)
                assert (a[2] is None) == (b[2] is None)
                if a[2] is not None:
                    assert np.array_equal(a[2], b[2], equal_nan=True)


def test_batched_equals_per_subgraph_path(gpu, gen):
    n, U, V = gen.random_regular(40, 6, 3)
    a = _lifted(gpu, "delta", n, U, V, k=2, n_samples=60, seed=4)
    gpu.set_lifted_mode(1)
    try:
        b = _lifted(gpu, "delta", n, U, V, k=2, n_samples=60, seed=4)
    finally:
        gpu.set_lifted_mode(0)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_delta1_batch_golden(gpu, gen):
    with open(os.path.join(GOLDEN, "delta1_batch16.json")) as f:
        gold = json.load(f)
    graphs = gen.delta1_batch_graphs(16)
    r = gpu.lifted_batch(graphs, k=1, precision=gpu.FP64)
    assert r.subgraphs == sum(g["items"] for g in gold)
    for (vals, cnts), rec, (n, U, V) in zip(r.histograms, gold, graphs):
        assert n == rec["n"] and U.size == rec["m"]
        assert hashlib.sha256(np.stack([U, V]).tobytes()).hexdigest() == rec["edges_sha256"]
        assert int(cnts.sum()) == rec["total"] and vals.size == rec["distinct"]
        assert _digest(vals, cnts) == rec["sha256"]


def _quantise(vals, cnts, eps, tol, flag_vals):
    """oracle bins of values, with boundary flags decided on flag_vals"""
    r = flag_vals / eps
    q = np.floor(r)
    bnd = ((r - q) * eps < tol) | ((q + 1 - r) * eps < tol)
    qo = np.floor(vals / eps).astype(np.int64)
    keys, inv = np.unique(qo[~bnd], return_inverse=True)
    return keys, np.bincount(inv, weights=cnts[~bnd]).astype(np.int64) if keys.size else np.zeros(0, np.int64)


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_quantised_histograms(gpu, ora, gen, precision):
    graphs = gen.delta1_batch_graphs(12)
    eps, tol = 1e-6, 1e-9
    r = gpu.lifted_batch(graphs, k=1, precision=precision, epsilon=eps, qtol=tol)
    for (hv, hc), (qk, qc), (n, U, V) in zip(r.histograms, r.quantised, graphs):
        ov, oc, _, _ = ora.lifted("delta", n, U, V, k=1, epsilon=eps)
        assert np.array_equal(hc, oc) and hv.size == ov.size
        if precision == 0:
            assert np.array_equal(hv, ov)
        else:
            assert np.max(np.abs(hv - ov)) <= 1e-9
        # every value counted once: table counts + boundary counts = total
        keys, cnt = _quantise(ov, oc, eps, tol, hv)
        assert np.array_equal(qk, keys) and np.array_equal(qc, cnt)
    tot = sum(int(c.sum()) for _, c in r.quantised) + sum(int(c.sum()) for _, c in r.boundary)
    assert tot == r.values


@pytest.mark.parametrize("kind,k", [("delta", 1), ("delta", 2), ("nabla", 1)])
def test_batch_parents_without_common_neighbours(gpu, ora, gen, kind, k):
    """Parents whose intersection index is empty (a path, a star, K3,3, an odd
    cycle) batched with one that has triangles: each graph's histogram equals
    the oracle's."""
    i32 = np.int32
    graphs = [
        (6, np.arange(5, dtype=i32), np.arange(1, 6, dtype=i32)),
        (6, np.zeros(5, i32), np.arange(1, 6, dtype=i32)),
        (6, np.repeat(np.arange(3, dtype=i32), 3), np.tile(np.arange(3, 6, dtype=i32), 3)),
        (7, np.arange(7, dtype=i32), ((np.arange(7) + 1) % 7).astype(i32)),
        gen.random_regular(12, 4, 5),
    ]
    r = gpu.lifted_batch(graphs, k=k, kind=kind)
    assert len(r.histograms) == len(graphs)
    for (hv, hc), (n, U, V) in zip(r.histograms, graphs):
        ov, oc, _, _ = ora.lifted(kind, n, U, V, k=k)
        assert np.array_equal(hv, ov) and np.array_equal(hc, oc), (kind, k, n, U.size)


def test_mixed_batch_quantised_after_fp64_redo(gpu, gen):
    """Items whose mixed-precision stop was ambiguous are fitted again in
    FP64; their mixed values' counts leave the quantised tables first, so the
    tables of every graph are those of its final (returned) values."""
    graphs = gen.delta1_batch_graphs(160)
    eps, tol = 1e-6, 1e-9
    r = gpu.lifted_batch(graphs, k=1, precision=1, epsilon=eps, qtol=tol)
    assert r.fp64_reruns > 0
    for (hv, hc), (qk, qc), (bk, bc) in zip(r.histograms, r.quantised, r.boundary):
        keys, cnt = _quantise(hv, hc, eps, tol, hv)
        assert np.array_equal(qk, keys) and np.array_equal(qc, cnt)
        q = hv / eps
        f = np.floor(q)
        near = ((q - f) * eps < tol) | ((f + 1 - q) * eps < tol)
        bkeys, binv = np.unique(f[near].astype(np.int64), return_inverse=True)
        assert np.array_equal(bk, bkeys)
        assert np.array_equal(bc, np.bincount(binv, weights=hc[near]).astype(np.int64) if bkeys.size else bc)
        assert np.all(qc > 0) and np.all(bc > 0)
    tot = sum(int(c.sum()) for _, c in r.quantised) + sum(int(c.sum()) for _, c in r.boundary)
    assert tot == r.values


def test_regular12_n256_nabla1_golden(gpu):
    z = np.load(os.path.join(GOLDEN, "regular12_n256.npz"))
    with open(os.path.join(GOLDEN, "regular12_n256_nabla1.json")) as f:
        gold = json.load(f)
    vals, cnts, _, num = _lifted(gpu, "nabla", 256, z["U"], z["V"], k=1)
    assert num == gold["items"] == 256
    assert int(cnts.sum()) == gold["total"] and vals.size == gold["distinct"]
    assert _digest(vals, cnts) == gold["sha256"]


def test_regular12_n256_delta2(gpu, ora):
    z = np.load(os.path.join(GOLDEN, "regular12_n256.npz"))
    U, V = z["U"], z["V"]
    # all 32,640 subgraphs: histogram total = live edges of every subgraph
    vals, cnts, _, num = _lifted(gpu, "delta", 256, U, V, k=2)
    assert num == math.comb(256, 2)
    assert int(cnts.sum()) == num * 1536 - 255 * 2 * 1536 + 1536
    # a sampled slice bit for bit against the oracle
    a = _lifted(gpu, "delta", 256, U, V, k=2, n_samples=150, seed=21, keep_multisets=True)
    b = ora.lifted("delta", 256, U, V, k=2, n_samples=150, seed=21, keep_multisets=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2], equal_nan=True)


def test_mixed_precision_lifted(gpu, ora, gen):
    n, U, V = gen.random_regular(64, 6, 17)
    gpu.set_precision(gpu.FP32_FP64)
    try:
        a = _lifted(gpu, "delta", n, U, V, k=1)
        st = gpu.last_lifted_stats()
    finally:
        gpu.set_precision(gpu.FP64)
    b = ora.lifted("delta", n, U, V, k=1)
    assert np.array_equal(a[1], b[1])
    assert np.max(np.abs(a[0] - b[0])) <= 1e-9
    assert st["subgraphs"] == n


def test_block_size_does_not_change_results(gpu, gen, monkeypatch):
    """The autotuner and forced block sizes only move work between threads:
    every sum is exactly rounded by one thread, so histograms are equal."""
    graphs = gen.delta1_batch_graphs(64)        # ~4.7k subgraphs: enough to calibrate
    ref = None
    for blk in ("64", "512", None):
        if blk:
            monkeypatch.setenv("DRESS_HIP_LIFT_BLOCK", blk)
        else:
            monkeypatch.delenv("DRESS_HIP_LIFT_BLOCK", raising=False)
        r = gpu.lifted_batch(graphs, k=1)
        cur = [(np.array(v), np.array(c)) for v, c in r.histograms]
        if ref is None:
            ref = cur
        else:
            for (v0, c0), (v1, c1) in zip(ref, cur):
                assert np.array_equal(v0, v1) and np.array_equal(c0, c1)


def test_quantised_tables_on_request(gpu, gen, monkeypatch):
    """Without qtol the floor(d/eps) tables stay empty and the exact
    histograms are those of a call with them; DRESS_HIP_LIFT_WORK=1 counts
    the subgraphs' work (iterations x live entries, half-edges, edges,
    vertices) without changing results."""
    graphs = gen.delta1_batch_graphs(20)
    a = gpu.lifted_batch(graphs, k=1)
    b = gpu.lifted_batch(graphs, k=1, qtol=1e-9)
    assert all(len(k) == 0 for k, _ in a.quantised) and all(len(k) == 0 for k, _ in a.boundary)
    assert sum(int(c.sum()) for _, c in b.quantised) + sum(int(c.sum()) for _, c in b.boundary) == b.values
    for (v0, c0), (v1, c1) in zip(a.histograms, b.histograms):
        assert np.array_equal(v0, v1) and np.array_equal(c0, c1)
    assert a.work == (0, 0, 0, 0)
    monkeypatch.setenv("DRESS_HIP_LIFT_WORK", "1")
    w = gpu.lifted_batch(graphs, k=1)
    for (v0, c0), (v1, c1) in zip(a.histograms, w.histograms):
        assert np.array_equal(v0, v1) and np.array_equal(c0, c1)
    T, S, E, N = w.work
    # every subgraph of a 6-regular graph minus one vertex: n-1 vertices,
    # 3n-6 edges, 2(3n-6) half-edges, each times its iterations (>= 1)
    lo_e = sum(n * (3 * n - 6) for n, _, _ in graphs)
    assert E >= lo_e and S == 2 * E and N >= sum(n * (n - 1) for n, _, _ in graphs) and T >= 0


@pytest.mark.parametrize("kind,k,precision", [("delta", 1, 0), ("delta", 1, 1), ("delta", 2, 0), ("nabla", 1, 0)])
def test_grouped_batch_equals_one_group(gpu, ora, gen, monkeypatch, kind, k, precision):
    """A batch run as groups of consecutive graphs (parents of later groups
    built on the host and histograms of earlier ones downloaded while the
    device works) returns what the same batch run as one group returns, with
    graphs that have no subgraph (k >= n) at group edges; every graph's
    histogram equals the oracle's in FP64."""
    i32 = np.int32
    graphs = list(gen.delta1_batch_graphs(40, n_lo=8, n_hi=40))
    graphs[0] = (2, np.zeros(1, i32), np.ones(1, i32))          # k >= n for k = 2
    graphs[17] = (1, np.zeros(0, i32), np.zeros(0, i32))
    graphs[39] = (2, np.zeros(1, i32), np.ones(1, i32))
    monkeypatch.setenv("DRESS_HIP_BATCH_GROUP_VALUES", "0")
    one = gpu.lifted_batch(graphs, k=k, kind=kind, precision=precision)
    for per_group in ("1", "60000"):                             # 8 groups, then groups of ~60k values
        monkeypatch.setenv("DRESS_HIP_BATCH_GROUP_VALUES", per_group)
        r = gpu.lifted_batch(graphs, k=k, kind=kind, precision=precision)
        assert (r.subgraphs, r.values) == (one.subgraphs, one.values)
        for (v0, c0), (v1, c1) in zip(one.histograms, r.histograms):
            assert np.array_equal(v0, v1) and np.array_equal(c0, c1)
    if precision == 0:
        for (hv, hc), (n, U, V) in zip(r.histograms, graphs):
            ov, oc, _, _ = ora.lifted(kind, n, U, V, k=k)
            assert np.array_equal(hv, ov) and np.array_equal(hc, oc), (kind, k, n)


def _complete_graph(n):
    iu = np.triu_indices(n, 1)
    return (n, iu[0].astype(np.int32), iu[1].astype(np.int32))


def test_grouped_batch_narrow_counts(gpu, ora, gen, monkeypatch):
    """A pipelined group's counts travel at the narrowest width that holds
    its largest count (1, 2 or 4 bytes) and are widened on host threads.
    Complete graphs, whose delta^1 subgraphs are all the same graph, put
    counts past 255 (K_12) and past 65,535 (K_64) into groups beside graphs
    of distinct values; every histogram equals the one-group call's and the
    int64 download's, and the oracle's."""
    graphs = list(gen.delta1_batch_graphs(12, n_lo=8, n_hi=30))
    graphs[3], graphs[7], graphs[10] = _complete_graph(12), _complete_graph(64), _complete_graph(5)
    oracle = [ora.lifted("delta", n, U, V, k=1)[:2] for n, U, V in graphs]
    assert 255 < oracle[3][1].max() <= 65535 and oracle[7][1].max() > 65535 and oracle[10][1].max() <= 255
    monkeypatch.setenv("DRESS_HIP_BATCH_GROUP_VALUES", "0")
    one = gpu.lifted_batch(graphs, k=1)
    monkeypatch.setenv("DRESS_HIP_BATCH_GROUP_VALUES", "1")                 # 8 groups
    runs = []
    for narrow in ("1", "0"):
        monkeypatch.setenv("DRESS_HIP_NARROW_COUNTS", narrow)
        for _ in range(2):                                                  # staging blocks reused
            runs.append(gpu.lifted_batch(graphs, k=1))
    for r in runs:
        for (v0, c0), (v1, c1), (ov, oc) in zip(one.histograms, r.histograms, oracle):
            assert np.array_equal(v0, v1) and np.array_equal(c0, c1)
            assert np.array_equal(v1, ov) and np.array_equal(c1, oc)


@pytest.mark.parametrize("m,width", [(5, 2), (40, 4)])
def test_large_flat_histogram_narrow_counts(gpu, ora, gen, m, width):
    """A single-graph histogram of 2^20 runs or more comes down as values and
    counts narrowed to the largest count's width, interleaved on the host:
    G(700, p) beside a complete K_m, whose edges keep one value in every
    subgraph that keeps K_m (a count past 255 for K_5, past 65,535 for
    K_40), equals the oracle's histogram."""
    n, U, V, _ = gen.gnp(700, 6.0, 3)
    iu = np.triu_indices(m, 1)
    U = np.concatenate([U, (iu[0] + n).astype(np.int32)])
    V = np.concatenate([V, (iu[1] + n).astype(np.int32)])
    ov, oc, _, _ = ora.lifted("delta", n + m, U, V, k=1)
    assert ov.size >= (1 << 20) and 1 << (8 * (width // 2)) < oc.max() < 1 << (8 * width)
    hv, hc = _lifted(gpu, "delta", n + m, U, V, k=1)[:2]
    assert np.array_equal(hv, ov) and np.array_equal(hc, oc)


def test_histogram_key_layouts_agree(gpu, ora, gen, monkeypatch):
    """The exact histogram sorts one 64-bit key per value holding (graph,
    value) when the values' keys span few enough bits to leave room for the
    graph ids, and otherwise sorts by value and then stably by graph: both
    give the same histograms, for a batch (many graphs) and for a reference-
    API call (one), and a weighted graph whose values span many binades
    equals the oracle."""
    graphs = gen.delta1_batch_graphs(24, n_lo=8, n_hi=40)
    n, U, V = gen.random_regular(30, 4, 2)
    rng = np.random.default_rng(3)
    W = 10.0 ** rng.uniform(-120, 120, U.size)                   # values over many binades
    res = []
    for packed in ("1", "0"):
        monkeypatch.setenv("DRESS_HIP_HIST_PACKED", packed)
        r = gpu.lifted_batch(graphs, k=1)
        res.append(([(np.array(v), np.array(c)) for v, c in r.histograms],
                    _lifted(gpu, "delta", n, U, V, k=2), _lifted(gpu, "delta", n, U, V, W, k=1)))
    for (v0, c0), (v1, c1) in zip(res[0][0], res[1][0]):
        assert np.array_equal(v0, v1) and np.array_equal(c0, c1)
    for j in (1, 2):
        assert np.array_equal(res[0][j][0], res[1][j][0]) and np.array_equal(res[0][j][1], res[1][j][1])
    ov, oc, _, _ = ora.lifted("delta", n, U, V, W, k=1)
    assert np.array_equal(res[0][2][0], ov) and np.array_equal(res[0][2][1], oc)
    assert ov.min() < 1e-30 < ov.max()


@pytest.mark.parametrize("fail_group", [0, 1, 4, 7])
def test_grouped_batch_failure_leaves_the_library_usable(gpu, gen, monkeypatch, fail_group):
    """A group whose host preparation fails (a host-side hook, no device
    fault) while earlier groups are queued or running ends the call with that
    error; the queued work finishes before its blocks are released, and the
    next call returns the same histograms as an undisturbed one."""
    graphs = gen.delta1_batch_graphs(48, n_lo=8, n_hi=40)
    monkeypatch.setenv("DRESS_HIP_BATCH_GROUP_VALUES", "1")                 # 8 groups
    ref = gpu.lifted_batch(graphs, k=1)
    monkeypatch.setenv("DRESS_HIP_BATCH_FAIL_GROUP", str(fail_group))
    with pytest.raises(gpu.HipError, match=f"group {fail_group} failed"):
        gpu.lifted_batch(graphs, k=1)
    monkeypatch.delenv("DRESS_HIP_BATCH_FAIL_GROUP")
    again = gpu.lifted_batch(graphs, k=1)
    assert (again.subgraphs, again.values) == (ref.subgraphs, ref.values)
    for (v0, c0), (v1, c1) in zip(ref.histograms, again.histograms):
        assert np.array_equal(v0, v1) and np.array_equal(c0, c1)


@pytest.mark.parametrize("kind,k,precision", [("delta", 1, 0), ("delta", 2, 0), ("nabla", 1, 1), ("delta", 1, 1)])
def test_histogram_merge_equals_radix(gpu, ora, gen, monkeypatch, kind, k, precision):
    """With at most 256 items per graph, the batched kernel leaves each
    item's values sorted (a bitonic network in LDS) and the histogram merges
    each graph's sorted runs; the radix-sorted histogram of the same batch is
    the same, in one group and in several, and FP64 equals the oracle."""
    graphs = gen.delta1_batch_graphs(30, n_lo=8, n_hi=22 if k == 2 else 60)
    res = []
    for merge, groups in (("1", "0"), ("1", "1"), ("0", "0")):
        monkeypatch.setenv("DRESS_HIP_HIST_MERGE", merge)
        monkeypatch.setenv("DRESS_HIP_BATCH_GROUP_VALUES", groups)
        r = gpu.lifted_batch(graphs, k=k, kind=kind, precision=precision)
        res.append([(np.array(v), np.array(c)) for v, c in r.histograms])
    for other in res[1:]:
        for (v0, c0), (v1, c1) in zip(res[0], other):
            assert np.array_equal(v0, v1) and np.array_equal(c0, c1)
    if precision == 0:
        for (hv, hc), (n, U, V) in zip(res[0], graphs):
            ov, oc, _, _ = ora.lifted(kind, n, U, V, k=k)
            assert np.array_equal(hv, ov) and np.array_equal(hc, oc)


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_grouped_weighted_variants_equal_oracle(gpu, ora, gen, monkeypatch, variant):
    """Weighted graphs of every variant (DIRECTED with repeated reciprocal
    edges of different weights, whose subgraphs get patched weight copies)
    in a pipelined batch of 8 groups: each graph's delta^1 histogram equals
    the one-group batch's and the oracle's."""
    i32 = np.int32
    rng = np.random.default_rng(11 + variant)
    graphs, weights = [], []
    for i in range(24):
        n, U, V = gen.random_regular(int(rng.integers(8, 30)) // 2 * 2, 4, 100 + i)
        if variant == 1 and i % 3 == 0:
            # a repeated reciprocal pair of another weight: u->v, v->u twice
            U = np.concatenate([U, V[:1], V[:1]]).astype(i32)
            V = np.concatenate([V, U[:1], U[:1]]).astype(i32)
        graphs.append((n, U, V))
        weights.append(rng.uniform(0.2, 2.0, U.size))
    res = []
    for groups in ("0", "1"):
        monkeypatch.setenv("DRESS_HIP_BATCH_GROUP_VALUES", groups)
        r = gpu.lifted_batch(graphs, k=1, variant=variant, weights=weights)
        res.append([(np.array(v), np.array(c)) for v, c in r.histograms])
    for (v0, c0), (v1, c1), (n, U, V), W in zip(res[0], res[1], graphs, weights):
        ov, oc, _, _ = ora.lifted("delta", n, U, V, W, variant=variant, k=1)
        assert np.array_equal(v0, ov) and np.array_equal(c0, oc), (variant, n)
        assert np.array_equal(v1, ov) and np.array_equal(c1, oc), (variant, n)


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("weighted", [False, True])
def test_parents_read_without_lds_copy_equal_oracle(gpu, ora, variant, weighted):
    """A parent of 1,536 edges (the 12-regular graph on 256 vertices) is
    run without the LDS copy of its structure (the copy would leave fewer
    than 16 waves per CU) and, its rows being near-regular, with the
    column-by-column copy of its rows; unweighted parents take the variant
    weight as a constant, weighted ones read it per term.  delta^1 and
    nabla^1 with multisets, bit for bit against the oracle."""
    z = np.load(os.path.join(GOLDEN, "regular12_n256.npz"))
    U, V = z["U"], z["V"]
    W = np.random.default_rng(40 + variant).uniform(0.1, 1.0, U.size) if weighted else None
    for kind in ("delta", "nabla"):
        a = _lifted(gpu, kind, 256, U, V, W, variant, k=1, keep_multisets=True)
        b = ora.lifted(kind, 256, U, V, W, variant=variant, k=1, keep_multisets=True)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (kind, variant, weighted)
        assert np.array_equal(a[2], b[2], equal_nan=True), (kind, variant, weighted)
