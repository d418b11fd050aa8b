"""Parity at the benchmark sizes, against digests of the compiled reference.

The fixtures (tests/golden/make_golden.py --full) hold SHA-256 digests of
libdress v0.8.3's own outputs at full size:

* the FP64 fit of the 4M-vertex Chung-Lu graph (edge values in input order,
  vertex values, iterations, final delta, 4,096 values at fixed positions);
* the exact delta^1 histogram of each of the 512 graphs of the batch;
* the exact delta^2 histogram of the 12-regular n = 256 graph (32,640
  subgraphs, 49.4M values).

FP64 results must equal them bit for bit.  FP32->FP64 results are held to
the north-star tolerance against the FP64 results just proven equal to the
reference: the same number of values, every value (every edge of every
subgraph, through the multisets, where the output keeps edge identity)
within 1e-9, and the floor(d/eps) histogram bit-exact once values within
1e-9 of a bin edge are set apart.
"""
import hashlib
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

TOL = 1e-9


def _sha(a, dtype=np.float64):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=dtype).tobytes()).hexdigest()


def _hist_sha(vals, cnts):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(vals, dtype=np.float64).tobytes())
    h.update(np.ascontiguousarray(cnts, dtype=np.int64).tobytes())
    return h.hexdigest()


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def _hexvals(xs):
    return np.array([float.fromhex(x) for x in xs])


# --------------------------------------------------------------- config 2

@pytest.fixture(scope="module")
def chung_lu_4m(gen):
    gold = _load("chung_lu_4m_fit.json")
    n, U, V = gen.chung_lu()
    assert n == gold["N"] and U.size == gold["E"]
    assert hashlib.sha256(np.stack([U, V]).tobytes()).hexdigest() == gold["edges_sha256"]
    return gold, n, U, V


def test_power_law_4m_fp64_equals_reference(gpu, chung_lu_4m):
    """dress_fit through the C ABI (host arrays in and out, dress.c:546-608):
    values, vertex norms, iterations and delta bit for bit."""
    gold, n, U, V = chung_lu_4m
    with gpu.Graph(n, U, V) as g:
        it, dl = g.fit(100, 1e-6, gpu.FP64)
        d = np.array(g.edge_dress)
        vd = np.array(g.vertex_dress)
    pick = np.array(gold["edge_pick"])
    err = float(np.max(np.abs(d[pick] - _hexvals(gold["edge_pick_values"]))))
    print(f"4M-vertex FP64 fit: iterations {it}, delta {dl!r}, max |diff| on 4,096 fixed edges {err:.3e}")
    assert it == gold["iterations"]
    assert dl == float.fromhex(gold["delta"])
    assert err == 0.0
    assert _sha(d) == gold["edge_dress_sha256"]
    assert _sha(vd) == gold["vertex_dress_sha256"]


def test_power_law_4m_mixed_precision_error(gpu, chung_lu_4m):
    """FP32->FP64 on the resident graph against the FP64 fit of the same
    graph (itself equal to the reference's): same stopping step, every edge
    within 1e-9; values in (0, 2] up to that tolerance."""
    gold, n, U, V = chung_lu_4m
    with gpu.Graph(n, U, V) as g:
        g.resident_reset()
        it64, dl64 = g.fit_resident(100, 1e-6, gpu.FP64)
        g.resident_download()
        d64 = np.array(g.edge_dress)
        g.set_edge_dress(np.ones(U.size))
        g.resident_reset()
        it32, dl32 = g.fit_resident(100, 1e-6, gpu.FP32_FP64)
        st = gpu.last_fit_stats()
        g.resident_download()
        d32 = np.array(g.edge_dress)
    assert _sha(d64) == gold["edge_dress_sha256"]
    err = float(np.max(np.abs(d64 - d32)))
    print(f"4M-vertex FP32->FP64 fit: {st.fp32_iterations} FP32 iterations, {st.fp64_reruns} FP64 reruns, "
          f"max edge error {err:.3e}")
    assert it32 == it64 == gold["iterations"]
    assert err <= TOL
    # the fixed point lies in (0, 2]; the FP64 values (the reference's) do,
    # and the mixed ones within the tolerance of them
    assert np.all(d64 > 0) and np.all(d64 <= 2.0)
    assert np.all(d32 > 0) and np.all(d32 <= 2.0 + TOL)


# --------------------------------------------------------------- config 3

def _expand(vals, cnts):
    return np.repeat(np.asarray(vals, dtype=np.float64), np.asarray(cnts, dtype=np.int64))


def _quantised(v64, vm, eps, tol):
    """floor(d / eps) histogram of the FP64 values v64 (one entry per value,
    sorted), leaving out the entries whose FP32->FP64 value vm (sorted, same
    length) lies within tol of a bin edge: what the device's table of vm
    must hold"""
    r = vm / eps
    q = np.floor(r)
    near = ((r - q) * eps < tol) | ((q + 1 - r) * eps < tol)
    keys, counts = np.unique(np.floor(v64[~near] / eps).astype(np.int64), return_counts=True)
    return keys, counts.astype(np.int64), int(near.sum())


@pytest.fixture(scope="module")
def batch512(gen):
    gold = _load("delta1_batch512.json")
    graphs = gen.delta1_batch_graphs(512)
    assert len(graphs) == len(gold) == 512
    for (n, U, V), rec in zip(graphs, gold):
        assert n == rec["n"] and U.size == rec["m"]
        assert hashlib.sha256(np.stack([U, V]).tobytes()).hexdigest() == rec["edges_sha256"]
    return gold, graphs


def test_delta1_batch512_fp64_equals_reference(gpu, batch512):
    """All 512 graphs, every vertex-deleted subgraph, in one batched call:
    each graph's exact histogram equals the reference's digest."""
    gold, graphs = batch512
    r = gpu.lifted_batch(graphs, k=1, precision=gpu.FP64)
    assert r.subgraphs == sum(g["items"] for g in gold)
    bad = [i for i, ((v, c), rec) in enumerate(zip(r.histograms, gold)) if _hist_sha(v, c) != rec["sha256"]]
    assert not bad, f"graphs whose histogram differs from the reference: {bad[:20]}"


def test_delta1_batch512_mixed_precision(gpu, batch512):
    """FP32->FP64: per graph the same number of values, within 1e-9 of the FP64
    (= reference) values, and the floor(d/eps) histogram bit-exact after the
    values within 1e-9 of a bin edge are set apart."""
    gold, graphs = batch512
    eps = 1e-6
    ref = gpu.lifted_batch(graphs, k=1, precision=gpu.FP64, epsilon=eps, qtol=TOL)
    mix = gpu.lifted_batch(graphs, k=1, precision=gpu.FP32_FP64, epsilon=eps, qtol=TOL)
    worst, apart = 0.0, 0
    for i, ((v64, c64), (vm, cm), (qk, qc), (bk, bc)) in enumerate(
            zip(ref.histograms, mix.histograms, mix.quantised, mix.boundary)):
        assert _hist_sha(v64, c64) == gold[i]["sha256"], i
        e64, em = _expand(v64, c64), _expand(vm, cm)
        assert e64.size == em.size == gold[i]["total"], i
        e = float(np.max(np.abs(em - e64))) if em.size else 0.0
        worst = max(worst, e)
        keys, counts, near = _quantised(e64, em, eps, TOL)
        assert np.array_equal(np.asarray(qk), keys) and np.array_equal(np.asarray(qc), counts), i
        assert int(np.asarray(bc).sum()) == near, i
        apart += near
    tot = sum(int(np.asarray(c).sum()) for _, c in mix.quantised) + apart
    assert tot == mix.values
    print(f"delta^1 batch of 512, FP32->FP64: max value error {worst:.3e}, {mix.fp64_reruns} FP64 reruns, "
          f"{apart} of {mix.values} values within 1e-9 of a bin edge")
    assert worst <= TOL


# --------------------------------------------------------------- config 4

@pytest.fixture(scope="module")
def regular12():
    z = np.load(os.path.join(GOLDEN, "regular12_n256.npz"))
    return 256, z["U"], z["V"]


def _lifted(gpu, kind, n, U, V, precision, **kw):
    gpu.set_precision(precision)
    try:
        with gpu.Graph(n, U, V) as g:
            return g.lifted(kind, **kw)
    finally:
        gpu.set_precision(gpu.FP64)


def test_regular12_delta2_fp64_equals_reference(gpu, regular12):
    """Every one of the 32,640 subgraphs: the full exact histogram equals
    the reference's digest."""
    gold = _load("regular12_n256_delta2.json")
    n, U, V = regular12
    vals, cnts, _, num = _lifted(gpu, "delta", n, U, V, gpu.FP64, k=2)
    assert num == gold["items"] == math.comb(256, 2)
    assert vals.size == gold["distinct"] and int(cnts.sum()) == gold["total"]
    assert _hist_sha(vals, cnts) == gold["sha256"]


@pytest.mark.parametrize("kind,k", [("delta", 2), ("nabla", 1)])
def test_regular12_mixed_precision_per_edge(gpu, regular12, kind, k):
    """FP32->FP64 delta^2 / nabla^1 with multisets: every edge of every
    subgraph within 1e-9 of the FP64 value, removed edges NaN in both, and
    the histograms' counts equal."""
    n, U, V = regular12
    a = _lifted(gpu, kind, n, U, V, gpu.FP64, k=k, keep_multisets=True)
    b = _lifted(gpu, kind, n, U, V, gpu.FP32_FP64, k=k, keep_multisets=True)
    if kind == "delta":
        assert _hist_sha(a[0], a[1]) == _load("regular12_n256_delta2.json")["sha256"]
    else:
        assert _hist_sha(a[0], a[1]) == _load("regular12_n256_nabla1.json")["sha256"]
    ma, mb = a[2], b[2]
    assert ma.shape == mb.shape and a[3] == b[3]
    assert np.array_equal(np.isnan(ma), np.isnan(mb))
    live = ~np.isnan(ma)
    err = float(np.max(np.abs(ma[live] - mb[live])))
    print(f"{kind}^{k} of the 12-regular n=256 graph, FP32->FP64: max edge error {err:.3e} "
          f"over {int(live.sum())} values")
    assert err <= TOL
    # exact histograms: the same number of values, within 1e-9 in sorted
    # order (two FP64 values that coincide need not coincide in FP32->FP64,
    # so the (value, count) pairs themselves may differ)
    ea, eb = _expand(a[0], a[1]), _expand(b[0], b[1])
    assert ea.size == eb.size == int(live.sum())
    assert float(np.max(np.abs(ea - eb))) <= TOL


def test_regular12_delta2_quantised_mixed(gpu, regular12):
    """delta^2 through the batched API: the FP32->FP64 floor(d/eps) table
    equals the FP64 (= reference) values' table once the values within 1e-9
    of a bin edge are set apart, and counts add up to every value."""
    n, U, V = regular12
    eps = 1e-6
    ref = gpu.lifted_batch([(n, U, V)], k=2, precision=gpu.FP64, epsilon=eps, qtol=TOL)
    mix = gpu.lifted_batch([(n, U, V)], k=2, precision=gpu.FP32_FP64, epsilon=eps, qtol=TOL)
    (v64, c64), (vm, cm) = ref.histograms[0], mix.histograms[0]
    assert _hist_sha(v64, c64) == _load("regular12_n256_delta2.json")["sha256"]
    e64, em = _expand(v64, c64), _expand(vm, cm)
    assert e64.size == em.size
    assert float(np.max(np.abs(e64 - em))) <= TOL
    keys, counts, near = _quantised(e64, em, eps, TOL)
    (qk, qc), (bk, bc) = mix.quantised[0], mix.boundary[0]
    assert np.array_equal(np.asarray(qk), keys) and np.array_equal(np.asarray(qc), counts)
    assert int(np.asarray(bc).sum()) == near
    assert int(np.asarray(qc).sum()) + near == mix.values == e64.size
