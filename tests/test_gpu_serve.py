"""The request server: results in a concurrent, batched mix equal the same
requests run alone through the server, bit for bit, the single-graph entry
points and the CPU oracle (bit for bit in FP64, within 1e-9 for
FP32->FP64); admission by footprint; a short synthetic stream end to end."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(gen):
    """fits (FP64 and FP32->FP64, weighted or not, two variants), delta^1
    small enough to be batched, one too large for the batched kernel, and a
    nabla^1."""
    out = []
    rng = np.random.default_rng(11)
    for i in range(14):
        n = int(rng.integers(20, 3000))
        _, U, V, W = gen.gnp(n, float(rng.uniform(3, 20)), 500 + i, weighted=bool(i % 2))
        out.append(dict(kind=0, n=n, U=U, V=V, W=W, variant=int(i % 3 == 2), precision=int(i % 4 == 3)))
    for i in range(9):
        n = int(rng.integers(16, 160))
        _, U, V, W = gen.gnp(n, float(rng.uniform(3, 10)), 900 + i, weighted=bool(i % 3 == 0))
        out.append(dict(kind=1, n=n, U=U, V=V, W=W, variant=0, precision=0))
    _, U, V, _ = gen.gnp(150, 130.0, 77)                      # ~9.7k edges: per-subgraph path
    out.append(dict(kind=1, n=150, U=U, V=V, W=None, variant=0, precision=0))
    _, U, V, _ = gen.gnp(40, 6.0, 78)
    out.append(dict(kind=2, n=40, U=U, V=V, W=None, variant=0, precision=0))
    return out


def _submit(srv, q):
    return srv.submit(q["n"], q["U"], q["V"], q["W"], kind=q["kind"], k=1, variant=q["variant"],
                      precision=q["precision"])


def _collect(srv, count, timeout=120.0):
    import time
    got = {}
    t_end = time.time() + timeout
    while len(got) < count and time.time() < t_end:
        r = srv.next(1.0)
        if r is not None:
            got[r.id] = r
    assert len(got) == count, f"{len(got)} of {count} requests finished"
    return got


def _same(a, b):
    assert a.status == 0 and b.status == 0, (a.error, b.error)
    if a.kind == 0:
        assert a.iterations == b.iterations and a.delta == b.delta
        assert np.array_equal(a.edge_dress, b.edge_dress) and np.array_equal(a.vertex_dress, b.vertex_dress)
    else:
        assert a.subgraphs == b.subgraphs
        assert np.array_equal(a.hist_values, b.hist_values) and np.array_equal(a.hist_counts, b.hist_counts)


def test_mix_equals_alone(gpu, ora, gen):
    from dress_hip import serve
    reqs = _requests(gen)
    with serve.Server(workers=4, budget_bytes=0, max_batch=8) as srv:
        ids = [_submit(srv, q) for q in reqs]
        mix = _collect(srv, len(ids))
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["lifted_requests"] > st["lifted_launches"]          # some launches were shared
    assert max(mix[i].batch_size for i in ids) > 1
    with serve.Server(workers=1, budget_bytes=0, max_batch=1) as srv:
        for q, i in zip(reqs, ids):
            j = _submit(srv, q)
            alone = _collect(srv, 1)[j]
            _same(mix[i], alone)
            assert alone.batch_size == 1
    # the oracle: bit for bit in FP64, within 1e-9 in FP32->FP64
    for q, i in zip(reqs, ids):
        r = mix[i]
        if q["kind"] == 0:
            o = ora.fit(q["n"], q["U"], q["V"], q["W"], variant=q["variant"], max_iterations=100, epsilon=1e-6)
            assert r.iterations == o["iterations"]
            if q["precision"] == 0:
                assert r.delta == o["delta"]
                assert np.array_equal(r.edge_dress, o["edge_dress"])
                assert np.array_equal(r.vertex_dress, o["vertex_dress"])
            else:
                assert np.max(np.abs(r.edge_dress - o["edge_dress"]), initial=0.0) <= 1e-9
        else:
            vals, cnts, _, num = ora.lifted("delta" if q["kind"] == 1 else "nabla", q["n"], q["U"], q["V"], q["W"],
                                            variant=q["variant"], k=1, use_omp=True)
            assert num == r.subgraphs
            assert np.array_equal(vals, r.hist_values) and np.array_equal(cnts, r.hist_counts)
    # and the single-graph entry points
    for q, i in zip(reqs, ids):
        r = mix[i]
        with gpu.Graph(q["n"], q["U"], q["V"], q["W"], None, q["variant"]) as g:
            if q["kind"] == 0:
                it, dl = g.fit(100, 1e-6, q["precision"])
                assert it == r.iterations
                if q["precision"] == 0:
                    assert dl == r.delta
                    assert np.array_equal(g.edge_dress, r.edge_dress)
                    assert np.array_equal(g.vertex_dress, r.vertex_dress)
                else:
                    # FP32->FP64: the batched kernel's FP32 phase rounds its
                    # products on its own; both stay within 1e-9 of FP64
                    assert np.max(np.abs(g.edge_dress - r.edge_dress), initial=0.0) <= 1e-9
            else:
                vals, cnts, _, num = g.lifted("delta" if q["kind"] == 1 else "nabla", 1)
                assert num == r.subgraphs
                assert np.array_equal(vals, r.hist_values) and np.array_equal(cnts, r.hist_counts)


def test_admission_by_footprint(gpu, gen):
    from dress_hip import serve
    _, U, V, _ = gen.gnp(2000, 10.0, 5)
    with serve.Server(workers=2, budget_bytes=4 << 20, max_batch=4) as srv:
        big = srv.submit(2000, U, V, kind=0)                     # ~8 MB estimated: over budget
        _, u, v, _ = gen.gnp(30, 4.0, 6)
        small = srv.submit(30, u, v, kind=0)
        got = _collect(srv, 2)
        st = srv.stats()
    assert got[big].status == -1 and "budget" in got[big].error
    assert got[small].status == 0
    assert st["rejected"] == 1 and st["peak_reserved_bytes"] <= 4 << 20


def test_short_stream(gpu):
    from dress_hip import serve
    rep = serve.run_stream(duration=2.0, rate=150.0, seed=1, workers=4, max_batch=16, drain=20.0,
                           generators_threads=2)
    assert rep["offered"]["requests"] > 200
    assert rep["completed"]["fit"] > 0 and rep["completed"]["delta"] > 0
    assert rep["failed"] == 0
    lat = rep["latency"]["all"]
    assert lat["n"] > 0 and 0 < lat["p50_ms"] <= lat["p99_ms"]


@pytest.mark.parametrize("precision", [0, 1])
def test_zero_iterations_served_and_batched(gpu, ora, gen, precision):
    """max_iterations = 0 runs no iteration (dress.c:546-608): every value
    stays 1.0, in FP64 and FP32->FP64, through the server's batched lane and
    through dress_hip_lifted_batch."""
    from dress_hip import serve
    graphs = [gen.random_regular(n, 6, 40 + n) for n in (16, 24, 30)]
    r = gpu.lifted_batch(graphs, k=1, max_iterations=0, precision=precision)
    for (hv, hc), (n, U, V) in zip(r.histograms, graphs):
        ov, oc, _, _ = ora.lifted("delta", n, U, V, k=1, max_iterations=0)
        assert np.array_equal(np.asarray(hv), ov) and np.array_equal(np.asarray(hc), oc)
        assert np.array_equal(ov, [1.0])
    with serve.Server(workers=1, budget_bytes=0, max_batch=8) as srv:
        ids = [srv.submit(n, U, V, kind=serve.DELTA, k=1, max_iterations=0, precision=precision)
               for n, U, V in graphs]
        got = _collect(srv, len(ids))
    for i, (n, U, V) in zip(ids, graphs):
        ov, oc, _, num = ora.lifted("delta", n, U, V, k=1, max_iterations=0)
        assert got[i].status == 0, got[i].error
        assert got[i].subgraphs == num
        assert np.array_equal(got[i].hist_values, ov) and np.array_equal(got[i].hist_counts, oc)


def test_device_error_stops_the_server(gpu, gen):
    """A request failing with an error that leaves the device unusable
    (injected on the host as HIP error 700, illegal address: no device fault
    is made) stops the server: queued requests are answered with the cause,
    new submissions are refused, the stats say so.  A failure that leaves
    the device usable (out of memory, 2) does not."""
    from dress_hip import serve
    from dress_hip._lib import HipError
    _, U, V, _ = gen.gnp(60, 6.0, 31)
    with serve.Server(workers=1, budget_bytes=0, max_batch=1) as srv:
        srv.inject_error(2)
        a = srv.submit(60, U, V, kind=serve.FIT)
        ra = _collect(srv, 1)[a]
        assert ra.status == -1 and "injected" in ra.error
        b = srv.submit(60, U, V, kind=serve.FIT)
        assert _collect(srv, 1)[b].status == 0
        assert srv.stats()["device_failed"] == 0
        srv.inject_error(700)
        # (the worker may stop the server before all six are in: later
        # submissions are then refused, which is what is tested below)
        ids = []
        for _ in range(6):
            try:
                ids.append(srv.submit(60, U, V, kind=serve.FIT))
            except HipError as e:
                assert "stopped" in str(e)
                break
        assert ids
        got = _collect(srv, len(ids))
        st = srv.stats()
        assert all(got[i].status == -1 for i in ids)
        assert any("injected" in got[i].error for i in ids)
        assert st["device_failed"] == 1 and st["queued"] == 0
        with pytest.raises(HipError, match="stopped"):
            srv.submit(60, U, V, kind=serve.FIT)


def test_large_delta_runs_in_chunks(gpu, ora, gen, monkeypatch):
    """A delta^1 request whose subgraph values exceed the server's chunk is
    planned into chunks (not refused for its footprint) and its histogram
    equals the oracle's; a request whose result exceeds the host budget is
    refused and counted apart."""
    from dress_hip import serve
    monkeypatch.setenv("DRESS_HIP_SERVE_CHUNK_VALUES", "20000")
    n, U, V, _ = gen.gnp(400, 12.0, 41)
    monkeypatch.setenv("DRESS_HIP_SERVE_HOST_BYTES", str(64 << 20))
    with serve.Server(workers=4, budget_bytes=2 << 30, max_batch=4) as srv:
        a = srv.submit(n, U, V, kind=serve.DELTA, k=1)
        n2, U2, V2, _ = gen.gnp(3000, 12.0, 42)                   # 3000 x 18k values: > 64 MB of pairs
        b = srv.submit(n2, U2, V2, kind=serve.DELTA, k=1)
        got = _collect(srv, 2)
        st = srv.stats()
    assert got[a].status == 0, got[a].error
    vals, cnts, _, num = ora.lifted("delta", n, U, V, k=1, use_omp=True)
    assert got[a].subgraphs == num
    assert np.array_equal(got[a].hist_values, vals) and np.array_equal(got[a].hist_counts, cnts)
    assert got[b].status == -1 and "host budget" in got[b].error
    assert st["rejected_host"] == 1 and st["rejected"] == 1
    assert st["chunked_requests"] == 1
