"""Serving front end for libdress_hip.so.

``Server`` wraps the native request server (dress_hip_server_*): fit and
delta^k / nabla^k requests of mixed sizes are admitted by their estimated
device-memory footprint, run concurrently on worker streams, and small
delta/nabla requests are packed into one batched launch with the others
of their kind waiting at that moment.

``run_stream`` drives it with the serving benchmark's synthetic stream:
Poisson arrivals at ``rate`` requests/s for ``duration`` s; a share
``fit_share`` (0.7) of plain fits and the rest delta^1; every request a
G(n, p) graph with n log-uniform in [20, 20000], mean degree uniform in
[3, 30] and a seed of its own; half carry float32 weights uniform in
(0.1, 1].  Latency is measured from the scheduled arrival to completion,
so time a request waits for its generator counts too.

    python -m dress_hip.serve --duration 60 --rate 2000
"""
import argparse
import ctypes
import json
import math
import threading
import time
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from . import _lib
from . import generators
from ._lib import HipError, ServeRequest, ServeResponse, ServeStats

FIT, DELTA, NABLA = 0, 1, 2
KIND_NAMES = {FIT: "fit", DELTA: "delta", NABLA: "nabla"}


def clock():
    """Seconds on the server's monotonic clock."""
    return _lib.lib().dress_hip_clock()


@dataclass
class Response:
    id: int
    kind: int
    status: int
    error: str
    N: int
    E: int
    iterations: int
    delta: float
    edge_dress: Optional[np.ndarray]
    vertex_dress: Optional[np.ndarray]
    hist_values: Optional[np.ndarray]
    hist_counts: Optional[np.ndarray]
    subgraphs: int
    batch_size: int
    footprint: int
    arrival: float
    admitted: float
    finished: float

    @property
    def latency(self):
        return self.finished - self.arrival


class Server:
    """The native request server on the current device."""

    def __init__(self, workers=8, budget_bytes=0, max_batch=64):
        L = _lib.lib()
        self._h = L.dress_hip_server_create(int(workers), int(budget_bytes), int(max_batch))
        if not self._h:
            raise HipError("dress_hip_server_create: " + _lib.last_error())

    def submit(self, n, U, V, W=None, kind=FIT, k=1, variant=0, max_iterations=100, epsilon=1e-6,
               precision=0, arrival=-1.0):
        """Queues a request (the arrays are copied); returns its id."""
        U = np.ascontiguousarray(U, dtype=np.int32)
        V = np.ascontiguousarray(V, dtype=np.int32)
        if U.shape != V.shape:
            raise ValueError("sources and targets differ in length")
        r = ServeRequest()
        r.kind, r.k, r.N, r.E = int(kind), int(k), int(n), int(U.size)
        r.U = U.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        r.V = V.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        if W is not None:
            W = np.ascontiguousarray(W, dtype=np.float64)
            if W.shape != U.shape:
                raise ValueError("weights differ in length from the edges")
            r.W = W.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        r.variant, r.max_iterations, r.epsilon = int(variant), int(max_iterations), float(epsilon)
        r.precision, r.arrival = int(precision), float(arrival)
        rid = _lib.lib().dress_hip_server_submit(self._h, ctypes.byref(r))
        if rid < 0:
            raise HipError("dress_hip_server_submit: " + _lib.last_error())
        return rid

    def next(self, timeout=0.1, keep_values=True) -> Optional[Response]:
        """The next finished request, or None after `timeout` seconds."""
        o = ServeResponse()
        L = _lib.lib()
        if not L.dress_hip_server_next(self._h, float(timeout), ctypes.byref(o)):
            return None
        try:
            ed = vd = hv = hc = None
            if keep_values and o.status == 0:
                if o.edge_dress:
                    ed = np.ctypeslib.as_array(o.edge_dress, shape=(o.E,)).copy() if o.E else np.zeros(0)
                    vd = np.ctypeslib.as_array(o.vertex_dress, shape=(o.N,)).copy() if o.N else np.zeros(0)
                if o.kind != FIT:
                    hv = np.array([o.hist[i].value for i in range(o.hist_size)], dtype=np.float64)
                    hc = np.array([o.hist[i].count for i in range(o.hist_size)], dtype=np.int64)
            return Response(o.id, o.kind, o.status, o.error.decode(errors="replace"), o.N, o.E, o.iterations,
                            o.delta, ed, vd, hv, hc, o.subgraphs, o.batch_size, o.footprint, o.arrival,
                            o.admitted, o.finished)
        finally:
            L.dress_hip_response_free(ctypes.byref(o))

    def stats(self):
        s = ServeStats()
        _lib.lib().dress_hip_server_stats(self._h, ctypes.byref(s))
        return {name: getattr(s, name) for name, _ in ServeStats._fields_}

    def inject_error(self, hip_error):
        """Failure-path test hook: the next finished request fails as if
        the device had reported HIP error code ``hip_error``."""
        _lib.lib().dress_hip_server_inject_error(self._h, int(hip_error))

    def close(self):
        if self._h:
            _lib.lib().dress_hip_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def request_stream(duration, rate, seed=0, fit_share=0.7, n_range=(20, 20000), deg_range=(3.0, 30.0),
                   weighted_share=0.5) -> List[Tuple[float, int, int, float, bool, int]]:
    """Arrivals of one run: (offset s, kind, n, mean degree, weighted,
    graph seed), Poisson at `rate`/s over [0, duration)."""
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    lo, hi = math.log(n_range[0]), math.log(n_range[1])
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        kind = FIT if rng.random() < fit_share else DELTA
        n = int(round(math.exp(rng.uniform(lo, hi))))
        d = float(rng.uniform(deg_range[0], deg_range[1]))
        w = bool(rng.random() < weighted_share)
        out.append((t, kind, n, d, w, (int(seed) << 32) + len(out)))


def _pct(x, q):
    return float(np.percentile(np.asarray(x), q)) * 1e3 if len(x) else None


def _lat(xs):
    return {"n": len(xs), "p50_ms": _pct(xs, 50), "p99_ms": _pct(xs, 99),
            "max_ms": (max(xs) * 1e3 if xs else None)}


def run_stream(duration=60.0, rate=2000.0, seed=0, workers=8, budget_bytes=0, max_batch=64, drain=10.0,
               fit_share=0.7, generators_threads=4, precision=0, log=None):
    """Runs the synthetic stream against a fresh Server; returns a report."""
    arrivals = request_stream(duration, rate, seed, fit_share)
    srv = Server(workers, budget_bytes, max_batch)
    done: List[Response] = []
    stop = threading.Event()

    def collect():
        while True:
            r = srv.next(0.05, keep_values=False)
            if r is not None:
                done.append(r)
            elif stop.is_set():
                return

    offered = {"fit": 0, "delta": 0}
    offered_edges = {"fit": 0, "delta": 0}
    lock = threading.Lock()
    t0 = clock() + 0.2

    def submitter(part):
        for t, kind, n, d, w, gs in arrivals[part::generators_threads]:
            wait = t0 + t - clock()
            if wait > 0:
                time.sleep(wait)
            _, U, V, W = generators.gnp(n, d, gs, w)
            srv.submit(n, U, V, W, kind=kind, k=1, precision=precision, arrival=t0 + t)
            with lock:
                offered[KIND_NAMES[kind]] += 1
                offered_edges[KIND_NAMES[kind]] += int(U.size)

    coll = threading.Thread(target=collect, daemon=True)
    coll.start()
    subs = [threading.Thread(target=submitter, args=(i,), daemon=True) for i in range(generators_threads)]
    for th in subs:
        th.start()
    depth = []
    while any(th.is_alive() for th in subs):
        st = srv.stats()
        depth.append((clock() - t0, st["queued"], st["running"]))
        time.sleep(0.25)
    for th in subs:
        th.join()
    t_arrivals = clock()
    while clock() < t_arrivals + drain:
        st = srv.stats()
        depth.append((clock() - t0, st["queued"], st["running"]))
        if st["queued"] + st["running"] == 0:
            break
        time.sleep(0.1)
    t_end = clock()
    time.sleep(0.1)
    st = srv.stats()
    with lock:
        finished = list(done)
    ended = {r.id for r in finished}
    stop.set()
    coll.join()
    srv.close()                               # cancels what is still queued or running
    elapsed = t_end - t0
    ok = [r for r in finished if r.status == 0 and r.finished <= t_end]
    by = {"fit": [r for r in ok if r.kind == FIT], "delta": [r for r in ok if r.kind == DELTA]}
    n_unfinished = len(arrivals) - len(ended)
    report = {
        "workload": f"Poisson {rate:g} req/s for {duration:g} s, {fit_share:.0%} fit / {1 - fit_share:.0%} delta^1, "
                    "G(n,p) n log-uniform [20,20000], mean degree U[3,30], half FP32-weighted",
        "workers": workers, "max_batch": max_batch,
        "duration_s": duration, "elapsed_s": elapsed,
        "offered": {"requests": len(arrivals), "per_s": len(arrivals) / duration, **offered},
        "completed": {"fit": len(by["fit"]), "delta": len(by["delta"]),
                      "per_s": len(ok) / elapsed if elapsed > 0 else None},
        "failed": int(st["failed"]), "rejected": int(st["rejected"]),
        "rejected_host": int(st["rejected_host"]), "chunked": int(st["chunked_requests"]),
        "device_failed": int(st["device_failed"]),
        "unfinished": n_unfinished,
        "fit_edge_updates_per_s": sum(r.E * (r.iterations + 1) for r in by["fit"]) / elapsed if elapsed > 0 else None,
        "delta1_subgraph_solves_per_s": sum(r.subgraphs for r in by["delta"]) / elapsed if elapsed > 0 else None,
        "latency": {"all": _lat([r.latency for r in ok]), "fit": _lat([r.latency for r in by["fit"]]),
                    "delta": _lat([r.latency for r in by["delta"]])},
        "queue": {"max": int(st["max_queued"]), "mean": float(np.mean([q for _, q, _ in depth])) if depth else 0.0,
                  "final": int(depth[-1][1]) if depth else 0},
        "memory": {"budget_GB": st["budget_bytes"] / 1e9, "peak_reserved_GB": st["peak_reserved_bytes"] / 1e9},
        "batching": {"lifted_launches": int(st["lifted_launches"]), "lifted_requests": int(st["lifted_requests"]),
                     "mean_batch": (st["lifted_requests"] / st["lifted_launches"]) if st["lifted_launches"] else 0.0},
    }
    return report


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--duration", type=float, default=60.0)
    ap.add_argument("--rate", type=float, default=2000.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--max-batch", type=int, default=64)
    ap.add_argument("--budget-gb", type=float, default=0.0)
    ap.add_argument("--drain", type=float, default=10.0)
    ap.add_argument("--fit-share", type=float, default=0.7)
    a = ap.parse_args()
    rep = run_stream(a.duration, a.rate, a.seed, a.workers, int(a.budget_gb * 1e9), a.max_batch, a.drain,
                     a.fit_share)
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()
