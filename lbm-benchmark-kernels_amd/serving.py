"""Continuous-batching permeability server over the multi-lattice batch.

Requests (a seeded porous sample and a storage tag, "fp64" or "fp32s")
arrive on a clock.  Between step intervals the server admits waiting
requests whose device bytes fit the budget -- first fit in arrival order,
a request passed over by later ones `max_bypass` times holds back the rest
until it fits -- builds their lattices (list build on the GPU) and adds
them to the batch of their storage type; every interval advances all admitted lattices
of both batches together (the two batches' intervals are launched before
either is waited for, so their work overlaps on the device), and a lattice
whose x momentum has settled (relative change per interval <= tol) or that
reached the step limit is answered and freed, which makes room for the next
requests.  No counterpart in the reference.

The admission and scheduling logic is independent of the GPU: `serve` takes
any backend with admit / launch / wait / release (``GpuBackend`` here; the
tests also drive it with a simulated one).
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

import workloads as W

PRECS = ("fp64", "fp32s")


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the stream starts
    sample: W.Sample
    prec: str = "fp64"


@dataclass
class Record:
    rid: int
    prec: str
    arrival: float
    admitted: float = math.nan
    finished: float = math.nan
    steps: int = 0
    converged: int = 0
    n_fluid: int = 0
    permeability: float = math.nan
    bytes: int = 0
    rejected: bool = False
    error: str = ""             # why the library refused the lattice, if it did
    nodes: np.ndarray | None = field(default=None, repr=False)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def poisson_arrivals(n: int, rate: float, seed: int) -> list:
    """Arrival times of a Poisson stream: exponential gaps of mean 1/rate."""
    rng = np.random.default_rng(seed)
    return list(np.cumsum(rng.exponential(1.0 / rate, n)))


def serving_samples(n: int, seed: int, edge=(32, 192), porosity=(0.5, 0.9), radius=(3.0, 8.0),
                    omega=(0.8, 1.6), xforce=1e-6, fp32_share=0.5):
    """Porous cubes of uniform edge, each with a storage tag."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        e = int(rng.integers(edge[0], edge[1] + 1))
        s = W.Sample(i, (e, e, e), float(rng.uniform(*radius)), float(rng.uniform(*porosity)),
                     int(rng.integers(1, 2 ** 31 - 1)), float(rng.uniform(*omega)), xforce)
        out.append((s, "fp32s" if rng.random() < fp32_share else "fp64"))
    return out


def device_bytes(sample: W.Sample, prec: str) -> int:
    """Device memory of one lattice, estimated from the sample alone (for
    backends that cannot size a request exactly): PDFs (19 values per
    node), full and run-compressed adjacency (72 + ~16 B per node),
    coordinates (12 B per node) and the cell -> node grid (4 B per cell);
    fluid nodes estimated from the target porosity.  GpuBackend.size gives
    the library's upper bound for the built geometry instead."""
    nodes = sample.cells * sample.porosity
    per_node = 19 * (8 if prec == "fp64" else 4) + 72 + 16 + 12
    return int(nodes * per_node + 4 * sample.cells)


def serve(requests, backend, budget: int, interval: int, tol: float, clock=time.perf_counter,
          sleep=time.sleep, max_bypass: int = 8):
    """Runs the request stream to completion; returns one Record per request.

    A request's device bytes come from backend.size(request) when the
    backend has it, else from device_bytes; a request whose lattice the
    library refuses to size (RuntimeError: more PDFs than its 32-bit indices
    address) is rejected with the library's message.  Waiting requests are admitted
    first fit in arrival order; a request that does not fit is passed over
    by later ones at most max_bypass times (0: strict arrival order), after
    which nothing behind it is admitted until it fits."""
    pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
    queue = []
    recs = {r.rid: Record(r.rid, r.prec, r.arrival) for r in requests}
    need_of, passed = {}, {}
    size = getattr(backend, "size", None)
    in_use = 0
    live = {}                                   # handle -> (request, bytes)
    t0 = clock()
    while pending or queue or live:
        now = clock() - t0
        while pending and pending[0].arrival <= now:
            queue.append(pending.popleft())
        # the next interval of the admitted lattices runs on the device while
        # the host builds the lattices of newly admitted requests
        running = bool(live)
        if running:
            backend.launch(interval)
        i, blocked = 0, []                      # first fit in arrival order
        while i < len(queue):
            r = queue[i]
            if r.rid not in need_of:
                try:
                    need_of[r.rid] = size(r) if size else device_bytes(r.sample, r.prec)
                except RuntimeError as e:
                    need_of[r.rid] = None
                    recs[r.rid].error = str(e)
                passed[r.rid] = 0
            need = need_of[r.rid]
            if need is None or need > budget:
                queue.pop(i)
                rec = recs[r.rid]
                rec.rejected, rec.bytes, rec.finished = True, need or 0, clock() - t0
                continue
            if in_use + need > budget:
                if passed[r.rid] >= max_bypass:
                    break                       # it has waited long enough: reserve room for it
                blocked.append(r.rid)
                i += 1
                continue
            queue.pop(i)
            h, n_fluid = backend.admit(r)
            in_use += need
            live[h] = (r, need)
            rec = recs[r.rid]
            rec.admitted, rec.bytes, rec.n_fluid = clock() - t0, need, n_fluid
            for b in blocked:
                passed[b] += 1
            if any(passed[b] >= max_bypass for b in blocked):
                break
        if not running:
            if not live and pending:
                sleep(max(0.0, pending[0].arrival - (clock() - t0)))
            continue
        for h, steps, converged, momentum_x, nodes in backend.wait(tol):
            r, need = live.pop(h)
            rec = recs[r.rid]
            rec.finished = clock() - t0
            rec.steps, rec.converged, rec.nodes = steps, converged, nodes
            s = r.sample
            rec.permeability = W.permeability(momentum_x, rec.n_fluid, s.cells, s.omega, s.xforce)
            backend.release(h)
            in_use -= need
    return [recs[r.rid] for r in requests]


class GpuBackend:
    """One LbmHipBatch per storage type; lattices built on admission."""

    def __init__(self, lbm, kernel_args=(), max_steps=20000, keep_nodes=False):
        self.lbm = lbm
        self.kernel_args = list(kernel_args)
        self.max_steps = max_steps
        self.keep_nodes = keep_nodes
        self.batches = {p: lbm.Batch(self.kernel_args + ["-prec", p]) for p in PRECS}
        self.where = {}                         # handle -> (prec, id)
        self.geos = {}                          # request id -> geometry built for sizing
        self.next = 0

    def _geometry(self, r: Request):
        if r.rid not in self.geos:
            s = r.sample
            self.geos[r.rid] = self.lbm.Geometry.from_string(s.geometry, s.dims)
        return self.geos[r.rid]

    def size(self, r: Request) -> int:
        """Device bytes of the request's lattice: the library's upper bound
        for its built geometry (exact fluid-node count and padding)."""
        return self.lbm.list_device_bytes(self._geometry(r), self.kernel_args + ["-prec", r.prec])

    def admit(self, r: Request):
        s = r.sample
        geo = self._geometry(r)
        del self.geos[r.rid]
        i = self.batches[r.prec].add(geo, s.omega, s.xforce, self.max_steps)
        h = self.next
        self.next += 1
        self.where[h] = (r.prec, i)
        return h, geo.n_fluid

    def launch(self, interval):
        self.running = {h: v for h, v in self.where.items()}
        for p in {q for q, _ in self.running.values()}:
            self.batches[p].launch(interval)

    def wait(self, tol):
        done = []
        for p in {q for q, _ in self.running.values()}:
            b = self.batches[p]
            b.wait(tol)
            for h, (q, i) in self.running.items():
                if q != p:
                    continue
                st = b.status(i)
                if not st.active:
                    nodes = b.dump(i) if self.keep_nodes else None
                    done.append((h, st.steps, st.converged, st.momentumX, nodes))
        return done

    def release(self, h):
        p, i = self.where.pop(h)
        self.batches[p].remove(i)

    def flups(self):
        return sum(b.stats()[1] for b in self.batches.values())

    def close(self):
        for b in self.batches.values():
            b.close()


class PlainGpuSolver:
    """The plain path for one request: its lattice alone in FP64 -- geometry,
    Init (list build), f = w, LbmHipRunSteady in checks of `interval` steps
    to `tol` or `max_steps`, Deinit."""

    def __init__(self, lbm, kernel_args=(), interval=100, tol=1e-6, max_steps=20000):
        self.lbm = lbm
        self.kernel_args = list(kernel_args)
        self.interval, self.tol, self.max_steps = interval, tol, max_steps
        self.node_updates = 0.0

    def __call__(self, r: Request):
        lbm, s = self.lbm, r.sample
        geo = lbm.Geometry.from_string(s.geometry, s.dims)
        with lbm.Kernel("list-aa-pv-soa", geo, self.kernel_args + ["-prec", "fp64", "-quiet"]) as k:
            k.set_initial_weights()
            k.set_shadow_download(False)
            res = k.run_steady(lbm.case_data(omega=s.omega, xforce=s.xforce, iterations=0), self.interval, self.tol,
                               self.max_steps)
        self.node_updates += geo.n_fluid * res.steps
        return res.steps, res.converged, W.permeability(res.momentumX, geo.n_fluid, s.cells, s.omega, s.xforce), \
            geo.n_fluid


def serve_plain(requests, solve, clock=time.perf_counter, sleep=time.sleep):
    """The plain one-request-at-a-time path, run on the arrival clock: the
    requests in arrival order, each solved alone by ``solve(request) ->
    (steps, converged, permeability, n_fluid)`` once it has arrived and the
    one before it is answered; every request is computed in FP64 whatever
    its tag."""
    t0 = clock()
    recs = []
    for r in sorted(requests, key=lambda q: q.arrival):
        now = clock() - t0
        if now < r.arrival:
            sleep(r.arrival - now)
        rec = Record(r.rid, "fp64", r.arrival, admitted=clock() - t0)
        rec.steps, rec.converged, rec.permeability, rec.n_fluid = solve(r)
        rec.finished = clock() - t0
        recs.append(rec)
    return recs


def fifo_latencies(arrivals, service):
    """Finish times of a single server taking requests in arrival order
    (Lindley's recursion): what serve_plain gives for these service times."""
    finish, t = [], -math.inf
    for a, s in zip(arrivals, service):
        t = max(t, a) + s
        finish.append(t)
    return finish


def summarize(records, flups=None, makespan=None):
    done = [r for r in records if not r.rejected]
    lat = [r.latency for r in done]
    span = makespan if makespan is not None else (max(r.finished for r in done) - min(r.arrival for r in records))
    out = {"requests": len(records), "rejected": len(records) - len(done),
           "p50_latency_s": round(W.percentile(lat, 50), 4), "p99_latency_s": round(W.percentile(lat, 99), 4),
           "solves_per_s": round(len(done) / span, 3), "makespan_s": round(span, 3),
           "converged": sum(r.converged for r in done)}
    if flups is not None:
        out["mflups"] = round(flups / span / 1e6, 1)
    return out
