"""CPU tests of the serving loop's admission and scheduling (a simulated
backend stands in for the GPU), of the workload generators and of the
permeability formula (against plane Poiseuille flow run on the oracle)."""
import math

import numpy as np
import pytest

import serving as S
import workloads as W


class SimBackend:
    """Each request needs a fixed number of steps; an interval costs time in
    proportion to the nodes it updates."""

    def __init__(self, clock, cost_per_node_step=1e-9):
        self.clock = clock
        self.cost = cost_per_node_step
        self.live = {}
        self.next = 0
        self.admit_order = []

    def admit(self, r):
        h = self.next
        self.next += 1
        n = int(r.sample.cells * r.sample.porosity)
        self.live[h] = [r, 0, 100 * (1 + r.rid % 5), n]
        self.admit_order.append(r.rid)
        return h, n

    def launch(self, interval):
        self.interval = interval
        self.running = list(self.live)
        self.clock.t += self.cost * interval * sum(v[3] for v in self.live.values())

    def wait(self, tol):
        done = []
        for h in self.running:
            v = self.live[h]
            v[1] += self.interval
            if v[1] >= v[2]:
                done.append((h, v[1], 1, 1e-3 * v[3], None))
        return done

    def release(self, h):
        del self.live[h]


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


def _requests(n, rate, seed=1, fp32_share=0.5):
    arr = S.poisson_arrivals(n, rate, seed)
    return [S.Request(i, a, s, p) for i, (a, (s, p)) in
            enumerate(zip(arr, S.serving_samples(n, seed, fp32_share=fp32_share)))]


def _peak_bytes(recs):
    ev = sorted([(r.admitted, r.bytes) for r in recs if not r.rejected] +
                [(r.finished, -r.bytes) for r in recs if not r.rejected], key=lambda e: (e[0], e[1]))
    cur = peak = 0
    for _, b in ev:
        cur += b
        peak = max(peak, cur)
    return peak


def _bypassed(recs):
    """Per request: how many later arrivals were admitted before it."""
    return [sum(1 for q in recs if q.arrival > r.arrival and q.admitted < r.admitted) for r in recs]


@pytest.mark.parametrize("rate", [10, 200])
@pytest.mark.parametrize("max_bypass", [0, 3, 8])
def test_budget_order_and_completion(rate, max_bypass):
    reqs = _requests(40, rate)
    budget = 3 * max(S.device_bytes(r.sample, r.prec) for r in reqs)
    clk = Clock()
    be = SimBackend(clk)
    recs = S.serve(reqs, be, budget, 100, 1e-6, clock=clk, sleep=clk.sleep, max_bypass=max_bypass)
    assert all(not r.rejected and r.finished >= r.admitted >= r.arrival for r in recs)
    assert _peak_bytes(recs) <= budget
    if max_bypass == 0:
        # strict arrival order: the head of the queue waits for room
        assert be.admit_order == [r.rid for r in sorted(reqs, key=lambda r: r.arrival)]
    # no request is overtaken by more than max_bypass later arrivals
    assert max(_bypassed(recs)) <= max_bypass
    assert [r.steps for r in recs] == [100 * (1 + r.rid % 5) for r in reqs]


def test_small_requests_pass_a_large_head():
    # a large request at the head of the queue that does not fit yet: the
    # small ones behind it are admitted first (up to max_bypass of them),
    # then it gets the room
    big = W.Sample(0, (120, 120, 120), 6.0, 0.8, 1, 1.0, 1e-6)
    small = [W.Sample(i, (32, 32, 32), 4.0, 0.8, 1 + i, 1.0, 1e-6) for i in range(1, 9)]
    running = W.Sample(9, (120, 120, 120), 6.0, 0.8, 99, 1.0, 1e-6)
    reqs = [S.Request(9, 0.0, running), S.Request(0, 0.001, big)] + \
           [S.Request(s.index, 0.002 + 0.001 * s.index, s) for s in small]
    nb = S.device_bytes(big, "fp64")
    budget = nb + S.device_bytes(running, "fp64") // 2      # the big one waits for the running one
    out = {}
    for mb in (0, 4):
        clk = Clock()
        be = SimBackend(clk)
        recs = {r.rid: r for r in S.serve(reqs, be, budget, 100, 1e-6, clock=clk, sleep=clk.sleep, max_bypass=mb)}
        assert _peak_bytes(list(recs.values())) <= budget
        out[mb] = (be.admit_order, recs)
    order0, _ = out[0]
    assert order0 == [9, 0] + [s.index for s in small]
    order4, recs4 = out[4]
    assert order4[:6] == [9, 1, 2, 3, 4, 0]         # four small ones pass, then the head is served
    assert max(_bypassed(list(recs4.values()))) <= 4
    assert np.mean([recs4[s.index].latency for s in small[:4]]) < np.mean([out[0][1][s.index].latency for s in small[:4]])


def test_continuous_batching_overlaps_requests():
    # with room for everything, a request arriving while others run is
    # admitted at the next interval boundary instead of waiting for them
    reqs = _requests(20, 100)
    clk = Clock()
    recs = S.serve(reqs, SimBackend(clk, 1e-11), 1 << 50, 100, 1e-6, clock=clk, sleep=clk.sleep)
    overl = sum(1 for a in recs for b in recs if a.rid < b.rid and a.admitted < b.finished and b.admitted < a.finished)
    assert overl > 0
    assert max(r.admitted - r.arrival for r in recs) < 0.5


def test_oversized_request_is_rejected():
    reqs = _requests(5, 50)
    budget = min(S.device_bytes(r.sample, r.prec) for r in reqs)
    clk = Clock()
    recs = S.serve(reqs, SimBackend(clk), budget, 100, 1e-6, clock=clk, sleep=clk.sleep)
    for r, q in zip(recs, reqs):
        assert r.rejected == (S.device_bytes(q.sample, q.prec) > budget)
    assert sum(not r.rejected for r in recs) >= 1


def test_request_the_library_refuses_is_rejected():
    """A backend whose size() raises (the library refusing a lattice too
    large for its 32-bit indices) gets that request rejected with the
    message; the others are served."""
    reqs = _requests(6, 50)

    class Refusing(SimBackend):
        def size(self, r):
            if r.rid == 2:
                raise RuntimeError("number of PDFs exceed 2^32 or empty dimensions.")
            return S.device_bytes(r.sample, r.prec)

    clk = Clock()
    be = Refusing(clk)
    recs = S.serve(reqs, be, 1 << 50, 100, 1e-6, clock=clk, sleep=clk.sleep)
    assert [r.rejected for r in recs] == [r.rid == 2 for r in reqs]
    assert "exceed 2^32" in recs[2].error and recs[2].bytes == 0
    assert 2 not in be.admit_order and len(be.admit_order) == 5
    assert all(r.steps > 0 for r in recs if not r.rejected)


def test_fifo_latencies_and_summary():
    fin = S.fifo_latencies([0.0, 0.5, 0.6, 5.0], [1.0, 1.0, 2.0, 0.5])
    assert fin == [1.0, 2.0, 4.0, 5.5]
    recs = [S.Record(i, "fp64", a, 0.0, f) for i, (a, f) in enumerate(zip([0.0, 0.5, 0.6, 5.0], fin))]
    s = S.summarize(recs)
    assert s["makespan_s"] == 5.5 and s["p50_latency_s"] == pytest.approx(W.percentile([1, 1.5, 3.4, 0.5], 50))


def test_plain_path_on_the_arrival_clock():
    """serve_plain on a simulated clock: requests taken in arrival order,
    each once it has arrived and the one before it is answered -- the finish
    times of Lindley's recursion over the same service times; every request
    computed in FP64."""
    arr = [0.0, 0.5, 0.6, 5.0, 5.1]
    service = [1.0, 1.0, 2.0, 0.5, 0.25]
    clk = Clock()
    sample = W.Sample(0, (8, 8, 8), 2.0, 0.8, 1, 1.0, 1e-6)
    reqs = [S.Request(i, a, sample, "fp32s" if i % 2 else "fp64") for i, a in enumerate(arr)]

    def solve(r):
        assert clk() >= r.arrival
        clk.t += service[r.rid]
        return 100 + r.rid, 1, 0.5, 10

    recs = S.serve_plain(list(reversed(reqs)), solve, clock=clk, sleep=clk.sleep)
    assert [r.rid for r in recs] == list(range(len(arr)))
    assert [r.finished for r in recs] == pytest.approx(S.fifo_latencies(arr, service))
    assert all(r.prec == "fp64" and r.arrival <= r.admitted < r.finished for r in recs)
    assert [r.steps for r in recs] == [100 + i for i in range(len(arr))]
    assert S.summarize(recs)["p50_latency_s"] == pytest.approx(W.percentile([1.0, 1.5, 3.4, 0.5, 0.65], 50))


def test_generators():
    a = W.batch_samples(200, 7)
    assert a == W.batch_samples(200, 7)
    for s in a:
        assert all(32 <= e <= 128 for e in s.dims) and 0.5 <= s.porosity <= 0.9
        assert 3 <= s.radius <= 8 and 0.8 <= s.omega <= 1.6
    logs = np.log([e for s in a for e in s.dims])
    assert abs(logs.mean() - (math.log(32) + math.log(128)) / 2) < 0.1    # log-uniform
    arr = S.poisson_arrivals(4000, 50.0, 3)
    assert np.all(np.diff(arr) > 0) and abs(np.mean(np.diff(arr)) - 1 / 50) < 0.002
    ss = S.serving_samples(300, 5)
    assert all(32 <= s.dims[0] <= 192 and s.dims[0] == s.dims[2] for s, _ in ss)
    assert 0.35 < np.mean([p == "fp32s" for _, p in ss]) < 0.65


def test_permeability_of_plane_poiseuille(oracle):
    # plates at z = 0 and z = nz-1 (solid), periodic x and y: the steady TRT
    # solution with Lambda = 3/16 is the parabola u(z) = F/(2 nu) z'(h - z'),
    # z' = z - 1/2 from the half-way wall, h fluid layers
    nx, ny, nz = 4, 4, 10
    mask = np.ones((nz, ny, nx), np.int32)
    mask[0], mask[-1] = 0, 0
    omega, F = 1.0, 1e-5
    d, _ = oracle.list_run(mask.reshape(-1), (nx, ny, nz), (1, 1, 0), 3000, omega, F, magic=3 / 16)
    jx = float((d @ oracle.CX.astype(float)).sum())
    h = nz - 2
    zp = np.arange(h) + 0.5
    k_exact = 0.5 * float((zp * (h - zp)).sum()) / nz
    k = W.permeability(jx, d.shape[0], nx * ny * nz, omega, F)
    assert k == pytest.approx(k_exact, rel=1e-6)


def test_device_bytes_refuses_what_init_refuses(lbm):
    """The size bound refuses the lattices Init refuses (host only, no mask
    read): more than 2^32 PDFs in the box, or a padded node count whose
    adjacency values do not fit 32 bits -- so the serving loop rejects such
    a request instead of admitting it and failing at Init."""
    import types
    ld = lbm.LatticeDesc()
    ld.Dims[0], ld.Dims[1], ld.Dims[2] = 1024, 1024, 256
    ld.nFluid = 1000
    with pytest.raises(RuntimeError, match=r"exceed 2\^32"):
        lbm.list_device_bytes(types.SimpleNamespace(ld=ld))
    # 226,000,000 all-fluid cells: 19 x 226e6 < 2^32, but padded to whole
    # 16 MiB planes (2,097,152 cells) the node count no longer fits
    ld.Dims[0], ld.Dims[1], ld.Dims[2] = 1000, 1000, 226
    ld.nFluid = 226_000_000
    assert lbm.list_device_bytes(types.SimpleNamespace(ld=ld), ["-pad", "no"]) > 19 * 8 * ld.nFluid
    with pytest.raises(RuntimeError, match="does not fit the 32-bit adjacency"):
        lbm.list_device_bytes(types.SimpleNamespace(ld=ld), ["-pad", "16777216+0"])


def test_device_bytes_bound_covers_the_lattice(lbm, oracle):
    """LbmHipListDeviceBytes (host only): at least the padded PDFs, the
    adjacency and the grid of the built geometry, more with FP64 than FP32,
    growing with the padding; the serving loop admits by it."""
    geo = lbm.Geometry.from_string("porous-3-0.7-5", (40, 36, 34))
    nf, cells = geo.n_fluid, int(np.prod(geo.dims))
    got = {}
    for prec, elem in (("fp64", 8), ("fp32s", 4)):
        b = lbm.list_device_bytes(geo, ["-prec", prec])
        nc = oracle.pad_cells(nf)
        assert b >= elem * 19 * nc + 72 * nf + 4 * cells
        assert b <= elem * 19 * nc + 3 * 72 * (nf + 255) + 12 * nf + 8 * cells + 13 * (2 << 20)
        got[prec] = b
    assert got["fp64"] > got["fp32s"]
    assert lbm.list_device_bytes(geo, ["-pad", "no"]) < got["fp64"]
    with pytest.raises(RuntimeError, match="fp64, fp32s or fp16s"):
        lbm.list_device_bytes(geo, ["-prec", "fp8"])
