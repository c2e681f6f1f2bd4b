"""GPU test of the continuous-batching server: every request's result is the
same alone and in the mix.  Requests of both storage types arrive while
others run, are admitted between intervals as memory allows, and their
lattices step together; each answer (steps to convergence, node views,
permeability) must equal, bit for bit, the same request solved alone through
the single-lattice LbmHipListRunSteady with the same interval and tolerance."""
import numpy as np
import pytest

import serving as S
import workloads as W

pytestmark = pytest.mark.gpu

INTERVAL, TOL, MAX = 100, 1e-5, 3000
KARGS = ["-lambda", "0.1875"]


def _alone(lbm, s, prec):
    geo = lbm.Geometry.from_string(s.geometry, s.dims)
    with lbm.Kernel("list-aa-pv-soa", geo, KARGS + ["-prec", prec, "-quiet"]) as k:
        k.set_initial_weights()
        r = k.run_steady(lbm.case_data(omega=s.omega, xforce=s.xforce, iterations=0), INTERVAL, TOL, MAX)
        return r.steps, r.converged, k.dump(), W.permeability(r.momentumX, geo.n_fluid, s.cells, s.omega, s.xforce)


@pytest.mark.parametrize("budget_slots", [100, 2])
def test_served_results_equal_alone(lbm, budget_slots):
    samples = S.serving_samples(6, 11, edge=(20, 36))
    arrivals = [0.0, 0.0, 0.01, 0.02, 0.05, 0.3]
    reqs = [S.Request(i, a, s, p) for i, (a, (s, p)) in enumerate(zip(arrivals, samples))]
    be = S.GpuBackend(lbm, KARGS, MAX, keep_nodes=True)
    budget = budget_slots * max(be.size(r) for r in reqs)
    try:
        recs = S.serve(reqs, be, budget, INTERVAL, TOL)
    finally:
        be.close()
    assert {r.prec for r in reqs} == {"fp64", "fp32s"}
    for rec, r in zip(recs, reqs):
        print(f"request {r.rid} {r.prec} dims={r.sample.dims} steps={rec.steps}", flush=True)
        steps, conv, nodes, perm = _alone(lbm, r.sample, r.prec)
        assert (rec.steps, rec.converged) == (steps, conv)
        assert np.array_equal(rec.nodes, nodes)
        assert rec.permeability == perm
    if budget_slots == 2:
        # never more than two lattices' worth resident at once
        ev = sorted([(x.admitted, x.bytes) for x in recs] + [(x.finished, -x.bytes) for x in recs],
                    key=lambda e: (e[0], e[1]))
        cur = peak = 0
        for _, b in ev:
            cur += b
            peak = max(peak, cur)
        assert peak <= budget


def test_c5_poisson_stream_at_500_per_s(lbm, oracle):
    """Serving at its configured shape: a Poisson stream at 500 requests/s of
    seeded porous cubes (edge U[32, 192], porosity U[0.5, 0.9], sphere radius
    U[3, 8], omega U[0.8, 1.6], F 1e-6), each tagged FP64 or shifted FP32, run
    to a relative x-momentum change < 1e-6 per 100 steps under a 200 GB
    budget.  Every answer equals the same request solved alone (steps, node
    views, permeability, bit for bit); the smallest FP64 and the smallest
    FP32 request equal the oracle at their step counts; p50/p99 latency and
    solves/s are reported."""
    n, interval, tol, limit = 12, 100, 1e-6, 20000
    samples = S.serving_samples(n, 2025)
    arrivals = S.poisson_arrivals(n, 500.0, 2525)
    reqs = [S.Request(i, a, s, p) for i, (a, (s, p)) in enumerate(zip(arrivals, samples))]
    assert {r.prec for r in reqs} == {"fp64", "fp32s"}
    print(f"c5 stream: {n} requests, edges {[r.sample.dims[0] for r in reqs]}", flush=True)
    be = S.GpuBackend(lbm, KARGS, limit, keep_nodes=True)
    try:
        recs = S.serve(reqs, be, int(200e9), interval, tol)
        summ = S.summarize(recs, be.flups())
    finally:
        be.close()
    print(f"c5 served: {summ}", flush=True)
    assert summ["rejected"] == 0 and summ["p99_latency_s"] >= summ["p50_latency_s"] > 0
    smallest = {p: min((r for r in reqs if r.prec == p), key=lambda r: r.sample.cells) for p in ("fp64", "fp32s")}
    for rec, r in zip(recs, reqs):
        s = r.sample
        print(f"c5 request {r.rid} {r.prec} edge {s.dims[0]} steps {rec.steps}", flush=True)
        assert rec.converged == 1 or rec.steps == limit
        geo = lbm.Geometry.from_string(s.geometry, s.dims)
        with lbm.Kernel("list-aa-pv-soa", geo, KARGS + ["-prec", r.prec, "-quiet"]) as k:
            k.set_initial_weights()
            a = k.run_steady(lbm.case_data(omega=s.omega, xforce=s.xforce, iterations=0), interval, tol, limit)
            assert (rec.steps, rec.converged) == (a.steps, a.converged)
            assert np.array_equal(rec.nodes, k.dump())
            assert rec.permeability == W.permeability(a.momentumX, geo.n_fluid, s.cells, s.omega, s.xforce)
        if r is smallest[r.prec]:
            m = geo.mask.reshape(-1)
            if r.prec == "fp64":
                want, _ = oracle.list_run(m, geo.dims, geo.periodic, rec.steps, s.omega, s.xforce, magic=0.1875,
                                          n_threads=8, vsize=1)
            else:
                want, _ = oracle.list_run_shifted(m, geo.dims, geo.periodic, rec.steps, s.omega, s.xforce, r.prec,
                                                  magic=0.1875)
            assert np.array_equal(rec.nodes, want), r.prec
        rec.nodes = None


def test_device_bytes_bound_holds_on_the_device(lbm):
    """The admission size (LbmHipListDeviceBytes) against what Init really
    takes from the device (hipMemGetInfo before and after), for a small and
    a large lattice in both storage types: never less."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")

    def free_bytes():
        f, t = C.c_size_t(), C.c_size_t()
        assert hip.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value

    warm = lbm.Geometry.from_string("box", (8, 8, 8), (1, 1, 0))
    with lbm.Kernel("list-aa-pv-soa", warm, ["-quiet"]):
        pass
    for gstr, dims in (("porous-3-0.7-5", (40, 36, 34)), ("porous-6-0.8-42", (160, 150, 140))):
        geo = lbm.Geometry.from_string(gstr, dims)
        for prec in ("fp64", "fp32s"):
            args = KARGS + ["-prec", prec, "-quiet", "-shadow", "lazy"]
            bound = lbm.list_device_bytes(geo, args)
            f0 = free_bytes()
            with lbm.Kernel("list-aa-pv-soa", geo, args):
                used = f0 - free_bytes()
            print(f"device bytes {dims} {prec}: used {used / 2**20:.1f} MiB, bound {bound / 2**20:.1f} MiB", flush=True)
            assert 0 < used <= bound
