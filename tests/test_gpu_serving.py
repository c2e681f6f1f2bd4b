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
    budget = budget_slots * max(S.device_bytes(r.sample, r.prec) for r in reqs)
    be = S.GpuBackend(lbm, KARGS, MAX, keep_nodes=True)
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
