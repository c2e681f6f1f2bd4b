"""GPU tests of the multi-lattice batch (LbmHipBatch*): many list-aa-pv-soa
domains of different sizes, geometries, omegas and forces stepped together by
one persistent launch per half-step (per half-step and group of domains),
replayed from captured HIP graphs.

Bar: every domain's node views after the batch equal, bit for bit, the same
lattice run alone through the plugin's Kernel() for the same number of steps
(and the oracle's result), whatever the launch shape; moments and the
convergence bookkeeping agree with the node views.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LAM = 3 / 16
DOMAINS = [("porous-2-0.7-5", (20, 18, 16), 1.2, 1e-5),
           ("channel", (14, 12, 10), 0.9, 2e-5),
           ("porous-3-0.75-9", (40, 36, 32), 1.5, 1e-5),
           ("box", (12, 12, 12), 1.0, 3e-5),
           ("blocks-2", (18, 16, 14), 1.3, 1e-5)]


def _geos(lbm):
    return [(lbm.Geometry.from_string(g, dims), om, F) for g, dims, om, F in DOMAINS]


def _alone(lbm, geo, om, F, runs, args):
    with lbm.Kernel("list-aa-pv-soa", geo, list(args)) as k:
        k.set_initial_weights()
        for s in runs:
            k.run(lbm.case_data(omega=om, xforce=F, iterations=s))
        return k.dump()


@pytest.mark.parametrize("prec", ["fp64", "fp32s", "fp16s"])
def test_batch_equals_single_lattice(lbm, oracle, prec):
    args = ["-prec", prec, "-lambda", repr(LAM)]
    geos = _geos(lbm)
    with lbm.Batch(args) as b:
        ids = [b.add(g, om, F) for g, om, F in geos]
        for _ in range(3):
            assert b.step(20) == len(ids)
        for i, (g, om, F) in zip(ids, geos):
            print(f"batch {prec} domain {i} dims={g.dims}", flush=True)
            got = b.dump(i)
            st = b.status(i)
            assert st.steps == 60 and st.active == 1
            # fp16s rounds the force increment into storage at every interval end
            runs = [20, 20, 20] if prec == "fp16s" else [60]
            assert np.array_equal(got, _alone(lbm, g, om, F, runs, args)), i
            m = g.mask.reshape(-1)
            if prec == "fp64":
                want, _ = oracle.list_run(m, g.dims, g.periodic, 60, om, F, magic=LAM, vsize=1)
            else:
                want, _ = oracle.list_run_shifted(m, g.dims, g.periodic, runs, om, F, prec, magic=LAM)
            assert np.array_equal(got, want), i


@pytest.mark.parametrize("prec", ["fp64", "fp32s", "fp16s"])
def test_batch_launch_shapes(lbm, prec):
    args = ["-prec", prec, "-lambda", repr(LAM)]
    geos = _geos(lbm)[:3]
    vmax = {"fp64": 1, "fp32s": 2, "fp16s": 4}[prec]
    want = [_alone(lbm, g, om, F, [40] if prec != "fp16s" else [20, 20], args) for g, om, F in geos]
    # fifth entry: groups of about that many MiB (one or two of these domains each)
    # seventh entry: odd-step chunks per XCD turn
    shapes = [(512, 1, vmax, 256), (256, 0, 1, 128), (512, 0, vmax, 512), (64, 0, 1, 64), (256, 1, vmax, 256),
              (512, 1, vmax, 256, 1), (256, 0, 1, 128, 2, 0), (512, 1, vmax, 256, 0, 0),
              (256, 1, vmax, 256, 0, 1, 128), (512, 0, 1, 128, 1, 1, 32), (256, 1, vmax, 512, 0, 1, 1)]
    for shape in shapes + ["auto"]:
        print(f"batch {prec} shape {shape}", flush=True)
        with lbm.Batch(args + ["-adj", "full"] if shape == (64, 0, 1, 64) else args) as b:
            ids = [b.add(g, om, F) for g, om, F in geos]
            if shape == "auto":
                t = b.autotune(2)
                assert t.evenMs > 0 and t.oddMs > 0 and t.pairMs > 0 and 1 <= t.vec <= vmax
                assert t.groupMB in (0, 64, 128, 192, 256, 384, 512, 768, 1024)
                assert t.oddGroup in (8, 32, 128)
            else:
                b.set_tuning(*shape)
            b.step(20)
            b.step(20)
            for i, w in zip(ids, want):
                assert np.array_equal(b.dump(i), w), (shape, i)


@pytest.mark.parametrize("group_mb", [0, 1])
def test_batch_membership_and_convergence(lbm, group_mb):
    args = ["-lambda", repr(LAM)]
    geos = _geos(lbm)
    with lbm.Batch(args) as b:
        if group_mb:
            b.set_tuning(512, 1, 1, 256, group_mb)      # about one domain per group
            assert b.tuning().groupMB == group_mb
        a = b.add(*geos[0], max_steps=40)
        c = b.add(*geos[1], max_steps=0)
        d = b.add(*geos[2], max_steps=100)
        assert b.step(20) == 3
        n0, c0, r0 = b.graph_stats()               # one graph per group, captured and replayed
        assert (c0, r0) == (n0, n0) and (n0 >= 2 if group_mb else n0 == 1)
        assert b.step(20) == 2                     # a reached its 40 steps
        assert b.graph_stats() == (n0, c0, 2 * n0)  # same ranges: replayed, nothing captured
        launched = 2 * n0
        assert (b.status(a).active, b.status(a).steps) == (0, 40)
        b.remove(c)                                # leaves mid-run
        e = b.add(*geos[3], max_steps=40)          # joins mid-run, reuses the free slot
        assert e == c
        for _ in range(2):
            b.step(20)
            launched += b.graph_stats()[0]
        assert b.graph_stats()[2] == launched
        st_d, st_e = b.status(d), b.status(e)
        assert (st_d.steps, st_e.steps, st_e.active) == (80, 40, 0)
        for i, (g, om, F), steps in ((a, geos[0], 40), (d, geos[2], 80), (e, geos[3], 40)):
            assert np.array_equal(b.dump(i), _alone(lbm, g, om, F, [steps], args)), i
        # moments of the status equal those of the node views
        dd = b.dump(d)
        m, j = b.moments(d)
        assert st_d.mass == pytest.approx(dd.sum(), rel=1e-12) and st_d.mass == pytest.approx(m, rel=1e-13)
        assert st_d.momentumX == pytest.approx(j, rel=1e-12)
        gpu_s, flups = b.stats()
        assert gpu_s > 0 and flups == sum(geo.n_fluid * s for (geo, _, _), s in
                                          ((geos[0], 40), (geos[1], 40), (geos[2], 80), (geos[3], 40)))
    # a loose tolerance: every domain converges before its limit
    with lbm.Batch(["-prec", "fp32s"]) as b:
        ids = [b.add(g, om, F, max_steps=40000) for g, om, F in geos[:2]]
        b.run(200, 1e-4)
        for i in ids:
            st = b.status(i)
            assert st.converged == 1 and st.active == 0 and 0 <= st.relChange <= 1e-4 and st.steps < 40000


def test_batch_with_an_empty_domain(lbm):
    """An all-solid domain among others: it holds no node, converges at its
    second check (no momentum to change) and leaves its neighbours' results
    bit-identical to the lattices run alone; ungrouped and one group per
    domain."""
    args = ["-lambda", repr(LAM)]
    geos = _geos(lbm)[:2]
    empty = lbm.Geometry.from_mask(np.zeros(6 * 5 * 4, dtype=np.int32), (6, 5, 4))
    want = [_alone(lbm, g, om, F, [40], args) for g, om, F in geos]
    for group_mb in (0, 1):
        with lbm.Batch(args) as b:
            if group_mb:
                b.set_tuning(512, 1, 1, 256, group_mb)
            a = b.add(*geos[0])
            e = b.add(empty, 1.0, 1e-5)
            c = b.add(*geos[1])
            assert b.step(20, 1e-6) == 3
            assert b.step(20, 1e-6) == 2
            st = b.status(e)
            assert (st.nFluid, st.active, st.converged, st.steps) == (0, 0, 1, 40)
            assert st.mass == 0.0 and st.momentumX == 0.0
            for i, w in zip((a, c), want):
                assert np.array_equal(b.dump(i), w), (group_mb, i)


def test_batch_setnode_guard(lbm):
    g, om, F = _geos(lbm)[0]
    with lbm.Batch() as b:
        i = b.add(g, om, F)
        b.step(10)
        kd = b.kernel_data(i)
        vals = (lbm.PdfT * 19)(*([0.05] * 19))
        lbm.hip_lib().LbmHipClearError()
        x, y, z = [int(v) for v in np.argwhere(g.mask != 0)[0][::-1]]
        kd.contents.SetNode(kd, x, y, z, vals)     # no current host copy: refused
        assert "call LbmHipSyncShadow first" in lbm.hip_lib().LbmHipLastError().decode()
        b.sync(i)
        lbm.hip_lib().LbmHipClearError()
        kd.contents.SetNode(kd, x, y, z, vals)
        assert lbm.hip_lib().LbmHipLastError().decode() == ""


@pytest.mark.parametrize("prec", ["fp32s", "fp16s"])
def test_batch_nodes_per_thread_change(lbm, prec):
    # with -pad no a lattice of odd size forces one node per thread in the even
    # step; when it leaves, the remaining lattices switch to 2 / 4 nodes per
    # thread and the replayed graph must follow
    args = ["-prec", prec, "-pad", "no", "-lambda", repr(LAM)]
    odd = (lbm.Geometry.from_string("porous-3-0.75-9", (21, 17, 13)), 1.1, 2e-5)
    even = (lbm.Geometry.from_string("porous-2-0.7-5", (20, 18, 16)), 1.3, 1e-5)
    assert odd[0].n_fluid % 2 == 1 and even[0].n_fluid % 4 == 0
    with lbm.Batch(args) as b:
        a = b.add(*odd, max_steps=20)
        c = b.add(*even)
        assert b.tuning().vec == 1
        b.step(20)
        assert b.status(a).active == 0 and b.tuning().vec == {"fp32s": 2, "fp16s": 4}[prec]
        b.step(20)
        b.step(20)
        runs = [20, 20, 20] if prec == "fp16s" else [60]
        assert np.array_equal(b.dump(c), _alone(lbm, even[0], even[1], even[2], runs, args))
        assert np.array_equal(b.dump(a), _alone(lbm, odd[0], odd[1], odd[2], [20], args))


def test_c4_batch_of_512_samples(lbm, oracle):
    """The batched workload at its configured size: 512 seeded porous
    samples (edges log-uniform in [32, 128] per axis, porosity U[0.5, 0.9],
    sphere radius U[3, 8], omega U[0.8, 1.6], F 1e-6), each to a relative
    x-momentum change <= 1e-6 per 100 steps or 5000 steps, in FP64 and in
    shifted FP32, autotuned.  Every domain ends converged or at its limit;
    FP32 permeabilities are within 1e-3 of FP64; every 64th domain equals its
    lattice run alone through Kernel() for the same steps, bit for bit; the
    smallest domain equals the oracle in both storage types."""
    import workloads as W

    samples = W.batch_samples(512, 2024)
    geos = [lbm.Geometry.from_string(s.geometry, s.dims) for s in samples]
    limit, check, tol = 5000, 100, 1e-6
    args = ["-lambda", repr(LAM)]
    stride = list(range(0, 512, 64))
    small = min(range(512), key=lambda i: geos[i].n_fluid)
    res = {}
    for prec in ("fp64", "fp32s"):
        print(f"c4 batch {prec}: 512 domains, {sum(g.n_fluid for g in geos)} nodes", flush=True)
        with lbm.Batch(args + ["-prec", prec]) as b:
            ids = [b.add(g, s.omega, s.xforce, limit) for g, s in zip(geos, samples)]
            b.autotune(2)
            active, rounds, launched = len(ids), 0, 0
            while active:
                active = b.step(check, tol)
                rounds += 1
                launched += b.graph_stats()[0]
                assert rounds <= limit // check + 1
            # every interval replayed from captured graphs, one per group
            slots, captures, replays = b.graph_stats()
            print(f"c4 batch {prec}: {rounds} intervals, {replays} graph replays, {captures} captures", flush=True)
            assert replays == launched and captures < replays
            st = [b.status(i) for i in ids]
            dumps = {j: b.dump(ids[j]) for j in stride + [small]}
        assert all(x.active == 0 for x in st)
        assert all(x.converged == 1 or x.steps == limit for x in st)
        assert all(x.converged == 0 or 0 <= x.relChange <= tol for x in st)
        res[prec] = ([W.permeability(x.momentumX, x.nFluid, s.cells, s.omega, s.xforce)
                      for x, s in zip(st, samples)], [x.steps for x in st], dumps)
        print(f"c4 batch {prec}: {sum(x.converged for x in st)} converged, median steps "
              f"{W.percentile([x.steps for x in st], 50)}", flush=True)
        for j in stride:
            g, s = geos[j], samples[j]
            with lbm.Kernel("list-aa-pv-soa", g, args + ["-prec", prec, "-quiet"]) as k:
                k.set_initial_weights()
                k.run(lbm.case_data(omega=s.omega, xforce=s.xforce, iterations=st[j].steps))
                assert np.array_equal(dumps[j], k.dump()), (prec, j)
        g, s = geos[small], samples[small]
        m = g.mask.reshape(-1)
        if prec == "fp64":
            want, _ = oracle.list_run(m, g.dims, g.periodic, st[small].steps, s.omega, s.xforce, magic=LAM,
                                      n_threads=8, vsize=1)
        else:
            want, _ = oracle.list_run_shifted(m, g.dims, g.periodic, st[small].steps, s.omega, s.xforce, prec,
                                              magic=LAM)
        assert np.array_equal(dumps[small], want), prec
    p64, p32 = np.array(res["fp64"][0]), np.array(res["fp32s"][0])
    err = np.abs(p32 - p64) / np.abs(p64)
    print(f"c4 permeability rel. err fp32s vs fp64: max {err.max():.3e} median {np.median(err):.3e}", flush=True)
    assert (p64 > 0).all() and err.max() <= 1e-3
