"""GPU parity of the shifted-storage variants of list-aa-pv-soa (-prec fp32s /
fp16s: deviations f - w stored in FP32 or as scaled binary16, collision in
FP32), the FP64 moment reductions and the steady-state driver.

Bars:
  * every -prec variant against the oracle's shifted restatement
    (orc_list_run_shifted), which follows the same operation order: node
    views bit-exact;
  * accuracy against the FP64 kernel on the 256^3 porous workload (the oracle
    for the variants is pinned against the FP64 oracle on the CPU):
    velocity relative L2 error and permeability relative error below the
    tolerances in ACC_TOL;
  * moment sums against sums of the node views, relative 1e-12 (mass) and
    1e-10 (momentum, a difference of nearly equal terms).
"""
import re

import numpy as np
import pytest

from conftest import blk_args, blk_tuple, golden, tiny_cases

pytestmark = pytest.mark.gpu

PRECS = ["fp32s", "fp16s"]
LAM = 3 / 16
# velocity relative L2 and permeability relative error vs FP64, 256^3 porous, 500 steps
ACC_TOL = {"fp32s": (1e-4, 1e-5), "fp16s": (1e-2, 1e-3)}
CX = np.array([0, 0, 1, -1, 1, 1, -1, -1, 0, 0, 1, -1, 0, 0, 0, 0, -1, 1, 0], dtype=float)


def _gpu(lbm, geo, runs, omega, F, args, nodes=None):
    with lbm.Kernel("list-aa-pv-soa", geo, list(args)) as k:
        if nodes is None:
            k.set_initial_weights()
        else:
            k.load(nodes)
        for steps in runs:
            k.run(lbm.case_data(omega=omega, xforce=F, iterations=steps))
        return k.dump(), k.list_info()


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("case", tiny_cases())
def test_tiny_shifted_bitexact(lbm, oracle, case, prec):
    g = golden(f"tiny_{case}.npz")
    dims, per = tuple(int(v) for v in g["dims"]), tuple(int(v) for v in g["periodic"])
    mask = g["mask"].astype(np.int32)
    args = blk_args(g)
    geo = lbm.Geometry.from_mask(mask, dims, per)
    steps, omega, F = int(g["steps"]), float(g["omega"]), float(g["xforce"])
    print(f"case {case} -prec {prec} dims={dims} args={args}", flush=True)
    got, info = _gpu(lbm, geo, [steps], omega, F, args + ["-prec", prec, "-lambda", repr(LAM)])
    assert info.prec == PRECS.index(prec) + 1
    assert info.bytesPerFlupEven == {"fp32s": 152.0, "fp16s": 76.0}[prec]
    want, _ = oracle.list_run_shifted(mask, dims, per, steps, omega, F, prec, magic=LAM, blk=blk_tuple(args))
    assert np.array_equal(got, want), f"max diff {np.abs(got - want).max()}"


@pytest.mark.parametrize("prec", PRECS)
def test_shifted_options_and_split_calls(lbm, oracle, prec):
    geo = lbm.Geometry.from_string("porous-3-0.75-9", (48, 40, 36))
    m = geo.mask.reshape(-1)
    runs = [30, 20, 10]
    want, _ = oracle.list_run_shifted(m, geo.dims, geo.periodic, runs, 1.25, 3e-5, prec, magic=LAM)
    for extra in ([], ["-adj", "full"], ["-nt", "off"], ["-even-cap", "off", "-bs", "128"],
                  ["-bs", "64", "-blk", "5"], ["-force", "on", "-adj", "full", "-nt", "off"],
                  ["-vec", "1"], ["-pad", "no"], ["-pad", "no", "-even-cap", "off", "-nt", "off"],
                  ["-even-cap", "force"], ["-even-cap", "force", "-pad", "no"]):
        print(f"porous48 -prec {prec} {extra}", flush=True)
        got, _ = _gpu(lbm, geo, runs, 1.25, 3e-5, ["-prec", prec, "-lambda", repr(LAM)] + extra)
        assert np.array_equal(got, want), extra
    got, _ = _gpu(lbm, geo, runs, 1.25, 3e-5, ["-prec", prec, "-lambda", repr(LAM), "-force", "off"])
    want_off, _ = oracle.list_run_shifted(m, geo.dims, geo.periodic, runs, 1.25, 3e-5, prec, magic=LAM,
                                          force_on=False)
    assert np.array_equal(got, want_off)


def test_fp16_scale_option(lbm, oracle):
    geo = lbm.Geometry.from_string("channel", (20, 14, 12), (1, 0, 0))
    m = geo.mask.reshape(-1)
    for scale, shift in (("1024", 10), ("65536", 16)):
        got, _ = _gpu(lbm, geo, [80], 1.4, 1e-4, ["-prec", "fp16s", "-fp16-scale", scale])
        want, _ = oracle.list_run_shifted(m, geo.dims, geo.periodic, 80, 1.4, 1e-4, "fp16s", f16_shift=shift)
        assert np.array_equal(got, want), scale


@pytest.mark.parametrize("prec", PRECS)
def test_shifted_setnode_getnode(lbm, oracle, prec):
    geo = lbm.Geometry.from_string("porous-2-0.7-5", (20, 18, 16))
    m = geo.mask.reshape(-1)
    rng = np.random.default_rng(3)
    nodes = np.tile(oracle.W, (geo.n_fluid, 1)) + 1e-3 * rng.standard_normal((geo.n_fluid, 19))
    # encode/decode only: the oracle's conversion of the same node views
    enc, _ = oracle.list_run_shifted(m, geo.dims, geo.periodic, [], 1.0, 0.0, prec, nodes=nodes)
    with lbm.Kernel("list-aa-pv-soa", geo, ["-prec", prec]) as k:
        k.load(nodes)
        assert np.array_equal(k.dump(), enc)
        k.run(lbm.case_data(iterations=0))          # upload and download unchanged
        assert np.array_equal(k.dump(), enc)
        k.run(lbm.case_data(omega=1.3, xforce=2e-5, iterations=40))
        got = k.dump()
    want, _ = oracle.list_run_shifted(m, geo.dims, geo.periodic, 40, 1.3, 2e-5, prec, nodes=nodes)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("prec", ["fp64"] + PRECS)
def test_moments_match_node_sums(lbm, prec):
    geo = lbm.Geometry.from_string("porous-3-0.75-9", (40, 36, 32))
    with lbm.Kernel("list-aa-pv-soa", geo, ["-prec", prec, "-lambda", repr(LAM)]) as k:
        k.set_initial_weights()
        m0, j0 = k.moments()
        assert m0 == pytest.approx(geo.n_fluid, rel=1e-14) and abs(j0) < 1e-12
        k.run(lbm.case_data(omega=1.2, xforce=1e-4, iterations=60))
        m, j = k.moments()
        d = k.dump()
    assert m == pytest.approx(d.sum(), rel=1e-12)
    assert j == pytest.approx((d @ CX).sum(), rel=1e-10)
    assert j > 0


def test_run_steady(lbm, oracle):
    geo = lbm.Geometry.from_string("channel", (14, 12, 10), (1, 0, 0))
    m = geo.mask.reshape(-1)
    cd = lbm.case_data(omega=1.1, xforce=1e-4, iterations=0)
    # tol 0 never stops early: chunks of 16 steps (15 rounded up) up to 100
    with lbm.Kernel("list-aa-pv-soa", geo) as k:
        k.set_initial_weights()
        r = k.run_steady(cd, 15, 0.0, 100)
        assert (r.steps, r.converged) == (100, 0)
        assert (r.mass, r.momentumX) == k.moments()
        got = k.dump()
    want, _ = oracle.list_run(m, geo.dims, geo.periodic, 100, 1.1, 1e-4, vsize=1)
    assert np.array_equal(got, want)
    # a loose tolerance stops at the first check that meets it
    with lbm.Kernel("list-aa-pv-soa", geo, ["-prec", "fp32s"]) as k:
        k.set_initial_weights()
        r = k.run_steady(cd, 200, 1e-4, 40000)
        assert r.converged == 1 and 0 <= r.relChange <= 1e-4 and r.steps % 200 == 0 and r.steps < 40000
        assert r.seconds > 0 and k.duration == pytest.approx(r.seconds)


def _ux_perm(d, n_cells, omega, F):
    ux = d @ CX - F / 2                      # node views hold the last step's force increment
    nu = (1 / omega - 0.5) / 3
    return ux, nu * ux.sum() / n_cells / F   # Darcy: k = nu * U_superficial / F


def test_porous256_shifted_accuracy(lbm):
    geo = lbm.Geometry.from_string("porous-6-0.8-42", (256, 256, 256))
    cells = 256 ** 3
    omega, F = 1.2, 1e-6
    print("porous256 fp64 500 steps", flush=True)
    ref, _ = _gpu(lbm, geo, [500], omega, F, ["-lambda", repr(LAM)])
    u_ref, k_ref = _ux_perm(ref, cells, omega, F)
    del ref
    assert k_ref > 0
    for prec in PRECS:
        print(f"porous256 {prec} 500 steps", flush=True)
        with lbm.Kernel("list-aa-pv-soa", geo, ["-prec", prec, "-lambda", repr(LAM)]) as k:
            k.set_initial_weights()
            k.run(lbm.case_data(omega=omega, xforce=F, iterations=500))
            mass, _ = k.moments()
            d = k.dump()
        u, kp = _ux_perm(d, cells, omega, F)
        del d
        vel = np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref)
        perm = abs(kp - k_ref) / k_ref
        print(f"  {prec}: velocity rel L2 {vel:.3e}, permeability rel {perm:.3e}, mass rel "
              f"{abs(mass - geo.n_fluid) / geo.n_fluid:.3e}", flush=True)
        assert vel < ACC_TOL[prec][0]
        assert perm < ACC_TOL[prec][1]
        assert abs(mass - geo.n_fluid) / geo.n_fluid < 1e-6


def test_porous256_steady_state_fp32(lbm):
    """The list-AA workload to steady state (relative x-momentum change
    < 1e-7 per 100 steps, tested in FP64 on the device) in FP64 and in
    shifted FP32: both converge, and FP32's velocity field and permeability
    stay within 1e-4 of FP64's (measured 5.3e-6 and 5.0e-6; the serving
    target is 1e-3)."""
    geo = lbm.Geometry.from_string("porous-6-0.8-42", (256, 256, 256))
    cells = 256 ** 3
    omega, F = 1.2, 1e-6
    res = {}
    for prec in ("fp64", "fp32s"):
        print(f"porous256 {prec} to steady state", flush=True)
        with lbm.Kernel("list-aa-pv-soa", geo, ["-prec", prec, "-lambda", repr(LAM)]) as k:
            k.set_initial_weights()
            r = k.run_steady(lbm.case_data(omega=omega, xforce=F, iterations=0), 100, 1e-7, 20000)
            u, kp = _ux_perm(k.dump(), cells, omega, F)
        print(f"  {prec}: {r.steps} steps, converged {r.converged}, permeability {kp:.6f}", flush=True)
        assert r.converged
        assert abs(r.mass - geo.n_fluid) / geo.n_fluid < 1e-10
        res[prec] = (u, kp)
    (u_ref, k_ref), (u, kp) = res["fp64"], res["fp32s"]
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) < 1e-4
    assert abs(kp - k_ref) / k_ref < 1e-4


def _channel64_steady(lbm, prec):
    """The small configuration's channel (64^3, x periodic, walls on y and
    z, aa-soa, omega 1.0, F 1e-5, Lambda 3/16) to a relative x-momentum
    change < 1e-7 per 100 steps; (converged, steps, velocity field,
    permeability)."""
    geo = lbm.Geometry.from_string("channel", (64, 64, 64), (1, 0, 0))
    omega, F = 1.0, 1e-5
    with lbm.Kernel("aa-soa", geo, ["-prec", prec, "-lambda", repr(LAM), "-quiet"]) as k:
        k.set_initial_weights()
        r = k.run_steady(lbm.case_data(omega=omega, xforce=F, iterations=0), 100, 1e-7, 80000)
        u, kp = _ux_perm(k.dump(), 64 ** 3, omega, F)
    print(f"channel64 {prec}: {r.steps} steps, converged {r.converged}, permeability {kp:.6e}", flush=True)
    return r.converged, r.steps, u, kp


def test_channel64_steady_state_precisions(lbm):
    """The precision configuration (shifted storage on the small
    configuration's channel, run to steady state): FP32 converges and stays
    within 1e-4 of FP64 in velocity and permeability (measured 1.2e-5).
    FP16 converges too but misses the 1e-3 permeability bar: binary16 keeps
    11 significant bits of each stored deviation (relative 4.9e-4, whatever
    the power-of-two scale), while the per-step force increment and the
    viscous part of the deviations are of the same order, so the rounding
    of every store biases the steady state -- by 4.5e-2 in permeability
    here (pinned between 1e-2 and 1e-1; the strict bar is the xfail below)."""
    c64, _, u64, k64 = _channel64_steady(lbm, "fp64")
    assert c64
    c32, _, u32, k32 = _channel64_steady(lbm, "fp32s")
    assert c32
    assert np.linalg.norm(u32 - u64) / np.linalg.norm(u64) < 1e-4 and abs(k32 - k64) / k64 < 1e-4
    c16, _, u16, k16 = _channel64_steady(lbm, "fp16s")
    assert c16
    err = abs(k16 - k64) / k64
    assert 1e-2 < err < 1e-1
    assert 1e-2 < np.linalg.norm(u16 - u64) / np.linalg.norm(u64) < 1e-1


@pytest.mark.xfail(strict=True, reason="binary16 storage of the deviations biases the channel's steady state "
                                       "by 4.5e-2 in permeability (11 significant bits per stored value)")
def test_channel64_fp16_meets_the_permeability_bar(lbm):
    _, _, _, k64 = _channel64_steady(lbm, "fp64")
    _, _, _, k16 = _channel64_steady(lbm, "fp16s")
    assert abs(k16 - k64) / k64 <= 1e-3


# ---- aa-soa with shifted storage ---------------------------------------
AA_CASES = ["fluid6", "blocks2_12", "box8", "channel10x9x7", "porous16"]


def _aa_gpu(lbm, geo, runs, omega, F, args, nodes=None):
    with lbm.Kernel("aa-soa", geo, list(args)) as k:
        if nodes is None:
            k.set_initial_weights()
        else:
            k.load(nodes)
        for steps in runs:
            k.run(lbm.case_data(omega=omega, xforce=F, iterations=steps))
        return k.dump()


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("case", AA_CASES)
def test_aa_shifted_bitexact(lbm, oracle, case, prec):
    """Against the oracle's shifted aa-soa restatement (fix-up copies around
    full-array steps; open faces through the ghost layer) -- node views
    bit-exact, across launch shapes and split Kernel() calls."""
    g = golden(f"tiny_{case}.npz")
    dims, per = tuple(int(v) for v in g["dims"]), tuple(int(v) for v in g["periodic"])
    mask = g["mask"].astype(np.int32)
    geo = lbm.Geometry.from_mask(mask, dims, per)
    omega, F = float(g["omega"]), float(g["xforce"])
    base = ["-prec", prec, "-lambda", repr(LAM)]
    want = oracle.aa_run_shifted(mask, dims, per, [60, 40], omega, F, prec, magic=LAM)
    for extra in ([], ["-vec", "1"], ["-nt", "on"], ["-even-cap", "force"], ["-even-cap", "force", "-vec", "1"],
                  ["-even-cap", "off"], ["-odd-pairs", "on"], ["-odd-pairs", "on", "-even-cap", "force"]):
        print(f"aa case {case} -prec {prec} {extra}", flush=True)
        got = _aa_gpu(lbm, geo, [60, 40], omega, F, base + extra)
        assert np.array_equal(got, want), (extra, np.abs(got - want).max())
    want_off = oracle.aa_run_shifted(mask, dims, per, 30, omega, F, prec, magic=LAM, force_on=False)
    got = _aa_gpu(lbm, geo, [30], omega, F, base + ["-force", "off"])
    assert np.array_equal(got, want_off)


@pytest.mark.parametrize("prec", PRECS)
def test_aa_shifted_long_rows_and_nodes(lbm, oracle, prec):
    """z rows longer than one block (300 cells), x periodic, y/z solid
    faces; random initial node views through SetNode; FP16 scale option."""
    geo = lbm.Geometry.from_string("porous-3-0.75-9", (24, 20, 300))
    m = geo.mask.reshape(-1)
    rng = np.random.default_rng(11)
    nodes = np.tile(oracle.W, (geo.n_fluid, 1)) + 1e-3 * rng.standard_normal((geo.n_fluid, 19))
    print(f"aa porous 24x20x300 -prec {prec}", flush=True)
    want = oracle.aa_run_shifted(m, geo.dims, geo.periodic, [40, 20], 1.25, 3e-5, prec, magic=LAM, nodes=nodes)
    for extra in ([], ["-vec", "1"], ["-odd-pairs", "on"]):
        got = _aa_gpu(lbm, geo, [40, 20], 1.25, 3e-5, ["-prec", prec, "-lambda", repr(LAM)] + extra, nodes=nodes)
        assert np.array_equal(got, want), extra
    # the list kernel gives the same bits here (no open faces)
    got_list, _ = _gpu(lbm, geo, [40, 20], 1.25, 3e-5, ["-prec", prec, "-lambda", repr(LAM)], nodes=nodes)
    assert np.array_equal(got_list, want)
    if prec == "fp16s":
        want = oracle.aa_run_shifted(m, geo.dims, geo.periodic, 30, 1.25, 3e-5, prec, magic=LAM, nodes=nodes,
                                     f16_shift=10)
        got = _aa_gpu(lbm, geo, [30], 1.25, 3e-5, ["-prec", prec, "-lambda", repr(LAM), "-fp16-scale", "1024"],
                      nodes=nodes)
        assert np.array_equal(got, want)


@pytest.mark.parametrize("prec", PRECS)
def test_aa_shifted_open_faces_ragged_rows(lbm, oracle, prec):
    """All-fluid boxes whose faces are open (ghost layer in use) and whose z
    rows are not a multiple of the even step's cells per thread: the groups
    that reach past the row end write the ghost values back unchanged."""
    for dims, per in (((9, 7, 7), (0, 0, 0)), ((11, 6, 10), (1, 0, 0)), ((5, 8, 13), (0, 1, 0))):
        mask = np.ones(int(np.prod(dims)), dtype=np.int32)
        geo = lbm.Geometry.from_mask(mask, dims, per)
        want = oracle.aa_run_shifted(mask, dims, per, [20, 10], 1.4, 5e-5, prec, magic=LAM)
        for extra in ([], ["-vec", "1"], ["-even-cap", "force"], ["-odd-pairs", "on"]):
            print(f"aa open {dims} {per} -prec {prec} {extra}", flush=True)
            got = _aa_gpu(lbm, geo, [20, 10], 1.4, 5e-5, ["-prec", prec, "-lambda", repr(LAM)] + extra)
            assert np.array_equal(got, want), (dims, extra)


@pytest.mark.parametrize("prec", PRECS)
def test_aa_shifted_rows_over_several_blocks(lbm, oracle, prec):
    """z rows of 1100 cells with scattered solid cells, z periodic: the narrow
    types' groups of z cells (FP32 pairs in both steps, FP16 groups of 4 in
    the even step) span several blocks per row."""
    dims, per = (5, 6, 1100), (0, 0, 1)
    rng = np.random.default_rng(5)
    mask = (rng.random(int(np.prod(dims))) > 0.05).astype(np.int32)
    geo = lbm.Geometry.from_mask(mask, dims, per)
    want = oracle.aa_run_shifted(mask, dims, per, [12, 8], 1.3, 2e-5, prec, magic=LAM)
    for extra in ([], ["-vec", "1"], ["-odd-pairs", "on"]):
        print(f"aa rows {dims} -prec {prec} {extra}", flush=True)
        got = _aa_gpu(lbm, geo, [12, 8], 1.3, 2e-5, ["-prec", prec, "-lambda", repr(LAM)] + extra)
        assert np.array_equal(got, want), extra


@pytest.mark.parametrize("prec", ["fp64"] + PRECS)
def test_aa_row_layouts(lbm, oracle, prec, capfd):
    """aa-soa lays its z rows out without ghost cells (FP64, FP32; rows
    padded to whole 128-byte lines) when no fluid cell can reach one (z
    periodic, or no fluid cell on either z face), and with ghost cells
    otherwise.  Rows of 64 and 71 cells with scattered solid cells: open z
    faces (ghost cells), z periodic and solid z faces (compact rows), each
    against the oracle, node views bit-exact."""
    cases = []
    for nz in (64, 71):
        dims = (4, 5, nz)
        rng = np.random.default_rng(3)
        base = (rng.random(int(np.prod(dims))) > 0.1).astype(np.int32)
        walls = base.reshape(nz, 5, 4).copy()
        walls[0] = 0
        walls[-1] = 0
        cases += [(dims, base, (0, 0, 0)), (dims, base, (0, 0, 1)), (dims, walls.reshape(-1), (1, 0, 0))]
    for dims, mask, per in cases:
        geo = lbm.Geometry.from_mask(mask, dims, per)
        if prec == "fp64":
            runs = [30]
            want = oracle.aa_run(mask, dims, per, 30, 1.3, 2e-5, magic=LAM)[0]
        else:
            runs = [20, 10]
            want = oracle.aa_run_shifted(mask, dims, per, runs, 1.3, 2e-5, prec, magic=LAM)
        print(f"aa rows {dims} {per} -prec {prec}", flush=True)
        capfd.readouterr()
        got = _aa_gpu(lbm, geo, runs, 1.3, 2e-5, ["-prec", prec, "-lambda", repr(LAM)])
        out = capfd.readouterr().out
        assert np.array_equal(got, want), per
        # the layout Init chose: compact rows unless a fluid cell reaches a z
        # ghost cell (open z faces here) or the storage is FP16
        line = 128 // {"fp64": 8, "fp32s": 4, "fp16s": 2}[prec]
        nz = dims[2]
        m = re.search(r"# z rows: (compact|with ghost cells), (\d+) values apart, z = 0 at value (\d+)", out)
        assert m, out
        if (per[2] or not mask.reshape(nz, -1)[[0, -1]].any()) and prec != "fp16s":
            want_layout = ("compact", -(-nz // line) * line, 0)
        else:
            want_layout = ("with ghost cells", -(-(line + nz + 1) // line) * line, line)
        assert (m.group(1), int(m.group(2)), int(m.group(3))) == want_layout, (dims, per, prec)


@pytest.mark.parametrize("prec", ["fp64"] + PRECS)
def test_aa_moments_match_node_sums(lbm, prec):
    """aa-soa's FP64 device sums against sums of its node views, and
    against the list kernel's on the same lattice (no open faces, so the two
    kernels hold the same node views)."""
    geo = lbm.Geometry.from_string("porous-3-0.75-9", (40, 36, 32))
    args = ["-prec", prec, "-lambda", repr(LAM)]
    with lbm.Kernel("aa-soa", geo, args) as k:
        k.set_initial_weights()
        m0, j0 = k.moments()
        assert m0 == pytest.approx(geo.n_fluid, rel=1e-14) and abs(j0) < 1e-12
        k.run(lbm.case_data(omega=1.2, xforce=1e-4, iterations=60))
        m, j = k.moments()
        d = k.dump()
    assert m == pytest.approx(d.sum(), rel=1e-12)
    assert j == pytest.approx((d @ CX).sum(), rel=1e-10)
    assert j > 0
    with lbm.Kernel("list-aa-pv-soa", geo, args) as k:
        k.set_initial_weights()
        k.run(lbm.case_data(omega=1.2, xforce=1e-4, iterations=60))
        ml, jl = k.moments()
        assert np.array_equal(k.dump(), d)
    assert ml == pytest.approx(m, rel=1e-13) and jl == pytest.approx(j, rel=1e-11)


def test_aa_run_steady(lbm, oracle):
    geo = lbm.Geometry.from_string("channel", (14, 12, 10), (1, 0, 0))
    m = geo.mask.reshape(-1)
    cd = lbm.case_data(omega=1.1, xforce=1e-4, iterations=0)
    with lbm.Kernel("aa-soa", geo) as k:
        k.set_initial_weights()
        r = k.run_steady(cd, 15, 0.0, 100)
        assert (r.steps, r.converged) == (100, 0)
        assert (r.mass, r.momentumX) == k.moments()
        got = k.dump()
    want, _ = oracle.aa_run(m, geo.dims, geo.periodic, 100, 1.1, 1e-4)
    assert np.array_equal(got, want)
    with lbm.Kernel("aa-soa", geo, ["-prec", "fp32s"]) as k:
        k.set_initial_weights()
        r = k.run_steady(cd, 200, 1e-4, 40000)
        assert r.converged == 1 and 0 <= r.relChange <= 1e-4 and r.steps % 200 == 0 and r.steps < 40000
        assert r.seconds > 0 and k.duration == pytest.approx(r.seconds)


@pytest.mark.parametrize("prec", PRECS)
def test_aa_shifted_large_lattice_defaults(lbm, prec):
    """Over 256 MiB of shifted storage aa-soa's even step defaults to
    non-temporal accesses at full occupancy: the same bits as the explicit
    cache-friendly launch."""
    geo = lbm.Geometry.from_string("channel", (80, 256, 256), (1, 0, 0))   # 309 MiB in FP16, 514 in FP32
    base = ["-prec", prec, "-lambda", repr(LAM)]
    print(f"aa 80x256x256 -prec {prec} defaults vs explicit", flush=True)
    got = _aa_gpu(lbm, geo, [20], 1.2, 1e-5, base)
    want = _aa_gpu(lbm, geo, [20], 1.2, 1e-5, base + ["-nt", "off", "-even-cap", "on"])
    assert np.array_equal(got, want)


def test_aa_duct256_shifted_moments(lbm):
    """aa-soa at 256^3 (duct channel, 200 steps): the FP64 device moments of
    the shifted variants against FP64 -- size-independent checks at the full
    size (mass conserved, x momentum within the storage type's accuracy)."""
    geo = lbm.Geometry.from_string("channel", (256, 256, 256), (1, 0, 0))
    res = {}
    for prec in ["fp64"] + PRECS:
        print(f"aa duct256 -prec {prec} 200 steps", flush=True)
        with lbm.Kernel("aa-soa", geo, ["-prec", prec, "-lambda", repr(LAM)]) as k:
            k.set_initial_weights()
            k.set_shadow_download(False)
            k.run(lbm.case_data(omega=1.0, xforce=1e-5, iterations=200))
            res[prec] = k.moments()
    m64, j64 = res["fp64"]
    assert abs(m64 - geo.n_fluid) / geo.n_fluid < 1e-12 and j64 > 0
    # measured: fp32s mass 9e-12, momentum 1.3e-8; fp16s mass 7.4e-6, momentum 1.0e-3
    for prec, (mtol, jtol) in {"fp32s": (1e-9, 1e-6), "fp16s": (2e-5, 2e-3)}.items():
        m, j = res[prec]
        print(f"  {prec}: mass rel {abs(m - geo.n_fluid) / geo.n_fluid:.2e}, momentum rel {abs(j - j64) / j64:.2e}",
              flush=True)
        assert abs(m - geo.n_fluid) / geo.n_fluid < mtol
        assert abs(j - j64) / j64 < jtol
