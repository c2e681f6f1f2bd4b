"""GPU parity: the HIP kernels through the plugin ABI against the oracle and
the reference's fixtures.

Bars:
  * list build (numbering, grid, adjacency values): bit-exact;
  * PDFs: bit-exact against the oracle evaluated with the scalar expression
    tree (the GPU uses that tree and no FMA), which for aa-soa is also
    bit-exact against the reference's own output (fixture hashes);
  * list-aa-pv-soa against the reference, whose AVX path uses another
    association tree for some nodes: relative max-norm <= 1e-12 after up to
    1000 steps, as written in REL_TOL.
"""
import hashlib

import numpy as np
import pytest

from conftest import DEGENERATE_SHAPES, blk_args as _blk_args, blk_tuple as _blk, golden, tiny_cases

pytestmark = pytest.mark.gpu

REL_TOL = 1e-12


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _relmax(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _run(lbm, name, geo, steps, omega, F, args=(), init="weights"):
    with lbm.Kernel(name, geo, list(args)) as k:
        cd = lbm.case_data(omega=omega, xforce=F, iterations=steps)
        if init == "weights":
            k.set_initial_weights()
        elif init == "verify":
            k.set_initial_verify(cd)
        k.run(cd)
        out = k.dump()
        extra = k.list_build() if name == "list-aa-pv-soa" else None
        t = k.timings()
    return out, extra, t


@pytest.mark.parametrize("case", tiny_cases())
def test_tiny_list(lbm, oracle, case):
    g = golden(f"tiny_{case}.npz")
    dims, per = tuple(int(v) for v in g["dims"]), tuple(int(v) for v in g["periodic"])
    mask = g["mask"].astype(np.int32)
    args = _blk_args(g)
    geo = lbm.Geometry.from_mask(mask, dims, per)
    print(f"case {case} list-aa-pv-soa dims={dims} args={args}", flush=True)
    steps, omega, F = int(g["steps"]), float(g["omega"]), float(g["xforce"])
    got, (adj, grid, coords), _ = _run(lbm, "list-aa-pv-soa", geo, steps, omega, F, args)
    fl = mask != 0
    assert np.array_equal(adj, g["adj"])
    assert np.array_equal(coords, g["coords"])
    assert np.array_equal(grid[fl], g["grid"][fl])
    assert (grid[~fl] == 0xffffffff).all()
    want, _ = oracle.list_run(mask, dims, per, steps, omega, F, blk=_blk(args), vsize=1)
    assert np.array_equal(got, want), f"max diff {np.abs(got - want).max()}"
    for key in ("list_t1", "list_t3"):
        assert _relmax(got, g[key]) <= REL_TOL


@pytest.mark.parametrize("case", tiny_cases())
def test_tiny_aa(lbm, case):
    g = golden(f"tiny_{case}.npz")
    dims, per = tuple(int(v) for v in g["dims"]), tuple(int(v) for v in g["periodic"])
    mask = g["mask"].astype(np.int32)
    geo = lbm.Geometry.from_mask(mask, dims, per)
    print(f"case {case} aa-soa dims={dims}", flush=True)
    got, _, _ = _run(lbm, "aa-soa", geo, int(g["steps"]), float(g["omega"]), float(g["xforce"]))
    # bit-exact against the reference's own aa-soa output
    assert np.array_equal(got, g["aa_t2"]), f"max diff {np.abs(got - g['aa_t2']).max()}"


@pytest.mark.parametrize("name", ["list-aa-pv-soa", "aa-soa"])
def test_verify16(lbm, oracle, name, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    g = golden("verify16.npz")
    geo = lbm.Geometry.from_string("box", (16, 16, 16), (1, 1, 0))
    print(f"verify16 {name}", flush=True)
    with lbm.Kernel(name, geo) as k:
        cd = lbm.case_data(omega=1.0, xforce=1e-5, iterations=1000)
        k.set_initial_verify(cd)
        k.run(cd)
        norm = k.verify(cd)
        got = k.dump()
        dens = k.density()
    if name == "aa-soa":
        assert _sha(got) == str(g["aa_sha"])
        assert norm == float(g["aa_norm"])
        assert dens == float(g["aa_density"])
    else:
        want, _ = oracle.list_run(geo.mask.reshape(-1), (16, 16, 16), (1, 1, 0), 1000, 1.0, 1e-5,
                                  vsize=1, init="verify")
        assert np.array_equal(got, want)
        assert _relmax(got[::13], g["list_t4_sample"]) <= REL_TOL
        assert abs(norm - float(g["list_norm"])) <= 1e-9
    assert f"{norm:e}" == "4.561374e-02"


@pytest.mark.parametrize("name", ["list-aa-pv-soa", "aa-soa"])
def test_verify64_12000_steps(lbm, name, tmp_path, monkeypatch):
    """The reference's analytic check (KernelVerifiy, src/Kernel.c:511-625)
    at the size where it passes: the -verify box at 64^3 (x and y periodic,
    walls at z), omega 1.0, F 1e-5, 12000 steps.  aa-soa is bit-identical to
    the reference's aa-soa and prints its deviation; list-aa-pv-soa is
    within 1e-12 of the reference's list kernel; both below the driver's
    0.1 threshold (src/Main.c:636)."""
    monkeypatch.chdir(tmp_path)
    g = golden("verify64_12000.npz")
    geo = lbm.Geometry.from_string("box", (64, 64, 64), (1, 1, 0))
    assert _sha(np.ascontiguousarray(geo.mask.reshape(-1))) == str(g["mask_sha"])
    print(f"verify64 {name} 12000 steps", flush=True)
    with lbm.Kernel(name, geo) as k:
        cd = lbm.case_data(omega=1.0, xforce=1e-5, iterations=int(g["steps"]))
        k.set_initial_verify(cd)
        k.run(cd)
        norm = k.verify(cd)
        got = k.dump()
    if name == "aa-soa":
        assert _sha(got) == str(g["aa_sha"])
        assert norm == float(g["aa_norm"])
    else:
        assert _relmax(got[::97], g["list_t8_sample"]) <= REL_TOL
        assert abs(norm - float(g["list_norm"])) <= 1e-9 * norm
    assert 0 < norm <= 0.1


def _duct_mean_velocity(F, nu, W, terms=200):
    """Mean velocity of laminar flow through a square duct of side W under a
    body force F (textbook series for the rectangular duct, a = b = W):
    F W^2 / (12 nu) * (1 - 192 / pi^5 * sum_{n odd} tanh(n pi / 2) / n^5)."""
    s = sum(np.tanh(n * np.pi / 2) / n ** 5 for n in range(1, 2 * terms, 2))
    return F * W * W / (12 * nu) * (1 - 192 / np.pi ** 5 * s)


@pytest.mark.parametrize("name", ["aa-soa", "list-aa-pv-soa"])
def test_duct64_steady_state_against_series(lbm, name):
    """The small configuration's square duct (64^3 channel: x periodic, walls
    on the y and z faces) at Lambda 3/16, omega 1.0, F 1e-5, run to steady
    state: half-way bounce-back puts the walls half a cell outside the 62
    fluid layers, so the mean velocity over the fluid cells must match the
    analytic square-duct series for W = 62 (velocity at the half step: node
    views carry the last force increment, u = sum_d c_x,d f_d - F/2)."""
    F, omega = 1e-5, 1.0
    geo = lbm.Geometry.from_string("channel", (64, 64, 64), (1, 0, 0))
    print(f"duct64 {name} to steady state", flush=True)
    with lbm.Kernel(name, geo, ["-lambda", "0.1875", "-quiet"]) as k:
        k.set_initial_weights()
        r = k.run_steady(lbm.case_data(omega=omega, xforce=F, iterations=0), 1000, 1e-10, 200000)
    assert r.converged == 1, r.relChange
    u = (r.momentumX - 0.5 * F * geo.n_fluid) / geo.n_fluid
    want = _duct_mean_velocity(F, (1 / omega - 0.5) / 3, 62)
    print(f"duct64 {name}: {r.steps} steps, mean u {u:.9e}, series {want:.9e}, rel {u / want - 1:.2e}", flush=True)
    assert abs(u / want - 1) <= 1e-3          # second-order wall error, 3.7e-3 at W = 16 and 1.1e-3 at W = 30


def test_channel64_aa_bitexact_vs_reference(lbm):
    g = golden("channel64.npz")
    geo = lbm.Geometry.from_string("channel", (64, 64, 64), (1, 0, 0))
    assert _sha(np.ascontiguousarray(geo.mask.reshape(-1))) == str(g["mask_sha"])
    print("channel64 aa-soa 1000 steps", flush=True)
    got, _, _ = _run(lbm, "aa-soa", geo, 1000, 1.0, 1e-5)
    assert _sha(got) == str(g["aa_sha"])
    assert got.sum() == pytest.approx(float(g["aa_mass"]), rel=1e-13)


def test_channel64_list(lbm, oracle):
    g = golden("channel64.npz")
    geo = lbm.Geometry.from_string("channel", (64, 64, 64), (1, 0, 0))
    print("channel64 list-aa-pv-soa 1000 steps", flush=True)
    got, _, _ = _run(lbm, "list-aa-pv-soa", geo, 1000, 1.0, 1e-5)
    assert _relmax(got[::97], g["list_t8_sample"]) <= REL_TOL
    want, _ = oracle.list_run(geo.mask.reshape(-1), (64, 64, 64), (1, 0, 0), 1000, 1.0, 1e-5, vsize=1)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("name", ["aa-soa", "list-aa-pv-soa"])
def test_channel64_small_configuration(lbm, oracle, name):
    """The small configuration exactly as stated: 64^3 channel (x periodic,
    walls on the y and z faces), rho 1 and u 0 at start (f = w), omega 1.0,
    Lambda 3/16, F 1e-5, 1000 steps.  The reference hard-codes Lambda 1/12
    (pinned by the fixture tests above); at 3/16 both kernels must equal the
    oracle bit for bit and conserve mass to 1e-12."""
    geo = lbm.Geometry.from_string("channel", (64, 64, 64), (1, 0, 0))
    m = geo.mask.reshape(-1)
    print(f"channel64 {name} Lambda 3/16 1000 steps", flush=True)
    got, _, _ = _run(lbm, name, geo, 1000, 1.0, 1e-5, ["-lambda", "0.1875"])
    if name == "aa-soa":
        want, _ = oracle.aa_run(m, geo.dims, geo.periodic, 1000, 1.0, 1e-5, magic=3 / 16)
    else:
        want, _ = oracle.list_run(m, geo.dims, geo.periodic, 1000, 1.0, 1e-5, magic=3 / 16, vsize=1)
    assert np.array_equal(got, want)
    assert abs(got.sum() - geo.n_fluid) / geo.n_fluid <= 1e-12


def test_porous256_list_build_and_steps(lbm):
    g = golden("porous256.npz")
    geo = lbm.Geometry.from_string("porous-6-0.8-42", (256, 256, 256))
    mask = np.ascontiguousarray(geo.mask.reshape(-1))
    assert _sha(mask) == str(g["mask_sha"])
    assert geo.n_fluid == int(g["n_fluid"])
    print(f"porous256 list-aa-pv-soa nFluid={geo.n_fluid}", flush=True)
    with lbm.Kernel("list-aa-pv-soa", geo) as k:
        info = k.list_info()
        assert info.nCells == int(g["n_cells"])
        adj, grid, coords = k.list_build()
        assert _sha(adj) == str(g["adj_sha"])
        assert _sha(np.where(mask != 0, grid, 0xffffffff).astype(np.uint32)) == str(g["grid_sha"])
        every = int(g["sample_every"])
        assert np.array_equal(coords[::every], g["coords_sample"])
        del adj, grid, coords
        k.set_initial_weights()
        k.run(lbm.case_data(omega=1.2, xforce=1e-6, iterations=20))
        got = k.dump()
    assert _relmax(got[::every], g["list_t8_sample"]) <= REL_TOL
    assert got.sum() == pytest.approx(float(g["list_mass"]), rel=1e-13)


def _x_plane_sums(dump, mask, nx):
    xs = np.flatnonzero(mask) % nx
    out = np.zeros((nx, dump.shape[1]))
    for d in range(dump.shape[1]):
        out[:, d] = np.bincount(xs, weights=dump[:, d], minlength=nx)
    return out


@pytest.mark.parametrize("lam", ["1/12", "3/16"])
def test_porous256_500_steps(lbm, lam):
    """The 256^3 porous workload at its full length (500 steps, omega 1.2,
    F 1e-6) against tests/golden/porous256_500.npz: bit-identical to the
    oracle's scalar tree (SHA-256 of all node views); at the reference's
    Lambda 1/12 also within 1e-12 of the reference's own kernel on every
    997th node and on the per-x-plane direction sums; mass conserved to
    1e-12 (src/BenchKernelD3Q19ListAaPv.c:83-142)."""
    g = golden("porous256_500.npz")
    geo = lbm.Geometry.from_string("porous-6-0.8-42", (256, 256, 256))
    mask = np.ascontiguousarray(geo.mask.reshape(-1))
    assert _sha(mask) == str(g["mask_sha"])
    tag, args = ("l12", []) if lam == "1/12" else ("l316", ["-lambda", "0.1875"])
    print(f"porous256 list-aa-pv-soa 500 steps Lambda {lam}", flush=True)
    with lbm.Kernel("list-aa-pv-soa", geo, args) as k:
        k.set_initial_weights()
        k.run(lbm.case_data(omega=float(g["omega"]), xforce=float(g["xforce"]), iterations=int(g["steps"])))
        d = k.dump()
    every = int(g["sample_every"])
    assert np.array_equal(d[::every], g[f"orc_{tag}_sample"])
    assert _sha(d) == str(g[f"orc_{tag}_sha"])
    m0 = float(g["mass0"])
    assert abs(d.sum() - m0) / m0 <= 1e-12
    if tag == "l12":
        assert _relmax(d[::every], g["ref_sample"]) <= REL_TOL
        xs = _x_plane_sums(d, mask, 256)
        assert np.abs(xs - g["ref_xsums"]).max() / np.abs(g["ref_xsums"]).max() <= REL_TOL
        assert abs(d.sum() - float(g["ref_mass"])) / m0 <= 1e-13
    # flow goes in +x under a +x force
    ux = d[:, [2, 4, 5, 10, 17]].sum(1) - d[:, [3, 6, 7, 11, 16]].sum(1)
    assert ux.mean() > 0


@pytest.mark.parametrize("name", ["list-aa-pv-soa", "aa-soa"])
def test_lambda_and_omega_variants(lbm, oracle, name):
    geo = lbm.Geometry.from_string("porous-2-0.7-5", (20, 18, 16))
    m = geo.mask.reshape(-1)
    for omega, lam in ((1.2, 3 / 16), (0.8, 0.25), (1.9, 1 / 6)):
        got, _, _ = _run(lbm, name, geo, 60, omega, 2e-5, ["-lambda", repr(lam)])
        want, _ = oracle.list_run(m, geo.dims, geo.periodic, 60, omega, 2e-5, magic=lam, vsize=1)
        assert np.array_equal(got, want), (omega, lam)


@pytest.mark.parametrize("name", ["list-aa-pv-soa", "aa-soa"])
def test_force_off_block_sizes_and_split_calls(lbm, oracle, name):
    geo = lbm.Geometry.from_string("channel", (14, 12, 10), (1, 0, 0))
    m = geo.mask.reshape(-1)
    want, _ = oracle.list_run(m, geo.dims, geo.periodic, 40, 1.1, 1e-4, vsize=1)
    for bs in ("64", "128", "512"):
        got, _, _ = _run(lbm, name, geo, 40, 1.1, 1e-4, ["-bs", bs] if name == "list-aa-pv-soa" else [])
        assert np.array_equal(got, want), bs
    # two Kernel() calls of 20 steps == one of 40 (force bookkeeping across calls);
    # 19 steps run as 20 (steps go in pairs)
    with lbm.Kernel(name, geo) as k:
        k.set_initial_weights()
        k.run(lbm.case_data(omega=1.1, xforce=1e-4, iterations=20))
        k.run(lbm.case_data(omega=1.1, xforce=1e-4, iterations=19))
        assert k.timings().steps == 20
        assert np.array_equal(k.dump(), want)
    got, _, _ = _run(lbm, name, geo, 40, 1.1, 1e-4, ["-force", "off"])
    want_off, _ = oracle.list_run(m, geo.dims, geo.periodic, 40, 1.1, 1e-4, vsize=1, force_on=False)
    assert np.array_equal(got, want_off)
    # per-launch timing events: off by default, on by argument or setter; same results
    got, _, t = _run(lbm, name, geo, 40, 1.1, 1e-4)
    assert t.loopMs > 0 and t.evenMsAvg == 0 and t.oddMsAvg == 0
    got, _, t = _run(lbm, name, geo, 40, 1.1, 1e-4, ["-launch-events", "on"])
    assert np.array_equal(got, want) and t.evenMsAvg > 0 and t.oddMsAvg > 0
    with lbm.Kernel(name, geo) as k:
        k.set_initial_weights()
        k.set_launch_events(True)
        k.run(lbm.case_data(omega=1.1, xforce=1e-4, iterations=40))
        assert k.timings().oddMsAvg > 0 and np.array_equal(k.dump(), want)


def test_edge_geometries(lbm, oracle):
    # all solid, one fluid cell, a 1-cell periodic line
    for mask, dims, per in [(np.zeros(27, np.int32), (3, 3, 3), (0, 0, 0)),
                            (np.eye(1, 27, 13, dtype=np.int32)[0], (3, 3, 3), (0, 0, 0)),
                            (np.ones(5, np.int32), (5, 1, 1), (1, 1, 1))]:
        geo = lbm.Geometry.from_mask(mask, dims, per)
        for name in ("list-aa-pv-soa", "aa-soa"):
            got, _, _ = _run(lbm, name, geo, 10, 1.0, 1e-3)
            want, _ = oracle.list_run(mask, dims, per, 10, 1.0, 1e-3, vsize=1)
            assert got.shape == want.shape
            assert np.array_equal(got, want), (name, dims)


def test_setnode_roundtrip_and_upload(lbm):
    geo = lbm.Geometry.from_string("box", (6, 5, 4), (1, 1, 0))
    rng = np.random.default_rng(0)
    for name in ("list-aa-pv-soa", "aa-soa"):
        with lbm.Kernel(name, geo) as k:
            v = rng.random((geo.n_fluid, 19))
            k.load(v)
            assert np.array_equal(k.dump(), v)
            assert geo.mask[2, 2, 1] != 0
            k.set_node(1, 2, 2, np.arange(19.0))
            assert np.array_equal(k.get_node(1, 2, 2), np.arange(19.0))
            # zero steps: data goes to the device and back unchanged
            k.run(lbm.case_data(iterations=0))
            assert np.array_equal(k.get_node(1, 2, 2), np.arange(19.0))
    with lbm.Kernel("list-aa-pv-soa", geo) as k:
        # (1, 2, 3) lies in the solid z = 3 wall: reported, not dereferenced
        k.set_node(1, 2, 3, np.arange(19.0))
        assert "is solid" in lbm.hip_lib().LbmHipLastError().decode()


@pytest.mark.parametrize("geom,dims", [("porous-3-0.75-9", (64, 48, 40)), ("blocks-3", (30, 24, 27))])
def test_adjacency_modes_and_streaming_hints(lbm, oracle, geom, dims):
    # run-compressed and full adjacency, with and without non-temporal
    # accesses, in either even-step block order, all give the oracle's
    # result bit for bit
    geo = lbm.Geometry.from_string(geom, dims)
    m = geo.mask.reshape(-1)
    want, _ = oracle.list_run(m, geo.dims, geo.periodic, 30, 1.25, 3e-5, vsize=1)
    for args in (["-adj", "rle"], ["-adj", "full"], ["-adj", "rle", "-nt", "off"], ["-adj", "full", "-nt", "off"],
                 ["-adj", "rle", "-bs", "128"], ["-adj", "rle", "-bs", "64", "-blk", "5"],
                 ["-adj", "rle", "-even-cap", "off"], ["-adj", "full", "-even-cap", "off", "-bs", "512"],
                 ["-adj", "rle", "-even-cap", "force"], ["-adj", "full", "-even-cap", "force", "-nt", "off"],
                 ["-adj", "rle", "-odd-order", "group"], ["-adj", "full", "-odd-order", "plain"],
                 ["-adj", "rle", "-odd-order", "band", "-bs", "64"], ["-adj", "rle", "-even-order", "down"],
                 ["-adj", "full", "-even-order", "down", "-nt", "on", "-even-cap", "force"]):
        with lbm.Kernel("list-aa-pv-soa", geo, args) as k:
            info = k.list_info()
            assert info.adjRle == (args[1] == "rle")
            assert info.nRunsWave >= info.nRuns
            k.set_initial_weights()
            k.run(lbm.case_data(omega=1.25, xforce=3e-5, iterations=30))
            got = k.dump()
        if "-blk" in args:
            w2, _ = oracle.list_run(m, geo.dims, geo.periodic, 30, 1.25, 3e-5, vsize=1, blk=(5, 5, 5))
            assert np.array_equal(got, w2), args
        else:
            assert np.array_equal(got, want), args


@pytest.mark.parametrize("args", [[], ["-nt", "on"], ["-even-cap", "force"], ["-nt", "on", "-even-cap", "force"],
                                  ["-even-cap", "off"]])
def test_aa_streaming_options(lbm, oracle, args):
    # the even step's launch shape (one 512-thread block over several y rows
    # per CU) and non-temporal accesses change no bit; odd Y and Z sizes
    geo = lbm.Geometry.from_string("porous-3-0.75-9", (40, 37, 70))
    m = geo.mask.reshape(-1)
    want, _ = oracle.list_run(m, geo.dims, geo.periodic, 30, 1.25, 3e-5, vsize=1)
    got, _, _ = _run(lbm, "aa-soa", geo, 30, 1.25, 3e-5, args)
    assert np.array_equal(got, want), args


@pytest.mark.parametrize("name", ["list-aa-pv-soa", "aa-soa"])
def test_large_lattice_default_launch(lbm, oracle, name):
    # above the size where the even step runs one 512-thread block per CU
    geo = lbm.Geometry.from_string("porous-4-0.85-5", (176, 128, 128))
    assert geo.n_fluid > 16 * 512 * 256
    m = geo.mask.reshape(-1)
    print(f"{name} 176x128x128 nFluid={geo.n_fluid}", flush=True)
    want, _ = oracle.list_run(m, geo.dims, geo.periodic, 10, 1.3, 2e-5, vsize=1)
    for args in ([], ["-even-cap", "off"]):
        got, _, _ = _run(lbm, name, geo, 10, 1.3, 2e-5, args)
        assert np.array_equal(got, want), args


@pytest.mark.parametrize("name", ["list-aa-pv-soa", "aa-soa"])
def test_index_range_limit(lbm, name):
    # 19 * cells must fit the 32-bit PDF indices (src/Main.c:399-406); the
    # check comes before the mask is read, so no mask is needed here
    ld = lbm.LatticeDesc()
    ld.Dims = (lbm.C.c_int * 3)(1000, 1000, 300)
    ld.nCells, ld.nFluid = 300_000_000, 300_000_000
    geo = lbm.Geometry.__new__(lbm.Geometry)
    geo.ld, geo._keep = ld, None
    with pytest.raises(RuntimeError, match="2\\^32"):
        lbm.Kernel(name, geo)


@pytest.mark.parametrize("dims,per", DEGENERATE_SHAPES)
def test_degenerate_shapes(lbm, oracle, dims, per):
    """Extents of one and two cells, open and periodic faces, all fluid, 20
    steps: both kernels bit-exact against the oracle (pinned to the
    reference on the same shapes) in FP64, and in shifted FP32 / FP16
    against the oracle's shifted restatements."""
    mask = np.ones(int(np.prod(dims)), dtype=np.int32)
    geo = lbm.Geometry.from_mask(mask, dims, per)
    wants = {("aa-soa", "fp64"): oracle.aa_run(mask, dims, per, 20, 1.3, 1e-5)[0],
             ("list-aa-pv-soa", "fp64"): oracle.list_run(mask, dims, per, 20, 1.3, 1e-5)[0]}
    for prec in ("fp32s", "fp16s"):
        wants[("aa-soa", prec)] = oracle.aa_run_shifted(mask, dims, per, 20, 1.3, 1e-5, prec)
        wants[("list-aa-pv-soa", prec)] = oracle.list_run_shifted(mask, dims, per, 20, 1.3, 1e-5, prec)[0]
    for (kernel, prec), want in wants.items():
        print(f"{kernel} {dims} {per} -prec {prec}", flush=True)
        with lbm.Kernel(kernel, geo, ["-prec", prec]) as k:
            k.set_initial_weights()
            k.run(lbm.case_data(omega=1.3, xforce=1e-5, iterations=20))
            got = k.dump()
        assert np.array_equal(got, want), (kernel, prec)


@pytest.mark.parametrize("name", ["list-aa-pv-soa", "aa-soa"])
def test_graph_replay_equals_plain_loop(lbm, name):
    """-graph on replays the captured step loop of a Kernel() call; every
    replay must equal the plain launch loop bit for bit, with host work
    (GetNode/SetNode through the host copy) between the calls and a new
    capture when omega or the step count change."""
    geo = lbm.Geometry.from_string("porous-3-0.7-5", (40, 36, 34))
    args = ["-lambda", "0.1875", "-quiet"]
    runs = [(20, 1.2), (20, 1.2), (20, 1.2), (30, 0.9), (30, 0.9), (20, 1.2)]
    print(f"graph replay {name}", flush=True)
    with lbm.Kernel(name, geo, args + ["-graph", "on"]) as kg, lbm.Kernel(name, geo, args) as kp:
        for k in (kg, kp):
            k.set_initial_weights()
        x, y, z = [int(v) for v in np.argwhere(geo.mask != 0)[5][::-1]]
        for i, (steps, omega) in enumerate(runs):
            for k in (kg, kp):
                k.run(lbm.case_data(omega=omega, xforce=2e-5, iterations=steps))
            got, want = kg.dump(), kp.dump()
            assert np.array_equal(got, want), i
            # host work between replays: nudge one node the same way in both
            for k in (kg, kp):
                v = k.get_node(x, y, z)
                k.set_node(x, y, z, [a * (1 + 1e-3) for a in v])
