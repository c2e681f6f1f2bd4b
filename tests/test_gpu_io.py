"""GPU: the list kernel's run-time output (VTK snapshots and statistics files
at the reference's points of a VTK_OUTPUT / STATISTICS build) equals, byte
for byte, the same output functions replayed on the CPU over the oracle's
states at those points.  (tests/test_io.py shows the functions themselves
write what the reference's write.)"""
import ctypes as C
import os

import numpy as np
import pytest

from test_io import node_view_kd, write_outputs

pytestmark = pytest.mark.gpu


def _in_dir(path, fn):
    old = os.getcwd()
    os.makedirs(path, exist_ok=True)
    os.chdir(path)
    try:
        return fn()
    finally:
        os.chdir(old)


def _files(where):
    out = {}
    for root, _, files in os.walk(where):
        for f in files:
            p = os.path.join(root, f)
            out[os.path.relpath(p, where)] = open(p, "rb").read()
    return out


def test_run_time_output_matches_replay(lbm, oracle, tmp_path):
    geo = lbm.Geometry.from_string("porous-2-0.6-3", (16, 14, 14), (1, 0, 0))
    steps, vmod, smod, omega, F = 24, 4, 6, 1.1, 2e-5
    H = lbm.host_lib()

    def gpu():
        with lbm.Kernel("list-aa-pv-soa", geo, ["-statistics"]) as k:
            k.set_initial_weights()
            cd = lbm.case_data(omega=omega, xforce=F, iterations=steps)
            cd.VtkOutput, cd.VtkModulus, cd.StatisticsModulus = 1, vmod, smod
            k.run(cd)
    _in_dir(str(tmp_path / "gpu"), gpu)
    got = _files(str(tmp_path / "gpu"))

    # replay: the reference's call sequence over the oracle's states
    m = geo.mask.reshape(-1)
    state = {}

    def at(s):
        if s not in state:
            state[s] = oracle.list_run(m, geo.dims, geo.periodic, s, omega, F, vsize=1)[0]
        return state[s]
    calls = [(0, "vtk", -1), (0, "stats", 0)]
    for it in range(0, steps, 2):
        if it % vmod == 0:
            calls.append((it + 2, "vtk", it))
        if it % 10 == 0 or it % smod == 0:
            calls.append((it + 2, "stats", it))
    calls += [(steps, "vtk", steps), (steps, "stats", steps)]
    cd = lbm.case_data(omega=omega, xforce=F, iterations=steps)
    cd.VtkOutput, cd.VtkModulus, cd.StatisticsModulus = 1, vmod, smod
    for fn in (H.VtkWrite, H.KernelStatistics):
        fn.argtypes = [C.c_void_p] * 3 + [C.c_int]
    where = str(tmp_path / "replay")
    for s, kind, it in calls:
        kd, keep = node_view_kd(lbm, geo, at(s))
        write_outputs(H.VtkWrite, H.KernelStatistics, lbm, geo, kd, cd, [(kind, it)], where)
        del keep
    want = _files(where)
    assert sorted(got) == sorted(want)
    assert "vtk/file-0012.vtk" in got and "vtk/file-0024.vtk" in got
    for name in want:
        assert got[name] == want[name], name


@pytest.mark.parametrize("geom,dims,per,steps,vmod,smod", [
    ("channel", (16, 12, 10), (1, 0, 0), 24, 4, 6),
    ("blocks-2", (14, 12, 10), (0, 0, 0), 12, 3, 4),      # open x faces: the ghost layer in the views
    ("box", (10, 8, 8), (1, 1, 0), 6, 1, 3),              # VtkModulus 1: the reference's odd-step snapshot at 1
])
def test_aa_run_time_output_matches_reference(lbm, tmp_path, geom, dims, per, steps, vmod, smod):
    """aa-soa writes the files of the reference's VTK_OUTPUT / STATISTICS build
    (VTK snapshots, ux-progress.dat, density-ux.dat) byte for byte: the
    reference's own aa-soa kernel, compiled with both into
    oracle/_ref/libref_io.so, runs the same case on the CPU."""
    from lbm_checkers import Reference
    if not Reference.available("io"):
        pytest.skip("oracle/_ref/libref_io.so not built (needs the reference sources)")
    ref = Reference("io")
    geo = lbm.Geometry.from_string(geom, dims, per)
    m = geo.mask.reshape(-1)
    omega, F = 1.1, 2e-5
    print(f"aa run-time output {geom} {dims} steps={steps} vtk every {vmod} statistics every {smod}", flush=True)

    def reference_run():
        ref.lib.ref_set_output(1, vmod, smod)
        try:
            want_nodes, _ = ref.aa_run(m, geo.dims, geo.periodic, steps, omega, F, init="weights")
        finally:
            ref.lib.ref_set_output(0, 1, 1)
        return want_nodes
    want_nodes = _in_dir(str(tmp_path / "ref"), reference_run)

    def gpu():
        with lbm.Kernel("aa-soa", geo, ["-statistics"]) as k:
            k.set_initial_weights()
            cd = lbm.case_data(omega=omega, xforce=F, iterations=steps)
            cd.VtkOutput, cd.VtkModulus, cd.StatisticsModulus = 1, vmod, smod
            k.run(cd)
            return k.dump()
    got_nodes = _in_dir(str(tmp_path / "gpu"), gpu)
    assert np.array_equal(got_nodes, want_nodes)
    got, want = _files(str(tmp_path / "gpu")), _files(str(tmp_path / "ref"))
    assert sorted(got) == sorted(want)
    assert "vtk/file--001.vtk" in got and f"vtk/file-{steps:04d}.vtk" in got
    if vmod == 1:
        assert "vtk/file-0001.vtk" in got
    for name in want:
        assert got[name] == want[name], name


def _vtk_fields(raw, n):
    head = raw.index(b"LOOKUP_TABLE default\n") + len(b"LOOKUP_TABLE default\n")
    p = raw.index(b"LOOKUP_TABLE default\n", head + n) + len(b"LOOKUP_TABLE default\n")
    rho = np.frombuffer(raw[p:p + 8 * n], dtype=">f8")
    q = raw.index(b"VECTORS VelocityVectors double\n", p + 8 * n) + len(b"VECTORS VelocityVectors double\n")
    return rho, np.frombuffer(raw[q:q + 24 * n], dtype=">f8").reshape(n, 3)


@pytest.mark.parametrize("name", ["aa-soa", "list-aa-pv-soa"])
def test_shifted_storage_run_time_output(lbm, tmp_path, name):
    """Run-time output with shifted FP32 storage writes the same files as
    FP64, with density and velocity within FP32 accuracy of the FP64 run."""
    geo = lbm.Geometry.from_string("channel", (16, 12, 10), (1, 0, 0))
    n = int(np.prod(geo.dims))
    steps, vmod = 12, 4

    def run(prec):
        def go():
            with lbm.Kernel(name, geo, ["-statistics", "-prec", prec]) as k:
                k.set_initial_weights()
                cd = lbm.case_data(omega=1.1, xforce=2e-5, iterations=steps)
                cd.VtkOutput, cd.VtkModulus, cd.StatisticsModulus = 1, vmod, 4
                k.run(cd)
        _in_dir(str(tmp_path / prec), go)
        return _files(str(tmp_path / prec))
    f64, f32 = run("fp64"), run("fp32s")
    assert sorted(f64) == sorted(f32) and any(k.startswith("vtk/") for k in f64)
    for name_ in f64:
        if not name_.startswith("vtk/"):
            continue
        r64, u64 = _vtk_fields(f64[name_], n)
        r32, u32 = _vtk_fields(f32[name_], n)
        assert np.abs(r32 - r64).max() < 1e-6, name_
        assert np.abs(u32 - u64).max() < 1e-7, name_
