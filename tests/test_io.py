"""Output counterparts (VTK snapshots, statistics files) against the
reference's own functions.

CPU: liblbmhost.so's VtkWrite / KernelStatistics and the reference's
(src/Vtk.c, src/Kernel.c, compiled into oracle/_ref/libref_io.so) are run on
the same node views, served through one KernelData's GetNode, and must write
identical bytes.  The VTK file is also decoded and checked against the node
views directly.
"""
import ctypes as C
import os
import struct

import numpy as np
import pytest

from conftest import ORACLE


def _ref_io():
    path = os.path.join(ORACLE, "_ref", "libref_io.so")
    if not os.path.exists(path):
        pytest.skip("oracle/_ref/libref_io.so not built (needs the reference sources)")
    L = C.CDLL(path, mode=os.RTLD_DEEPBIND)
    return L


def node_view_kd(lbm, geo, nodes):
    """A KernelData whose GetNode serves `nodes` (mask order) -- for calling
    output functions on a given state."""
    dims = geo.dims
    m = geo.mask.reshape(dims[2], dims[1], dims[0])
    idx = np.full(m.shape, -1, dtype=np.int64)
    idx[m != 0] = np.arange(int(np.count_nonzero(m)))

    def get(kd, x, y, z, out):
        k = idx[z, y, x]
        for d in range(19):
            out[d] = float(nodes[k, d])

    kd = lbm.KernelData()
    cb = lbm._Node(get)
    kd.GetNode = cb
    for i in range(3):
        kd.Dims[i] = kd.GlobalDims[i] = dims[i]
    return kd, cb


def write_outputs(fn_vtk, fn_stats, lbm, geo, kd, cd, calls, where):
    """Runs the (kind, iteration) calls in directory `where`; returns
    {file name: bytes}."""
    old = os.getcwd()
    os.makedirs(where, exist_ok=True)
    os.chdir(where)
    try:
        for kind, it in calls:
            if kind == "vtk":
                fn_vtk(C.byref(geo.ld), C.byref(kd), C.byref(cd), it)
            else:
                fn_stats(C.byref(kd), C.byref(geo.ld), C.byref(cd), it)
    finally:
        os.chdir(old)
    out = {}
    for root, _, files in os.walk(where):
        for f in files:
            p = os.path.join(root, f)
            out[os.path.relpath(p, where)] = open(p, "rb").read()
    return out


def _state(lbm, oracle, geom, dims, per, steps, F=2e-5):
    geo = lbm.Geometry.from_string(geom, dims, per)
    nodes, _ = oracle.list_run(geo.mask.reshape(-1), geo.dims, geo.periodic, steps, 1.1, F, vsize=1)
    return geo, nodes


@pytest.mark.parametrize("geom,dims,per", [("box", (8, 6, 5), (1, 1, 0)), ("porous-2-0.6-3", (16, 14, 14), (1, 0, 0)),
                                          ("channel", (7, 8, 6), (1, 0, 0))])
def test_output_files_match_reference(lbm, oracle, tmp_path, geom, dims, per):
    ref = _ref_io()
    for fn in (ref.VtkWrite, lbm.host_lib().VtkWrite):
        fn.argtypes = [C.c_void_p] * 3 + [C.c_int]
    for fn in (ref.KernelStatistics, lbm.host_lib().KernelStatistics):
        fn.argtypes = [C.c_void_p] * 3 + [C.c_int]
    geo, nodes = _state(lbm, oracle, geom, dims, per, 30)
    kd, keep = node_view_kd(lbm, geo, nodes)
    cd = lbm.case_data(omega=1.1, xforce=2e-5, iterations=30)
    cd.StatisticsModulus = 4
    calls = [("vtk", -1), ("stats", 0), ("vtk", 0), ("stats", 0), ("stats", 4), ("stats", 10), ("vtk", 20),
             ("stats", 20), ("stats", 30), ("vtk", 30)]
    want = write_outputs(ref.VtkWrite, ref.KernelStatistics, lbm, geo, kd, cd, calls, str(tmp_path / "ref"))
    got = write_outputs(lbm.host_lib().VtkWrite, lbm.host_lib().KernelStatistics, lbm, geo, kd, cd, calls,
                        str(tmp_path / "mine"))
    assert sorted(want) == sorted(got) == sorted(["density-ux.dat", "ux-progress.dat", "vtk/file--001.vtk",
                                                  "vtk/file-0000.vtk", "vtk/file-0020.vtk", "vtk/file-0030.vtk"])
    for name in want:
        assert got[name] == want[name], name
    del keep


def test_vtk_content_decodes_to_node_views(lbm, oracle, tmp_path):
    geo, nodes = _state(lbm, oracle, "porous-2-0.6-3", (16, 14, 14), (1, 0, 0), 20)
    kd, keep = node_view_kd(lbm, geo, nodes)
    F = 2e-5
    cd = lbm.case_data(omega=1.1, xforce=F, iterations=20)
    got = write_outputs(lbm.host_lib().VtkWrite, None, lbm, geo, kd, cd, [("vtk", 20)], str(tmp_path))
    raw = got["vtk/file-0020.vtk"]
    nx, ny, nz = geo.dims
    n = nx * ny * nz
    head = raw.index(b"LOOKUP_TABLE default\n") + len(b"LOOKUP_TABLE default\n")
    assert raw[:head].startswith(b"# vtk DataFile Version 1.0\nComment: lid driven cavity, iteration   20\nBINARY\n")
    types = np.frombuffer(raw[head:head + n], dtype=np.uint8)
    assert np.array_equal(types, geo.mask.reshape(-1).astype(np.uint8))
    p = raw.index(b"LOOKUP_TABLE default\n", head + n) + len(b"LOOKUP_TABLE default\n")
    rho = np.frombuffer(raw[p:p + 8 * n], dtype=">f8")
    q = raw.index(b"VECTORS VelocityVectors double\n", p + 8 * n) + len(b"VECTORS VelocityVectors double\n")
    vel = np.frombuffer(raw[q:q + 24 * n], dtype=">f8").reshape(n, 3)
    assert q + 24 * n == len(raw)
    fl = geo.mask.reshape(-1) != 0
    rho_seq = np.array([sum(float(v) for v in row) for row in nodes])       # direction order
    assert np.array_equal(rho[fl], rho_seq) and (rho[~fl] == 0).all()
    E, W = [2, 4, 5, 10, 17], [3, 6, 7, 11, 16]
    ux = []
    for row in nodes:                        # E + NE + SE + TE + BE - W - NW - SW - TW - BW
        u = float(row[E[0]])
        for d in E[1:]:
            u += float(row[d])
        for d in W:
            u -= float(row[d])
        ux.append(u + 0.5 * F)
    assert np.array_equal(vel[fl, 0], np.array(ux))
    assert (vel[~fl] == 0).all()
    del keep


@pytest.mark.parametrize("dims", [(8, 6, 10), (6, 7, 9)])
def test_flow_profile_matches_reference(lbm, oracle, tmp_path, dims):
    """KernelVerifiy (src/Kernel.c:511-625): flow-profile.dat byte for byte
    and the same deviation, on the -verify geometry (x and y periodic, walls
    at z) with an even and an odd nY."""
    ref = _ref_io()
    fns = (ref.KernelVerifiy, lbm.host_lib().KernelVerifiy)
    for fn in fns:
        fn.argtypes = [C.c_void_p] * 4
        fn.restype = None
    F = 2e-5
    geo, nodes = _state(lbm, oracle, "box", dims, (1, 1, 0), 40, F=F)
    kd, keep = node_view_kd(lbm, geo, nodes)
    cd = lbm.case_data(omega=1.1, xforce=F, iterations=40)
    files, norms = [], []
    old = os.getcwd()
    for i, fn in enumerate(fns):
        where = tmp_path / str(i)
        where.mkdir()
        os.chdir(where)
        try:
            norm = C.c_double(-1.0)
            fn(C.byref(geo.ld), C.byref(kd), C.byref(cd), C.byref(norm))
        finally:
            os.chdir(old)
        files.append((where / "flow-profile.dat").read_bytes())
        norms.append(norm.value)
    assert files[0].count(b"\n") == dims[2] + 4
    assert files[1] == files[0]
    assert norms[1] == norms[0] > 0
    del keep
