"""LbmHipListBuildNoAlloc -- the list build into caller device buffers with
its scratch in a static device array -- against the list build of
D3Q19ListAaPvInit_PushSoAHip (LbmHipListCopyBuild), bit for bit, on two calls
in a row, the second on a smaller lattice (the scratch keeps the first
call's block counts).  Device buffers are allocated here by the test
through the HIP runtime; the entry point itself allocates nothing."""
import ctypes as C

import numpy as np
import pytest


class _Hip:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(1, nbytes)) == 0
        self.bufs.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0      # hipMemcpyHostToDevice
        return p

    def download(self, p, count, dtype=np.uint32):
        out = np.empty(count, dtype=dtype)
        if count:
            assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0   # hipMemcpyDeviceToHost
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


CASES = [  # geometry, dims, periodic, -blk words, blk triple
    ("porous-3-0.7-11", (40, 36, 34), (1, 0, 0), [], (0, 0, 0)),
    ("channel", (12, 10, 9), (1, 0, 0), ["-blk", "3"], (3, 3, 3)),
]


def _build_no_alloc(lbm, hip, geo, blk, n_cells, adj_stride, capacity):
    L = hip_lib = lbm.hip_lib()
    dims = geo.dims
    ncell = int(np.prod(dims))
    d_mask = hip.upload(np.asarray(geo.mask.reshape(-1), dtype=np.int32))
    d_grid = hip.alloc(4 * ncell)
    d_coords = hip.alloc(4 * 3 * capacity)
    d_adj = hip.alloc(4 * 18 * adj_stride)
    nf = C.c_int(-1)
    I3 = C.c_int * 3
    rc = hip_lib.LbmHipListBuildNoAlloc(d_mask, I3(*dims), I3(*geo.periodic), I3(*blk), n_cells, adj_stride,
                                        capacity, d_grid, d_coords, d_adj, C.byref(nf))
    assert rc == 0, L.LbmHipLastError().decode()
    n = nf.value
    grid = hip.download(d_grid, ncell)
    coords = hip.download(d_coords, 3 * n).reshape(n, 3)
    adj = hip.download(d_adj, 18 * adj_stride).reshape(18, adj_stride)[:, :n].T.copy()
    return n, adj, grid, coords


@pytest.mark.gpu
def test_list_build_no_alloc_matches_init(lbm):
    lbm.hip_lib().LbmHipSetErrorMode(lbm.LBM_HIP_ERRORS_RETURN)
    hip = _Hip()
    try:
        for gstr, dims, per, kargs, blk in CASES:        # the second call is the smaller lattice
            geo = lbm.Geometry.from_string(gstr, dims, per)
            print(f"no-alloc list build {gstr} {dims} blk={blk}", flush=True)
            with lbm.Kernel("list-aa-pv-soa", geo, ["-quiet", *kargs]) as k:
                info = k.list_info()
                want_adj, want_grid, want_coords = k.list_build()
            n, adj, grid, coords = _build_no_alloc(lbm, hip, geo, blk, info.nCells, info.adjStride, info.nFluid)
            assert n == info.nFluid == geo.n_fluid
            assert np.array_equal(grid, want_grid)
            assert np.array_equal(coords, want_coords.reshape(n, 3))
            assert np.array_equal(adj, want_adj.reshape(n, 18))
    finally:
        hip.free()


@pytest.mark.gpu
def test_list_build_no_alloc_rejects_what_does_not_fit(lbm):
    L = lbm.hip_lib()
    L.LbmHipSetErrorMode(lbm.LBM_HIP_ERRORS_RETURN)
    hip = _Hip()
    I3 = C.c_int * 3
    try:
        d = hip.alloc(64)
        nf = C.c_int(-1)
        # more cells than the 1 MiB scratch covers: refused before any launch
        L.LbmHipClearError()
        rc = L.LbmHipListBuildNoAlloc(d, I3(512, 512, 300), I3(0, 0, 0), I3(0, 0, 0), 1 << 20, 1 << 20, 1 << 20,
                                      d, d, d, C.byref(nf))
        assert rc == -1 and nf.value == 0
        assert "scratch limit" in L.LbmHipLastError().decode()
        # more fluid cells than the caller has room for: counted, nothing written
        geo = lbm.Geometry.from_string("box", (8, 8, 8), (1, 1, 0))
        d_mask = hip.upload(np.asarray(geo.mask.reshape(-1), dtype=np.int32))
        d_grid = hip.alloc(4 * 512)
        grid0 = hip.download(d_grid, 512)
        L.LbmHipClearError()
        rc = L.LbmHipListBuildNoAlloc(d_mask, I3(8, 8, 8), I3(1, 1, 0), None, 1024, 1024, geo.n_fluid - 1,
                                      d_grid, d, d, C.byref(nf))
        assert rc == -1 and nf.value == 0
        assert "room for" in L.LbmHipLastError().decode()
        assert np.array_equal(hip.download(d_grid, 512), grid0)
    finally:
        L.LbmHipClearError()
        hip.free()
