"""ctypes access to the checker libraries built by oracle/Makefile.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by the smoke check in
__graft_entry__.py and by the cpu_baseline leg of bench.py -- never by the
product package.

* ``Oracle`` wraps ``oracle/liblbmoracle.so``: the C restatement of the
  reference's list build, list-AA and aa kernels, forcing, initial state and
  verification norm (``oracle/lbm_oracle.c``).
* ``Reference`` wraps ``oracle/_ref/libref_verif.so`` / ``libref_bench.so``:
  the reference's own sources compiled unmodified plus the glue in
  ``oracle/ref_harness_*.c``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NDIR = 19
NLINK = 18

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


def _i3(v):
    return (C.c_int * 3)(*[int(x) for x in v])


class _ListT(C.Structure):
    _fields_ = [("dims", C.c_int * 3), ("nFluid", C.c_int), ("nCells", C.c_int),
                ("mask", C.c_void_p), ("grid", C.c_void_p), ("adj", C.c_void_p),
                ("pdf", C.c_void_p), ("iteration", C.c_int)]


class _AaT(C.Structure):
    _fields_ = [("dims", C.c_int * 3), ("gdims", C.c_int * 3), ("periodic", C.c_int * 3),
                ("mask", C.c_void_p), ("pdf", C.POINTER(C.c_double)),
                ("bbSrc", C.c_void_p), ("bbDst", C.c_void_p), ("nBB", C.c_int),
                ("iteration", C.c_int)]


class _PadT(C.Structure):
    _fields_ = [("nEntries", C.c_int), ("modulus", C.c_int * 16), ("offset", C.c_int * 16)]


class _RefResult(C.Structure):
    _fields_ = [("nFluid", C.c_int), ("nCells", C.c_int), ("nBounceBack", C.c_int),
                ("nConsecNodes", C.c_int), ("duration", C.c_double),
                ("verifyNorm", C.c_double), ("density", C.c_double)]


@dataclass
class ListBuild:
    n_fluid: int
    n_cells: int
    grid: np.ndarray      # uint32 per cell, 0xffffffff for obstacles
    coords: np.ndarray    # uint32 (nFluid, 3)
    adj: np.ndarray       # uint32 (nFluid, 18), node-major


class Oracle:
    """The C restatement (oracle/lbm_oracle.c)."""

    def __init__(self, path: str | None = None):
        path = path or os.path.join(HERE, "liblbmoracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run make -C oracle)")
        L = self.lib = C.CDLL(path)
        L.orc_pad_parse.argtypes = [C.c_char_p, C.POINTER(_PadT)]
        L.orc_pad_cells.argtypes = [C.c_int, C.c_int, C.POINTER(_PadT)]
        L.orc_geometry.argtypes = [C.c_int, C.c_int, C.c_int * 3, C.c_int * 3, _i32p, C.c_int * 3]
        L.orc_list_number.argtypes = [_i32p, C.c_int * 3, C.c_int * 3, _u32p, _u32p]
        L.orc_list_adjacency.argtypes = [_i32p, C.c_int * 3, C.c_int * 3, _u32p, _u32p,
                                         C.c_int, C.c_int, _u32p]
        L.orc_list_runs.argtypes = [_u32p, C.c_int, C.c_int, C.c_void_p]
        L.orc_trt_collide.argtypes = [_f64p, _f64p, C.c_double, C.c_double, C.c_int]
        L.orc_omega_odd.argtypes = [C.c_double, C.c_double]
        L.orc_omega_odd.restype = C.c_double
        L.orc_list_run.argtypes = [C.POINTER(_ListT), C.c_int, C.c_double, C.c_double,
                                   C.c_double, C.c_int, C.c_int, C.c_int]
        L.orc_list_dump.argtypes = [C.POINTER(_ListT), _f64p]
        L.orc_list_set_node.argtypes = [C.POINTER(_ListT), C.c_int, C.c_int, C.c_int, _f64p]
        L.orc_aa_init.argtypes = [C.POINTER(_AaT), _i32p, C.c_int * 3, C.c_int * 3]
        L.orc_aa_free.argtypes = [C.POINTER(_AaT)]
        L.orc_aa_run.argtypes = [C.POINTER(_AaT), C.c_int, C.c_double, C.c_double,
                                 C.c_double, C.c_int]
        L.orc_aa_dump.argtypes = [C.POINTER(_AaT), _f64p]
        L.orc_aa_set_node.argtypes = [C.POINTER(_AaT), C.c_int, C.c_int, C.c_int, _f64p]
        L.orc_initial_density_node.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int,
                                               C.c_int, _f64p]
        L.orc_initial_velocity_node.argtypes = [C.c_double, _f64p]
        L.orc_force_term.argtypes = [C.c_int, C.c_double]
        L.orc_force_term.restype = C.c_double
        L.orc_verify_norm.argtypes = [_f64p, _i32p, C.c_int, C.c_int, C.c_double, C.c_double]
        L.orc_verify_norm.restype = C.c_double
        L.orc_f16_round.argtypes = [C.c_float]
        L.orc_f16_round.restype = C.c_float
        L.orc_list_run_shifted.argtypes = [C.POINTER(_ListT), _f32p, C.c_int, C.c_int, C.c_int,
                                           C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_shifted_from_nodes.argtypes = [C.POINTER(_ListT), _f32p, C.c_int, C.c_int, _f64p]
        L.orc_shifted_to_nodes.argtypes = [C.POINTER(_ListT), _f32p, C.c_int, C.c_int, _f64p]
        L.orc_aa_run_shifted.argtypes = [C.POINTER(_AaT), _f32p, C.c_int, C.c_int, C.c_int,
                                         C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_aa_run_shifted.restype = C.c_int
        L.orc_aa_shifted_from_nodes.argtypes = [C.POINTER(_AaT), _f32p, C.c_int, C.c_int, _f64p]
        L.orc_aa_shifted_to_nodes.argtypes = [C.POINTER(_AaT), _f32p, C.c_int, C.c_int, _f64p]
        self.CX = np.ctypeslib.as_array((C.c_int * NDIR).in_dll(L, "ORC_CX"))
        self.CY = np.ctypeslib.as_array((C.c_int * NDIR).in_dll(L, "ORC_CY"))
        self.CZ = np.ctypeslib.as_array((C.c_int * NDIR).in_dll(L, "ORC_CZ"))
        self.INV = np.ctypeslib.as_array((C.c_int * NDIR).in_dll(L, "ORC_INV"))
        self.W = np.ctypeslib.as_array((C.c_double * NDIR).in_dll(L, "ORC_W"))

    # ---- padding ---------------------------------------------------------
    def pad_parse(self, s: str):
        p = _PadT()
        rc = self.lib.orc_pad_parse(s.encode(), C.byref(p))
        return rc, [(p.modulus[i], p.offset[i]) for i in range(p.nEntries)]

    def pad_cells(self, n: int, s: str | None = None, cell_bytes: int = 8) -> int:
        if s is None:
            return self.lib.orc_pad_cells(n, cell_bytes, None)
        p = _PadT()
        rc = self.lib.orc_pad_parse(s.encode(), C.byref(p))
        if rc != 0:
            raise ValueError(f"bad padding string {s!r} ({rc})")
        return self.lib.orc_pad_cells(n, cell_bytes, C.byref(p))

    # ---- geometry ----------------------------------------------------------
    def geometry(self, gtype: int, dims, periodic=(0, 0, 0), block: int = 8):
        mask = np.zeros(int(np.prod(dims)), dtype=np.int32)
        pf = (C.c_int * 3)()
        self.lib.orc_geometry(gtype, block, _i3(dims), _i3(periodic), mask, pf)
        return mask, tuple(pf)

    # ---- list build --------------------------------------------------------
    def list_build(self, mask, dims, periodic, blk=(0, 0, 0), n_cells=None) -> ListBuild:
        mask = np.ascontiguousarray(mask, dtype=np.int32)
        ncell = int(np.prod(dims))
        grid = np.zeros(ncell, dtype=np.uint32)
        nf_max = int(np.count_nonzero(mask))
        coords = np.zeros(max(nf_max, 1) * 3, dtype=np.uint32)
        nf = self.lib.orc_list_number(mask, _i3(dims), _i3(blk), grid, coords)
        if n_cells is None:
            n_cells = self.pad_cells(nf)
        adj = np.zeros(max(nf, 1) * NLINK, dtype=np.uint32)
        self.lib.orc_list_adjacency(mask, _i3(dims), _i3(periodic), grid, coords, nf, n_cells, adj)
        return ListBuild(nf, n_cells, grid, coords[:3 * nf].reshape(nf, 3), adj[:NLINK * nf].reshape(nf, NLINK))

    def list_runs(self, adj, n_fluid, n_threads=1):
        adj = np.ascontiguousarray(adj, dtype=np.uint32).reshape(-1)
        return self.lib.orc_list_runs(adj, n_fluid, n_threads, None)

    def collide(self, pre, omega, magic=1.0 / 12.0, tree=0):
        pre = np.ascontiguousarray(pre, dtype=np.float64)
        post = np.zeros(NDIR)
        self.lib.orc_trt_collide(pre, post, omega, self.lib.orc_omega_odd(omega, magic), tree)
        return post

    def omega_odd(self, omega, magic=1.0 / 12.0):
        return self.lib.orc_omega_odd(omega, magic)

    def force_term(self, d, F):
        return self.lib.orc_force_term(d, F)

    # ---- full runs ---------------------------------------------------------
    def init_values(self, mode: str, omega: float, rho: float = 1.0, x: int = 0, nx: int = 2):
        v = np.zeros(NDIR)
        if mode == "weights":
            return self.W * rho
        if mode == "density":
            self.lib.orc_initial_density_node(1.0, 1.0, omega, x, nx, v)
            return v
        if mode == "velocity":
            self.lib.orc_initial_velocity_node(omega, v)
            return v
        raise ValueError(mode)

    def list_run(self, mask, dims, periodic, iterations, omega, xforce, *, magic=1.0 / 12.0,
                 force_on=True, blk=(0, 0, 0), n_threads=1, vsize=1, init="weights", rho=1.0,
                 build: ListBuild | None = None):
        """Returns (dump[nFluid,19] in mask order, ListBuild)."""
        mask = np.ascontiguousarray(mask, dtype=np.int32)
        b = build or self.list_build(mask, dims, periodic, blk)
        pdf = np.zeros(NDIR * b.n_cells, dtype=np.float64)
        adj = np.ascontiguousarray(b.adj.reshape(-1))
        st = _ListT()
        st.dims = _i3(dims)
        st.nFluid, st.nCells = b.n_fluid, b.n_cells
        st.mask = mask.ctypes.data
        st.grid = b.grid.ctypes.data
        st.adj = adj.ctypes.data
        st.pdf = pdf.ctypes.data
        st.iteration = -1
        self._init_state(st, mask, dims, init, omega, rho, self.lib.orc_list_set_node)
        self.lib.orc_list_run(C.byref(st), iterations, omega, magic, xforce, int(force_on),
                              n_threads, vsize)
        out = np.zeros((b.n_fluid, NDIR))
        self.lib.orc_list_dump(C.byref(st), out)
        return out, b

    PREC = {"fp32s": 1, "fp16s": 2}

    def f16_round(self, v: float) -> float:
        return float(self.lib.orc_f16_round(v))

    def list_run_shifted(self, mask, dims, periodic, runs, omega, xforce, prec, *, f16_shift=12,
                         magic=1.0 / 12.0, force_on=True, blk=(0, 0, 0), nodes=None,
                         build: ListBuild | None = None):
        """Shifted-storage list-AA (prec "fp32s" or "fp16s").  ``runs`` is the
        step count of each Kernel() call, run back to back; ``nodes`` the
        initial node views in mask order (default f = w).  Returns the node
        views after the last run (nFluid x 19, mask order) and the build."""
        mask = np.ascontiguousarray(mask, dtype=np.int32)
        b = build or self.list_build(mask, dims, periodic, blk)
        code = self.PREC[prec]
        store = np.zeros(NDIR * b.n_cells, dtype=np.float32)
        adj = np.ascontiguousarray(b.adj.reshape(-1))
        st = _ListT()
        st.dims = _i3(dims)
        st.nFluid, st.nCells = b.n_fluid, b.n_cells
        st.mask = mask.ctypes.data
        st.grid = b.grid.ctypes.data
        st.adj = adj.ctypes.data
        st.iteration = -1
        if nodes is None:
            nodes = np.tile(self.W, (b.n_fluid, 1))
        self.lib.orc_shifted_from_nodes(C.byref(st), store, code, f16_shift,
                                        np.ascontiguousarray(nodes, dtype=np.float64))
        for steps in ([runs] if isinstance(runs, int) else runs):
            self.lib.orc_list_run_shifted(C.byref(st), store, code, f16_shift, int(steps), omega,
                                          magic, xforce, int(force_on))
        out = np.zeros((b.n_fluid, NDIR))
        self.lib.orc_shifted_to_nodes(C.byref(st), store, code, f16_shift, out)
        return out, b

    def list_run_seconds(self, mask, dims, periodic, iterations, omega, xforce, *, magic=1.0 / 12.0,
                         force_on=False):
        """Seconds of orc_list_run alone (scalar tree, OpenMP) from all-zero
        PDFs -- the same work per step as any state; the CPU timing of the
        restatement."""
        import time
        mask = np.ascontiguousarray(mask, dtype=np.int32)
        b = self.list_build(mask, dims, periodic)
        pdf = np.zeros(NDIR * b.n_cells, dtype=np.float64)
        adj = np.ascontiguousarray(b.adj.reshape(-1))
        st = _ListT()
        st.dims = _i3(dims)
        st.nFluid, st.nCells = b.n_fluid, b.n_cells
        st.mask = mask.ctypes.data
        st.grid = b.grid.ctypes.data
        st.adj = adj.ctypes.data
        st.pdf = pdf.ctypes.data
        st.iteration = -1
        t0 = time.perf_counter()
        self.lib.orc_list_run(C.byref(st), iterations, omega, magic, xforce, int(force_on), 1, 1)
        return time.perf_counter() - t0

    def aa_run(self, mask, dims, periodic, iterations, omega, xforce, *, magic=1.0 / 12.0,
               force_on=True, init="weights", rho=1.0):
        mask = np.ascontiguousarray(mask, dtype=np.int32)
        st = _AaT()
        if self.lib.orc_aa_init(C.byref(st), mask, _i3(dims), _i3(periodic)) != 0:
            raise MemoryError("orc_aa_init")
        try:
            self._init_state(st, mask, dims, init, omega, rho, self.lib.orc_aa_set_node)
            self.lib.orc_aa_run(C.byref(st), iterations, omega, magic, xforce, int(force_on))
            out = np.zeros((int(np.count_nonzero(mask)), NDIR))
            self.lib.orc_aa_dump(C.byref(st), out)
            return out, st.nBB
        finally:
            self.lib.orc_aa_free(C.byref(st))

    def aa_run_shifted(self, mask, dims, periodic, runs, omega, xforce, prec, *, f16_shift=12,
                       magic=1.0 / 12.0, force_on=True, nodes=None):
        """Shifted-storage aa-soa (prec "fp32s" or "fp16s"), runs and initial
        node views as in list_run_shifted.  Returns the node views after the
        last run (nFluid x 19, mask order)."""
        mask = np.ascontiguousarray(mask, dtype=np.int32)
        nf = int(np.count_nonzero(mask))
        code = self.PREC[prec]
        st = _AaT()
        if self.lib.orc_aa_init(C.byref(st), mask, _i3(dims), _i3(periodic)) != 0:
            raise MemoryError("orc_aa_init")
        try:
            store = np.zeros(NDIR * int(np.prod(np.asarray(dims) + 2)), dtype=np.float32)
            if nodes is None:
                nodes = np.tile(self.W, (nf, 1))
            self.lib.orc_aa_shifted_from_nodes(C.byref(st), store, code, f16_shift,
                                               np.ascontiguousarray(nodes, dtype=np.float64))
            for steps in ([runs] if isinstance(runs, int) else runs):
                if self.lib.orc_aa_run_shifted(C.byref(st), store, code, f16_shift, int(steps), omega,
                                               magic, xforce, int(force_on)) != 0:
                    raise MemoryError("orc_aa_run_shifted")
            out = np.zeros((nf, NDIR))
            self.lib.orc_aa_shifted_to_nodes(C.byref(st), store, code, f16_shift, out)
            return out
        finally:
            self.lib.orc_aa_free(C.byref(st))

    def _init_state(self, st, mask, dims, init, omega, rho, setter):
        if init == "none":
            return
        nx, ny, nz = dims
        m = mask.reshape(nz, ny, nx)
        zs, ys, xs = np.nonzero(m)
        if init == "weights":
            v = np.ascontiguousarray(self.W * rho)
            for z, y, x in zip(zs, ys, xs):
                setter(C.byref(st), int(x), int(y), int(z), v)
        elif init == "verify":
            # KernelSetInitialDensity on every non-solid cell, then
            # KernelSetInitialVelocity on FLUID cells (src/Main.c:562-567)
            for z, y, x in zip(zs, ys, xs):
                v = self.init_values("density", omega, x=int(x), nx=nx)
                setter(C.byref(st), int(x), int(y), int(z), v)
            v = self.init_values("velocity", omega)
            for z, y, x in zip(zs, ys, xs):
                if m[z, y, x] == 1:
                    setter(C.byref(st), int(x), int(y), int(z), v)
        else:
            raise ValueError(init)

    def verify_norm(self, dump, mask, dims, omega, xforce):
        """KernelVerifiy's deviation from a node-view dump (mask order)."""
        nx, ny, nz = dims
        m = np.asarray(mask).reshape(nz, ny, nx)
        idx = np.full(m.shape, -1, dtype=np.int64)
        idx[m != 0] = np.arange(int(np.count_nonzero(m)))
        x = nx // 2
        ux = np.zeros(ny * nz)
        fl = np.zeros(ny * nz, dtype=np.int32)
        E, W = [2, 4, 5, 10, 17], [3, 6, 7, 11, 16]
        for y in range(ny):
            for z in range(nz):
                k = idx[z, y, x]
                if k >= 0:
                    p = dump[k]
                    # left-to-right sum as src/Kernel.c:488-489
                    s = p[2] + p[4] + p[5] + p[10] + p[17] - p[3] - p[6] - p[7] - p[11] - p[16]
                    ux[y * nz + z] = s
                    fl[y * nz + z] = 1
        del E, W
        return self.lib.orc_verify_norm(ux, fl, ny, nz, omega, xforce)


class Reference:
    """The reference's own sources, compiled by oracle/Makefile into oracle/_ref."""

    def __init__(self, variant: str = "verif"):
        path = os.path.join(HERE, "_ref", f"libref_{variant}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        # the reference's internal calls must reach its own functions, never
        # same-named ones of a library loaded before (liblbmhost.so)
        L = self.lib = C.CDLL(path, mode=os.RTLD_DEEPBIND)
        L.ref_list_run.argtypes = [_i32p, C.c_int * 3, C.c_int * 3, C.c_int, C.c_int, C.c_void_p,
                                   C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int,
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.POINTER(_RefResult)]
        L.ref_aa_run.argtypes = [_i32p, C.c_int * 3, C.c_int * 3, C.c_int, C.c_int, C.c_void_p,
                                 C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int,
                                 C.c_void_p, C.POINTER(_RefResult)]
        L.ref_geometry.argtypes = [C.c_char_p, C.c_int * 3, C.c_int * 3, _i32p, C.c_int * 3]
        L.ref_pad_cells.argtypes = [C.c_int, C.c_char_p]
        L.ref_set_output.argtypes = [C.c_int, C.c_int, C.c_int]

    @staticmethod
    def available(variant: str = "verif") -> bool:
        return os.path.exists(os.path.join(HERE, "_ref", f"libref_{variant}.so"))

    @staticmethod
    def _args(words):
        if not words:
            return 0, None, None
        arr = (C.c_char_p * len(words))(*[w.encode() for w in words])
        return len(words), C.cast(arr, C.c_void_p), arr

    _INIT = {"none": 0, "verify": 1, "weights": 2}

    def list_run(self, mask, dims, periodic, iterations, omega, xforce, *, n_threads=1,
                 kernel_args=(), init="weights", rho=1.0, verify=False, want_adj=True):
        mask = np.ascontiguousarray(mask, dtype=np.int32)
        nf = int(np.count_nonzero(mask))
        dump = np.zeros((nf, NDIR))
        adj = np.zeros((nf, NLINK), dtype=np.uint32)
        grid = np.zeros(mask.size, dtype=np.uint32)
        coords = np.zeros((nf, 3), dtype=np.uint32)
        res = _RefResult()
        n, argp, keep = self._args(list(kernel_args))
        self.lib.ref_list_run(mask, _i3(dims), _i3(periodic), n_threads, n, argp, omega, xforce,
                              iterations, self._INIT[init], rho, int(verify),
                              dump.ctypes.data, adj.ctypes.data if want_adj else None,
                              grid.ctypes.data, coords.ctypes.data, C.byref(res))
        del keep
        return dump, ListBuild(nf, res.nCells, grid, coords, adj), res

    def aa_run(self, mask, dims, periodic, iterations, omega, xforce, *, n_threads=1,
               init="weights", rho=1.0, verify=False):
        mask = np.ascontiguousarray(mask, dtype=np.int32)
        nf = int(np.count_nonzero(mask))
        dump = np.zeros((nf, NDIR))
        res = _RefResult()
        self.lib.ref_aa_run(mask, _i3(dims), _i3(periodic), n_threads, 0, None, omega, xforce,
                            iterations, self._INIT[init], rho, int(verify), dump.ctypes.data,
                            C.byref(res))
        return dump, res

    def geometry(self, gtype: str, dims, periodic=(0, 0, 0)):
        mask = np.zeros(int(np.prod(dims)), dtype=np.int32)
        ldp = (C.c_int * 3)()
        self.lib.ref_geometry(gtype.encode(), _i3(dims), _i3(periodic), mask, ldp)
        return mask, tuple(ldp)

    def pad_cells(self, n: int, s: str | None = None) -> int:
        return self.lib.ref_pad_cells(n, s.encode() if s is not None else None)
