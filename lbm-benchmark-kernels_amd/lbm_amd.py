"""Python mirror of the kernel-plugin interface of liblbmhip.so / liblbmhost.so.

The reference is driven from C: its driver builds a ``LatticeDesc``, looks a
kernel up by name in a ``KernelFunctions`` table, calls ``Init``, optionally
initialises the lattice through ``SetNode`` or ``KernelSetInitialDensity``,
calls ``kd->Kernel`` and reads results through ``GetNode``
(src/Main.c:474-592).  This module exposes the same objects through ctypes,
with the same names and argument meanings, so tests and the benchmark read
like that sequence.  Structures mirror ``include/lbm_kernel_abi.h`` field for
field.

There is no fallback: if the HIP library is not built, or no GPU is present,
``Init`` fails with the library's own error.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NDIR = 19
NLINK = 18

PdfT = C.c_double


class LatticeDesc(C.Structure):
    _fields_ = [("Dims", C.c_int * 3), ("Lattice", C.POINTER(C.c_int)), ("nCells", C.c_int),
                ("nFluid", C.c_int), ("nObst", C.c_int), ("nInlet", C.c_int), ("nOutlet", C.c_int),
                ("PeriodicX", C.c_int), ("PeriodicY", C.c_int), ("PeriodicZ", C.c_int),
                ("Name", C.c_char_p)]


class CaseData(C.Structure):
    _fields_ = [("Omega", PdfT), ("RhoIn", PdfT), ("RhoOut", PdfT), ("Ux", PdfT), ("Uy", PdfT),
                ("Uz", PdfT), ("XForce", PdfT), ("MaxIterations", C.c_int), ("VtkOutput", C.c_int),
                ("VtkModulus", C.c_int), ("StatisticsModulus", C.c_int)]


class KernelData(C.Structure):
    pass


_BCGet = C.CFUNCTYPE(None, C.POINTER(KernelData), C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(PdfT))
_BCSet = C.CFUNCTYPE(None, C.POINTER(KernelData), C.c_int, C.c_int, C.c_int, C.c_int, PdfT)
_Node = C.CFUNCTYPE(None, C.POINTER(KernelData), C.c_int, C.c_int, C.c_int, C.POINTER(PdfT))
_Kern = C.CFUNCTYPE(None, C.POINTER(LatticeDesc), C.POINTER(KernelData), C.POINTER(CaseData))

KernelData._fields_ = [
    ("Pdfs", C.POINTER(PdfT) * 2), ("SrcPdfs", C.POINTER(PdfT)), ("DstPdfs", C.POINTER(PdfT)),
    ("PdfsActive", C.POINTER(PdfT)), ("Dims", C.c_int * 3), ("GlobalDims", C.c_int * 3),
    ("Offsets", C.c_int * 3), ("ObstIndices", C.POINTER(C.c_int)), ("nObstIndices", C.c_int),
    ("BounceBackPdfsSrc", C.POINTER(C.c_int)), ("BounceBackPdfsDst", C.POINTER(C.c_int)),
    ("nBounceBackPdfs", C.c_int), ("LoopBalance", C.c_double), ("Duration", C.c_double),
    ("BoundaryConditionsGetPdf", _BCGet), ("BoundaryConditionsSetPdf", _BCSet),
    ("GetNode", _Node), ("SetNode", _Node), ("Kernel", _Kern)]


class Parameters(C.Structure):
    _fields_ = [("nArgs", C.c_int), ("Args", C.POINTER(C.c_char_p)),
                ("nKernelArgs", C.c_int), ("KernelArgs", C.POINTER(C.c_char_p))]


_Init = C.CFUNCTYPE(None, C.POINTER(LatticeDesc), C.POINTER(C.POINTER(KernelData)), C.POINTER(Parameters))
_Deinit = C.CFUNCTYPE(None, C.POINTER(LatticeDesc), C.POINTER(C.POINTER(KernelData)))


class KernelFunctions(C.Structure):
    _fields_ = [("Name", C.c_char * 128), ("Init", _Init), ("Deinit", _Deinit)]


class LbmHipTimings(C.Structure):
    _fields_ = [("steps", C.c_int), ("loopMs", C.c_double), ("evenMsAvg", C.c_double),
                ("oddMsAvg", C.c_double), ("evenMsMin", C.c_double), ("oddMsMin", C.c_double),
                ("uploadMs", C.c_double), ("downloadMs", C.c_double)]


class LbmHipListInfo(C.Structure):
    _fields_ = [("nFluid", C.c_int), ("nCells", C.c_int), ("adjStride", C.c_int), ("nRuns", C.c_int),
                ("dims", C.c_int * 3), ("periodic", C.c_int * 3),
                ("bytesPerFlupEven", C.c_double), ("bytesPerFlupOdd", C.c_double),
                ("nRunsWave", C.c_int), ("adjRle", C.c_int), ("prec", C.c_int)]


class LbmHipSteady(C.Structure):
    _fields_ = [("steps", C.c_int), ("converged", C.c_int), ("relChange", C.c_double),
                ("mass", C.c_double), ("momentumX", C.c_double), ("seconds", C.c_double)]


class LbmHipBatchStatus(C.Structure):
    _fields_ = [("active", C.c_int), ("converged", C.c_int), ("steps", C.c_int), ("nFluid", C.c_int),
                ("relChange", C.c_double), ("mass", C.c_double), ("momentumX", C.c_double),
                ("seconds", C.c_double)]


class LbmHipBatchTuning(C.Structure):
    _fields_ = [("evenBlock", C.c_int), ("evenCap", C.c_int), ("vec", C.c_int), ("oddBlock", C.c_int),
                ("evenMs", C.c_double), ("oddMs", C.c_double), ("groupMB", C.c_int), ("pairMs", C.c_double),
                ("nontemporal", C.c_int), ("oddGroup", C.c_int)]


class LbmHipStreamResult(C.Structure):
    _fields_ = [("copyGBs", C.c_double), ("update19GBs", C.c_double), ("copyBytes", C.c_double),
                ("update19Bytes", C.c_double)]


class GeoPorous(C.Structure):
    _fields_ = [("Radius", C.c_double), ("Porosity", C.c_double), ("Seed", C.c_uint64),
                ("nSpheres", C.c_int)]


GEO_TYPES = {"box": 0, "channel": 1, "pipe": 2, "blocks": 3, "fluid": 4, "porous": 5}
LBM_HIP_ERRORS_EXIT, LBM_HIP_ERRORS_RETURN = 0, 1

# Symbols each library exports (declared in include/lbm_hip.h, include/lbm_host.h).
HIP_SYMBOLS = ["D3Q19ListAaPvInit_PushSoAHip", "D3Q19ListAaPvDeinit_PushSoAHip",
               "D3Q19AaInit_AaSoAHip", "D3Q19AaDeinit_AaSoAHip", "g_lbmHipKernels",
               "g_lbmHipKernelCount", "LbmHipSetErrorMode", "LbmHipLastError", "LbmHipClearError",
               "LbmHipGetTimings", "LbmHipSetLaunchEvents", "LbmHipSetShadowDownload",
               "LbmHipSyncShadow", "LbmHipListGetInfo", "LbmHipListCopyBuild", "LbmHipDumpNodes",
               "LbmHipLoadNodes", "LbmHipDeviceInfo", "LbmHipStreamBench", "LbmHipListMoments",
               "LbmHipListRunSteady", "LbmHipMoments", "LbmHipRunSteady", "LbmHipBatchCreate", "LbmHipBatchDestroy", "LbmHipBatchAdd",
               "LbmHipBatchRemove", "LbmHipBatchKernelData", "LbmHipBatchStatusOf", "LbmHipBatchStep",
               "LbmHipBatchRun", "LbmHipBatchSync", "LbmHipBatchGetTuning", "LbmHipBatchSetTuning",
               "LbmHipBatchAutotune", "LbmHipBatchStats", "LbmHipBatchGraphStats", "LbmHipBatchLaunch", "LbmHipBatchWait",
               "LbmHipListBuildNoAlloc", "LbmHipListDeviceBytes"]
HOST_SYMBOLS = ["D3Q19_X", "D3Q19_Y", "D3Q19_Z", "D3Q19_INV", "D3Q19_NAMES", "KernelDensity", "VtkWrite",
                "KernelStatistics", "KernelStatisticsAdv", "KernelSetInitialDensity",
                "KernelSetInitialVelocity", "KernelVerifiy", "KernelAddBodyForce", "GeoCreateByType",
                "GeoCreateByStr", "GeoFree"]


class LibraryMissing(RuntimeError):
    pass


_libs: dict = {}


def lib_path(name: str) -> str:
    return os.path.join(HERE, name)


def host_lib() -> C.CDLL:
    if "host" not in _libs:
        p = lib_path("liblbmhost.so")
        if not os.path.exists(p):
            raise LibraryMissing(f"{p} is not built; run __graft_entry__.build()")
        # global: liblbmhip.so's kernels call the program's VtkWrite and
        # KernelStatistics (weak references resolved when it loads)
        L = C.CDLL(p, mode=C.RTLD_GLOBAL)
        L.GeoCreateByStr.argtypes = [C.c_char_p, C.c_int * 3, C.c_int * 3, C.POINTER(LatticeDesc)]
        L.GeoCreateByType.argtypes = [C.c_int, C.c_void_p, C.c_int * 3, C.c_int * 3, C.POINTER(LatticeDesc)]
        L.GeoFree.argtypes = [C.POINTER(LatticeDesc)]
        for fn in ("KernelSetInitialDensity", "KernelSetInitialVelocity"):
            getattr(L, fn).argtypes = [C.POINTER(LatticeDesc), C.POINTER(KernelData), C.POINTER(CaseData)]
        L.KernelAddBodyForce.argtypes = [C.POINTER(KernelData), C.POINTER(LatticeDesc), C.POINTER(CaseData)]
        L.KernelVerifiy.argtypes = [C.POINTER(LatticeDesc), C.POINTER(KernelData), C.POINTER(CaseData),
                                    C.POINTER(PdfT)]
        L.KernelDensity.argtypes = [C.POINTER(KernelData), C.POINTER(LatticeDesc)]
        L.KernelDensity.restype = PdfT
        L.KernelStatisticsAdv.argtypes = [C.POINTER(KernelData), C.POINTER(LatticeDesc),
                                          C.POINTER(CaseData), C.c_int, C.c_int]
        L.KernelStatistics.argtypes = [C.POINTER(KernelData), C.POINTER(LatticeDesc), C.POINTER(CaseData), C.c_int]
        L.VtkWrite.argtypes = [C.POINTER(LatticeDesc), C.POINTER(KernelData), C.POINTER(CaseData), C.c_int]
        _libs["host"] = L
    return _libs["host"]


def hip_lib() -> C.CDLL:
    if "hip" not in _libs:
        p = lib_path("liblbmhip.so")
        if not os.path.exists(p):
            raise LibraryMissing(f"{p} is not built; run __graft_entry__.build()")
        host_lib()                      # first: provides the output functions
        L = C.CDLL(p)
        KD = C.POINTER(KernelData)
        L.LbmHipSetErrorMode.argtypes = [C.c_int]
        L.LbmHipLastError.restype = C.c_char_p
        L.LbmHipGetTimings.argtypes = [KD, C.POINTER(LbmHipTimings)]
        L.LbmHipSetLaunchEvents.argtypes = [KD, C.c_int]
        L.LbmHipSetShadowDownload.argtypes = [KD, C.c_int]
        L.LbmHipSyncShadow.argtypes = [KD]
        L.LbmHipListGetInfo.argtypes = [KD, C.POINTER(LbmHipListInfo)]
        L.LbmHipListCopyBuild.argtypes = [KD, C.c_void_p, C.c_void_p, C.c_void_p]
        L.LbmHipDumpNodes.argtypes = [KD, C.POINTER(LatticeDesc), C.c_void_p]
        L.LbmHipLoadNodes.argtypes = [KD, C.POINTER(LatticeDesc), C.c_void_p]
        L.LbmHipDeviceInfo.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.LbmHipStreamBench.argtypes = [C.c_uint64, C.c_uint32, C.c_int, C.POINTER(LbmHipStreamResult)]
        L.LbmHipListMoments.argtypes = [KD, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.LbmHipMoments.argtypes = [KD, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.LbmHipBatchCreate.argtypes = [C.POINTER(Parameters)]
        L.LbmHipBatchCreate.restype = C.c_void_p
        L.LbmHipBatchDestroy.argtypes = [C.c_void_p]
        L.LbmHipBatchAdd.argtypes = [C.c_void_p, C.POINTER(LatticeDesc), C.c_double, C.c_double, C.c_int]
        L.LbmHipBatchRemove.argtypes = [C.c_void_p, C.c_int]
        L.LbmHipBatchKernelData.argtypes = [C.c_void_p, C.c_int]
        L.LbmHipBatchKernelData.restype = KD
        L.LbmHipBatchStatusOf.argtypes = [C.c_void_p, C.c_int, C.POINTER(LbmHipBatchStatus)]
        L.LbmHipBatchStep.argtypes = [C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_int)]
        L.LbmHipBatchRun.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.LbmHipBatchLaunch.argtypes = [C.c_void_p, C.c_int]
        L.LbmHipBatchWait.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_int)]
        L.LbmHipBatchSync.argtypes = [C.c_void_p, C.c_int]
        L.LbmHipBatchGetTuning.argtypes = [C.c_void_p, C.POINTER(LbmHipBatchTuning)]
        L.LbmHipBatchSetTuning.argtypes = [C.c_void_p, C.POINTER(LbmHipBatchTuning)]
        L.LbmHipBatchAutotune.argtypes = [C.c_void_p, C.c_int, C.POINTER(LbmHipBatchTuning)]
        L.LbmHipBatchStats.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.LbmHipBatchGraphStats.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_longlong),
                                            C.POINTER(C.c_longlong)]
        L.LbmHipListRunSteady.argtypes = [KD, C.POINTER(CaseData), C.c_int, C.c_double, C.c_int,
                                          C.POINTER(LbmHipSteady)]
        L.LbmHipRunSteady.argtypes = [KD, C.POINTER(CaseData), C.c_int, C.c_double, C.c_int,
                                      C.POINTER(LbmHipSteady)]
        I3 = C.c_int * 3
        L.LbmHipListBuildNoAlloc.argtypes = [C.c_void_p, I3, I3, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
        L.LbmHipListDeviceBytes.argtypes = [C.POINTER(LatticeDesc), C.POINTER(Parameters), C.POINTER(C.c_uint64)]
        _libs["hip"] = L
    return _libs["hip"]


def list_device_bytes(geo: "Geometry", kernel_args: Sequence[str] = ()) -> int:
    """Upper bound of the device bytes a list-aa-pv-soa Init takes for geo
    (LbmHipListDeviceBytes; host-only, no device needed)."""
    L = hip_lib()
    words = [w.encode() for w in kernel_args]
    argv = (C.c_char_p * max(1, len(words)))(*words)
    params = Parameters(0, None, len(words), argv)
    out = C.c_uint64(0)
    L.LbmHipSetErrorMode(LBM_HIP_ERRORS_RETURN)
    L.LbmHipClearError()
    if L.LbmHipListDeviceBytes(C.byref(geo.ld), C.byref(params), C.byref(out)) != 0:
        raise RuntimeError(L.LbmHipLastError().decode())
    return int(out.value)


def kernel_table() -> dict:
    """Name -> KernelFunctions of liblbmhip.so's table (src/KernelFunctions.h:56 layout)."""
    L = hip_lib()
    n = C.c_int.in_dll(L, "g_lbmHipKernelCount").value
    tab = (KernelFunctions * n).in_dll(L, "g_lbmHipKernels")
    return {tab[i].Name.decode(): tab[i] for i in range(n)}


def stream_bench(n_doubles: int, n_nodes: int, reps: int = 10) -> dict:
    """Copy and update-19 bandwidth yardsticks of the device (GB/s)."""
    r = LbmHipStreamResult()
    if hip_lib().LbmHipStreamBench(n_doubles, n_nodes, reps, C.byref(r)) != 0:
        raise RuntimeError(hip_lib().LbmHipLastError().decode())
    return {"copy_gbs": r.copyGBs, "update19_gbs": r.update19GBs, "copy_bytes": r.copyBytes,
            "update19_bytes": r.update19Bytes}


def device_info():
    name = C.create_string_buffer(256)
    cus = C.c_int(0)
    clk = C.c_double(0.0)
    rc = hip_lib().LbmHipDeviceInfo(name, 256, C.byref(cus), C.byref(clk))
    if rc != 0:
        return None
    return {"name": name.value.decode(), "compute_units": cus.value}


# --------------------------------------------------------------------------
class Geometry:
    """A LatticeDesc whose mask lives in a numpy array.  ``mask`` is a
    (nZ, nY, nX) view of that array and keeps it alive on its own, so it stays
    valid after the Geometry is gone."""

    def __init__(self, ld: LatticeDesc, keep: np.ndarray):
        self.ld = ld
        self._keep = keep
        nx, ny, nz = self.dims
        self.mask = keep.reshape(nz, ny, nx)

    @property
    def dims(self):
        return tuple(self.ld.Dims)

    @property
    def periodic(self):
        return (self.ld.PeriodicX, self.ld.PeriodicY, self.ld.PeriodicZ)

    @property
    def n_fluid(self):
        return self.ld.nFluid

    @classmethod
    def from_string(cls, gtype: str, dims: Sequence[int], periodic=(0, 0, 0)) -> "Geometry":
        """GeoCreateByStr of liblbmhost.so (src/Geometry.c:37 semantics); the
        mask is copied into numpy memory and the C allocation freed."""
        ld = LatticeDesc()
        host_lib().GeoCreateByStr(gtype.encode(), (C.c_int * 3)(*dims), (C.c_int * 3)(*periodic),
                                  C.byref(ld))
        n = int(ld.Dims[0]) * int(ld.Dims[1]) * int(ld.Dims[2])
        m = np.ctypeslib.as_array(ld.Lattice, shape=(n,)).copy()
        host_lib().GeoFree(C.byref(ld))
        ld.Lattice = m.ctypes.data_as(C.POINTER(C.c_int))
        return cls(ld, m)

    @classmethod
    def from_mask(cls, mask: np.ndarray, dims: Sequence[int], periodic=(0, 0, 0), name=b"mask"):
        """LatticeDesc over a caller mask (int32, x fastest), counts as the reference computes them."""
        m = np.ascontiguousarray(np.asarray(mask, dtype=np.int32).reshape(-1))
        ld = LatticeDesc()
        ld.Dims = (C.c_int * 3)(*dims)
        ld.Lattice = m.ctypes.data_as(C.POINTER(C.c_int))
        ld.nCells = int(m.size)
        ld.nFluid = int(np.count_nonzero(m))
        ld.nObst = int(m.size - ld.nFluid)
        ld.nInlet = int(np.count_nonzero(m == 2))
        ld.nOutlet = int(np.count_nonzero(m == 4))
        ld.PeriodicX, ld.PeriodicY, ld.PeriodicZ = (int(p) for p in periodic)
        ld.Name = name
        return cls(ld, m)


def case_data(omega=1.0, xforce=1e-5, iterations=10, rho_in=1.0, rho_out=1.0) -> CaseData:
    cd = CaseData()
    cd.Omega, cd.RhoIn, cd.RhoOut, cd.XForce = omega, rho_in, rho_out, xforce
    cd.MaxIterations = iterations
    cd.VtkModulus = 100
    cd.StatisticsModulus = 100
    return cd


class Kernel:
    """One Init'ed kernel of liblbmhip.so (KernelData* plus its table entry)."""

    def __init__(self, name: str, geo: Geometry, kernel_args: Sequence[str] = (), errors: str = "return"):
        L = hip_lib()
        L.LbmHipSetErrorMode(LBM_HIP_ERRORS_RETURN if errors == "return" else LBM_HIP_ERRORS_EXIT)
        L.LbmHipClearError()
        table = kernel_table()
        if name not in table:
            raise KeyError(f'requested kernel "{name}" not found')
        self.name = name
        self.kf = table[name]
        self.geo = geo
        words = [w.encode() for w in kernel_args]
        self._argv = (C.c_char_p * max(1, len(words)))(*words)
        self.params = Parameters(0, None, len(words), self._argv)
        self.kd = C.POINTER(KernelData)()
        self.kf.Init(C.byref(geo.ld), C.byref(self.kd), C.byref(self.params))
        if not self.kd:
            raise RuntimeError(L.LbmHipLastError().decode() or "kernel initialisation failed")

    # -- reference-style calls ------------------------------------------------
    def run(self, cd: CaseData):
        L = hip_lib()
        L.LbmHipClearError()
        self.kd.contents.Kernel(C.byref(self.geo.ld), self.kd, C.byref(cd))
        err = L.LbmHipLastError().decode()
        if err:
            raise RuntimeError(err)

    @property
    def duration(self) -> float:
        return self.kd.contents.Duration

    def get_node(self, x, y, z):
        buf = (PdfT * NDIR)()
        self.kd.contents.GetNode(self.kd, x, y, z, buf)
        return np.array(buf[:])

    def set_node(self, x, y, z, values):
        buf = (PdfT * NDIR)(*[float(v) for v in values])
        self.kd.contents.SetNode(self.kd, x, y, z, buf)

    def set_initial_verify(self, cd: CaseData):
        """KernelSetInitialDensity + KernelSetInitialVelocity (src/Main.c:562-567)."""
        H = host_lib()
        H.KernelSetInitialDensity(C.byref(self.geo.ld), self.kd, C.byref(cd))
        H.KernelSetInitialVelocity(C.byref(self.geo.ld), self.kd, C.byref(cd))

    def set_initial_weights(self, rho=1.0):
        w = np.array([1 / 18] * 4 + [1 / 36] * 4 + [1 / 18] + [1 / 36] * 4 + [1 / 18] + [1 / 36] * 4 + [1 / 3])
        vals = np.tile(w * rho, (self.geo.n_fluid, 1))
        self.load(vals)

    def verify(self, cd: CaseData) -> float:
        out = PdfT(-1.0)
        host_lib().KernelVerifiy(C.byref(self.geo.ld), self.kd, C.byref(cd), C.byref(out))
        return out.value

    def density(self) -> float:
        return host_lib().KernelDensity(self.kd, C.byref(self.geo.ld))

    # -- bulk access ----------------------------------------------------------
    def dump(self) -> np.ndarray:
        out = np.zeros((self.geo.n_fluid, NDIR))
        if hip_lib().LbmHipDumpNodes(self.kd, C.byref(self.geo.ld), out.ctypes.data) != 0:
            raise RuntimeError(hip_lib().LbmHipLastError().decode())
        return out

    def load(self, values: np.ndarray):
        v = np.ascontiguousarray(values, dtype=np.float64)
        assert v.shape == (self.geo.n_fluid, NDIR)
        if hip_lib().LbmHipLoadNodes(self.kd, C.byref(self.geo.ld), v.ctypes.data) != 0:
            raise RuntimeError(hip_lib().LbmHipLastError().decode())

    def timings(self) -> LbmHipTimings:
        t = LbmHipTimings()
        hip_lib().LbmHipGetTimings(self.kd, C.byref(t))
        return t

    def set_launch_events(self, on: bool):
        hip_lib().LbmHipSetLaunchEvents(self.kd, int(on))

    def set_shadow_download(self, on: bool):
        hip_lib().LbmHipSetShadowDownload(self.kd, int(on))

    def list_info(self) -> LbmHipListInfo:
        info = LbmHipListInfo()
        if hip_lib().LbmHipListGetInfo(self.kd, C.byref(info)) != 0:
            raise RuntimeError(hip_lib().LbmHipLastError().decode())
        return info

    def list_build(self):
        """(adj[nFluid,18] node-major, grid[cells], coords[nFluid,3]) as the reference stores them."""
        info = self.list_info()
        nf = info.nFluid
        adj = np.zeros((nf, NLINK), dtype=np.uint32)
        grid = np.zeros(int(np.prod(self.geo.dims)), dtype=np.uint32)
        coords = np.zeros((nf, 3), dtype=np.uint32)
        if hip_lib().LbmHipListCopyBuild(self.kd, adj.ctypes.data, grid.ctypes.data, coords.ctypes.data) != 0:
            raise RuntimeError(hip_lib().LbmHipLastError().decode())
        return adj, grid, coords

    def moments(self):
        """(mass, x momentum) of the device lattice, FP64 sums on the GPU."""
        m, j = C.c_double(), C.c_double()
        if hip_lib().LbmHipMoments(self.kd, C.byref(m), C.byref(j)) != 0:
            raise RuntimeError(hip_lib().LbmHipLastError().decode())
        return m.value, j.value

    def run_steady(self, cd: CaseData, check_every: int, tol: float, max_steps: int) -> LbmHipSteady:
        r = LbmHipSteady()
        L = hip_lib()
        L.LbmHipClearError()
        if L.LbmHipRunSteady(self.kd, C.byref(cd), check_every, tol, max_steps, C.byref(r)) != 0:
            raise RuntimeError(L.LbmHipLastError().decode())
        return r

    def close(self):
        if self.kd:
            self.kf.Deinit(C.byref(self.geo.ld), C.byref(self.kd))
            self.kd = C.POINTER(KernelData)()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many list-aa-pv-soa lattices stepped together (LbmHipBatch* of liblbmhip.so)."""

    def __init__(self, kernel_args: Sequence[str] = (), errors: str = "return"):
        L = hip_lib()
        L.LbmHipSetErrorMode(LBM_HIP_ERRORS_RETURN if errors == "return" else LBM_HIP_ERRORS_EXIT)
        L.LbmHipClearError()
        words = [w.encode() for w in kernel_args]
        self._argv = (C.c_char_p * max(1, len(words)))(*words)
        self.params = Parameters(0, None, len(words), self._argv)
        self.h = L.LbmHipBatchCreate(C.byref(self.params))
        if not self.h:
            raise RuntimeError(L.LbmHipLastError().decode() or "batch creation failed")
        self.geos = {}

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(hip_lib().LbmHipLastError().decode())

    def add(self, geo: Geometry, omega: float, xforce: float, max_steps: int = 0) -> int:
        L = hip_lib()
        L.LbmHipClearError()
        i = L.LbmHipBatchAdd(self.h, C.byref(geo.ld), omega, xforce, max_steps)
        if i < 0:
            raise RuntimeError(L.LbmHipLastError().decode())
        self.geos[i] = geo
        return i

    def remove(self, i: int):
        self._check(hip_lib().LbmHipBatchRemove(self.h, i))
        self.geos.pop(i, None)

    def step(self, steps: int, tol: float = 0.0) -> int:
        n = C.c_int()
        self._check(hip_lib().LbmHipBatchStep(self.h, steps, tol, C.byref(n)))
        return n.value

    def launch(self, steps: int):
        self._check(hip_lib().LbmHipBatchLaunch(self.h, steps))

    def wait(self, tol: float = 0.0) -> int:
        n = C.c_int()
        self._check(hip_lib().LbmHipBatchWait(self.h, tol, C.byref(n)))
        return n.value

    def run(self, check_every: int, tol: float):
        self._check(hip_lib().LbmHipBatchRun(self.h, check_every, tol))

    def status(self, i: int) -> LbmHipBatchStatus:
        st = LbmHipBatchStatus()
        self._check(hip_lib().LbmHipBatchStatusOf(self.h, i, C.byref(st)))
        return st

    def kernel_data(self, i: int):
        return hip_lib().LbmHipBatchKernelData(self.h, i)

    def sync(self, i: int = -1):
        self._check(hip_lib().LbmHipBatchSync(self.h, i))

    def dump(self, i: int) -> np.ndarray:
        """Node views of domain i (mask order), after copying it to the host."""
        self.sync(i)
        geo = self.geos[i]
        out = np.zeros((geo.n_fluid, NDIR))
        self._check(hip_lib().LbmHipDumpNodes(self.kernel_data(i), C.byref(geo.ld), out.ctypes.data))
        return out

    def moments(self, i: int):
        m, j = C.c_double(), C.c_double()
        self._check(hip_lib().LbmHipListMoments(self.kernel_data(i), C.byref(m), C.byref(j)))
        return m.value, j.value

    def tuning(self) -> LbmHipBatchTuning:
        t = LbmHipBatchTuning()
        self._check(hip_lib().LbmHipBatchGetTuning(self.h, C.byref(t)))
        return t

    def set_tuning(self, even_block: int, even_cap: int, vec: int, odd_block: int, group_mb: int = 0,
                   nontemporal: int = 1, odd_group: int = 8):
        t = LbmHipBatchTuning(even_block, even_cap, vec, odd_block, 0.0, 0.0, group_mb, 0.0, nontemporal, odd_group)
        self._check(hip_lib().LbmHipBatchSetTuning(self.h, C.byref(t)))

    def autotune(self, pairs: int = 4) -> LbmHipBatchTuning:
        t = LbmHipBatchTuning()
        self._check(hip_lib().LbmHipBatchAutotune(self.h, pairs, C.byref(t)))
        return t

    def stats(self):
        g, f = C.c_double(), C.c_double()
        self._check(hip_lib().LbmHipBatchStats(self.h, C.byref(g), C.byref(f)))
        return g.value, f.value

    def graph_stats(self):
        """(launch ranges of the last interval, graphs captured, graph replays)."""
        n, c, r = C.c_int(), C.c_longlong(), C.c_longlong()
        self._check(hip_lib().LbmHipBatchGraphStats(self.h, C.byref(n), C.byref(c), C.byref(r)))
        return n.value, c.value, r.value

    def close(self):
        if self.h:
            hip_lib().LbmHipBatchDestroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
