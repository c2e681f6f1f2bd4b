"""The drop-in boundary without a GPU: libraries load, export every symbol
the headers declare, struct layouts match the reference's, and the entry
points fail loudly (no CPU fallback) when no device is present.  One GPU
test checks a limit Init enforces before it allocates anything."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import PKG, ROOT

PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include HDR
#define OFS(T, m) printf(#T "." #m " %zu\n", offsetof(T, m));
int main(void) {
  printf("LatticeDesc %zu\n", sizeof(LatticeDesc));
  OFS(LatticeDesc, Lattice) OFS(LatticeDesc, nCells) OFS(LatticeDesc, PeriodicZ) OFS(LatticeDesc, Name)
  printf("CaseData %zu\n", sizeof(CaseData));
  OFS(CaseData, XForce) OFS(CaseData, MaxIterations) OFS(CaseData, StatisticsModulus)
  printf("KernelData %zu\n", sizeof(KernelData));
  OFS(KernelData, PdfsActive) OFS(KernelData, Dims) OFS(KernelData, ObstIndices) OFS(KernelData, nBounceBackPdfs)
  OFS(KernelData, LoopBalance) OFS(KernelData, Duration) OFS(KernelData, BoundaryConditionsGetPdf)
  OFS(KernelData, GetNode) OFS(KernelData, SetNode) OFS(KernelData, Kernel)
  printf("Parameters %zu\n", sizeof(Parameters));
  OFS(Parameters, KernelArgs)
  printf("KernelFunctions %zu\n", sizeof(KernelFunctions));
  OFS(KernelFunctions, Init) OFS(KernelFunctions, Deinit)
  return 0;
}
"""


def _probe(tmp_path, header, incs, defines=()):
    src = tmp_path / "probe.c"
    src.write_text(PROBE.replace("HDR", f'"{header}"'))
    exe = tmp_path / "probe"
    # the reference header also defines its kernel table; nothing in it is called
    cmd = ["gcc", "-std=c99", *[f"-I{i}" for i in incs], *defines, str(src), "-o", str(exe),
           "-Wl,--unresolved-symbols=ignore-all"]
    subprocess.run(cmd, check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return dict(line.rsplit(" ", 1) for line in out.strip().splitlines())


def test_layout_matches_ctypes_mirror(tmp_path, lbm):
    got = _probe(tmp_path, "lbm_kernel_abi.h", [os.path.join(ROOT, "include")])
    for name, T in [("LatticeDesc", lbm.LatticeDesc), ("CaseData", lbm.CaseData), ("KernelData", lbm.KernelData),
                    ("Parameters", lbm.Parameters), ("KernelFunctions", lbm.KernelFunctions)]:
        assert int(got[name]) == C.sizeof(T), name
        for key, v in got.items():
            if key.startswith(name + "."):
                assert int(v) == getattr(T, key.split(".")[1]).offset, key


def test_layout_matches_reference_headers(tmp_path):
    ref = "/root/reference/src"
    if not os.path.exists(os.path.join(ref, "Kernel.h")):
        pytest.skip("reference headers not present")
    mine = _probe(tmp_path, "lbm_kernel_abi.h", [os.path.join(ROOT, "include")])
    (tmp_path / "r").mkdir()
    theirs = _probe(tmp_path / "r", "KernelFunctions.h", [ref],
                    ["-DPRECISION_DP", "-DDATA_LAYOUT_SOA", "-DPROP_MODEL_PUSH", "-D_POSIX_C_SOURCE=200112L",
                     "-DVECTOR_AVX"])
    assert mine == theirs


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    funcs = set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", txt, flags=re.M))
    data = set(re.findall(r"^extern\s+[\w\s\*]+?\b(\w+)\s*(?:\[[^\]]*\])?;", txt, flags=re.M))
    return {f for f in funcs | data if not f.startswith(("typedef", "enum"))} - {"void", "int", "const"}


def _exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(PKG, lib)], check=True,
                         capture_output=True, text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


@pytest.mark.parametrize("header,lib", [("lbm_hip.h", "liblbmhip.so"), ("lbm_host.h", "liblbmhost.so")])
def test_exports_every_declared_symbol(header, lib):
    declared = _declared(header)
    assert declared, header
    missing = declared - _exported(lib)
    assert not missing, f"{lib} lacks {sorted(missing)}"


def test_symbol_lists_in_mirror(lbm):
    assert set(lbm.HIP_SYMBOLS) <= _exported("liblbmhip.so")
    assert set(lbm.HOST_SYMBOLS) <= _exported("liblbmhost.so")
    # the HIP library must not clash with the reference's Kernel.o/Geometry.o
    # when both are linked into the reference driver
    clash = {"KernelDensity", "KernelVerifiy", "GeoCreateByStr", "PadCells", "D3Q19_X", "MemAlloc", "VtkWrite",
             "KernelStatistics"}
    assert not (clash & _exported("liblbmhip.so"))


def test_kernel_table(lbm):
    assert list(lbm.kernel_table()) == ["list-aa-pv-soa", "aa-soa"]


def _no_gpu(lbm):
    return lbm.device_info() is None


def test_init_fails_loudly_without_device(lbm):
    if not _no_gpu(lbm):
        pytest.skip("a GPU is present")
    geo = lbm.Geometry.from_string("box", (8, 8, 8), (1, 1, 0))
    for name in ("list-aa-pv-soa", "aa-soa"):
        with pytest.raises(RuntimeError, match="no HIP device"):
            lbm.Kernel(name, geo)


@pytest.mark.parametrize("args,msg", [
    (["-foo"], "unknown kernel parameter"),
    (["-blk"], "requires a parameter"),
    (["-blk", "-3"], "must be in"),
    (["-pad", "64+64"], "equal or larger than modulus"),
    (["-pad", "x"], "invalid characters"),
    (["-force", "maybe"], "on or off"),
    (["-bs", "100"], "multiple of 64"),
    (["-lambda", "-1"], "positive"),
    (["-prec", "fp8"], "fp64, fp32s or fp16s"),
    (["-prec"], "fp64, fp32s or fp16s"),
    (["-fp16-scale", "3000"], "power of two"),
    (["-fp16-scale", "0"], "must be in"),
    (["-vec", "2"], "auto or 1"),
    (["-even-cap", "maybe"], "on, off or force"),
    (["-odd-order", "zigzag"], "band, group or plain"),
    (["-even-order", "sideways"], "down or up"),
    (["-launch-events", "maybe"], "on or off"),
])
def test_kernel_argument_errors(lbm, args, msg):
    geo = lbm.Geometry.from_string("box", (8, 8, 8), (1, 1, 0))
    with pytest.raises(RuntimeError, match=msg):
        lbm.Kernel("list-aa-pv-soa", geo, args)


@pytest.mark.parametrize("args", [["-pad", "auto"], ["-adj", "full"], ["-shadow", "lazy"],
                                  ["-odd-order", "band"], ["-even-order", "up"]])
def test_aa_rejects_list_only_argument(lbm, args):
    geo = lbm.Geometry.from_string("box", (8, 8, 8), (1, 1, 0))
    with pytest.raises(RuntimeError, match="unknown kernel parameter"):
        lbm.Kernel("aa-soa", geo, args)


def test_list_rejects_aa_only_argument(lbm):
    geo = lbm.Geometry.from_string("box", (8, 8, 8), (1, 1, 0))
    with pytest.raises(RuntimeError, match="unknown kernel parameter"):
        lbm.Kernel("list-aa-pv-soa", geo, ["-odd-pairs", "on"])


@pytest.mark.parametrize("args,msg", [
    (["-prec", "fp8"], "fp64, fp32s or fp16s"),
    (["-fp16-scale", "3000"], "power of two"),
    (["-even-cap", "maybe"], "on, off or force"),
    (["-vec", "2"], "auto or 1"),
    (["-odd-pairs", "sometimes"], "auto, on or off"),
])
def test_aa_argument_errors(lbm, args, msg):
    geo = lbm.Geometry.from_string("box", (8, 8, 8), (1, 1, 0))
    with pytest.raises(RuntimeError, match=msg):
        lbm.Kernel("aa-soa", geo, args)


@pytest.mark.gpu
def test_aa_rejects_grid_extents_over_limit(lbm):
    """aa-soa launches y and x as grid dimensions: a y extent of 65536 is
    refused at Init, before any allocation or launch."""
    dims = (1, 65536, 1)
    geo = lbm.Geometry.from_mask(np.ones(65536, dtype=np.int32), dims)
    with pytest.raises(RuntimeError, match="65535"):
        lbm.Kernel("aa-soa", geo)


@pytest.mark.gpu
def test_init_refuses_more_than_2_32_pdfs(lbm):
    """Both kernels refuse a box of more than 2^32 PDFs at Init, before any
    allocation (the reference's check, src/Main.c:399-406): 1024 x 1024 x 217
    all-fluid cells, 19 x 227.5e6 PDFs."""
    dims = (1024, 1024, 217)
    geo = lbm.Geometry.from_mask(np.ones(int(np.prod(dims)), dtype=np.int32), dims, (1, 0, 0))
    for name in ("list-aa-pv-soa", "aa-soa"):
        with pytest.raises(RuntimeError, match=r"exceed 2\^32"):
            lbm.Kernel(name, geo)
    del geo


def test_list_build_no_alloc_refuses_oversized_boxes(lbm):
    """LbmHipListBuildNoAlloc checks the box before it touches the device: a
    box of more than 2^32 PDFs, or a node count whose adjacency does not fit
    32 bits, is refused with the reference's message and *nFluid = 0.  The
    pointers are placeholders: the checks come first, nothing is read."""
    L = lbm.hip_lib()
    L.LbmHipSetErrorMode(lbm.LBM_HIP_ERRORS_RETURN)
    I3 = C.c_int * 3
    n = C.c_int(-1)
    fake = C.c_void_p(4096)
    for dims, n_cells, msg in (((1024, 1024, 256), 1000, "exceed 2^32"),
                               ((1000, 1000, 226), 226_492_416, "out of range")):
        L.LbmHipClearError()
        rc = L.LbmHipListBuildNoAlloc(fake, I3(*dims), I3(1, 0, 0), None, n_cells, 1024, 1000, fake, fake, fake,
                                      C.byref(n))
        assert rc == -1 and n.value == 0, dims
        assert msg in L.LbmHipLastError().decode(), dims


def test_batch_fails_loudly_without_device(lbm):
    if not _no_gpu(lbm):
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError, match="no HIP device"):
        lbm.Batch(["-prec", "fp32s"])


def test_batch_argument_errors(lbm):
    with pytest.raises(RuntimeError, match="fp64, fp32s or fp16s"):
        lbm.Batch(["-prec", "fp8"])
    with pytest.raises(RuntimeError, match="unknown kernel parameter"):
        lbm.Batch(["-foo"])
