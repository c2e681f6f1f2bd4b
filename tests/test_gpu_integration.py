"""The drop-in binding exercised for real: the reference's own driver.

oracle/Makefile (target ``integration``) compiles the reference's src/Main.c
and every other object of its driver, unmodified, and registers the two
kernels of liblbmhip.so in the driver's table g_kernels[] as
``list-aa-pv-soa-hip`` and ``aa-soa-hip`` (oracle/hip_kernel_table.h,
force-included into Main.c), linked against liblbmhip.so.  These tests run
that binary the way the reference's own test protocol does
(src/test-verification.sh:52-112): ``-verify -kernel K -t T -pin 0..T-1`` for
T = 1..5, and again with ``-- -blk 7`` when ``-kernel K -- -h`` lists a
blocking parameter; each run must exit 0 with the reference's deviation
from the analytic channel profile (src/Main.c:631-642).
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "lbmbenchk-ref-hip")
HIP_KERNELS = ["list-aa-pv-soa-hip", "aa-soa-hip"]
# KernelVerifiy's deviation for the -verify preset (16^3 box, 1000 steps),
# the reference CPU kernels' own value (src/Kernel.c:511-625)
DEVIATION = "4.561374e-02"

needs_driver = pytest.mark.skipif(not os.path.exists(DRIVER), reason="reference driver not built (oracle/Makefile integration)")


def _run(args, cwd, timeout=120):
    p = subprocess.run([DRIVER, *args], cwd=cwd, capture_output=True, text=True, timeout=timeout)
    return p.returncode, p.stdout + p.stderr


@needs_driver
def test_reference_driver_lists_hip_kernels(tmp_path):
    rc, out = _run(["-list"], tmp_path)
    assert rc == 0, out
    names = out.split("Available kernels to benchmark:")[1].split()
    # the reference's table first, in its order, then the two registered ones
    assert names[0] == "list-aa-pv-soa" and "aa-soa" in names
    assert names[-2:] == HIP_KERNELS


@pytest.mark.gpu
@needs_driver
@pytest.mark.parametrize("kernel", HIP_KERNELS)
def test_reference_verification_protocol(kernel, tmp_path):
    rc, usage = _run(["-kernel", kernel, "--", "-h"], tmp_path)
    start = usage.find("Kernel parameters:")
    blocking = start >= 0 and "-blk" in usage[start:]
    assert blocking, usage
    for threads in range(1, 6):
        pin = ",".join(str(t) for t in range(threads))
        for kargs in ([], ["--", "-blk", "7"]):
            args = ["-verify", "-kernel", kernel, "-t", str(threads), "-pin", pin, *kargs]
            print("reference driver", " ".join(args), flush=True)
            rc, out = _run(args, tmp_path)
            assert rc == 0, out
            m = re.search(r"deviation from analytical solution: (\S+)", out)
            assert m and m.group(1) == DEVIATION, out
            assert "# VERIFICATION SUCCEEDED." in out
            assert f"kernel: {kernel}" in out


@pytest.mark.gpu
@needs_driver
def test_reference_driver_hip_matches_reference_kernel(tmp_path):
    # the same binary, the same -verify case: the reference's CPU kernel and
    # the registered GPU kernel print the same deviation and density line
    outs = {}
    for kernel in ("list-aa-pv-soa", "list-aa-pv-soa-hip"):
        rc, out = _run(["-verify", "-kernel", kernel, "-t", "2"], tmp_path)
        assert rc == 0, out
        outs[kernel] = out
    dev = {k: re.search(r"deviation from analytical solution: (\S+)", o).group(1) for k, o in outs.items()}
    assert dev["list-aa-pv-soa"] == dev["list-aa-pv-soa-hip"] == DEVIATION
    # the driver's KernelStatisticsAdv prints the mean density line
    dens = {k: re.findall(r"density.*", o) for k, o in outs.items()}
    assert dens["list-aa-pv-soa"] and dens["list-aa-pv-soa"] == dens["list-aa-pv-soa-hip"]
