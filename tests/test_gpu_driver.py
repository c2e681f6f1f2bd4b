"""The reference's own test protocol (src/test-verification.sh:52-112): run
every kernel with -verify and require exit code 0, i.e. the analytic-profile
deviation <= 0.1 -- here through the lbmbenchk-hip driver on the GPU."""
import os
import subprocess

import pytest

from conftest import PKG

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(PKG, "lbmbenchk-hip")


@pytest.mark.parametrize("kernel", ["list-aa-pv-soa", "aa-soa"])
@pytest.mark.parametrize("kargs", [[], ["--", "-blk", "7"]])
def test_verify(tmp_path, kernel, kargs):
    # -blk is a loop-blocking hint of the reference's CPU kernels (both take it); results do not change
    print(f"driver -verify -kernel {kernel} {' '.join(kargs)}", flush=True)
    r = subprocess.run([DRIVER, "-verify", "-kernel", kernel, *kargs], capture_output=True, text=True,
                       cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "# VERIFICATION: deviation from analytical solution: 4.561374e-02" in r.stdout
    assert "VERIFICATION SUCCEEDED" in r.stdout
    p_line = [line for line in r.stdout.splitlines() if line.startswith("P:")]
    assert len(p_line) == 1 and "MFLUP/s" in p_line[0]
    assert (tmp_path / "flow-profile.dat").exists()
    assert (tmp_path / "density-ux.dat").exists()


@pytest.mark.parametrize("kernel", ["list-aa-pv-soa", "aa-soa"])
@pytest.mark.parametrize("prec", ["fp32s", "fp16s"])
def test_verify_shifted_storage(tmp_path, kernel, prec):
    # the reference's verification also passes with shifted FP32 / FP16 storage
    print(f"driver -verify -kernel {kernel} -- -prec {prec}", flush=True)
    r = subprocess.run([DRIVER, "-verify", "-kernel", kernel, "--", "-prec", prec], capture_output=True, text=True,
                       cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "VERIFICATION SUCCEEDED" in r.stdout
    dev = float(next(ln for ln in r.stdout.splitlines() if "deviation from analytical" in ln).split()[-1])
    # FP64: 4.561374e-02; FP32 storage within 1e-6 of it, FP16 (measured 3.42e-02) below the 0.1 bar
    assert abs(dev - 4.561374e-02) < 1e-6 if prec == "fp32s" else 0.02 < dev < 0.05
    p_line = next(ln for ln in r.stdout.splitlines() if ln.startswith("P:"))
    assert p_line.rstrip().endswith(prec)


def test_driver_porous_run(tmp_path):
    r = subprocess.run([DRIVER, "-dims", "64x64x64", "-geometry", "porous-4-0.8-1", "-iterations", "100",
                        "-omega", "1.2", "-x-force", "1e-6", "-kernel", "list-aa-pv-soa"],
                       capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "porous medium" in r.stdout


def test_verify_with_run_time_output(tmp_path):
    # snapshots and statistics during the run leave the result unchanged
    r = subprocess.run([DRIVER, "-verify", "-kernel", "list-aa-pv-soa", "-vtk", "500", "-statistics", "100"],
                       capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "# VERIFICATION: deviation from analytical solution: 4.561374e-02" in r.stdout
    for it in ("-001", "0000", "0500", "1000"):
        assert (tmp_path / "vtk" / f"file-{it}.vtk").exists(), it
    rows = [ln.split() for ln in open(tmp_path / "ux-progress.dat") if not ln.startswith("#")]
    iters = [int(x[0]) for x in rows]
    assert iters[0] == 0 and iters[-1] == 1000 and 500 in iters and 990 in iters
    aa = tmp_path / "aa"
    aa.mkdir()
    r = subprocess.run([DRIVER, "-verify", "-kernel", "aa-soa", "-vtk", "500", "-statistics", "100"],
                       capture_output=True, text=True, cwd=aa, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "# VERIFICATION: deviation from analytical solution: 4.561374e-02" in r.stdout
    for it in ("-001", "0000", "0500", "1000"):
        assert (aa / "vtk" / f"file-{it}.vtk").exists(), it
    iters = [int(ln.split()[0]) for ln in open(aa / "ux-progress.dat") if not ln.startswith("#")]
    assert iters[0] == 0 and 500 in iters and 990 in iters
