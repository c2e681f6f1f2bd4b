"""Host side on CPU: geometry generators against the oracle and the live
reference, the porous generator's contract, and the driver's command line
and error paths (which exit before touching a device), and bench.py's
refusal of runs on several GPUs."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import PKG

DRIVER = os.path.join(PKG, "lbmbenchk-hip")

CASES = [("box", 0, (9, 8, 7), (1, 1, 0)), ("channel", 1, (10, 6, 5), (0, 0, 0)),
         ("channel", 1, (10, 6, 5), (1, 0, 0)), ("pipe", 2, (8, 12, 10), (0, 0, 0)),
         ("blocks-2", 3, (12, 10, 9), (0, 0, 0)), ("blocks", 3, (16, 16, 16), (0, 1, 0)),
         ("fluid", 4, (5, 5, 5), (0, 0, 0))]


@pytest.mark.parametrize("gstr,gtype,dims,per", CASES)
def test_geometry_matches_oracle(lbm, oracle, gstr, gtype, dims, per):
    geo = lbm.Geometry.from_string(gstr, dims, per)
    block = int(gstr.split("-")[1]) if "-" in gstr else 8
    want, _ = oracle.geometry(gtype, dims, per, block=block)
    assert np.array_equal(geo.mask.reshape(-1), want)
    assert geo.periodic == tuple(per)                 # stored flags are the requested ones
    assert geo.n_fluid == int(np.count_nonzero(want))


@pytest.mark.parametrize("gstr,gtype,dims,per", CASES)
def test_geometry_matches_reference(lbm, reference, gstr, gtype, dims, per):
    geo = lbm.Geometry.from_string(gstr, dims, per)
    want, ldp = reference.geometry(gstr, dims, per)
    assert np.array_equal(geo.mask.reshape(-1), want)
    assert geo.periodic == ldp


def test_porous_contract(lbm):
    g = lbm.Geometry.from_string("porous-4-0.75-11", (40, 32, 24))
    m = g.mask
    X, Y, Z = g.dims
    assert g.periodic == (1, 0, 0)
    assert (m[:, 0, :] == 0).all() and (m[:, -1, :] == 0).all()     # y walls
    assert (m[0] == 0).all() and (m[-1] == 0).all()                 # z walls
    por = g.n_fluid / (X * Y * Z)
    assert por <= 0.75 and por > 0.70                                # stops right after crossing
    # deterministic and seed-dependent
    g2 = lbm.Geometry.from_string("porous-4-0.75-11", (40, 32, 24))
    g3 = lbm.Geometry.from_string("porous-4-0.75-12", (40, 32, 24))
    assert np.array_equal(g.mask, g2.mask)
    assert not np.array_equal(g.mask, g3.mask)


def test_porous_spheres_are_spheres(lbm):
    # one sphere: porosity target just below the empty channel's
    X = Y = Z = 30
    empty = lbm.Geometry.from_string("channel", (X, Y, Z), (1, 0, 0)).n_fluid
    target = (empty - 1) / (X * Y * Z)
    g = lbm.Geometry.from_string(f"porous-5-{target!r}-3", (X, Y, Z))
    base = lbm.Geometry.from_string("channel", (X, Y, Z), (1, 0, 0)).mask
    solid = (base != 0) & (g.mask == 0)
    zs, ys, xs = np.nonzero(solid)
    # all newly solid cells lie within radius 5 of one point (x periodic)
    assert len(xs) >= 1
    cx = np.angle(np.exp(2j * np.pi * xs / X).mean()) % (2 * np.pi) * X / (2 * np.pi)
    dx = (xs - cx + X / 2) % X - X / 2
    r = np.sqrt(dx ** 2 + (ys - ys.mean()) ** 2 + (zs - zs.mean()) ** 2)
    assert r.max() <= 5.0 + 1.0


def _driver(*args, cwd=None):
    return subprocess.run([DRIVER, *args], capture_output=True, text=True, cwd=cwd, timeout=60)


def test_driver_list():
    r = _driver("-list")
    assert r.returncode == 0
    assert "list-aa-pv-soa" in r.stdout and "aa-soa" in r.stdout


@pytest.mark.parametrize("args,msg", [
    (["-kernel", "push-soa"], 'requested kernel "push-soa" not found'),
    (["-iterations", "0"], "number of iterations must be > 0"),
    (["-dims", "4x4"], "invalid dimensions"),
    (["-dims"], "requires a parameter"),
    (["-bogus"], "unknown parameter"),
    (["-geometry", "sphere"], "unknown geometry"),
    (["-geometry", "blocks-9", "-dims", "16x16x16"], "block size for geometry"),
    (["-dims", "2048x2048x1024"], "number of PDFs exceed"),
])
def test_driver_errors(tmp_path, args, msg):
    r = _driver(*args, cwd=tmp_path)
    assert r.returncode == 1
    assert msg in r.stdout


def test_driver_without_device(tmp_path, lbm):
    if lbm.device_info() is not None:
        pytest.skip("a GPU is present")
    r = _driver("-verify", cwd=tmp_path)
    assert r.returncode == 1
    assert "no HIP device" in r.stdout


def test_mask_outlives_geometry(lbm):
    # the mask view keeps its memory alive after the Geometry object is gone
    import gc
    m = lbm.Geometry.from_string("channel", (12, 10, 8), (1, 0, 0)).mask
    gc.collect()
    junk = [lbm.Geometry.from_string("box", (12, 10, 8)) for _ in range(20)]   # reuse freed memory
    want = lbm.Geometry.from_string("channel", (12, 10, 8), (1, 0, 0)).mask
    assert np.array_equal(m, want) and m.shape == (8, 10, 12)
    del junk


@pytest.mark.parametrize("argv,env", [(["--gpus", "2"], {}), (["--gpus", "8"], {}), ([], {"WORLD_SIZE": "2"})])
def test_bench_refuses_several_gpus(argv, env):
    """bench.py measures one GPU: asked for more it exits with status 2 and a
    message, before it touches a device, and prints no JSON line."""
    root = os.path.dirname(PKG)
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py")] + argv, capture_output=True, text=True,
                       env={**os.environ, **env}, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""
