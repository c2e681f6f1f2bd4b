"""Pins the oracle's shifted-storage variants (FP32 / scaled FP16 deviations
f - w, collision in FP32).  The reference has no such variant, so they are
pinned against what fixes them otherwise:

  * the binary16 rounding against numpy's float16 conversion (IEEE
    round-to-nearest-even), every tie between adjacent binary16 values
    included;
  * whole runs against the FP64 list oracle (itself pinned bit-exact to the
    reference): the velocity field must agree to the tolerances below, which
    a wrong constant fails (the Lambda check shows the margin);
  * exact properties: the rest state is a fixed point, and in FP32 splitting
    a run into Kernel() calls changes no bit.
"""
import numpy as np
import pytest

from conftest import golden

# velocity relative L2 error vs FP64 after the runs below (measured maxima:
# 5.0e-6 FP32, 4.6e-4 FP16 on these cases)
VEL_TOL = {"fp32s": 2e-5, "fp16s": 5e-3}
MASS_TOL = {"fp32s": 1e-8, "fp16s": 2e-5}
CASES = ["porous16", "channel10x9x7"]
RUNS = [(1.2, 3 / 16, 1e-5, 200), (1.0, 1 / 12, 1e-5, 1000), (1.9, 3 / 16, 1e-4, 400)]


def _case(name):
    g = golden(f"tiny_{name}.npz")
    return g["mask"].astype(np.int32), tuple(int(v) for v in g["dims"]), tuple(int(v) for v in g["periodic"])


def test_f16_round_matches_numpy(oracle):
    rng = np.random.default_rng(7)
    v = rng.standard_normal(20000).astype(np.float32) * (10.0 ** rng.integers(-9, 5, 20000)).astype(np.float32)
    h = np.arange(0, 0x7bff, dtype=np.uint16).view(np.float16).astype(np.float64)
    ties = ((h[:-1] + h[1:]) / 2).astype(np.float32)       # exact in float32
    edge = np.array([65504, 65519.996, 65520, 1e6, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26, 1e-9, 0.0, -0.0,
                     -65520, -2.0 ** -25], dtype=np.float32)
    vals = np.concatenate([v, ties, -ties, edge])
    with np.errstate(over="ignore"):
        want = vals.astype(np.float16).astype(np.float32)
    got = np.array([oracle.f16_round(float(x)) for x in vals], dtype=np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("prec", ["fp32s", "fp16s"])
def test_rest_state_is_fixed_point(oracle, prec):
    mask, dims, per = _case("porous16")
    out, b = oracle.list_run_shifted(mask, dims, per, 20, 1.2, 0.0, prec, magic=3 / 16)
    assert np.array_equal(out, np.tile(oracle.W, (b.n_fluid, 1)))


@pytest.mark.parametrize("prec", ["fp32s", "fp16s"])
@pytest.mark.parametrize("case", CASES)
def test_shifted_tracks_fp64(oracle, case, prec):
    mask, dims, per = _case(case)
    c = oracle.CX.astype(float)
    for omega, magic, F, steps in RUNS:
        ref, b = oracle.list_run(mask, dims, per, steps, omega, F, magic=magic)
        got, _ = oracle.list_run_shifted(mask, dims, per, steps, omega, F, prec, magic=magic, build=b)
        u_ref, u = ref @ c, got @ c
        err = np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref)
        assert err < VEL_TOL[prec], (omega, magic, F, err)
        assert abs(got.sum() - ref.sum()) / ref.sum() < MASS_TOL[prec]
        if case == "porous16" and magic == 3 / 16:
            # the tolerance separates a wrong relaxation constant
            bad, _ = oracle.list_run_shifted(mask, dims, per, steps, omega, F, prec, magic=1 / 12, build=b)
            assert np.linalg.norm(bad @ c - u_ref) / np.linalg.norm(u_ref) > 3 * VEL_TOL[prec]


def test_fp32_split_runs_are_bit_identical(oracle):
    mask, dims, per = _case("porous16")
    whole, b = oracle.list_run_shifted(mask, dims, per, 120, 1.2, 1e-5, "fp32s", magic=3 / 16)
    split, _ = oracle.list_run_shifted(mask, dims, per, [40, 60, 20], 1.2, 1e-5, "fp32s", magic=3 / 16, build=b)
    assert np.array_equal(whole, split)


# aa-soa over shifted storage.  Where no fluid cell touches a non-periodic
# face the aa and list boundary handling coincide (bounce-back and periodic
# wrap only), and the two restatements -- fix-up copies around full-array
# steps against adjacency-list pulls -- must give the same bits.
NO_OPEN = ["box8", "channel10x9x7", "pipe12x10x10", "porous16"]
OPEN = ["fluid6", "blocks2_12"]


@pytest.mark.parametrize("prec", ["fp32s", "fp16s"])
@pytest.mark.parametrize("case", NO_OPEN)
def test_aa_shifted_equals_list_shifted(oracle, case, prec):
    mask, dims, per = _case(case)
    want, _ = oracle.list_run_shifted(mask, dims, per, [60, 40], 1.3, 2e-5, prec, magic=3 / 16)
    got = oracle.aa_run_shifted(mask, dims, per, [60, 40], 1.3, 2e-5, prec, magic=3 / 16)
    assert np.array_equal(got, want)
    off = oracle.aa_run_shifted(mask, dims, per, 30, 1.3, 2e-5, prec, magic=3 / 16, force_on=False)
    want_off, _ = oracle.list_run_shifted(mask, dims, per, 30, 1.3, 2e-5, prec, magic=3 / 16, force_on=False)
    assert np.array_equal(off, want_off)


# Open faces: the ghost layer starts at f = 0 and the cells next to it sit
# far from f = w, so the stored deviations are not small; the node views are
# held to an absolute bound instead (measured maxima 5.4e-7 FP32, 8.2e-5
# FP16), the mass to a looser relative one (measured 1.7e-6, 5.5e-5).  A
# ghost layer started at f = w instead misses by about 0.05.
ABS_TOL = {"fp32s": 1e-6, "fp16s": 1.5e-4}
OPEN_MASS_TOL = {"fp32s": 5e-6, "fp16s": 1.5e-4}


@pytest.mark.parametrize("prec", ["fp32s", "fp16s"])
@pytest.mark.parametrize("case", OPEN + ["porous16"])
def test_aa_shifted_tracks_fp64(oracle, case, prec):
    mask, dims, per = _case(case)
    c = oracle.CX.astype(float)
    for omega, magic, F, steps in RUNS:
        ref, _ = oracle.aa_run(mask, dims, per, steps, omega, F, magic=magic)
        got = oracle.aa_run_shifted(mask, dims, per, steps, omega, F, prec, magic=magic)
        mass_tol = OPEN_MASS_TOL if case in OPEN else MASS_TOL
        assert abs(got.sum() - ref.sum()) / ref.sum() < mass_tol[prec]
        if case in OPEN:
            assert np.abs(got - ref).max() < ABS_TOL[prec], (case, omega, magic, F)
        else:
            err = np.linalg.norm(got @ c - ref @ c) / np.linalg.norm(ref @ c)
            assert err < VEL_TOL[prec], (case, omega, magic, F, err)


def test_aa_shifted_rest_state_and_split(oracle):
    mask, dims, per = _case("porous16")
    for prec in ("fp32s", "fp16s"):
        out = oracle.aa_run_shifted(mask, dims, per, 20, 1.2, 0.0, prec, magic=3 / 16)
        assert np.array_equal(out, np.tile(oracle.W, (out.shape[0], 1)))
    mask, dims, per = _case("blocks2_12")
    whole = oracle.aa_run_shifted(mask, dims, per, 120, 1.2, 1e-5, "fp32s", magic=3 / 16)
    split = oracle.aa_run_shifted(mask, dims, per, [40, 60, 20], 1.2, 1e-5, "fp32s", magic=3 / 16)
    assert np.array_equal(whole, split)
