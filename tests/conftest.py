"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on CPU."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "lbm-benchmark-kernels_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")

# Degenerate all-fluid lattices (dims, periodic): each extent of 1 or 2 once,
# open faces, and periodic wrap onto the cell itself.
DEGENERATE_SHAPES = [((1, 4, 5), (1, 0, 0)), ((3, 1, 4), (0, 1, 0)), ((2, 3, 1), (0, 0, 1)), ((1, 1, 6), (1, 1, 0)),
                     ((2, 2, 2), (0, 0, 0)), ((1, 1, 1), (1, 1, 1)), ((5, 1, 1), (0, 0, 0))]
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def lbm():
    import lbm_amd
    return lbm_amd


@pytest.fixture(scope="session")
def oracle():
    from lbm_checkers import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from lbm_checkers import Reference
    if not Reference.available("verif"):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    return Reference("verif")


def golden(name: str):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip(f"fixture {name} not generated")
    return np.load(path, allow_pickle=False)


def tiny_cases():
    return sorted(f[5:-4] for f in os.listdir(GOLDEN) if f.startswith("tiny_") and f.endswith(".npz"))


def blk_args(g):
    """Kernel arguments stored in a tiny fixture (node-numbering blocks,
    padding)."""
    return [str(a) for a in g["kernel_args"] if str(a)]


def blk_tuple(args):
    """The (bx, by, bz) of -blk / -blk-x/y/z words, 0 = unblocked."""
    blk = [0, 0, 0]
    it = iter(args)
    for a in it:
        v = next(it)
        if a == "-blk":
            blk = [int(v)] * 3
        elif a.startswith("-blk-"):
            blk["xyz".index(a[-1])] = int(v)
    return tuple(blk)


def pad_arg(args):
    """The -pad string among the words, or None (the reference's auto)."""
    it = iter(args)
    for a in it:
        v = next(it)
        if a == "-pad":
            return v
    return None
