"""Synthetic workloads of the batched and serving forms: seeded porous
samples and the quantities a permeability solve reports.

A sample is a porous cube (or box) of solid spheres as lbm_amd.Geometry
builds it ("porous-<radius>-<porosity>-<seed>": x periodic, y/z faces
solid), with its own omega and the body force F along x.  Permeability from
Darcy's law, k = nu * U / F, U the superficial velocity (x momentum over all
cells) and nu = (1/omega - 1/2) / 3 the lattice viscosity.  The node views
hold the force increment of the last step (the reference adds it after every
step, src/Kernel.c:759-803), so the velocity at the half step is
u = sum_d c_d f_d - F/2; with it, plane Poiseuille flow gives the analytic
permeability (tests/test_serving.py).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class Sample:
    index: int
    dims: tuple
    radius: float
    porosity: float
    seed: int
    omega: float
    xforce: float

    @property
    def geometry(self) -> str:
        return f"porous-{self.radius:.3f}-{self.porosity:.4f}-{self.seed}"

    @property
    def cells(self) -> int:
        return int(np.prod(self.dims))


def log_uniform_edges(rng, lo, hi, n=3):
    return tuple(int(round(math.exp(v))) for v in rng.uniform(math.log(lo), math.log(hi), n))


def batch_samples(n=512, seed=2024, edge=(32, 128), porosity=(0.5, 0.9), radius=(3.0, 8.0),
                  omega=(0.8, 1.6), xforce=1e-6):
    """n samples: edge lengths log-uniform per axis, porosity, sphere radius
    and omega uniform, from one seeded generator."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        dims = log_uniform_edges(rng, *edge)
        out.append(Sample(i, dims, float(rng.uniform(*radius)), float(rng.uniform(*porosity)),
                          int(rng.integers(1, 2 ** 31 - 1)), float(rng.uniform(*omega)), xforce))
    return out


def viscosity(omega: float) -> float:
    return (1.0 / omega - 0.5) / 3.0


def permeability(momentum_x: float, n_fluid: int, cells: int, omega: float, xforce: float) -> float:
    """Darcy permeability from the x momentum sum over the fluid nodes."""
    u_sup = (momentum_x - 0.5 * xforce * n_fluid) / cells
    return viscosity(omega) * u_sup / xforce


def percentile(values, q):
    v = sorted(values)
    if not v:
        return float("nan")
    k = (len(v) - 1) * q / 100.0
    lo, hi = math.floor(k), math.ceil(k)
    return v[lo] + (v[hi] - v[lo]) * (k - lo)
