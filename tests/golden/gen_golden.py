"""Writes the golden fixtures of tests/golden/ from the reference build.

Runs the reference's own list-aa-pv-soa and aa-soa kernels (compiled from
its unmodified sources into oracle/_ref/ by oracle/Makefile, driven by the
glue in oracle/ref_harness_*.c) on small geometries and on the two
full-size workloads (64^3 channel, 256^3 porous medium), and stores inputs
and outputs:

  tiny_<case>.npz   full node views (nFluid x 19, mask order) after 100
                    steps, list build (adjacency node-major, grid, coords)
  verify16.npz      the reference's -verify case: SHA-256 of the node views,
                    every 13th node, density, deviation norm
  channel64.npz     64^3 channel (periodic x), omega 1.0, F 1e-5, 1000 steps:
                    SHA-256 + every 97th node + mass, both kernels
  porous256.npz     256^3 porous medium (porous-6-0.8-42): mask SHA-256,
                    adjacency SHA-256 + sampled rows, node views after 20
                    steps at omega 1.2, F 1e-6 (SHA-256 + samples + mass)
  porous256_500.npz the same lattice after 500 steps (omega 1.2, F 1e-6):
                    the reference's kernel at its Lambda 1/12 (8 threads)
                    and the oracle's scalar tree at Lambda 1/12 and 3/16,
                    each as SHA-256, every 997th node, mass and the sums of
                    every direction over every x plane; plus the oracle's
                    largest relative deviation from the reference over the
                    whole field
  verify64_12000.npz the reference's analytic check at the size where it
                    passes (64^3 -verify box, 12000 steps,
                    omega 1.0, F 1e-5, KernelSetInitialDensity/Velocity):
                    both kernels' deviation from the Poiseuille profile,
                    SHA-256 of the node views, every 97th node

Usage: python tests/golden/gen_golden.py [--skip-large] [--only NAME] [--tiny NAME ...]
Needs oracle/_ref (built from the reference sources) and liblbmhost.so
(for the porous generator).  The reference is never imported here as
Python; it is compiled C called through ctypes.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "lbm-benchmark-kernels_amd"))
from lbm_checkers import Oracle, Reference  # noqa: E402
import lbm_amd  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

TINY = [
    # name, geometry string, dims, periodic flags, list kernel args
    ("box8", "box", (8, 8, 8), (1, 1, 0), ()),
    ("box8_blk3", "box", (8, 8, 8), (1, 1, 0), ("-blk", "3")),
    ("channel10x9x7", "channel", (10, 9, 7), (1, 0, 0), ()),
    ("channel10x9x7_blk", "channel", (10, 9, 7), (1, 0, 0), ("-blk-x", "4", "-blk-z", "3")),
    ("blocks2_12", "blocks-2", (12, 12, 12), (0, 0, 0), ()),
    ("pipe12x10x10", "pipe", (12, 10, 10), (1, 0, 0), ()),
    ("porous16", "porous-2-0.7-7", (16, 14, 12), (0, 0, 0), ()),
    ("fluid6", "fluid", (6, 5, 4), (0, 0, 0), ()),
    # the reference's -pad strings: node counts and adjacency values follow them
    ("porous16_pad", "porous-2-0.7-7", (16, 14, 12), (0, 0, 0), ("-pad", "64+8,1024+128")),
    ("channel10x9x7_padno_blk3", "channel", (10, 9, 7), (1, 0, 0), ("-pad", "no", "-blk", "3")),
]


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def geometry(gstr, dims, periodic):
    g = lbm_amd.Geometry.from_string(gstr, dims, periodic)
    return np.array(g.mask.reshape(-1), dtype=np.int32), g.periodic, g.n_fluid


def x_plane_sums(dump, mask, nx):
    """Sum of every direction over the fluid nodes of every x plane
    (nx x 19), from a node-view dump in mask order."""
    xs = np.flatnonzero(mask) % nx
    out = np.zeros((nx, dump.shape[1]))
    for d in range(dump.shape[1]):
        out[:, d] = np.bincount(xs, weights=dump[:, d], minlength=nx)
    return out


def porous256_500(ref, manifest):
    dims, steps, omega, force, every = (256, 256, 256), 500, 1.2, 1e-6, 997
    mask, ldp, nf = geometry("porous-6-0.8-42", dims, (0, 0, 0))
    t0 = time.time()
    out = {"mask_sha": sha(mask), "n_fluid": nf, "steps": steps, "omega": omega, "xforce": force,
           "sample_every": every}
    dref, _, rl = ref.list_run(mask, dims, ldp, steps, omega, force, n_threads=8, want_adj=False)
    orc = Oracle()
    for tag, lam in (("l12", 1.0 / 12.0), ("l316", 3.0 / 16.0)):
        d, _ = orc.list_run(mask, dims, ldp, steps, omega, force, magic=lam, n_threads=8, vsize=1)
        out[f"orc_{tag}_sha"] = sha(d)
        out[f"orc_{tag}_sample"] = d[::every]
        out[f"orc_{tag}_mass"] = float(d.sum())
        out[f"orc_{tag}_xsums"] = x_plane_sums(d, mask, dims[0])
        if tag == "l12":
            out["orc_vs_ref_relmax"] = float(np.abs(d - dref).max() / np.abs(dref).max())
        del d
    out["ref_sha"] = sha(dref)
    out["ref_sample"] = dref[::every]
    out["ref_mass"] = float(dref.sum())
    out["ref_xsums"] = x_plane_sums(dref, mask, dims[0])
    out["mass0"] = float(nf)       # f = w: one unit of mass per node
    np.savez_compressed(os.path.join(OUT, "porous256_500.npz"), **out)
    manifest["porous256_500"] = {"nFluid": nf, "steps": steps, "orc_vs_ref_relmax": out["orc_vs_ref_relmax"],
                                 "secs": time.time() - t0}
    print("porous256_500", nf, out["orc_vs_ref_relmax"], time.time() - t0, flush=True)


def verify64(ref, manifest):
    dims, steps = (64, 64, 64), 12000
    mask, ldp, nf = geometry("box", dims, (1, 1, 0))
    t0 = time.time()
    dl, _, rl = ref.list_run(mask, dims, ldp, steps, 1.0, 1e-5, n_threads=8, init="verify", verify=True,
                             want_adj=False)
    da, ra = ref.aa_run(mask, dims, ldp, steps, 1.0, 1e-5, n_threads=8, init="verify", verify=True)
    np.savez_compressed(os.path.join(OUT, "verify64_12000.npz"), mask_sha=sha(mask), n_fluid=nf, steps=steps,
                        list_t8_sha=sha(dl), list_t8_sample=dl[::97], list_norm=rl.verifyNorm,
                        aa_sha=sha(da), aa_sample=da[::97], aa_norm=ra.verifyNorm)
    manifest["verify64_12000"] = {"nFluid": nf, "list_norm": rl.verifyNorm, "aa_norm": ra.verifyNorm,
                                  "secs": time.time() - t0}
    print("verify64_12000", rl.verifyNorm, ra.verifyNorm, time.time() - t0, flush=True)


def tiny(ref, manifest, name, gstr, dims, per, kargs):
    """A tiny case: full node dumps of the reference's list kernel (1 and 3
    threads) and aa-soa after 100 steps, with its list build."""
    mask, ldp, nf = geometry(gstr, dims, per)
    steps, omega, force = 100, 1.3, 1e-5
    dl1, b, r1 = ref.list_run(mask, dims, ldp, steps, omega, force, n_threads=1, kernel_args=kargs)
    dl3, _, r3 = ref.list_run(mask, dims, ldp, steps, omega, force, n_threads=3, kernel_args=kargs)
    da, ra = ref.aa_run(mask, dims, ldp, steps, omega, force, n_threads=2)
    np.savez_compressed(os.path.join(OUT, f"tiny_{name}.npz"),
                        mask=mask.astype(np.int8), dims=np.array(dims), periodic=np.array(ldp),
                        kernel_args=np.array(list(kargs) or [""]), steps=steps, omega=omega, xforce=force,
                        list_t1=dl1, list_t3=dl3, aa_t2=da, adj=b.adj, grid=b.grid, coords=b.coords,
                        n_cells=r1.nCells, n_bb=ra.nBounceBack, runs_t1=r1.nConsecNodes,
                        runs_t3=r3.nConsecNodes)
    manifest[f"tiny_{name}"] = {"geometry": gstr, "dims": dims, "periodic": ldp, "nFluid": nf,
                                "nCells": r1.nCells}
    print("tiny", name, nf, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--skip-large", action="store_true")
    ap.add_argument("--only", choices=["porous256_500", "verify64_12000"])
    ap.add_argument("--tiny", action="append", choices=[t[0] for t in TINY],
                    help="write only this tiny case (repeatable)")
    args = ap.parse_args()
    ref = Reference("verif")
    with open(os.path.join(OUT, "manifest.json")) as f:
        manifest = json.load(f)
    cwd = os.getcwd()
    tmp = tempfile.mkdtemp()
    os.chdir(tmp)          # the reference's verification writes flow-profile.dat
    try:
        if args.only:
            {"porous256_500": porous256_500, "verify64_12000": verify64}[args.only](ref, manifest)
            write_manifest(manifest)
            return
        if args.tiny:
            for entry in TINY:
                if entry[0] in args.tiny:
                    tiny(ref, manifest, *entry)
            write_manifest(manifest)
            return
        for entry in TINY:
            tiny(ref, manifest, *entry)

        # the reference's -verify preset (src/Main.c:232-265)
        dims = (16, 16, 16)
        mask, ldp, nf = geometry("box", dims, (1, 1, 0))
        dl, _, rl = ref.list_run(mask, dims, ldp, 1000, 1.0, 1e-5, n_threads=4, init="verify", verify=True)
        da, ra = ref.aa_run(mask, dims, ldp, 1000, 1.0, 1e-5, n_threads=3, init="verify", verify=True)
        np.savez_compressed(os.path.join(OUT, "verify16.npz"), list_t4_sha=sha(dl), aa_sha=sha(da),
                            list_t4_sample=dl[::13], aa_sample=da[::13], list_norm=rl.verifyNorm,
                            aa_norm=ra.verifyNorm, list_density=rl.density, aa_density=ra.density)
        manifest["verify16"] = {"list_norm": rl.verifyNorm, "aa_norm": ra.verifyNorm}
        print("verify16", rl.verifyNorm, ra.verifyNorm, flush=True)

        if not args.skip_large:
            dims = (64, 64, 64)
            mask, ldp, nf = geometry("channel", dims, (1, 0, 0))
            t0 = time.time()
            da, ra = ref.aa_run(mask, dims, ldp, 1000, 1.0, 1e-5, n_threads=8)
            dl, _, rl = ref.list_run(mask, dims, ldp, 1000, 1.0, 1e-5, n_threads=8, want_adj=False)
            np.savez_compressed(os.path.join(OUT, "channel64.npz"), mask_sha=sha(mask), n_fluid=nf,
                                aa_sha=sha(da), aa_sample=da[::97], aa_mass=float(da.sum()),
                                list_t8_sha=sha(dl), list_t8_sample=dl[::97], list_mass=float(dl.sum()),
                                aa_duration=ra.duration, list_duration=rl.duration)
            manifest["channel64"] = {"nFluid": nf, "aa_mass": float(da.sum()), "secs": time.time() - t0}
            print("channel64", nf, time.time() - t0, flush=True)
            del da, dl

            dims = (256, 256, 256)
            mask, ldp, nf = geometry("porous-6-0.8-42", dims, (0, 0, 0))
            t0 = time.time()
            dl, b, rl = ref.list_run(mask, dims, ldp, 20, 1.2, 1e-6, n_threads=8)
            every = 9973
            np.savez_compressed(os.path.join(OUT, "porous256.npz"), mask_sha=sha(mask), n_fluid=nf,
                                periodic=np.array(ldp), n_cells=rl.nCells, adj_sha=sha(b.adj),
                                adj_sample=b.adj[::every], coords_sample=b.coords[::every],
                                grid_sha=sha(np.where(mask != 0, b.grid, 0xffffffff).astype(np.uint32)),
                                list_t8_sha=sha(dl), list_t8_sample=dl[::every], list_mass=float(dl.sum()),
                                runs_t8=rl.nConsecNodes, duration=rl.duration, sample_every=every)
            manifest["porous256"] = {"nFluid": nf, "nCells": rl.nCells, "runs_t8": rl.nConsecNodes,
                                     "secs": time.time() - t0}
            print("porous256", nf, rl.nCells, time.time() - t0, flush=True)
            del dl, b
            porous256_500(ref, manifest)
            verify64(ref, manifest)
    finally:
        os.chdir(cwd)
    write_manifest(manifest)


def write_manifest(manifest):
    with open(os.path.join(OUT, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, default=lambda o: list(o) if isinstance(o, tuple) else str(o))


if __name__ == "__main__":
    main()
