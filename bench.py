"""Benchmark: D3Q19 TRT list-AA (list-aa-pv-soa) on a 256^3 porous medium, FP64.

Workload (the list-AA FP64 configuration the headline MFLUP/s is quoted on):
256x256x256 cells, solid spheres of radius 6 with centres from a SplitMix64
stream (seed 42) until porosity <= 0.80, x periodic, y/z faces solid;
omega 1.2, TRT Lambda 3/16, Luo body force 1e-6 in x, initial state
f_i = w_i (rho 1, u 0).  Synthetic input, generated here.

A "step" is one LBM timestep over all fluid nodes.  The lattice, list and
adjacency are built on and resident in HBM before timing; the timed region
is exactly --steps timesteps (rounded up to even: AA steps come in
even/odd pairs), measured with HIP events recorded on the kernel's stream
around the step loop and around every launch.  value = nFluid * steps /
time / 1e6 (src/Main.c:594).

--prec fp32s / fp16s run the shifted-storage variants (deviations f - w in
FP32 or scaled binary16, collision in FP32) on the same workload: 4 or 2 B
per value instead of 8; the default and headline is FP64.

roofline: the odd-step kernel (the longer of the two).  Algorithmic bytes
per fluid node: 19 loads + 19 stores of 8 B = 304 B plus the adjacency it
reads -- run-compressed by default: 72 B per run of consecutive adjacency
plus 12 B per 64-node wavefront (the library reports the exact figure);
376 B with -adj full.  Even step: 304 B.  achieved = bytes/node * nFluid /
mean odd-launch time.  traffic: HBM bytes per odd launch from rocprofv3
FETCH_SIZE + WRITE_SIZE scaled by a copy kernel of known size
(profiles/list_odd_porous256_traffic.json), null if absent.  The copy and
update-19 bandwidths of the device are measured in the same run as
yardsticks (the reference rates its kernels against update-19 bandwidth).

cpu_baseline: the reference's own list-aa-pv-soa (benchmark build, compiled
from its sources into oracle/_ref) on the same geometry, OpenMP on the host
cores, a bounded sample of steps; falls back to the C restatement.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lbm-benchmark-kernels_amd"))

GEOMETRY = "porous-6-0.8-42"
DIMS = (256, 256, 256)
OMEGA, LAMBDA, XFORCE = 1.2, 0.1875, 1e-6
PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "list_odd_porous256_traffic.json")


def cpu_baseline(mask, dims, periodic, n_fluid, steps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    from lbm_checkers import Oracle, Reference

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    mask = np.ascontiguousarray(mask, dtype=np.int32)
    if Reference.available("bench"):
        ref = Reference("bench")
        t0 = time.time()
        _, _, res = ref.list_run(mask, dims, periodic, steps, OMEGA, XFORCE, n_threads=threads,
                                 init="weights", want_adj=False)
        wall = time.time() - t0
        return {"value": n_fluid * steps / res.duration / 1e6, "unit": "MFLUP/s", "cores": threads,
                "kind": "reference",
                "sample": f"reference list-aa-pv-soa benchmark build (gcc -O3 -mavx, OpenMP {threads} threads), "
                          f"same 256^3 porous geometry, {steps} steps timed by the reference's own kernel timer "
                          f"(list build + run {wall:.1f} s wall)"}
    orc = Oracle()
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    t0 = time.time()
    orc.list_run(mask, dims, periodic, steps, OMEGA, XFORCE, magic=LAMBDA, force_on=False, vsize=1)
    dt = time.time() - t0
    return {"value": n_fluid * steps / dt / 1e6, "unit": "MFLUP/s", "cores": threads, "kind": "port",
            "sample": f"C restatement, {steps} steps incl. list build, OpenMP {threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--prec", choices=["fp64", "fp32s", "fp16s"], default="fp64",
                    help="PDF storage: fp64 (the headline), or deviations f - w in FP32 / scaled FP16")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU (its share of the lattice at the real shape); "
              "runs on several GPUs are not part of this build.", file=sys.stderr)
        sys.exit(2)

    # the C libraries print their reports on stdout; keep stdout for the JSON line
    sys.stdout.flush()
    json_fd = os.dup(1)
    os.dup2(2, 1)

    import lbm_amd as lbm

    steps = args.steps + (args.steps & 1)
    warm = args.warmup + (args.warmup & 1)
    geo = lbm.Geometry.from_string(GEOMETRY, DIMS)
    nf = geo.n_fluid
    print(f"bench: {GEOMETRY} {DIMS} nFluid={nf} steps={steps} warmup={warm}", file=sys.stderr, flush=True)
    k = lbm.Kernel("list-aa-pv-soa", geo, ["-lambda", repr(LAMBDA), "-prec", args.prec])
    info = k.list_info()
    b_even = info.bytesPerFlupEven
    k.set_initial_weights()
    k.set_shadow_download(False)
    cd = lbm.case_data(omega=OMEGA, xforce=XFORCE, iterations=warm)
    if warm:
        k.run(cd)
    cd.MaxIterations = steps
    t0 = time.perf_counter()
    k.run(cd)
    wall = time.perf_counter() - t0
    t = k.timings()
    loop_s = t.loopMs / 1e3
    value = nf * steps / loop_s / 1e6
    b_odd = info.bytesPerFlupOdd
    odd_gbs = b_odd * nf / (t.oddMsAvg / 1e3) / 1e9
    even_gbs = b_even * nf / (t.evenMsAvg / 1e3) / 1e9
    print(f"bench: {value:.1f} MFLUP/s loop {t.loopMs:.2f} ms wall {wall * 1e3:.2f} ms even {t.evenMsAvg:.4f} ms "
          f"({even_gbs:.0f} GB/s) odd {t.oddMsAvg:.4f} ms ({odd_gbs:.0f} GB/s) runs={info.nRuns}",
          file=sys.stderr, flush=True)
    mask = geo.mask.reshape(-1).copy()
    periodic = geo.periodic
    runs_wave = info.nRunsWave
    k.close()
    yard = lbm.stream_bench(1 << 27, nf, reps=10)
    print(f"bench: yardsticks copy {yard['copy_gbs']:.0f} GB/s update-19 {yard['update19_gbs']:.0f} GB/s",
          file=sys.stderr, flush=True)

    traffic = None
    if args.prec == "fp64" and os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("bytes_per_launch")

    base = None
    if not args.no_cpu_baseline:
        print("bench: cpu baseline ...", file=sys.stderr, flush=True)
        base = cpu_baseline(mask, DIMS, periodic, nf, args.cpu_steps + (args.cpu_steps & 1))

    dev = lbm.device_info() or {}
    label = {"fp64": "FP64", "fp32s": "FP32 shifted storage", "fp16s": "FP16 shifted storage, FP32 compute"}[args.prec]
    line = {
        "metric": f"MFLUP/s (D3Q19 TRT list-AA {label}, 256^3 porous)",
        "value": round(value, 2),
        "unit": "MFLUP/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": t.loopMs / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32s": "f32", "fp16s": "f16 storage / f32 compute"}[args.prec],
        "data": "synthetic",
        "config": {"workload": "list-aa-pv-soa, 256x256x256 porous medium (spheres r=6, porosity<=0.80, seed 42), "
                               f"x periodic, omega 1.2, Lambda 3/16, Luo force 1e-6, {label}",
                   "n_fluid": nf, "n_cells_padded": info.nCells, "device": dev.get("name")},
        "roofline": {"bound": "hbm", "kernel": "k_list_odd (odd AA step, run-compressed adjacency)",
                     "achieved": round(odd_gbs, 1), "peak": PEAK_GBS, "unit": "GB/s",
                     "frac": round(odd_gbs / PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_flup": round(b_odd, 3), "odd_ms_avg": t.oddMsAvg, "runs_per_wave_cut": runs_wave,
                     "even_kernel": {"achieved": round(even_gbs, 1), "frac": round(even_gbs / PEAK_GBS, 4),
                                     "bytes_per_flup": b_even, "ms_avg": t.evenMsAvg},
                     "step_average_gbs": round((b_even + b_odd) / 2 * nf * steps / loop_s / 1e9, 1),
                     "yardsticks_gbs": {"copy": round(yard["copy_gbs"], 1), "update19": round(yard["update19_gbs"], 1)},
                     "frac_of_copy": round(odd_gbs / yard["copy_gbs"], 4),
                     "list_aa_340B_equivalent_frac": round(340 * nf * steps / loop_s / 1e9 / PEAK_GBS, 4)},
        "cpu_baseline": base,
    }
    os.dup2(json_fd, 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
