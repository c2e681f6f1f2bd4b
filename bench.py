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
per fluid node: the list-AA odd-step balance, 19 loads + 19 stores of 8 B
plus 18 int32 adjacency entries = 376 B (doc/main.rst:347), the adjacency
compression not credited; even step 304 B (the two average 340 B/FLUP).
achieved = 376 B * nFluid / mean odd-launch time.  The bytes the kernel
needs with the run-compressed adjacency (304 B plus 72 B per run and 12 B
per 64-node wavefront, reported by the library: 317.8 B here) give the
"moved" figures beside it.  traffic: HBM bytes per odd launch from rocprofv3
FETCH_SIZE + WRITE_SIZE scaled by a copy kernel of known size
(profiles/list_odd_porous256_traffic.json), null if absent; hbm_pair: the
measured bytes of an even + odd pair from the same file over this run's
even + odd launch time (the achieved HBM rate of a step pair).  The copy and
update-19 bandwidths of the device are measured in the same run as
yardsticks (the reference rates its kernels against update-19 bandwidth).

cpu_baseline: the reference's own list-aa-pv-soa (benchmark build, compiled
from its sources into oracle/_ref) on the same geometry, OpenMP on the host
cores, a bounded sample of steps, unpinned and pinned one thread per core,
with the host topology; falls back to the C restatement.

solves / serving (FP64 runs, unless --no-solves): the other half of the
metric, timed in the same run.  solves: 512 seeded porous samples (edges
log-uniform in [32,128]) to steady state in one shifted-FP32 batch against
the plain per-request FP64 path on every sample, setup counted on both sides,
and the FP32 permeability error against those FP64 answers.  serving: a
Poisson stream (500 req/s, edges U[32,192], mixed and all-FP32 tags) through
the continuous-batching server, p50/p99 latency and solves/s against the
plain path queued in arrival order.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
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
    """The reference's own list-aa-pv-soa on the host cores: the benchmark
    build (no forcing, the reference's default) is the baseline value; the
    VERIFICATION build (forcing after every step) and the C restatement
    (oracle, scalar tree) are timed beside it on the same cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    from lbm_checkers import Oracle, Reference

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    mask = np.ascontiguousarray(mask, dtype=np.int32)
    orc = Oracle()
    psteps = max(2, steps // 5 + (steps // 5) % 2)       # the restatement is slower: a smaller sample
    port = n_fluid * psteps / orc.list_run_seconds(mask, dims, periodic, psteps, OMEGA, XFORCE, magic=LAMBDA) / 1e6
    if not Reference.available("bench"):
        return {"value": port, "unit": "MFLUP/s", "cores": threads, "kind": "port",
                "sample": f"C restatement (scalar tree, OpenMP {threads} threads), same 256^3 porous geometry, "
                          f"{psteps} steps, no forcing"}
    t0 = time.time()
    _, _, res = Reference("bench").list_run(mask, dims, periodic, steps, OMEGA, XFORCE, n_threads=threads,
                                            init="weights", want_adj=False)
    wall = time.time() - t0
    value = n_fluid * steps / res.duration / 1e6
    out = {"value": value, "unit": "MFLUP/s", "cores": threads, "kind": "reference",
           "sample": f"reference list-aa-pv-soa benchmark build (gcc -O3 -mavx, OpenMP {threads} threads, no forcing), "
                     f"same 256^3 porous geometry, {steps} steps timed by the reference's own kernel timer "
                     f"(list build + run {wall:.1f} s wall)",
           "port_value": round(port, 1), "port_steps": psteps, "port_over_reference": round(port / value, 3)}
    if Reference.available("verif"):
        _, _, rv = Reference("verif").list_run(mask, dims, periodic, steps, OMEGA, XFORCE, n_threads=threads,
                                               init="weights", want_adj=False)
        out["verification_build_value"] = round(n_fluid * steps / rv.duration / 1e6, 1)
    # the same benchmark build with its threads pinned one per core (a child
    # process: the OpenMP runtime reads the binding when it starts)
    env = dict(os.environ, OMP_PLACES="cores", OMP_PROC_BIND="close", OMP_NUM_THREADS=str(threads))
    try:
        child = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-child", str(steps)],
                               env=env, capture_output=True, text=True, timeout=300)
        out["pinned_value"] = round(json.loads(child.stdout.strip().splitlines()[-1])["value"], 1)
        out["pinned_env"] = "OMP_PLACES=cores OMP_PROC_BIND=close"
    except (subprocess.SubprocessError, ValueError, IndexError, KeyError) as e:
        out["pinned_value"] = None
        out["pinned_error"] = str(e)[:200]
    out["host"] = host_topology()
    return out


def host_topology():
    """CPU model, sockets, physical cores and logical CPUs of the host (from
    /proc/cpuinfo) and the CPUs this process may run on."""
    info = {"logical_cpus": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        cores, sockets, model = set(), set(), None
        phys = core = None
        with open("/proc/cpuinfo") as f:
            for line in f:
                key, _, val = line.partition(":")
                key, val = key.strip(), val.strip()
                if key == "model name" and model is None:
                    model = val
                elif key == "physical id":
                    phys = val
                elif key == "core id":
                    core = val
                elif not key and phys is not None:
                    cores.add((phys, core))
                    sockets.add(phys)
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
            sockets.add(phys)
        info.update({"cpu_model": model, "sockets": len(sockets) or None, "physical_cores": len(cores) or None})
    except OSError:
        pass
    return info


def cpu_baseline_child(steps):
    """Child of cpu_baseline: the reference's benchmark build on the workload,
    its MFLUP/s as one JSON line (threads and binding from the environment)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import lbm_amd as lbm
    from lbm_checkers import Reference

    geo = lbm.Geometry.from_string(GEOMETRY, DIMS)
    mask = np.ascontiguousarray(geo.mask.reshape(-1), dtype=np.int32)
    threads = int(os.environ.get("OMP_NUM_THREADS", "1"))
    sys.stdout.flush()
    keep = os.dup(1)
    os.dup2(2, 1)                      # the C libraries' report lines go to stderr
    _, _, res = Reference("bench").list_run(mask, DIMS, geo.periodic, steps, OMEGA, XFORCE, n_threads=threads,
                                            init="weights", want_adj=False)
    os.write(keep, (json.dumps({"value": geo.n_fluid * steps / res.duration / 1e6}) + "\n").encode())


def _log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def _common(metric, value, unit, steps, workload, extra):
    line = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": 0,
            "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic", "config": {"workload": workload}}
    line.update(extra)
    return line


def _batch_solve(lbm, W, samples, geos, prec, kargs, args):
    """All samples in one batch of storage type prec, to the stopping rule;
    setup (list builds, autotuning) and stepping timed separately."""
    with lbm.Batch(kargs + ["-prec", prec]) as b:
        t0 = time.perf_counter()
        ids = [b.add(g, s.omega, s.xforce, args.max_steps) for g, s in zip(geos, samples)]
        tune = b.autotune(4)
        setup = time.perf_counter() - t0
        _log(f"{prec}: {len(ids)} domains set up and tuned in {setup:.1f} s; even {tune.evenBlock}/cap "
             f"{tune.evenCap}/vec {tune.vec}/nt {tune.nontemporal} odd {tune.oddBlock}/xcd group {tune.oddGroup} "
             f"groups {tune.groupMB} MiB "
             f"(pair {tune.pairMs:.3f} ms)")
        t0 = time.perf_counter()
        active, rounds = len(ids), 0
        while active:
            active = b.step(args.check, args.tol)
            rounds += 1
            if rounds % 10 == 0:
                _log(f"{prec}: {rounds * args.check} steps, {active} active")
        wall = time.perf_counter() - t0
        gpu_s, flups = b.stats()
        _, captures, replays = b.graph_stats()
        st = [b.status(i) for i in ids]
    perm = [W.permeability(x.momentumX, x.nFluid, s.cells, s.omega, s.xforce) for x, s in zip(st, samples)]
    rec = {"solves_per_s": round(len(ids) / (setup + wall), 3), "wall_s": round(setup + wall, 3),
           "run_s": round(wall, 3), "gpu_s": round(gpu_s, 3), "setup_s": round(setup, 2),
           "mflups": round(flups / wall / 1e6, 1), "intervals": rounds,
           "graph_replays": replays, "graph_captures": captures, "converged": sum(x.converged for x in st),
           "steps_median": W.percentile([x.steps for x in st], 50), "steps_max": max(x.steps for x in st),
           "tuning": {"even_block": tune.evenBlock, "even_cap": tune.evenCap, "vec": tune.vec,
                      "odd_block": tune.oddBlock, "even_ms": round(tune.evenMs, 4), "odd_ms": round(tune.oddMs, 4),
                      "group_mib": tune.groupMB, "pair_ms": round(tune.pairMs, 4), "nontemporal": tune.nontemporal,
                      "odd_group": tune.oddGroup}}
    _log(f"{prec}: {rec['solves_per_s']} solves/s ({rec['run_s']} s run + {rec['setup_s']} s setup), "
         f"{rec['mflups']} MFLUP/s, {rec['converged']} converged")
    return rec, perm, [x.steps for x in st]


def _per_request_fp64(lbm, W, samples, geos, kargs, args, stride=1):
    """The plain path: each sample alone through Init + LbmHipRunSteady +
    Deinit, FP64, same stopping rule; Init (list build, allocation) is timed
    with the run, as the batch's setup is."""
    sub = list(range(0, len(samples), max(1, stride)))
    t_all, t_run, perm = 0.0, 0.0, {}
    for i in sub:
        s, g = samples[i], geos[i]
        t0 = time.perf_counter()
        with lbm.Kernel("list-aa-pv-soa", g, kargs + ["-quiet"]) as k:
            k.set_initial_weights()
            k.set_shadow_download(False)
            t1 = time.perf_counter()
            r = k.run_steady(lbm.case_data(omega=s.omega, xforce=s.xforce, iterations=0), args.check, args.tol,
                             args.max_steps)
            t_run += time.perf_counter() - t1
        t_all += time.perf_counter() - t0
        perm[i] = W.permeability(r.momentumX, g.n_fluid, s.cells, s.omega, s.xforce)
    rate = len(sub) / t_all
    _log(f"per-request fp64: {len(sub)} samples in {t_all:.1f} s (Init + run + Deinit; run alone {t_run:.1f} s) "
         f"-> {rate:.2f} solves/s")
    return rate, t_all, perm, t_run


def solves_record(args, lbm):
    """The batched-solve half of the metric on the 512-sample workload: the
    shifted-FP32 batch against the plain per-request FP64 path on every
    sample (setup counted on both sides), and the FP32 permeability error
    against those FP64 answers."""
    import workloads as W

    samples = W.batch_samples(args.samples, args.seed)
    geos = [lbm.Geometry.from_string(s.geometry, s.dims) for s in samples]
    kargs = ["-lambda", repr(LAMBDA)]
    b32, perm32, _ = _batch_solve(lbm, W, samples, geos, "fp32s", kargs, args)
    base, t_base, perm64, t_run = _per_request_fp64(lbm, W, samples, geos, kargs, args)
    err = [abs(perm32[i] - p) / abs(p) for i, p in perm64.items()]
    run_only = len(samples) / b32["run_s"] / (len(perm64) / t_run)
    return {"workload": f"{len(samples)} porous samples (edges log-uniform [32,128], porosity U[0.5,0.9], radius "
                        f"U[3,8], omega U[0.8,1.6], F 1e-6, Lambda 3/16, seed {args.seed}) to relative x-momentum "
                        f"change < {args.tol:g} per {args.check} steps or {args.max_steps} steps",
            "batch_fp32s_solves_per_s": b32["solves_per_s"], "per_request_fp64_solves_per_s": round(base, 3),
            "ratio": round(b32["solves_per_s"] / base, 3),
            "ratio_run_only": round(run_only, 3),
            "ratio_note": "ratio: wall time with setup on both sides (batch: list builds + autotuning; per request: "
                          "Init + initial state + Deinit); ratio_run_only: stepping alone on both sides",
            "permeability_rel_err_fp32s_vs_fp64": {"max": max(err), "median": W.percentile(err, 50)},
            "batch_fp32s": b32, "per_request_fp64_s": round(t_base, 2), "per_request_fp64_run_s": round(t_run, 2)}


def batch_mode(args, lbm):
    """Many seeded porous samples of uneven size solved to steady state (or the
    step limit) in one batch, FP64 and shifted FP32, against the plain
    per-request FP64 path on every sample (setup counted on both sides)."""
    import workloads as W

    samples = W.batch_samples(args.samples, args.seed)
    t0 = time.perf_counter()
    geos = [lbm.Geometry.from_string(s.geometry, s.dims) for s in samples]
    _log(f"{len(samples)} samples generated in {time.perf_counter() - t0:.1f} s, "
         f"{sum(g.n_fluid for g in geos) / 1e6:.1f} M fluid nodes")
    kargs = ["-lambda", repr(LAMBDA)]
    res, perms = {}, {}
    for prec in ("fp64", "fp32s"):
        res[prec], perms[prec], _ = _batch_solve(lbm, W, samples, geos, prec, kargs, args)
    err = [abs(a - b) / abs(b) for a, b in zip(perms["fp32s"], perms["fp64"])]
    base, t_base, _, t_run = _per_request_fp64(lbm, W, samples, geos, kargs, args, args.baseline_stride)
    v = res["fp32s"]["solves_per_s"]
    return _common(
        f"solves/s ({len(samples)} porous samples, batched list-AA, shifted FP32)", v, "solves/s", args.max_steps,
        f"{len(samples)} porous samples, edges log-uniform in [32,128], porosity U[0.5,0.9], sphere radius U[3,8], "
        f"omega U[0.8,1.6], F 1e-6, Lambda 3/16, seed {args.seed}; to relative x-momentum change < {args.tol:g} "
        f"per {args.check} steps or {args.max_steps} steps",
        {"batch": res,
         "per_request_fp64": {"solves_per_s": round(base, 3), "samples_timed": len(range(0, len(samples),
                                                                                          args.baseline_stride)),
                              "seconds": round(t_base, 2), "run_seconds": round(t_run, 2),
                              "note": "every %d-th sample alone through Init + LbmHipRunSteady + Deinit"
                                      % args.baseline_stride},
         "speedup_fp32s_batch_vs_per_request_fp64": round(v / base, 2),
         "speedup_fp64_batch_vs_per_request_fp64": round(res["fp64"]["solves_per_s"] / base, 2),
         "permeability_rel_err_fp32s_vs_fp64": {"max": max(err), "p99": W.percentile(err, 99),
                                                "median": W.percentile(err, 50)}})


def precision_mode(args, lbm):
    """Steady states in FP64, shifted FP32 and shifted FP16, velocity-field
    and permeability error against FP64 of the same kernel:
      * the list-AA workload itself (256^3 porous, list-aa-pv-soa, omega
        1.2, F 1e-6) until the relative x-momentum change per check is
        below 1e-7 -- FP16 stops at 1.5x the FP64 step count if it does not
        get there (its rounding floor);
      * the 64^3 channel (x periodic, walls on y and z) with aa-soa (the
        small configuration's kernel) and with list-aa-pv-soa."""
    import numpy as np
    import workloads as W

    tol = 1e-7
    cx = np.array([0, 0, 1, -1, 1, 1, -1, -1, 0, 0, 1, -1, 0, 0, 0, 0, -1, 1, 0], dtype=float)
    pgeo = lbm.Geometry.from_string(GEOMETRY, DIMS)
    porous, pref = {}, None
    for prec in ("fp64", "fp32s", "fp16s"):
        cap = args.porous_max_steps if pref is None else int(1.5 * pref[2]) + args.check
        with lbm.Kernel("list-aa-pv-soa", pgeo, ["-lambda", repr(LAMBDA), "-prec", prec, "-quiet"]) as k:
            k.set_initial_weights()
            r = k.run_steady(lbm.case_data(omega=OMEGA, xforce=XFORCE, iterations=0), args.check, tol, cap)
            u = k.dump() @ cx - XFORCE / 2
        perm = W.permeability(r.momentumX, pgeo.n_fluid, int(np.prod(DIMS)), OMEGA, XFORCE)
        rec = {"steps": r.steps, "converged": r.converged, "rel_change": r.relChange,
               "mflups": round(pgeo.n_fluid * r.steps / r.seconds / 1e6, 1), "permeability": perm,
               "mass_rel_drift": abs(r.mass - pgeo.n_fluid) / pgeo.n_fluid}
        if pref is None:
            pref = (u, perm, r.steps)
        else:
            rec["velocity_rel_l2_vs_fp64"] = float(np.linalg.norm(u - pref[0]) / np.linalg.norm(pref[0]))
            rec["permeability_rel_err_vs_fp64"] = abs(perm - pref[1]) / abs(pref[1])
        del u
        porous[prec] = rec
        _log(f"list-aa-pv-soa 256^3 porous {prec}: {rec}")
    del pref

    dims, omega, F = (64, 64, 64), 1.0, 1e-5
    geo = lbm.Geometry.from_string("channel", dims, (1, 0, 0))   # -periodic-x, as the reference's driver takes it
    out, fields = {}, {}
    for kernel in ("aa-soa", "list-aa-pv-soa"):
        out[kernel], ref = {}, None
        for prec in ("fp64", "fp32s", "fp16s"):
            with lbm.Kernel(kernel, geo, ["-lambda", repr(LAMBDA), "-prec", prec, "-quiet"]) as k:
                k.set_initial_weights()
                r = k.run_steady(lbm.case_data(omega=omega, xforce=F, iterations=0), args.check, tol, 200000)
                d = k.dump()
            u = d @ cx - F / 2
            fields[(kernel, prec)] = u
            perm = W.permeability(r.momentumX, geo.n_fluid, int(np.prod(dims)), omega, F)
            rec = {"steps": r.steps, "converged": r.converged, "rel_change": r.relChange,
                   "mflups": round(geo.n_fluid * r.steps / r.seconds / 1e6, 1), "permeability": perm,
                   "mass_rel_drift": abs(r.mass - geo.n_fluid) / geo.n_fluid}
            if ref is None:
                ref = (u, perm)
            else:
                rec["velocity_rel_l2_vs_fp64"] = float(np.linalg.norm(u - ref[0]) / np.linalg.norm(ref[0]))
                rec["permeability_rel_err_vs_fp64"] = abs(perm - ref[1]) / abs(ref[1])
            out[kernel][prec] = rec
            _log(f"{kernel} {prec}: {rec}")
    # no open faces here: both kernels reach the same bits in every storage type
    out["aa_equals_list"] = {p: bool(np.array_equal(fields[("aa-soa", p)], fields[("list-aa-pv-soa", p)]))
                             for p in ("fp64", "fp32s", "fp16s")}
    _log(f"aa-soa == list-aa-pv-soa velocity fields: {out['aa_equals_list']}")
    out["porous256_list"] = porous
    return _common("velocity rel. L2 vs FP64 at steady state (256^3 porous, list-aa-pv-soa, shifted FP32)",
                   porous["fp32s"]["velocity_rel_l2_vs_fp64"], "rel", porous["fp64"]["steps"],
                   f"list-aa-pv-soa, {GEOMETRY} {DIMS[0]}^3, omega {OMEGA}, Lambda 3/16, F {XFORCE:g}, to relative "
                   f"x-momentum change < {tol:g} per {args.check} steps; also aa-soa and list-aa-pv-soa on the 64^3 "
                   "channel (x periodic, walls y/z), omega 1.0, F 1e-5",
                   {"higher_is_better": False, "precision": out})


def serving_record(lbm, n=16, rate=500.0, seed=2025, tol=1e-6, interval=100, max_steps=20000, budget=200e9):
    """The serving half of the metric on a short stream of the configured
    shape: n Poisson arrivals at `rate`, porous cubes of edge U[32,192],
    FP64 or shifted-FP32 tag, through the continuous-batching server under
    `budget` bytes; against the plain path run on the same arrival clock
    (each request alone in FP64, in arrival order)."""
    import serving as S

    samples = S.serving_samples(n, seed)
    kargs = ["-lambda", repr(LAMBDA)]
    arr = S.poisson_arrivals(n, rate, seed + int(rate))
    reqs = [S.Request(i, a, s, p) for i, (a, (s, p)) in enumerate(zip(arr, samples))]
    solver = S.PlainGpuSolver(lbm, kargs, interval, tol, max_steps)
    plain_recs = S.serve_plain(reqs, solver)
    plain = S.summarize(plain_recs, solver.node_updates)
    perm64 = {r.rid: r.permeability for r in plain_recs}
    _log(f"plain fp64 path: {plain}")
    be = S.GpuBackend(lbm, kargs, max_steps)
    try:
        recs = S.serve(reqs, be, int(budget), interval, tol)
        served = S.summarize(recs, be.flups())
    finally:
        be.close()
    err = [abs(r.permeability - perm64[r.rid]) / abs(perm64[r.rid]) for r in recs if r.prec == "fp32s" and not r.rejected]
    served = {k: (float(v) if hasattr(v, "item") else v) for k, v in served.items()}
    served["permeability_rel_err_fp32s_vs_fp64_max"] = max(err) if err else None
    # the same stream with every request tagged shifted FP32
    reqs32 = [S.Request(i, a, s, "fp32s") for i, (a, (s, _)) in enumerate(zip(arr, samples))]
    be = S.GpuBackend(lbm, kargs, max_steps)
    try:
        recs32 = S.serve(reqs32, be, int(budget), interval, tol)
        served32 = S.summarize(recs32, be.flups())
    finally:
        be.close()
    served32 = {k: (float(v) if hasattr(v, "item") else v) for k, v in served32.items()}
    served32["permeability_rel_err_vs_fp64_max"] = max(abs(r.permeability - perm64[r.rid]) / abs(perm64[r.rid])
                                                       for r in recs32)
    plain = {k: (float(v) if hasattr(v, "item") else v) for k, v in plain.items()}
    _log(f"serving: mixed {served} | all fp32s {served32} | plain {plain}")
    return {"workload": f"{n} requests, Poisson {rate:g}/s (seed {seed}), porous cubes of edge U[32,192], "
                        f"porosity U[0.5,0.9], radius U[3,8], omega U[0.8,1.6], F 1e-6, FP64 or shifted-FP32 tag, "
                        f"to relative x-momentum change < {tol:g} per {interval} steps, {budget / 1e9:g} GB budget; "
                        f"plain path run on the same arrival clock, one request at a time in FP64",
            "served": served, "served_all_fp32s": served32, "plain_fp64": plain,
            "solves_speedup": round(served["solves_per_s"] / plain["solves_per_s"], 3),
            "solves_speedup_all_fp32s": round(served32["solves_per_s"] / plain["solves_per_s"], 3)}


def serve_mode(args, lbm):
    """Poisson streams of permeability requests (porous cubes, edge U[32,192],
    FP64 or shifted-FP32 tag) through the continuous-batching server, one
    stream per rate; the plain path solves the same requests one at a time in
    FP64 on the same arrival clock."""
    import serving as S

    rates = [float(r) for r in args.rates.split(",")]
    samples = S.serving_samples(args.requests, args.seed)
    kargs = ["-lambda", repr(LAMBDA), "-pad", "16384+128"]
    max_steps, tol, interval = 20000, args.tol, args.check
    out, perm64 = {}, {}
    for rate in rates:
        arr = S.poisson_arrivals(len(samples), rate, args.seed + int(rate))
        reqs = [S.Request(i, a, s, p) for i, (a, (s, p)) in enumerate(zip(arr, samples))]
        solver = S.PlainGpuSolver(lbm, kargs, interval, tol, max_steps)
        plain_recs = S.serve_plain(reqs, solver)
        plain = S.summarize(plain_recs, solver.node_updates)
        perm64 = {r.rid: r.permeability for r in plain_recs}
        be = S.GpuBackend(lbm, kargs, max_steps)
        try:
            recs = S.serve(reqs, be, int(args.budget_gb * 1e9), interval, tol)
            summ = S.summarize(recs, be.flups())
        finally:
            be.close()
        err = [abs(r.permeability - perm64[r.rid]) / abs(perm64[r.rid]) for r in recs
               if r.prec == "fp32s" and not r.rejected]
        summ["permeability_rel_err_fp32s_vs_fp64_max"] = max(err) if err else None
        summ["fp32s_requests"] = len(err)
        out[str(int(rate))] = {"served": summ, "plain_fp64": plain,
                               "solves_speedup": round(summ["solves_per_s"] / plain["solves_per_s"], 2)}
        _log(f"rate {rate:g}/s: served {summ} | plain {plain}")
    # the same requests at the top rate, all tagged shifted-FP32
    rate = max(rates)
    arr = S.poisson_arrivals(len(samples), rate, args.seed + int(rate))
    reqs = [S.Request(i, a, s, "fp32s") for i, (a, (s, _)) in enumerate(zip(arr, samples))]
    be = S.GpuBackend(lbm, kargs, max_steps)
    try:
        recs = S.serve(reqs, be, int(args.budget_gb * 1e9), interval, tol)
        summ = S.summarize(recs, be.flups())
    finally:
        be.close()
    err = [abs(r.permeability - perm64[r.rid]) / abs(perm64[r.rid]) for r in recs]
    summ["permeability_rel_err_vs_fp64_max"] = max(err)
    plain = out[str(int(rate))]["plain_fp64"]
    out["fp32s_only_at_top_rate"] = {"served": summ, "solves_speedup": round(summ["solves_per_s"] / plain["solves_per_s"], 2)}
    _log(f"rate {rate:g}/s, all fp32s: {summ}")
    top = out[str(int(max(rates)))]
    return _common(f"solves/s (served permeability requests, Poisson {max(rates):g}/s)",
                   top["served"]["solves_per_s"], "solves/s", max_steps,
                   f"{len(samples)} requests per rate: porous cubes, edge U[32,192], porosity U[0.5,0.9], sphere "
                   f"radius U[3,8], omega U[0.8,1.6], F 1e-6, FP64 or shifted-FP32 tag (50/50), to relative x-momentum "
                   f"change < {tol:g} per {interval} steps (at most {max_steps}); device budget {args.budget_gb:g} GB",
                   {"p99_latency_s": top["served"]["p99_latency_s"], "serving": out})


def aa_mode(args, lbm):
    """aa-soa (full-array AA kernel) on the duct channel: 64^3 x 1000 steps
    (the small configuration the reference also runs on its CPU) and 256^3,
    the latter also with shifted FP32 / FP16 storage.  Throughput from a run
    without per-launch events, even/odd launch averages from a second run
    with them."""
    out = {}
    for edge, steps, prec in ((64, 1000, "fp64"), (256, 200, "fp64"), (256, 200, "fp32s"), (256, 200, "fp16s")):
        geo = lbm.Geometry.from_string("channel", (edge, edge, edge), (1, 0, 0))
        kargs = ["-lambda", repr(LAMBDA), "-quiet", "-prec", prec] + args.kernel_args.split()
        with lbm.Kernel("aa-soa", geo, kargs) as k:
            k.set_initial_weights()
            k.set_shadow_download(False)
            k.run(lbm.case_data(omega=1.0, xforce=1e-5, iterations=20))
            k.run(lbm.case_data(omega=1.0, xforce=1e-5, iterations=steps))
            loop_ms = k.timings().loopMs
            k.set_launch_events(True)
            k.run(lbm.case_data(omega=1.0, xforce=1e-5, iterations=steps))
            t = k.timings()
        nf = geo.n_fluid
        pdf_b = 2 * 19 * {"fp64": 8, "fp32s": 4, "fp16s": 2}[prec]
        be, bo = pdf_b + 1 / 8, pdf_b + 4   # PDFs plus the fluid bit (even) or the cell's info word (odd)
        key = f"{edge}^3" + ("" if prec == "fp64" else f" {prec}")
        out[key] = {"n_fluid": nf, "steps": steps, "mflups": round(nf * steps / (loop_ms / 1e3) / 1e6, 1),
                    "mflups_with_launch_events": round(nf * steps / (t.loopMs / 1e3) / 1e6, 1),
                    "even_ms": round(t.evenMsAvg, 5), "odd_ms": round(t.oddMsAvg, 5),
                    "even_gbs": round(be * nf / (t.evenMsAvg / 1e3) / 1e9, 1),
                    "odd_gbs": round(bo * nf / (t.oddMsAvg / 1e3) / 1e9, 1)}
        _log(f"aa-soa {key}: {out[key]}")
    return _common("MFLUP/s (D3Q19 TRT aa-soa FP64, 256^3 duct channel)", out["256^3"]["mflups"], "MFLUP/s", 200,
                   "aa-soa, channel (x periodic, walls y/z), omega 1.0, Lambda 3/16, F 1e-5, FP64, f = w at start",
                   {"dtype": "f64", "aa": out})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--prec", choices=["fp64", "fp32s", "fp16s"], default="fp64",
                    help="PDF storage: fp64 (the headline), or deviations f - w in FP32 / scaled FP16")
    ap.add_argument("--mode", choices=["list", "batch", "precision", "serve", "aa"], default="list",
                    help="list: the 256^3 headline; batch: many porous samples in one batch; "
                         "precision: the 64^3 channel to steady state in every storage type; "
                         "serve: Poisson request streams through the continuous-batching server")
    ap.add_argument("--kernel-args", default="", help="extra kernel arguments (list and aa modes), space separated")
    ap.add_argument("--rates", default="10,50,200,500", help="serve: arrival rates (requests/s)")
    ap.add_argument("--requests", type=int, default=32, help="serve: requests per rate")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="serve: device memory budget (GB)")
    ap.add_argument("--samples", type=int, default=512, help="batch: number of porous samples")
    ap.add_argument("--seed", type=int, default=2024, help="batch: sample generator seed")
    ap.add_argument("--max-steps", type=int, default=5000, help="batch: step limit per sample")
    ap.add_argument("--check", type=int, default=100, help="batch/precision: steps between convergence checks")
    ap.add_argument("--porous-max-steps", type=int, default=60000,
                    help="precision: step limit of the FP64 256^3 porous steady state")
    ap.add_argument("--tol", type=float, default=1e-6, help="batch: relative change of the x momentum per check")
    ap.add_argument("--baseline-stride", type=int, default=1,
                    help="batch: every n-th sample is also solved alone (the per-request FP64 path)")
    ap.add_argument("--cpu-baseline-child", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--no-solves", action="store_true",
                    help="list: leave out the batched-solve and serving records (512-sample batch and a Poisson "
                         "request stream, each against the per-request FP64 path)")
    args = ap.parse_args()
    if args.cpu_baseline_child:
        cpu_baseline_child(args.cpu_baseline_child)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU (its share of the lattice at the real shape); "
              "runs on several GPUs are not part of this build.", file=sys.stderr)
        sys.exit(2)

    # the C libraries print their reports on stdout; keep stdout for the JSON line
    sys.stdout.flush()
    json_fd = os.dup(1)
    os.dup2(2, 1)

    import lbm_amd as lbm

    if args.mode != "list":
        line = {"batch": batch_mode, "precision": precision_mode, "serve": serve_mode, "aa": aa_mode}[args.mode](args, lbm)
        os.dup2(json_fd, 1)
        print(json.dumps(line), flush=True)
        return

    steps = args.steps + (args.steps & 1)
    warm = args.warmup + (args.warmup & 1)
    geo = lbm.Geometry.from_string(GEOMETRY, DIMS)
    nf = geo.n_fluid
    print(f"bench: {GEOMETRY} {DIMS} nFluid={nf} steps={steps} warmup={warm}", file=sys.stderr, flush=True)
    k = lbm.Kernel("list-aa-pv-soa", geo, ["-lambda", repr(LAMBDA), "-prec", args.prec] + args.kernel_args.split())
    info = k.list_info()
    b_even = info.bytesPerFlupEven
    k.set_initial_weights()
    k.set_shadow_download(False)
    k.set_launch_events(True)       # even/odd launch averages for the roofline, over the timed steps
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
    b_moved = info.bytesPerFlupOdd              # with the run-compressed adjacency
    b_odd = b_even + 18 * 4                     # list-AA odd balance: full int32 adjacency
    odd_gbs = b_odd * nf / (t.oddMsAvg / 1e3) / 1e9
    moved_gbs = b_moved * nf / (t.oddMsAvg / 1e3) / 1e9
    even_gbs = b_even * nf / (t.evenMsAvg / 1e3) / 1e9
    print(f"bench: {value:.1f} MFLUP/s loop {t.loopMs:.2f} ms wall {wall * 1e3:.2f} ms even {t.evenMsAvg:.4f} ms "
          f"({even_gbs:.0f} GB/s) odd {t.oddMsAvg:.4f} ms ({odd_gbs:.0f} GB/s at {b_odd} B, {moved_gbs:.0f} GB/s "
          f"at {b_moved:.1f} B) runs={info.nRuns}", file=sys.stderr, flush=True)
    mask = geo.mask.reshape(-1).copy()
    periodic = geo.periodic
    runs_wave = info.nRunsWave
    k.close()
    yard = lbm.stream_bench(1 << 27, nf, reps=10)
    print(f"bench: yardsticks copy {yard['copy_gbs']:.0f} GB/s update-19 {yard['update19_gbs']:.0f} GB/s",
          file=sys.stderr, flush=True)

    traffic, pair = None, None
    if args.prec == "fp64" and os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            tf = json.load(f)
        traffic = tf.get("bytes_per_launch")
        if "even" in tf and "odd" in tf:
            # measured HBM bytes of an even + odd pair over this run's pair time
            pb = tf["even"]["bytes_per_launch"] + tf["odd"]["bytes_per_launch"]
            pair_gbs = pb / ((t.evenMsAvg + t.oddMsAvg) / 1e3) / 1e9
            pair = {"bytes": pb, "achieved_gbs": round(pair_gbs, 1), "frac": round(pair_gbs / PEAK_GBS, 4)}

    base = None
    if not args.no_cpu_baseline:
        print("bench: cpu baseline ...", file=sys.stderr, flush=True)
        base = cpu_baseline(mask, DIMS, periodic, nf, args.cpu_steps + (args.cpu_steps & 1))

    solves = serving = None
    if not args.no_solves and args.prec == "fp64":
        print("bench: batched solves ...", file=sys.stderr, flush=True)
        solves = solves_record(args, lbm)
        print("bench: served stream ...", file=sys.stderr, flush=True)
        serving = serving_record(lbm)

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
                     "scope": "frac is the dominant kernel's (the odd step); step_average_frac is the whole step's",
                     "achieved": round(odd_gbs, 1), "peak": PEAK_GBS, "unit": "GB/s",
                     "frac": round(odd_gbs / PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT) if traffic is not None else None,
                     "hbm_pair": pair,
                     "bytes_per_flup": b_odd,
                     "bytes_basis": "list-AA odd balance: 19 loads + 19 stores + 18 int32 adjacency entries per "
                                    "node, adjacency compression not credited",
                     "moved": {"bytes_per_flup": round(b_moved, 3), "achieved": round(moved_gbs, 1),
                               "frac": round(moved_gbs / PEAK_GBS, 4)},
                     "odd_ms_avg": t.oddMsAvg, "runs_per_wave_cut": runs_wave,
                     "even_kernel": {"achieved": round(even_gbs, 1), "frac": round(even_gbs / PEAK_GBS, 4),
                                     "bytes_per_flup": b_even, "ms_avg": t.evenMsAvg},
                     "step_average_gbs": round((b_even + b_odd) / 2 * nf * steps / loop_s / 1e9, 1),
                     "yardsticks_gbs": {"copy": round(yard["copy_gbs"], 1), "update19": round(yard["update19_gbs"], 1)},
                     "moved_frac_of_copy": round(moved_gbs / yard["copy_gbs"], 4),
                     "step_average_frac": round((b_even + b_odd) / 2 * nf * steps / loop_s / 1e9 / PEAK_GBS, 4)},
        "cpu_baseline": base,
        "solves": solves,
        "serving": serving,
    }
    os.dup2(json_fd, 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
