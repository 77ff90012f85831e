#!/usr/bin/env python3
"""Swarm log-likelihood throughput on one MI355X.

Workload: methanol-like training set, N=2000 points, d=12 ALF features
(5 distances, 4 angles, 3 dihedrals), kernel rbf on the 9 distances/angles +
periodic on the 3 dihedrals, constant mean, marginal likelihood, nugget 1e-10,
swarm of 50 particles with theta ~ U[0, 1).  One step = one swarm evaluation
(50 log-likelihoods: covariance build, batched Cholesky, likelihood) through
the C ABI's batched path, theta uploaded and ll / INFO read back every step as
a PSO iteration needs them.

Prints one JSON line (rank 0).  The CPU baseline times the reference's host
algorithm (precomputed pair tensors, dgemv + exp, LAPACK dpotrf / dpotrs from
scipy's OpenBLAS, OpenMP over particles) on a bounded sample of the same
workload.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

# Kernel arguments in device memory: the ~40 dependent launches of one
# evaluation start sooner (+1.2-1.8 % evaluations/s at the default workload,
# three interleaved runs).  Read when the HIP runtime initialises, so set
# before torch or the library touch the GPU; a value already in the
# environment is kept.
os.environ.setdefault("HIP_FORCE_DEV_KERNARG", "1")

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "ferebus-v7_amd"))
sys.path.insert(0, str(REPO))



# committed rocprofv3 summaries, per factorisation path (PROFILES: the
# methanol-like N=2000 swarm-50 workload on the three-launch lock-step
# schedule, FB_CHAIN=0; CHAINED_PROFILES: the default two-launch schedule,
# keyed on (workload, path, stream groups) at the default sizes)
CHAINED_PROFILES = {
    ("methanol", "lockstep-kc16-per-panel", 1): (REPO / "profiles" / "chained_factor_per_panel_methanol_n2000_s50.md",
                                                 "lockstep_factor"),
    ("methanol", "lockstep-kc16-per-panel", 2): (REPO / "profiles" /
                                                 "chained_factor_per_panel_groups2_methanol_n2000_s50.md",
                                                 "lockstep_factor"),
    ("methanol", "lockstep-kc16", 2): (REPO / "profiles" / "chained_factor_groups2_methanol_n2000_s50.md",
                                       "lockstep_factor"),
    ("ibuprofen", "lockstep-kc16-per-panel", 2): (REPO / "profiles" / "ibuprofen_build_and_factor_n6000_s64.md",
                                                  "lockstep_factor"),
    ("ibuprofen", "lockstep-kc16", 2): (REPO / "profiles" / "ibuprofen_chained_groups2_n6000_s64.md", "lockstep_factor"),
    ("nma", "lockstep-kc16-per-panel", 2): (REPO / "profiles" / "nma_build_and_factor_n3000_s32.md", "lockstep_factor"),
    ("water", "lockstep-kc16", 1): (REPO / "profiles" / "water_build_and_factor_n500_s20.md", "lockstep_factor"),
}
PROFILES = {
    "persistent": (REPO / "profiles" / "dag_factor_kernel_methanol_n2000_s50.md", "dag_factor_kernel"),
    "lockstep-kc16": (REPO / "profiles" / "lockstep_factor_methanol_n2000_s50.md", "lockstep_factor"),
    "lockstep-kc16-per-panel": (REPO / "profiles" / "lockstep_factor_per_panel_methanol_n2000_s50.md",
                                "lockstep_factor"),
}


def profiled_traffic(path: str, chained: bool = False, groups: int = 1, workload: str = "methanol"):
    """(HBM bytes per factorisation, source) from the committed rocprofv3
    counter summary of that path (FETCH_SIZE + WRITE_SIZE, copy-kernel
    calibrated), or (None, None)."""
    if chained:
        table, key_ = CHAINED_PROFILES, (workload, path, groups)
    elif workload == "methanol":
        table, key_ = PROFILES, path
    else:
        return None, None
    if key_ not in table:
        return None, None
    f, key = table[key_]
    try:
        for line in f.read_text().splitlines():
            if line.startswith(f"traffic_bytes_per_launch {key}:"):
                return float(line.split(":")[1]), str(f.relative_to(REPO))
    except OSError:
        pass
    return None, None


# --workload: (synthetic generator, default N, default swarm, seed, config text)
WORKLOADS = {
    "methanol": ("methanol", 2000, 50, 0, "methanol-like N={n} d={d} swarm={S} rbf(9)+periodic(3) constant mean, "
                                          "marginal, nugget 1e-10, theta~U[0,1)"),
    "water": ("water", 500, 20, 0, "water-O1-like N={n} d={d} swarm={S} rbf[1,2]+periodic[3] constant mean, "
                                   "marginal, nugget 1e-10, theta~U[1e-3,1)"),
    "nma": ("nma", 3000, 32, 3, "NMA-like N={n} d={d} swarm={S} ARD rbf(30)+periodic(9) constant mean, marginal, "
                                "optimised nugget~U[1e-12,1e-4), theta~U[0,1)"),
    "ibuprofen": ("ibuprofen", 6000, 64, 4, "ibuprofen-like N={n} d={d} swarm={S} rbf(63)+periodic(30) linear mean, "
                                            "concentrated, nugget 1e-10, theta~U[0,1)"),
}


def potrf_flops(n: int) -> float:
    """LAPACK dpotrf operation count."""
    return n ** 3 / 3.0 + n ** 2 / 2.0 + n / 6.0


def uneven_run(steps: int = 3):
    """Batched uneven training sets: 10 glycine-like GPs (d=24, N ~ U[1000,
    4000]), 40 particles each -- 400 factorisations of mixed size per PSO
    iteration, evaluated by one fb_multi_swarm_loglik call.  Timed four
    ways: the call as autotuned (its first call times the schedules below
    and keeps the fastest), the merged chained schedule pinned
    (FB_MULTI_SCHEDULE=2: every GP's matrices in each lock-step launch, one
    hipGraph per iteration), the persistent single launch pinned (=1), and
    the ten handles' own evaluations called one after another; the results
    are bit-identical."""
    import torch
    from ferebus_amd import MultiSwarm, SwarmLikelihood, compile_kernel, fit_mean, synthetic
    ws = synthetic.glycine_gps(seed=2)
    rng = np.random.default_rng(9)
    gps, thetas = [], []
    for w in ws:
        prog = compile_kernel(w.kernel, w.kernels, w.x.shape[1])
        gps.append(SwarmLikelihood(w.x, w.y, prog, mean=fit_mean(w.mean, w.x, w.y), noise=w.noise,
                                   likelihood=w.likelihood, max_swarm=40))
        thetas.append(rng.uniform(0.0, 1.0, (40, gps[-1].nparams)))
    multi = MultiSwarm(gps)
    pinned = {}
    for sched in ("2", "1"):
        os.environ["FB_MULTI_SCHEDULE"] = sched
        try:
            pinned[sched] = MultiSwarm(gps)
        finally:
            del os.environ["FB_MULTI_SCHEDULE"]
    multi.swarm_loglik(thetas)               # tunes every handle and the schedule
    for mm in pinned.values():
        mm.swarm_loglik(thetas)

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            out = fn()
        return (time.perf_counter() - t) / steps, out

    t_multi, (lls, infos) = timed(lambda: multi.swarm_loglik(thetas))
    t_merged, (llm, _) = timed(lambda: pinned["2"].swarm_loglik(thetas))
    t_one, (lls1, _) = timed(lambda: pinned["1"].swarm_loglik(thetas))
    t_seq, _ = timed(lambda: [gp.swarm_loglik(th) for gp, th in zip(gps, thetas)])
    same = all(np.array_equal(gp.swarm_loglik(th)[0], ll, equal_nan=True) and
               np.array_equal(ll, l1, equal_nan=True) and np.array_equal(ll, l2, equal_nan=True)
               for gp, th, ll, l1, l2 in zip(gps, thetas, lls, lls1, llm))
    # one PSO iteration of the ten swarms (selection, update, evaluation):
    # one fb_multi_pso_step against each swarm's own fb_pso_step in turn
    from ferebus_amd.pso import DeviceSwarm
    prng = np.random.default_rng(5)
    sws = [DeviceSwarm(gp, 40, 0.789, 1.494, 1.494) for gp in gps]
    for sw, th in zip(sws, thetas):
        sw.init(th, prng.uniform(-0.05, 0.05, th.shape))
    rs = [(prng.random(th.shape), prng.random(th.shape)) for th in thetas]
    pso_steps = {g: (sws[g],) + rs[g] for g in range(len(gps))}
    multi.pso_step(pso_steps)                # captures the iteration
    for sw, r in zip(sws, rs):
        sw.step(*r)
    t_pso_multi, _ = timed(lambda: multi.pso_step(pso_steps))
    t_pso_seq, _ = timed(lambda: [sw.step(*r) for sw, r in zip(sws, rs)])
    for sw in sws:
        sw.close()
    out = {"gps": len(gps), "N": [gp.n for gp in gps], "particles": 40 * len(gps),
           "ms_batched_call": round(1e3 * t_multi, 3), "schedule": multi.schedule(),
           "ms_pso_iteration_batched": round(1e3 * t_pso_multi, 3),
           "ms_pso_iteration_per_gp": round(1e3 * t_pso_seq, 3),
           "ms_merged_lockstep": round(1e3 * t_merged, 3), "ms_one_launch": round(1e3 * t_one, 3),
           "ms_per_gp_evaluations": round(1e3 * t_seq, 3),
           "evals_per_s_batched_call": round(40 * len(gps) / t_multi, 1),
           "per_gp_paths": sorted({gp.factor_path(40) for gp in gps}), "bit_identical": bool(same)}
    for mm in pinned.values():
        mm.close()
    multi.close()
    for gp in gps:
        gp.close()
    return out


def nma_converge(cap: int):
    """The same comparison on the NMA-like gradient workload: N=3000, d=30,
    ARD rbf (30) + periodic (9 dihedrals) + optimised nugget searched in
    [1e-12, 1e-4], swarm 32; its schedule, 5 L-BFGS steps on gbest every 10
    iterations, first."""
    from ferebus_amd import SwarmLikelihood, compile_kernel, fit_mean, synthetic
    w = synthetic.nma(n=3000, seed=3)
    prog = compile_kernel(w.kernel, w.kernels, w.x.shape[1])
    gp = SwarmLikelihood(w.x, w.y, prog, mean=fit_mean(w.mean, w.x, w.y), noise=w.noise, optimise_noise=True,
                         likelihood=w.likelihood, max_swarm=32)
    try:
        out = converge_runs(gp, 32, cap, schedules=((10, 5), (10, 20)), noise_max=w.extra["noise_search"][1])
    finally:
        gp.close()
    out["workload"] = "NMA-like N=3000 d=30, ARD rbf + periodic + optimised nugget in [1e-12, 1e-4], swarm 32"
    return out


def converge_runs(gp, S, cap, schedules=((10, 20), (10, 5)), **settings):
    """Wall time to the converged logL of the swarm search on this workload.

    Plain PSO (the reference's optimiser, seed 1) runs until the relative
    change of its best cost stays <= 1e-8 for 50 iterations (pso.f90:619-636
    with the examples' tolerance / stall settings) or `cap` iterations; its
    final cost is the target.  The same PSO with 5 L-BFGS steps on gbest
    every 10 iterations (adjoint gradient) then runs until its best cost
    reaches that target.  Both run on the device-resident swarm, one graph
    replay per iteration, inputs resident; times exclude set-up."""
    from ferebus_amd.lbfgs import LbfgsSettings
    from ferebus_amd.pso import PsoSettings, run_pso

    base = dict(swarm_size=S, iterations=cap, relative_change=True, tolerance=1e-8, stall_iterations=50, seed=1,
                **settings)
    plain = run_pso(gp, PsoSettings(**base))
    target = plain.best_cost
    out = {
        "criterion": "relative change of best cost <= 1e-8 for 50 iterations (plain PSO), "
                     f"cap {cap} iterations; each polished run stops at the plain run's final logL",
        "plain_pso": {"seconds": round(plain.seconds, 3), "iterations": plain.iterations,
                      "logL": -plain.best_cost, "evaluations": plain.evaluations,
                      "ms_per_iteration": round(1e3 * plain.seconds / max(plain.iterations, 1), 3)},
        "lbfgs_polished_pso": [],
    }
    # polishing schedules: every 10 iterations, 20 or 5 L-BFGS steps (trial
    # steps evaluated 4 at a time)
    for every, steps in schedules:
        pol = run_pso(gp, PsoSettings(**base, lbfgs_every=every, lbfgs=LbfgsSettings(steps=steps)),
                      target_cost=target)
        reached = next((t for it, t, c in pol.history if c <= target), None)
        out["lbfgs_polished_pso"].append({
            "schedule": f"{steps} L-BFGS steps on gbest every {every} iterations",
            "seconds_to_plain_logL": None if reached is None else round(reached, 3),
            "time_ratio_over_plain": None if reached is None else round(reached / plain.seconds, 4),
            "iterations": pol.iterations, "logL": -pol.best_cost, "polish_calls": pol.polish_calls,
            "gradient_evals": pol.polish_evals, "polish_improvements": pol.polish_improvements})
    out["time_ratio_polished_over_plain"] = out["lbfgs_polished_pso"][0]["time_ratio_over_plain"]
    return out


def host_info():
    """CPU model, logical CPUs and this process's CPU share (affinity mask,
    cgroup quota, OMP_NUM_THREADS): the CPU baseline runs one OpenMP thread
    per CPU of the share."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    logical = os.cpu_count() or 1
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else logical
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            share = min(share, max(1, int(int(quota) / int(period))))
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        share = min(share, int(omp))
    return {"cpu_model": model, "logical_cpus": logical, "cpu_share": share, "threads_used": share}


def lapack_info(path):
    """The host LAPACK the CPU baseline opened: library file and OpenBLAS
    build configuration."""
    import ctypes
    desc = os.path.basename(path)
    try:
        lib = ctypes.CDLL(path)
        for sym in ("scipy_openblas_get_config", "openblas_get_config"):
            if hasattr(lib, sym):
                fn = getattr(lib, sym)
                fn.restype = ctypes.c_char_p
                desc += " (" + fn().decode().strip() + ")"
                break
    except OSError:
        pass
    return desc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=tuple(WORKLOADS), default="methanol",
                    help="methanol: the headline workload (N=2000, d=12, swarm 50); water: N=500, d=3, swarm 20; "
                         "nma: N=3000, d=30, optimised nugget, swarm 32; ibuprofen: N=6000, d=93, linear mean, "
                         "concentrated likelihood, swarm 64")
    ap.add_argument("--n", type=int, default=None, help="training points (default: the workload's)")
    ap.add_argument("--swarm", type=int, default=None, help="particles (default: the workload's)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="bound on the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-converge", action="store_true",
                    help="skip the wall-time-to-converged-logL runs (plain PSO vs L-BFGS-polished PSO)")
    ap.add_argument("--converge-cap", type=int, default=1000,
                    help="iteration cap of each PSO run")
    ap.add_argument("--no-uneven", action="store_true",
                    help="skip the batched-uneven run (10 glycine-like GPs, N in [1000, 4000], swarm 40 each)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run 5 device-to-device copies of this many bytes (a known-size "
                         "copy kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the swarm evaluation is not sharded across GPUs "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    from ferebus_amd import SwarmLikelihood, compile_kernel, device_name, fit_mean, synthetic

    if args.copy_bytes:
        src = torch.ones(args.copy_bytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": args.copy_bytes, "copies": 5}))
        return

    make, n_default, s_default, seed, desc = WORKLOADS[args.workload]
    if args.n is None:
        args.n = n_default
    if args.swarm is None:
        args.swarm = s_default
    w = getattr(synthetic, make)(n=args.n, seed=seed)
    d = w.x.shape[1]
    prog = compile_kernel(w.kernel, w.kernels, d)
    mfit = fit_mean(w.mean, w.x, w.y)
    S = args.swarm
    gp = SwarmLikelihood(w.x, w.y, prog, mean=mfit, noise=w.noise, optimise_noise=w.optimise_noise,
                         likelihood=w.likelihood, max_swarm=S)
    rng = np.random.default_rng(1)
    thetas = [rng.uniform(w.search_min, w.search_max, (S, gp.nparams)) for _ in range(4)]
    if w.optimise_noise:   # the nugget, last, drawn from its own search box
        for th in thetas:
            th[:, -1] = rng.uniform(*w.extra.get("noise_search", (1e-12, 1e-6)), S)

    torch.cuda.init()
    stream = torch.cuda.Stream()              # default priority
    gp.set_stream(stream.cuda_stream)

    # warm-up (captures the graph of the batched pipeline)
    for k in range(max(args.warmup, 1)):
        ll, info = gp.swarm_loglik(thetas[k % 4])
    fails = int(np.sum(info != 0))

    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for k in range(args.steps):
        ll, info = gp.swarm_loglik(thetas[k % 4])
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_time(e1)
    ms_per_step = gpu_ms / args.steps
    value = S * args.steps / (gpu_ms / 1e3)

    # times from HIP events on the launch stream, in separate passes (the
    # events split the pipeline, so these passes are not the timed one): the
    # phases (build, factorisation span, likelihood) with no events between
    # the factorisation's launches, then every launch bracketed
    prof_steps = 3
    kinds = ("build", "factor", "update", "solve_factor", "update_solve", "likelihood", "swarm", "factor_span")
    gp.profile(2)
    for k in range(prof_steps):
        gp.swarm_loglik(thetas[k % 4])
    span = {k: gp.profile_read(k) for k in kinds}
    gp.profile(1)
    for k in range(prof_steps):
        gp.swarm_loglik(thetas[k % 4])
    prof = {k: gp.profile_read(k) for k in kinds}
    gp.profile(False)
    n = args.n
    groups = gp.factor_groups(S)
    persistent = prof["factor"][1] > 0
    if persistent:
        # one persistent launch (dag_factor_kernel, after a counter reset) per evaluation
        fact_ms = span["factor"][0] / prof_steps
        fact_kernel = "dag_factor_kernel (whole batched Cholesky of the swarm, one persistent launch)"
    else:
        fact_ms = span["factor_span"][0] / prof_steps
    # chained lock-step schedule: no third launch per panel
    chained = not persistent and prof["update_solve"][1] == 0
    launches = (prof["update"][1] + prof["solve_factor"][1] + prof["update_solve"][1]) // prof_steps
    grouped = (f", the swarm in {groups} groups of particles on concurrent streams" if groups > 1 else "")
    if chained:
        fact_kernel = ("chained lock-step batched Cholesky: per 128-column panel panel_chain_kernel (update, "
                       "diagonal chain handed over between work-groups) and solve_rows_kernel where the next "
                       f"update launch does not solve the rows ({launches} launches per evaluation{grouped}); "
                       "duration = the factorisation's span from the end of the build to the start of the "
                       "likelihood")
    elif not persistent:
        fact_kernel = ("lock-step batched Cholesky: per 128-column panel panel_update_factor_kernel, "
                       f"trsm_factor_kernel, update_trsm_kernel ({launches} launches per evaluation{grouped}); "
                       "duration = the factorisation's span")
    fact_flops = S * potrf_flops(n)
    factor_path = gp.factor_path(S)
    default_size = (n, S) == (n_default, s_default)
    traffic, traffic_src = (profiled_traffic(factor_path, chained, groups, args.workload)
                            if default_size else (None, None))
    build_ms = span["build"][0] / prof_steps
    build_bytes = S * 8.0 * n * (n + 1) / 2.0 + 8.0 * n * d
    achieved = fact_flops / (fact_ms * 1e-3) / 1e12

    result = {
        "metric": f"logL evals/sec at N={n} d={d} swarm {S} (FP64)",
        "value": round(value, 2),
        "unit": "logL evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": (f"synthetic (seeded {w.name} geometry: r1, r2 and angle, Morse-potential target)"
                 if args.workload == "water" else f"synthetic (seeded {w.name} ALF features, Gaussian-bump target)"),
        "config": {"workload": desc.format(n=n, d=d, S=S),
                   "path": "batched (hipGraph replay per swarm evaluation, "
                           + ("persistent factorisation)" if persistent else "lock-step factorisation)"),
                   "factor_path": factor_path,
                   "stream_groups": groups,
                   "schedule": "persistent" if persistent else ("chained (2 launches per panel)" if chained
                                                               else "3 launches per panel"),
                   "failed_particles_last_step": fails},
        "roofline": {
            "kernel": fact_kernel,
            "bound": "mfma",
            "achieved": round(achieved, 3),
            "peak": 78.6,
            "unit": "TFLOP/s",
            "frac": round(achieved / 78.6, 4),
            "traffic": traffic,
            "algorithmic_flops_per_launch": fact_flops,
            "traffic_source": None if traffic is None else "rocprofv3 FETCH_SIZE + WRITE_SIZE, " + traffic_src,
            "avg_duration_ms": round(fact_ms, 4),
        },
        "phases_ms_per_step": {k: round(span[k][0] / prof_steps, 4) for k in ("build", "factor_span", "likelihood",
                                                                               "swarm")},
        "kernels_ms_per_step": {k: round(v[0] / prof_steps, 4) for k, v in prof.items()},
        "kernel_launches_per_step": {k: v[1] // prof_steps for k, v in prof.items()},
        "build_hbm_gbs": round(build_bytes / (build_ms * 1e-3) / 1e9, 1) if build_ms > 0 else None,
        "wall_s": round(wall, 3),
        "device": device_name(),
    }

    if not args.no_plain:
        t = time.perf_counter()
        pl, pinfo = gp.particle_loglik(thetas[0])
        plain_s = time.perf_counter() - t
        result["plain_path"] = {"evals_per_s": round(S / plain_s, 2),
                                "batched_over_plain": round(value / (S / plain_s), 2)}

    if not args.no_converge:
        result["converged"] = converge_runs(gp, S, args.converge_cap)
        result["converged_nma"] = nma_converge(args.converge_cap)

    if not args.no_uneven:
        result["batched_uneven"] = uneven_run()

    if not args.no_cpu:
        import oracle
        host = host_info()
        cores = host["threads_used"]
        spec = oracle.Spec.from_program(prog)
        lapack_path = oracle.scipy_openblas_path()
        bl = oracle.Baseline(w.x, w.y, mfit.value, spec, noise=w.noise, optimise_noise=w.optimise_noise,
                             likelihood=w.likelihood, lapack_path=lapack_path)
        pool = np.concatenate([np.atleast_2d(tt) for tt in thetas])
        bl.swarm(pool[:cores], threads=cores)               # first touch of the tensors, untimed
        done, t = 0, time.perf_counter()
        while True:
            bl.swarm(pool[(done + np.arange(cores)) % len(pool)], threads=cores)
            done += cores
            if time.perf_counter() - t > args.cpu_seconds:
                break
        cpu_s = time.perf_counter() - t
        result["cpu_baseline"] = {
            "value": round(done / cpu_s, 3), "unit": "logL evals/s", "cores": cores,
            "kind": "port",
            "sample": f"{done} particles of the same N={n} workload ({cpu_s:.1f} s), batches of {cores}, one "
                      f"LAPACK thread per particle (reference host algorithm: pair tensors + dgemv + exp, "
                      f"dpotrf / dpotrs)",
            "cpu_model": host["cpu_model"], "logical_cpus": host["logical_cpus"],
            "cpu_share": host["cpu_share"],
            "lapack": (lapack_info(lapack_path) if bl.lapack else "the oracle's own dpotf2 (no host LAPACK)"),
        }
        bl.close()
    gp.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
