"""Benchmark: simulated days per hour of the TRiSK shallow-water CRK4 step on
one MI355X, at 655,362 cells (icosahedral level 8).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one CRK4 time step (4 fused RK stages + beta reduction + update) of
the level-8 mesh (655,362 cells, 1,966,080 edges, 1,310,720 vertices), dt =
90 s, 4 passive tracers, seeded synthetic initial state, all inputs resident
in HBM.  Prints one JSON line.  The CPU baseline times the FP64 CPU path in
Fortran (tmcore_amd/fortran/tm_swm_cpu_mod.F90, OpenMP) on a bounded sample of
the same workload on this host, with the oracle's C restatement beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time


ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU as lscpu names it (model name, sockets, cores per socket),
    else /proc/cpuinfo's model name."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        f = dict(line.split(":", 1) for line in out.splitlines() if ":" in line)
        f = {k.strip(): v.strip() for k, v in f.items()}
        if "Model name" in f:
            return (f"{f['Model name']} ({f.get('Socket(s)', '?')} socket(s) x {f.get('Core(s) per socket', '?')} "
                    f"cores x {f.get('Thread(s) per core', '?')} threads; {f.get('CPU(s)', '?')} CPUs)")
    except (OSError, subprocess.SubprocessError, ValueError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(mesh, fit, state, dt, budget_s=8.0):
    """The FP64 CPU path in Fortran (tmcore_amd/fortran/tm_swm_cpu_mod.F90,
    amdflang -O3 -fopenmp, bit-identical to the oracle) on a bounded number of
    steps of the same workload: once on one thread, once on every host thread
    available to this process (OpenMP over the element loops, at most 16), as
    the reference ran with GCC auto-parallelisation.  The oracle's C
    restatement (gcc -O2, OpenMP) is timed beside it on all threads."""
    from oracle.oracle import Oracle, set_threads
    from tmcore_amd.cpu_path import cpu_run
    gd, u, ghs, q = state
    o = Oracle(mesh, fit, dt=dt)
    iap = o.initial_iap(gd, u)
    nthreads = min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count())
    days_h = lambda n, el: (n * dt / 86400.0) / (el / 3600.0)
    res = {}
    for threads in (1, nthreads):
        t1 = cpu_run(mesh, fit, gd, iap, ghs, q, dt, 1, threads)[-1]
        n = int(max(2, min(30, budget_s / max(t1, 1e-3))))
        el = cpu_run(mesh, fit, gd, iap, ghs, q, dt, n, threads)[-1]
        res[threads] = (n, el, days_h(n, el))
    set_threads(nthreads)
    t0 = time.perf_counter()
    o.run(1, gd, iap, ghs, q)
    n_c = int(max(2, min(30, budget_s / max(time.perf_counter() - t0, 1e-3))))
    t0 = time.perf_counter()
    o.run(n_c, gd, iap, ghs, q)
    el_c = time.perf_counter() - t0
    n, el, v = res[nthreads]
    n1, el1, v1 = res[1]
    return {"value": v, "unit": "sim-days/hour", "cores": nthreads, "kind": "port",
            "language": "Fortran (amdflang -O3 -fopenmp -ffp-contract=off), bit-identical to the oracle",
            "sample": f"{n} CRK4 steps (dt={dt:g} s, {q.shape[0]} tracers) of the same level-{mesh.level} workload, "
                      f"{el:.1f} s wall on {nthreads} threads; single thread: {n1} steps in {el1:.1f} s",
            "single_thread_value": v1, "s_per_step": el / n, "s_per_step_single_thread": el1 / n1,
            "c_port": {"value": days_h(n_c, el_c), "s_per_step": el_c / n_c, "threads": nthreads,
                       "build": "gcc -O2 -ffp-contract=off -fopenmp (oracle/)"},
            "cpu_model": cpu_model(), "nproc": os.cpu_count(), "threads": nthreads}


PROFILE_MD = os.path.join(ROOT, "profiles", "trisk_step_L8_fp64_rocprof.md")


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (scaled by the copy-probe calibration
    the summary states), or None when the summary has no row for it."""
    import re
    try:
        text = open(PROFILE_MD).read()
    except OSError:
        return None
    in_table = False
    for line in text.splitlines():
        if "HBM MB / launch" in line:
            in_table = True
            continue
        if in_table and not line.startswith("|"):
            in_table = False
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if in_table and len(cells) == 5 and re.sub(r"<.*", "", cells[0]) == kernel:
            try:
                return float(cells[4]) * 1e6
            except ValueError:
                return None
    return None


def ensemble_bench(days=2.0):
    from tmcore_amd.ensemble_bench import config4
    return config4(days)


def precision_judge(mesh, fit, seed=2, dt=90.0, days=1, judge_level=9, judge_dt=37.5):
    """FP64 and FP32 forms on the benchmark mesh against an FP64 judge on the
    next finer mesh: both forms run `days` simulated days from the seeded
    state; the judge runs the same initial recipe on level `judge_level` and
    is coarse-grained to the benchmark mesh by area-weighted cell means.
    Returns each form's relative L2 thickness error against the judge, its
    sim-days/hour, and the FP32 speed-up at that (equal) error."""
    import numpy as np
    from tmcore_amd.coarse import CoarseGrainer
    from tmcore_amd.closure_train import l2_thickness
    from tmcore_amd.mesh import RADIUS, Mesh, seeded_state
    from tmcore_amd.swm import SwmGpu
    fm = Mesh(judge_level)
    gd, u, ghs, _ = seeded_state(fm, seed)
    g = SwmGpu(fm, fm.fit(), dt=judge_dt, precision=64)
    g.upload(gd, ghs, u=u)
    n = int(round(days * 86400.0 / judge_dt))
    for done in range(0, n, 1152):
        g.step_async(min(1152, n - done))
        g.sync()
    judge = CoarseGrainer(fm, mesh, velocity=False).cell_mean(g.download()["gd"])
    g.close()
    area = mesh.areaCell * RADIUS ** 2
    gd, u, ghs, _ = seeded_state(mesh, seed)
    out = {"judge": f"level {judge_level} ({fm.nCells} cells), FP64, dt {judge_dt:g} s, coarse-grained to level "
                    f"{mesh.level} by area-weighted cell means", "days": days}
    n = int(round(days * 86400.0 / dt))
    for prec in (64, 32):
        g = SwmGpu(mesh, fit, dt=dt, precision=prec)
        g.upload(gd, ghs, u=u)
        g.step_async(n)
        g.sync()
        ms = g.last_elapsed_ms()
        out[f"fp{prec}"] = {"l2_h_rel_vs_judge": l2_thickness(g.download()["gd"], judge, area),
                            "sim_days_per_hour": round(days / (ms / 3.6e6), 1)}
        g.close()
    out["fp32_speedup_at_equal_error"] = round(out["fp32"]["sim_days_per_hour"] / out["fp64"]["sim_days_per_hour"], 3)
    return out


def equal_error_cost(err, curve):
    """Sim-days/hour the plain method needs for error `err`, interpolated
    log-log between the two points of curve [(sim_days_per_hour, error), ...]
    (plain runs at increasing resolution) that bracket it; None outside."""
    import math
    pts = sorted(curve, key=lambda p: -p[1])  # decreasing error
    for (s0, e0), (s1, e1) in zip(pts, pts[1:]):
        if e1 <= err <= e0 and e0 > e1:
            w = (math.log(e0) - math.log(err)) / (math.log(e0) - math.log(e1))
            return math.exp(math.log(s0) + w * (math.log(s1) - math.log(s0)))
    return None


CLOSURE_WEIGHTS = os.path.join(ROOT, "tmcore_amd", "data", "closure_L6.npz")


def closure_bench(days=1, seed=6000):
    """Closure form: level 6 (40,962 cells, dt 300 s) with the trained closure
    attached.  Speed: one simulated day of the plain and the corrected
    level-6 model and of plain level 7 and 8, timed with HIP events.  Accuracy: held-out seed
    6000 over `days` days against the level-9 FP32 truth coarse-grained to
    level 6, beside plain level-7 (dt 150 s) and level-8 (dt 90 s) runs of
    the same initial state, each compared with the truth at its own level and
    both mapped to level 6 (relative L2 of the thickness)."""
    from tmcore_amd.closure_experiment import Evaluator
    from tmcore_amd.closure_train import Pipeline
    from tmcore_amd.mesh import seeded_state
    if not os.path.exists(CLOSURE_WEIGHTS):
        return {"skipped": f"no trained weights at {os.path.relpath(CLOSURE_WEIGHTS, ROOT)}"}
    p = Pipeline(9, 6, 300.0, 8, log=log)
    ev = Evaluator(p, CLOSURE_WEIGHTS, "7:150,8:90")
    try:
        speed = {}
        for name, g in (("L6_plain", ev.plain), ("L6_closure", ev.corr), ("L7_plain", ev.others["L7"][1]),
                        ("L8_plain", ev.others["L8"][1])):
            gd, u, ghs, _ = seeded_state(g.mesh, seed)
            g.upload(gd, ghs, u=u)
            n = int(round(86400.0 / g.opts.dt))
            g.step_async(20)
            g.sync()
            g.step_async(n)
            g.sync()
            ms = g.last_elapsed_ms()
            speed[name] = {"ms_per_step": round(ms / n, 5), "sim_days_per_hour": round(1.0 / (ms / 3.6e6), 1)}
        errs, _ = ev.run_seed(seed, days)
    finally:
        ev.close()
        p.close()
    final = {k: v[-1] for k, v in errs.items()}
    # the closure's error placed on the plain method's error-versus-cost curve
    curve = [(speed[f"L{L}_plain"]["sim_days_per_hour"], final[f"L{L}_plain"]) for L in (6, 7, 8)]
    e_c, s_c = final["L6_closure"], speed["L6_closure"]["sim_days_per_hour"]
    s_p = equal_error_cost(e_c, curve)
    eq = {"error": e_c, "closure_sim_days_per_hour": s_c,
          "plain_sim_days_per_hour_at_equal_error": round(s_p, 1) if s_p else None,
          "closure_speedup_at_equal_error": round(s_c / s_p, 3) if s_p else None,
          "plain_curve": [{"level": L, "sim_days_per_hour": c[0], "error": c[1]} for L, c in zip((6, 7, 8), curve)],
          "interpolation": "log(sim-days/hour) linear in log(error) between the plain runs that bracket it"}
    return {"speed": speed, f"l2_h_rel_day{days}_seed{seed}": final, "equal_error": eq,
            "trained_on": ev.meta.get("seeds_all", ev.meta["seeds"]), "deployment_gain": ev.meta.get("gain", 1.0),
            "weights": os.path.relpath(CLOSURE_WEIGHTS, ROOT)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=960)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--dt", type=float, default=90.0)
    ap.add_argument("--ntracers", type=int, default=4)
    ap.add_argument("--profile-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-days", type=float, default=2.0)
    ap.add_argument("--no-closure", action="store_true")
    ap.add_argument("--no-judge", action="store_true", help="skip the FP64 level-9 judge of the FP64/FP32 forms")
    ap.add_argument("--closure-days", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by Hilbert-ordered ranges with point-to-point halo exchange)",
              file=sys.stderr)
        sys.exit(2)

    from tmcore_amd.mesh import Mesh, seeded_state
    from tmcore_amd.swm import SwmGpu

    t0 = time.perf_counter()
    mesh = Mesh(args.level)
    fit = mesh.fit()
    seeds = [200 + i for i in range(1, args.ntracers + 1)]
    gd, u, ghs, q = seeded_state(mesh, 2, seeds)
    log(f"mesh L{args.level}: {mesh.nCells} cells, {mesh.nEdges} edges, {mesh.nVertices} vertices "
        f"({time.perf_counter() - t0:.1f} s incl. fit)")

    def run_form(precision):
        g = SwmGpu(mesh, fit, dt=args.dt, ntracers=args.ntracers, precision=precision)
        g.upload(gd, ghs, u=u, tracers=q)
        g.step_async(args.warmup)
        g.sync()
        t0 = time.perf_counter()
        g.step_async(args.steps)
        g.sync()
        wall = time.perf_counter() - t0
        ev_ms = g.last_elapsed_ms()
        st = g.download()
        # a second download of the same state into the same (already touched)
        # arrays, timed: what a host with preallocated state arrays sees
        t0 = time.perf_counter()
        g.download(out=st)
        down = time.perf_counter() - t0
        # a second upload of the initial state, timed: the host-buffer boundary
        # (validation, renumbering, host-to-device copies, the iap_u prologue)
        t0 = time.perf_counter()
        g.upload(gd, ghs, u=u, tracers=q)
        up = time.perf_counter() - t0
        return g, wall, ev_ms, st, up, down

    g64, wall, ev_ms, st64, up64, down64 = run_form(64)
    ms_per_step = 1000.0 * wall / args.steps
    sim_days_h = (args.steps * args.dt / 86400.0) / (wall / 3600.0)
    log(f"FP64: {ms_per_step:.4f} ms/step wall, {ev_ms / args.steps:.4f} ms/step events, "
        f"{sim_days_h:.1f} sim-days/h")
    # the same run with the state handed over in host memory: one upload
    # before it and one download (gd, u, iap_u, tracers) after it
    host_boundary = {
        "upload_ms": round(1000.0 * up64, 3), "download_ms": round(1000.0 * down64, 3),
        "sim_days_per_hour_incl_transfers": round((args.steps * args.dt / 86400.0) /
                                                  ((wall + up64 + down64) / 3600.0), 2),
        "ms_per_step_with_a_download_every_step": round(ms_per_step + 1000.0 * down64, 4)}
    log(f"FP64 host boundary: {host_boundary}")
    sb = g64.step_bytes()
    step_gbs = sb["step"] / (ev_ms / args.steps * 1e-3) / 1e9

    prof = g64.profile(args.profile_steps)
    tot = {k: v[0] * v[1] for k, v in prof.items()}
    dom = max(tot, key=tot.get)
    avg_ms, nl, kbytes = prof[dom]
    achieved = kbytes / (avg_ms * 1e-3) / 1e9
    kernels = {k: {"avg_us": round(v[0] * 1e3, 2), "launches_per_step": v[1] // max(args.profile_steps, 1),
                   "GBps": round(v[2] / (v[0] * 1e-3) / 1e9, 1) if v[0] > 0 else None}
               for k, v in prof.items()}
    log(f"kernels: {json.dumps(kernels)}")
    g64.close()

    from tmcore_amd.swm import hbm_copy_gbps
    copy_gbs = hbm_copy_gbps(2.0 * 2**30, 20)
    log(f"HBM copy probe (2 GiB, read+write): {copy_gbs:.1f} GB/s")

    forms = {"fp64": {"sim_days_per_hour": round(sim_days_h, 2), "ms_per_step": round(ms_per_step, 4),
                      "host_boundary": host_boundary}}
    if not args.no_fp32:
        g32, wall32, ev32, st32, _, _ = run_form(32)
        g32.close()
        from numpy.linalg import norm
        h64 = (st64["gd"] + ghs)
        h32 = (st32["gd"] + ghs)
        forms["fp32"] = {"sim_days_per_hour": round((args.steps * args.dt / 86400.0) / (wall32 / 3600.0), 2),
                         "ms_per_step": round(1000.0 * wall32 / args.steps, 4),
                         "l2_height_rel_err_vs_fp64": float(norm(h32 - h64) / norm(h64)),
                         "after_steps": args.warmup + args.steps}
        log(f"FP32: {forms['fp32']}")

    if not args.no_judge and not args.no_fp32:
        forms["equal_error_vs_fp64_judge"] = precision_judge(mesh, fit, dt=args.dt)
        log(f"judge: {forms['equal_error_vs_fp64_judge']}")

    if not args.no_ensemble:
        forms["ensemble_fp32"] = ensemble_bench(args.ensemble_days)
        log(f"ensemble: {forms['ensemble_fp32']}")

    if not args.no_closure:
        forms["closure_L6"] = closure_bench(args.closure_days)
        log(f"closure: {forms['closure_L6']}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(mesh, fit, (gd, u, ghs, q), args.dt)
        log(f"CPU baseline: {cpu}")

    traffic = profiled_traffic(dom) if args.level == 8 and args.ntracers == 4 else None
    out = {
        "metric": "sim-days/hour at 655k cells: FP64 vs FP32 vs closure at equal L2 height error",
        "value": round(sim_days_h, 3),
        "unit": "sim-days/hour",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded solid-body flow + 8 Gaussian bumps + Gaussian mountain, 4 smoothed U[0,1] tracers)",
        "config": {"workload": f"TRiSK SWM CRK4 step, icosahedral Voronoi level {args.level} "
                               f"({mesh.nCells} cells, {mesh.nEdges} edges, {mesh.nVertices} vertices), "
                               f"dt={args.dt:g} s, {args.ntracers} tracers, pv_scheme 3 order 3, FP64",
                   "cells": mesh.nCells, "dt_s": args.dt, "ntracers": args.ntracers, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": os.path.relpath(PROFILE_MD, ROOT) if traffic else None,
                     "bytes_per_launch": kbytes, "avg_launch_us": round(avg_ms * 1e3, 2)},
        "hbm_copy_probe_GBps": round(copy_gbs, 1),
        "step_algorithmic_GBps": round(step_gbs, 1),
        "step_algorithmic_frac": round(step_gbs / HBM_PEAK_GBS, 4),
        "kernels": kernels,
        "forms": forms,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
