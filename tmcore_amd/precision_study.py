"""FP32 precision form against the FP64 device run on one mesh.

    python -m tmcore_amd.precision_study [--level 10] [--dt 22.5] [--steps 960] [--seed 3]

Both forms start from the same seeded state (seeded_state(mesh, seed)) and
run the same number of CRK4 steps; printed as one JSON line: ms/step and
sim-days/hour of each form, the relative L2 differences of thickness h = gd/g
(area-weighted) and normal velocity u (areaEdge-weighted) of FP32 against
FP64, and the drift of the four invariants over the run in each form.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from .mesh import RADIUS, Mesh, seeded_state
from .swm import SwmGpu


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--level", type=int, default=10)
    ap.add_argument("--dt", type=float, default=22.5)
    ap.add_argument("--steps", type=int, default=960)
    ap.add_argument("--seed", type=int, default=3)
    args = ap.parse_args()
    t0 = time.perf_counter()
    m = Mesh(args.level)
    fit = m.fit()
    gd, u, ghs, _ = seeded_state(m, args.seed)
    log(f"mesh L{args.level}: {m.nCells} cells, {m.nEdges} edges ({time.perf_counter() - t0:.1f} s incl. fit)")
    aC = m.areaCell * RADIUS ** 2
    aE = m.dvEdge * m.dcEdge * RADIUS ** 2
    out = {"level": args.level, "cells": m.nCells, "dt_s": args.dt, "steps": args.steps, "seed": args.seed}
    res = {}
    for prec in (64, 32):
        g = SwmGpu(m, fit, dt=args.dt, precision=prec)
        g.upload(gd, ghs, u=u)
        inv0 = g.invariants()
        done, ms = 0, 0.0
        while done < args.steps:
            n = min(args.steps - done, 240)
            g.step_async(n)
            g.sync()
            ms += g.last_elapsed_ms()
            done += n
            log(f"FP{prec}: {done}/{args.steps} steps")
        inv1 = g.invariants()
        st = g.download()
        res[prec] = st
        drift = (inv1 - inv0)
        drift[:3] /= inv0[:3]
        out[f"fp{prec}"] = {"ms_per_step": ms / args.steps,
                            "sim_days_per_hour": (args.steps * args.dt / 86400.0) / (ms / 3.6e6),
                            "invariant_drift": {"mass_rel": drift[0], "energy_rel": drift[1],
                                                "enstrophy_rel": drift[2], "abs_vorticity_abs": drift[3]}}
        g.close()
        log(json.dumps(out[f"fp{prec}"]))
    d64, d32 = res[64], res[32]
    out["l2_h_rel_fp32_vs_fp64"] = float(np.sqrt(np.sum(aC * (d32["gd"] - d64["gd"]) ** 2) /
                                                 np.sum(aC * d64["gd"] ** 2)))
    out["l2_u_rel_fp32_vs_fp64"] = float(np.sqrt(np.sum(aE * (d32["u"] - d64["u"]) ** 2) /
                                                 np.sum(aE * d64["u"] ** 2)))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
