"""Batched-ensemble benchmark (one persistent launch for every member).

    python -m tmcore_amd.ensemble_bench [--days 2]
    python -m tmcore_amd.ensemble_bench --group 6:128:300 [--steps 20] [--tuning 64,1,4] [--slices 8]

128 members on level 6 (dt 300 s), 32 on level 7 (dt 150 s) and 8 on level 8
(dt 90 s), FP32 state with FP64 invariants, member m seeded 1000 + m, advanced
`days` simulated days after the autotuner has picked chunk size, workgroups
per CU and members per thread (member slices per level: the library's
default unless given).  Prints one JSON line: member-days per hour of
the whole run and the step latency of each level group run alone.
"""
from __future__ import annotations

import argparse
import json

import numpy as np

from .ensemble import Ensemble
from .mesh import Mesh, seeded_state

SPEC = [(6, 128, 300.0), (7, 32, 150.0), (8, 8, 90.0)]
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def step_bytes(mesh, fit, members, ntracers=0, rs=4.0, fsz=4.0):
    """Compulsory HBM bytes of one step of a level group: per stage, each mesh
    constant once (shared by the members) plus 7 state values per prognostic
    element and member (stage in, base, accumulator read and write, previous
    and current tendency, next state); the finalize 3 per element and member.
    The single-model model of tm_step_bytes with FP32 state and the FP32 copies
    of the edge weights and lengths."""
    nC, nE, nV = mesh.nCells, mesh.nEdges, mesh.nVertices
    nf = fit.nFit2Vertices.astype(np.int64)
    v1, v2 = mesh.verticesOnEdge[:, 0] - 1, mesh.verticesOnEdge[:, 1] - 1
    sum_eoe, sum_eoc = int(mesh.nEdgesOnEdge.sum()), int(mesh.nEdgesOnCell.sum())
    sum_fit_e, sum_fit_v = int((nf[v1] - 1).sum() + (nf[v2] - 1).sum()), int((nf - 1).sum())
    cst = (nE * (8 + 8 + 2 * fsz) + sum_eoe * (4 + fsz) + sum_fit_e * fsz + nV * (12 + 12 + 24 + 8 + 8) +
           sum_fit_v * 4 + sum_eoc * 4 + nC * (8 + 8))
    prog = (nE + nC * (1 + ntracers)) * members
    return 4 * (cst + 7 * rs * prog) + 3 * rs * prog


def config4(days=2.0, spec=SPEC, slices=None):
    meshes = [Mesh(L) for L, _, _ in spec]
    fits = [m.fit() for m in meshes]
    ens = Ensemble([(m, f, M, dt) for m, f, (L, M, dt) in zip(meshes, fits, spec)], precision=32, slices=slices)
    seed = 1000
    for li, (m, (L, M, dt)) in enumerate(zip(meshes, spec)):
        for k in range(M):
            gd, u, ghs, _ = seeded_state(m, seed)
            seed += 1
            ens.upload(li, k, gd, u, ghs)
    (chunk, bpc, width), tick_ms = ens.autotune(ticks=2)
    steps = [int(round(days * 86400 / dt)) for _, _, dt in spec]
    ms = ens.run(steps)
    member_days = sum(M for _, M, _ in spec) * days
    inv = np.array([ens.invariants(li, k) for li, (_, M, _) in enumerate(spec) for k in range(M)])
    out = {"member_days_per_hour": member_days / (ms / 3.6e6), "run_s": ms / 1e3, "member_days": member_days,
           "members": {f"L{L}": M for L, M, _ in spec}, "steps": {f"L{L}": s for (L, _, _), s in zip(spec, steps)},
           "slices": {f"L{L}": s for (L, _, _), s in zip(spec, ens.slices)},
           "autotune": {"chunk": chunk, "blocks_per_cu": bpc, "members_per_thread": width,
                        "ms_per_tick": round(tick_ms, 4)},
           "finite_invariants": bool(np.isfinite(inv).all())}
    # roofline: compulsory bytes of the whole run over its time, and of each
    # level group's step over that group's step latency alone
    sb = [step_bytes(m, f, M) for m, f, (L, M, dt) in zip(meshes, fits, spec)]
    run_bytes = sum(b * n for b, n in zip(sb, steps))
    achieved = run_bytes / (ms * 1e-3) / 1e9
    # per-size step latency: each level group alone, 20 steps
    lat, frac = {}, {}
    for li, (L, M, dt) in enumerate(spec):
        ns = [0] * len(spec)
        ns[li] = 20
        t = ens.run(ns) / 20
        lat[f"L{L}x{M}"] = round(t, 4)
        frac[f"L{L}x{M}"] = round(sb[li] / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    out["step_latency_ms_alone"] = lat
    out["roofline"] = {"bound": "hbm", "kernel": "k_ens_run", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_run": run_bytes,
                       "model": "per step and level: mesh constants once per stage + 7 FP32 state values per "
                                "prognostic element and member per stage + 3 in the finalize",
                       "frac_alone": frac}
    ens.close()
    return out


def group_steps(spec, steps, tuning=None, slices=None):
    """One level group alone (spec = (level, members, dt)) for `steps` steps
    after two warm-up steps, at the given (chunk, blocks_per_cu,
    members_per_thread) or the autotuner's choice and the given member slices
    or the library's default: ms per step and the compulsory-bytes roofline
    of the step."""
    L, M, dt = spec
    m = Mesh(L)
    f = m.fit()
    ens = Ensemble([(m, f, M, dt)], precision=32, slices=None if slices is None else [slices])
    used = ens.slices[0]
    for k in range(M):
        gd, u, ghs, _ = seeded_state(m, 1000 + k)
        ens.upload(0, k, gd, u, ghs)
    tuned = None
    if tuning:
        ens.set_tuning(*tuning)
    else:
        tuned, _ = ens.autotune(ticks=2)
    ens.run([2])
    ms = ens.run([steps]) / steps
    ens.close()
    b = step_bytes(m, f, M)
    return {"group": f"L{L}x{M}", "ms_per_step": round(ms, 4), "bytes_per_step": b, "tuning": tuning or tuned,
            "slices": used,
            "GBps": round(b / (ms * 1e-3) / 1e9, 1), "frac": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--days", type=float, default=2.0)
    ap.add_argument("--group", default=None, help="level:members:dt of one group run alone (e.g. 6:128:300)")
    ap.add_argument("--steps", type=int, default=20, help="steps of --group")
    ap.add_argument("--tuning", default=None, help="chunk,blocks_per_cu,members_per_thread (default: autotune)")
    ap.add_argument("--slices", default=None,
                    help="member slices: one number for --group, comma-separated per level otherwise")
    args = ap.parse_args()
    sl = [int(x) for x in args.slices.split(",")] if args.slices else None
    if args.group:
        L, M, dt = args.group.split(":")
        tuning = tuple(int(x) for x in args.tuning.split(",")) if args.tuning else None
        print(json.dumps(group_steps((int(L), int(M), float(dt)), args.steps, tuning, sl[0] if sl else None)),
              flush=True)
    else:
        print(json.dumps(config4(args.days, slices=sl)), flush=True)


if __name__ == "__main__":
    main()
