"""Batched-ensemble benchmark (one persistent launch for every member).

    python -m tmcore_amd.ensemble_bench [--days 2]

128 members on level 6 (dt 300 s), 32 on level 7 (dt 150 s) and 8 on level 8
(dt 90 s), FP32 state with FP64 invariants, member m seeded 1000 + m, advanced
`days` simulated days after the autotuner has picked chunk size, workgroups
per CU and members per thread.  Prints one JSON line: member-days per hour of
the whole run and the step latency of each level group run alone.
"""
from __future__ import annotations

import argparse
import json

import numpy as np

from .ensemble import Ensemble
from .mesh import Mesh, seeded_state

SPEC = [(6, 128, 300.0), (7, 32, 150.0), (8, 8, 90.0)]


def config4(days=2.0, spec=SPEC):
    meshes = [Mesh(L) for L, _, _ in spec]
    ens = Ensemble([(m, m.fit(), M, dt) for m, (L, M, dt) in zip(meshes, spec)], precision=32)
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
           "autotune": {"chunk": chunk, "blocks_per_cu": bpc, "members_per_thread": width,
                        "ms_per_tick": round(tick_ms, 4)},
           "finite_invariants": bool(np.isfinite(inv).all())}
    # per-size step latency: each level group alone, 20 steps
    lat = {}
    for li, (L, M, dt) in enumerate(spec):
        ns = [0] * len(spec)
        ns[li] = 20
        lat[f"L{L}x{M}"] = round(ens.run(ns) / 20, 4)
    out["step_latency_ms_alone"] = lat
    ens.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--days", type=float, default=2.0)
    args = ap.parse_args()
    print(json.dumps(config4(args.days)), flush=True)


if __name__ == "__main__":
    main()
