"""Closure experiment driver: train on coarse-grained fine-mesh truth, then
evaluate multi-day forecasts against it.

    python -m tmcore_amd.closure_experiment train --seeds 5000:5100 --out W.npz [--init W0.npz]
    python -m tmcore_amd.closure_experiment eval --weights W.npz --seeds 6000:6004 --days 10 --out R.json

Training samples are pairs at coarse steps --samples of every initial state
(seeded_state(fine, seed)); the last --holdout seeds' samples are kept for
validation.  Progress lines go to stderr as each initial state or epoch ends.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from .closure import Closure, Mlp
from .closure_train import STEPS_PER_DAY, Pipeline, l2_thickness, load, save
from .coarse import CoarseGrainer
from .mesh import RADIUS, Mesh, seeded_state
from .swm import SwmGpu


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def seed_range(s):
    a, b = s.split(":")
    return list(range(int(a), int(b)))


def online_check(p: Pipeline, mlp: Mlp, seed: int, hours: int = 24, every_h: int = 3, alphas=(1.0, 0.5),
                 intervals=None, lumped=(True,)):
    """Coarse runs from the coarse-grained truth of `seed`, plain (alpha 0) and
    with the closure outputs scaled by alpha, the MLP run every `interval`
    steps (default: the training horizon) and held in between; thickness error
    against the truth every `every_h` hours (NaN after a run stops)."""
    steps = int(round(every_h * 3600.0 / p.dt))
    n = hours // every_h
    gd, u, ghs, _ = seeded_state(p.fm, seed)
    p.fine.upload(gd, ghs, u=u)
    truth = [p._cg_now()]
    for _ in range(n):
        p._advance_fine(steps * p.nsub)
        truth.append(p._cg_now())
    ghs_c = p.cg.cell_mean(ghs)
    g = SwmGpu(p.cm, p.cm.fit(), dt=p.dt)
    clo = Closure(g, mlp)
    sh, sc, out = mlp.get_scales()
    res = {}
    runs = [(0.0, 1, False)] + [(a, iv, lm) for a in alphas for iv in (intervals or [p.horizon])
                                for lm in (lumped if iv > 1 else (False,))]
    for alpha, iv, lm in runs:
        if alpha == 0.0:
            clo.detach()
        else:
            mlp.set_scales(sh, sc, out * alpha)
            clo.attach()
            clo.set_interval(iv, lm)
        g.upload(truth[0][0], ghs_c, u=truth[0][1])
        errs = [0.0]
        for k in range(n):
            try:
                g.step(steps)
                errs.append(l2_thickness(g.download()["gd"], truth[k + 1][0], p.area))
            except Exception:
                errs += [float("nan")] * (n - k)
                break
        key = "plain" if alpha == 0.0 else f"alpha {alpha:g} interval {iv}" + (" lumped" if lm else "")
        res[key] = errs
        log(f"online seed {seed} {key}: " + " ".join(f"{x:.3e}" for x in errs))
    mlp.set_scales(sh, sc, out)
    g.close()
    return res


def cmd_train(args):
    p = Pipeline(args.fine_level, args.coarse_level, args.dt, args.nsub, log=log, horizon=args.horizon)
    seeds = seed_range(args.seeds)
    samples = [int(k) for k in args.samples.split(",")]

    def sample_steps(seed):
        """The fixed sample steps, or (--window-end) n samples spaced by the
        horizon in a window whose end is drawn from [n*horizon, window_end]
        per seed (seeded), so the initial states cover the whole range."""
        if args.window_end <= 0:
            return samples
        n, h = len(samples), args.horizon
        end = int(np.random.default_rng(seed).integers(n * h, args.window_end + 1))
        return [end - (n - j) * h for j in range(n)]
    mlp = Mlp(seed=args.init_seed)
    clo = Closure(p.coarse, mlp)
    t0 = time.perf_counter()
    if args.init:
        meta0 = load(args.init, mlp)
        p.tg_scale = meta0["scales"]["rms_t_gd"]
        scales = meta0["scales"]
        pending = []
    else:
        pending = [(a, b, g) for s in seeds[:2] for _, a, b, g in p.truth_pairs(s, sample_steps(s))]
        masked = tuple(int(f) for f in args.mask_fields.split(",") if f != "")
        scales = p.fit_scales(mlp, pending, masked)
    # zero-output network with the same scales: the loss of the plain method
    zero = Mlp(seed=0)
    zp = zero.params
    zp[-(128 * 2 + 2):] = 0.0
    zero.params = zp
    zero.set_scales(*mlp.get_scales())
    for a, b, g in pending:
        p.add(clo, a, b, g)
    used = seeds[:2] if not args.init else []
    for s in seeds[2:] if not args.init else seeds:
        if time.perf_counter() - t0 > args.data_budget:
            log(f"data budget of {args.data_budget:g} s reached: stopping before seed {s}")
            break
        for _, a, b, g in p.truth_pairs(s, sample_steps(s)):
            p.add(clo, a, b, g)
        used.append(s)
        log(f"seed {s}: {clo.size()} samples, {time.perf_counter() - t0:.0f} s")
    seeds = used
    n = clo.size()
    nval = args.holdout * len(samples)
    ntrain = n - nval
    data_s = time.perf_counter() - t0
    clo0 = Closure(p.coarse, zero)
    plain_val = clo0.eval(ntrain, nval).tolist()
    plain_train = clo0.eval(0, ntrain).tolist()
    log(f"{ntrain} training / {nval} validation samples; plain-method loss train {plain_train} val {plain_val}")
    hist = []
    t1 = time.perf_counter()
    for ep in range(args.epochs):
        lr = args.lr * 0.5 * (1 + math.cos(math.pi * ep / args.epochs))
        h = clo.train(ntrain, 1, lr, seed=args.shuffle_seed + ep)[0].tolist()
        v = clo.eval(ntrain, nval).tolist()
        hist.append({"epoch": ep, "lr": lr, "train": h, "val": v})
        log(f"epoch {ep}: lr {lr:.2e} train {h} val {v}")
    train_s = time.perf_counter() - t1
    online = (online_check(p, mlp, seeds[-1], args.online_hours, 6, (1.0, 0.5), [p.horizon], (True,))
              if args.online_hours > 0 else None)
    meta = {"fine_level": args.fine_level, "coarse_level": args.coarse_level, "dt": args.dt, "nsub": args.nsub,
            "horizon": args.horizon,
            "seeds": [seeds[0], seeds[-1] + 1], "samples": samples, "window_end": args.window_end,
            "masked_fields": args.mask_fields, "ntrain": ntrain, "nval": nval,
            "scales": scales, "plain_loss": {"train": plain_train, "val": plain_val}, "history": hist,
            "data_seconds": data_s, "train_seconds": train_s, "init": args.init, "online_check": online,
            "previous": (meta0 if args.init else None)}
    save(args.out, mlp, meta)
    log(f"saved {args.out}")


class Evaluator:
    """Plain and corrected coarse runs, and plain finer runs, against the
    coarse-grained truth of held-out initial states."""

    def __init__(self, p: Pipeline, weights: str, compare: str = "7:150,8:90"):
        self.p = p
        self.mlp = Mlp(seed=0)
        self.meta = load(weights, self.mlp)
        cm = p.cm
        self.plain = SwmGpu(cm, cm.fit(), dt=p.dt)
        self.corr = SwmGpu(cm, cm.fit(), dt=p.dt)
        self.closure = Closure(self.corr, self.mlp)
        self.closure.attach()
        self.closure.set_interval(int(self.meta.get("horizon", 1)), lumped=True)
        self.closure.set_gain(float(self.meta.get("gain", 1.0)))  # the deployment gain stored with the weights
        # finer plain runs: compared with the truth filtered to their own level
        # (fine -> Lr), both then mapped to the coarse grid (Lr -> coarse), so
        # the partition noise of the Lr -> coarse map applies to run and truth alike
        self.others = {}
        for spec in filter(None, compare.split(",")):
            L, dt = spec.split(":")
            m = Mesh(int(L))
            self.others[f"L{L}"] = (m, SwmGpu(m, m.fit(), dt=float(dt)), CoarseGrainer(m, cm, velocity=False),
                                    CoarseGrainer(p.fm, m, velocity=False))

    def close(self):
        """Releases every device model and the network (the pipeline stays open)."""
        self.closure.detach()
        for g in [self.plain, self.corr] + [o[1] for o in self.others.values()]:
            g.close()
        self.mlp.close()

    def run_seed(self, seed, days):
        """{run: [error at day 0..days]}, {run: wall seconds of its days}."""
        p, area = self.p, self.p.area
        truth, ghs_c, mapped = p.truth_daily(seed, days, [o[3] for o in self.others.values()])
        gd_t = [t[0] for t in truth]
        r, walls = {}, {}
        L = p.cm_level
        for name, model in (("plain", self.plain), ("closure", self.corr)):
            t0 = time.perf_counter()
            gds = p.coarse_run(model, truth[0], ghs_c, days)
            walls[f"L{L}_{name}"] = time.perf_counter() - t0
            r[f"L{L}_{name}"] = [l2_thickness(g, t, area) for g, t in zip(gds, gd_t)]
        for k, (name, (m, model, cg, _)) in enumerate(self.others.items()):
            gd, u, ghs, _ = seeded_state(m, seed)
            model.upload(gd, ghs, u=u)
            per_day = int(round(STEPS_PER_DAY / model.opts.dt))

            def err(gd_r, d):
                return float(np.sqrt(np.sum(area * cg.cell_mean(gd_r - mapped[k][d]) ** 2)
                                     / np.sum(area * gd_t[d] ** 2)))
            errs = [err(gd, 0)]
            t0 = time.perf_counter()
            for d in range(days):
                model.step(per_day)
                errs.append(err(model.download()["gd"], d + 1))
            walls[f"{name}_plain"] = time.perf_counter() - t0
            r[f"{name}_plain"] = errs
        return r, walls


def cmd_eval(args):
    p = Pipeline(args.fine_level, args.coarse_level, args.dt, args.nsub, log=log)
    ev = Evaluator(p, args.weights, args.compare)
    res = {"weights": args.weights, "days": args.days, "seeds": {}}
    walls = {}
    for s in seed_range(args.seeds):
        r, w = ev.run_seed(s, args.days)
        for k, v in w.items():
            walls.setdefault(k, []).append(v)
        res["seeds"][s] = r
        log(f"seed {s}: final-day errors " + ", ".join(f"{k} {v[-1]:.3e}" for k, v in r.items()))
        with open(args.out, "w") as f:  # partial results survive an interrupted run
            json.dump(res, f, indent=1)
    keys = list(next(iter(res["seeds"].values())).keys())
    res["mean"] = {k: np.nanmean([res["seeds"][s][k] for s in res["seeds"]], axis=0).tolist() for k in keys}
    res["speed_sim_days_per_hour_wall"] = {k: args.days / (np.mean(v) / 3600.0) for k, v in walls.items()}
    meta = ev.meta
    res["train_meta"] = {k: meta[k] for k in ("seeds", "samples", "ntrain", "nval", "plain_loss")}
    res["train_meta"]["final_epoch"] = meta["history"][-1] if meta["history"] else None
    with open(args.out, "w") as f:
        json.dump(res, f, indent=1)
    log("mean errors by day: " + json.dumps({k: [round(x, 6) for x in v] for k, v in res["mean"].items()}))
    log(f"wrote {args.out}")


def cmd_check(args):
    """online_check of trained weights on one held-out seed."""
    mlp = Mlp(seed=0)
    meta = load(args.weights, mlp)
    p = Pipeline(args.fine_level, args.coarse_level, args.dt, args.nsub, log=log, horizon=int(meta.get("horizon", 1)))
    alphas = tuple(float(a) for a in args.alphas.split(","))
    intervals = [int(i) for i in args.intervals.split(",")]
    res = online_check(p, mlp, int(args.seeds.split(":")[0]), args.hours, 3, alphas, intervals, (False, True))
    with open(args.out, "w") as f:
        json.dump(res, f, indent=1)


def cmd_profile(args):
    """A short closure workload for kernel profiling: level-6 steps with the
    closure attached (features, MFMA forward, correction) and two training
    epochs over synthetic samples (forward, loss, backward, Adam)."""
    m = Mesh(args.coarse_level)
    g = SwmGpu(m, m.fit(), dt=args.dt)
    gd, u, ghs, _ = seeded_state(m, 6000)
    g.upload(gd, ghs, u=u)
    mlp = Mlp(seed=1)
    sh, sc = np.zeros(40), np.ones(40)
    sh[10:20] = np.mean(gd)
    sc[:10], sc[10:20], sc[20:30], sc[30:40] = 1 / 11.4, 1 / 3900.0, 1 / 1e-5, 1 / 1e-6
    mlp.set_scales(sh, sc, np.array([1.2e-5, 2.7e-3 * np.mean(m.areaCell) * RADIUS / np.mean(m.dvEdge)]))
    clo = Closure(g, mlp)
    rng = np.random.default_rng(0)
    for k in range(8):
        clo.add_sample(rng.standard_normal(m.nEdges) * 1.2e-5, rng.standard_normal(m.nCells) * 2.7e-3, 2.7e-3)
    clo.attach()
    g.step_async(args.steps)
    g.sync()
    log(f"closure steps: {g.last_elapsed_ms() / args.steps:.4f} ms/step")
    t0 = time.perf_counter()
    clo.train(8, 2, 1e-4)
    log(f"training: {(time.perf_counter() - t0) / 16 * 1e3:.3f} ms per sample step")


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    pr = sub.add_parser("profile")
    pr.add_argument("--coarse-level", type=int, default=6)
    pr.add_argument("--dt", type=float, default=300.0)
    pr.add_argument("--steps", type=int, default=50)
    for name in ("train", "eval", "check"):
        a = sub.add_parser(name)
        a.add_argument("--fine-level", type=int, default=9)
        a.add_argument("--coarse-level", type=int, default=6)
        a.add_argument("--dt", type=float, default=300.0)
        a.add_argument("--nsub", type=int, default=8)
        a.add_argument("--seeds", required=True)
        a.add_argument("--out", required=True)
    t = sub.choices["train"]
    t.add_argument("--samples", default="0,15,30,45,60,75,90,105")
    t.add_argument("--horizon", type=int, default=12, help="coarse steps of the target tendency")
    t.add_argument("--epochs", type=int, default=30)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--holdout", type=int, default=5)
    t.add_argument("--init", default=None)
    t.add_argument("--init-seed", type=int, default=0)
    t.add_argument("--shuffle-seed", type=int, default=0)
    t.add_argument("--data-budget", type=float, default=1e9, help="seconds of truth generation before training")
    t.add_argument("--online-hours", type=int, default=24, help="online check on the last (held-out) seed")
    t.add_argument("--mask-fields", default="", help="input fields (0 u, 1 gd_edge, 2 vorticity, 3 divergence) not shown")
    t.add_argument("--window-end", type=int, default=0,
                   help="random per-seed sample windows ending by this coarse step (0: the fixed --samples)")
    c = sub.choices["check"]
    c.add_argument("--weights", required=True)
    c.add_argument("--hours", type=int, default=24)
    c.add_argument("--alphas", default="1,0.5")
    c.add_argument("--intervals", default="1,12")
    e = sub.choices["eval"]
    e.add_argument("--weights", required=True)
    e.add_argument("--days", type=int, default=10)
    e.add_argument("--compare", default="7:150,8:90")
    args = ap.parse_args()
    {"train": cmd_train, "eval": cmd_eval, "check": cmd_check, "profile": cmd_profile}[args.cmd](args)


if __name__ == "__main__":
    main()
