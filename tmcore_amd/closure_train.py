"""Training and evaluation pipeline of the learned closure.

Truth: a fine icosahedral model (level 9 by default, dt = dt_coarse / nsub,
FP32 form) started from seeded initial states, coarse-grained onto the coarse
mesh (level 6) by ``CoarseGrainer``.  A training sample is a pair of
coarse-grained truth states a = CG(X(t)), b = CG(X(t + n dt)) and the mean
tendency error of n plain coarse steps from a (n = horizon):
    t_u  = (iap(b) - iap(Plain^n(a))) / (n dt sqrt(gd_edge(a)))
    t_gd = (gd(b) - gd(Plain^n(a))) / (n dt)
With n = 1 this is the one-step error; a horizon of an hour averages out the
fast gravity-wave part of the one-step error, which does not accumulate in a
forecast and which a closure fitted to it feeds back on.
held on the device (``Closure.add_sample``) and fitted by the library's
trainer.  Evaluation runs plain and corrected coarse models for several days
from held-out initial states beside the coarse-grained truth and, for the
cost/accuracy comparison, plain runs on finer meshes coarse-grained the same
way; the error measure is the area-weighted relative L2 of the thickness
h = gd / g.
"""
from __future__ import annotations

import json
import time

import numpy as np

from .closure import NIN, NOUT, Closure, Mlp
from .coarse import CoarseGrainer
from .mesh import RADIUS, Mesh, seeded_state
from .swm import SwmGpu

STEPS_PER_DAY = 86400.0


def save(path, mlp: Mlp, meta: dict):
    sh, sc, out = mlp.get_scales()
    np.savez(path, params=mlp.params, in_shift=sh, in_scale=sc, out_scale=out, meta=np.array(json.dumps(meta)))


def export_raw(path, npz_path):
    """The raw weights file the Fortran driver reads (closure_file): little-endian
    float32 [parameters, 40 input shifts, 40 input scales, 2 output scales]."""
    z = np.load(npz_path, allow_pickle=False)
    np.concatenate([z["params"], z["in_shift"], z["in_scale"], z["out_scale"]]).astype("<f4").tofile(path)


def load(path, mlp: Mlp) -> dict:
    z = np.load(path, allow_pickle=False)
    mlp.params = z["params"]
    mlp.set_scales(z["in_shift"], z["in_scale"], z["out_scale"])
    return json.loads(str(z["meta"]))


def l2_thickness(gd, gd_ref, area):
    return float(np.sqrt(np.sum(area * (gd - gd_ref) ** 2) / np.sum(area * gd_ref ** 2)))


class Pipeline:
    def __init__(self, fine_level=9, coarse_level=6, dt=300.0, nsub=8, fine_precision=32, log=print, horizon=1):
        self.log = log
        t0 = time.perf_counter()
        self.fm, self.cm = Mesh(fine_level), Mesh(coarse_level)
        self.cg = CoarseGrainer(self.fm, self.cm)
        self.dt, self.nsub = dt, nsub
        self.horizon = horizon
        self.cm_level = coarse_level
        self.fine = SwmGpu(self.fm, self.fm.fit(), dt=dt / nsub, precision=fine_precision)
        self.coarse = SwmGpu(self.cm, self.cm.fit(), dt=dt)
        coe = self.cm.cellsOnEdge - 1
        self.c1, self.c2 = coe[:, 0], coe[:, 1]
        self.area = self.cm.areaCell * RADIUS ** 2
        self.dv = self.cm.dvEdge * RADIUS
        self.log(f"pipeline: L{fine_level} ({self.fm.nCells} cells, dt {dt / nsub:g} s, FP{fine_precision}) -> "
                 f"L{coarse_level} ({self.cm.nCells} cells, dt {dt:g} s), setup {time.perf_counter() - t0:.1f} s")

    def close(self):
        """Releases the fine and coarse device models."""
        self.fine.close()
        self.coarse.close()

    # ---------------------------------------------------------------- truth
    def _advance_fine(self, nsteps):
        while nsteps > 0:
            n = min(nsteps, 4096)
            self.fine.step_async(n)
            self.fine.sync()
            nsteps -= n

    def _cg_now(self):
        st = self.fine.download()
        return self.cg.cell_mean(st["gd"]), self.cg.edge_velocity(st["u"])

    def truth_pairs(self, seed, sample_steps):
        """Coarse-grained truth pairs (k, a, b, ghs_c) at coarse steps k -> k + horizon."""
        gd, u, ghs, _ = seeded_state(self.fm, seed)
        self.fine.upload(gd, ghs, u=u)
        ghs_c = self.cg.cell_mean(ghs)
        cur = 0
        for k in sorted(sample_steps):
            if k < cur:
                raise ValueError("sample steps closer than the horizon")
            self._advance_fine(self.nsub * (k - cur))
            a = self._cg_now()
            self._advance_fine(self.nsub * self.horizon)
            b = self._cg_now()
            cur = k + self.horizon
            yield k, a, b, ghs_c

    def targets(self, a, b, ghs_c):
        g = self.coarse
        g.upload(a[0], ghs_c, u=a[1])
        g.step(self.horizon)
        p = g.download()
        g.upload(b[0], ghs_c, u=b[1])
        iap_b = g.download()["iap_u"]
        gde = 0.5 * (a[0][self.c1] + a[0][self.c2])
        T = self.horizon * self.dt
        t_u = (iap_b - p["iap_u"]) / (T * np.sqrt(gde))
        t_gd = (b[0] - p["gd"]) / T
        return t_u, t_gd

    # ------------------------------------------------------------ training
    def fit_scales(self, mlp: Mlp, pairs, masked_fields=()):
        """Input normalisation per field and output scales from the first pairs;
        fields in masked_fields get a zero input scale (the network does not
        see them)."""
        clo = Closure(self.coarse, mlp)
        mlp.set_scales(np.zeros(NIN), np.ones(NIN), np.ones(NOUT))
        xs, tu, tg = [], [], []
        for a, b, ghs_c in pairs:
            t_u, t_gd = self.targets(a, b, ghs_c)
            self.coarse.upload(a[0], ghs_c, u=a[1])
            xs.append(clo.features().astype(np.float64))
            tu.append(t_u)
            tg.append(t_gd)
        X = np.concatenate(xs)
        sh, sc = np.zeros(NIN), np.ones(NIN)
        for f in range(4):
            blk = X[:, 10 * f:10 * f + 10]
            sh[10 * f:10 * f + 10] = blk.mean()
            sc[10 * f:10 * f + 10] = 0.0 if f in masked_fields else 1.0 / max(blk.std(), 1e-300)
        rms_u = float(np.sqrt(np.mean(np.concatenate(tu) ** 2)))
        rms_g = float(np.sqrt(np.mean(np.concatenate(tg) ** 2)))
        s1 = rms_g * np.mean(self.area) / np.mean(self.dv)
        mlp.set_scales(sh, sc, np.array([rms_u, s1]))
        self.tg_scale = rms_g
        self.log(f"scales: field means {sh[::10].tolist()} inverse stds {sc[::10].tolist()} "
                 f"rms t_u {rms_u:.3e} m/s^2 rms t_gd {rms_g:.3e} m^2/s^3")
        return {"in_shift": sh.tolist(), "in_scale": sc.tolist(), "rms_t_u": rms_u, "rms_t_gd": rms_g}

    def add(self, clo: Closure, a, b, ghs_c):
        t_u, t_gd = self.targets(a, b, ghs_c)
        self.coarse.upload(a[0], ghs_c, u=a[1])
        clo.add_sample(t_u, t_gd, self.tg_scale)

    # ---------------------------------------------------------- evaluation
    def truth_daily(self, seed, days, maps=()):
        """Coarse-grained truth (gd, u) at day 0..days, the coarse ghs, and the
        fine thickness at each day through every cell map in `maps`."""
        gd, u, ghs, _ = seeded_state(self.fm, seed)
        self.fine.upload(gd, ghs, u=u)
        out, mapped = [], [[] for _ in maps]
        per_day = int(round(STEPS_PER_DAY / (self.dt / self.nsub)))
        for d in range(days + 1):
            if d:
                self._advance_fine(per_day)
                self.log(f"  truth seed {seed} day {d}")
            st = self.fine.download()
            out.append((self.cg.cell_mean(st["gd"]), self.cg.edge_velocity(st["u"])))
            for k, mp in enumerate(maps):
                mapped[k].append(mp.cell_mean(st["gd"]))
        return out, self.cg.cell_mean(ghs), mapped

    def coarse_run(self, model: SwmGpu, a, ghs_c, days):
        """Daily thickness of a coarse run from a (NaN once it fails)."""
        model.upload(a[0], ghs_c, u=a[1])
        per_day = int(round(STEPS_PER_DAY / model.opts.dt))
        res = [a[0]]
        ok = True
        for _ in range(days):
            if ok:
                try:
                    model.step(per_day)
                    res.append(model.download()["gd"])
                    continue
                except Exception as e:  # blow-up: non-finite invariants
                    self.log(f"  run stopped: {e}")
                    ok = False
            res.append(np.full(self.cm.nCells, np.nan))
        return res
