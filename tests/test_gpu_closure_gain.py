"""The trained closure (tmcore_amd/data/closure_L6.npz) against the plain
method on a held-out initial state: level-6 runs (dt 300 s) from the level-9
truth coarse-grained to level 6, one simulated day, relative L2 thickness
error against the coarse-grained truth.  The closure must beat the plain
method it corrects, and keep total mass."""
import os

import numpy as np
import pytest

from tmcore_amd.closure import Closure, Mlp
from tmcore_amd.closure_train import Pipeline, l2_thickness, load
from tmcore_amd.mesh import seeded_state
from tmcore_amd.swm import SwmGpu

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tmcore_amd", "data",
                       "closure_L6.npz")


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="no trained closure weights")
def test_trained_closure_beats_plain_on_held_out_state():
    seed, hours = 6000, 24
    mlp = Mlp(seed=0)
    meta = load(WEIGHTS, mlp)
    p = Pipeline(9, 6, 300.0, 8, log=lambda *a: print(*a, flush=True), horizon=int(meta["horizon"]))
    gd, u, ghs, _ = seeded_state(p.fm, seed)
    p.fine.upload(gd, ghs, u=u)
    a = p._cg_now()
    ghs_c = p.cg.cell_mean(ghs)
    p._advance_fine(int(hours * 3600 / 37.5))
    truth = p._cg_now()
    errs = {}
    for name in ("plain", "closure"):
        g = SwmGpu(p.cm, p.cm.fit(), dt=300.0)
        if name == "closure":
            clo = Closure(g, mlp)
            clo.attach()
            clo.set_interval(int(meta["horizon"]), lumped=True)
        g.upload(a[0], ghs_c, u=a[1])
        inv, _ = g.step(int(hours * 12))
        st = g.download()
        errs[name] = l2_thickness(st["gd"], truth[0], p.area)
        print(f"{name}: relative L2 thickness error after {hours} h = {errs[name]:.4e}", flush=True)
        # mass of the corrected run equals the mass of the start state to round-off
        m0 = np.sum(p.area * a[0])
        assert abs(np.sum(p.area * st["gd"]) / m0 - 1) <= 1e-12
        g.close()
    assert errs["closure"] < 0.97 * errs["plain"], errs
    mlp.close()
