"""Learned closure on the device (libtmgpu tm_mlp_* / tm_closure_*) against
the FP64 closure oracle (oracle/mlp_oracle.py).

Bars: the MLP forward/backward agree with the oracle that rounds where the
device rounds (BF16 operands, FP32 MFMA accumulation here vs exact there) to
1e-3 relative L2 for the outputs and 1e-2 for gradients, and with the
unrounded FP64 oracle to 3e-2 (the BF16 budget); the inputs agree with the
oracle's to FP32 rounding; a closure whose outputs are zero leaves the step
bitwise equal to the plain step; the corrected step's increment agrees with
the oracle's to 5e-3 and keeps total mass to round-off.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as mo
from oracle.oracle import Oracle
from tmcore_amd.closure import NPARAMS, Closure, Mlp
from tmcore_amd.mesh import RADIUS, Mesh, seeded_state
from tmcore_amd.swm import SwmGpu

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)


_M = {}


def mesh(level):
    if level not in _M:
        _M[level] = Mesh(level)
    return _M[level]


@pytest.mark.parametrize("n", [1, 17, 1000])
def test_mlp_forward_matches_oracle(n):
    mlp = Mlp(seed=1)
    p = mlp.params.astype(np.float64)
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, 40)).astype(np.float32)
    print(f"mlp forward n={n}", flush=True)
    Y = mlp.forward(X)
    Yb, _ = mo.forward(p, X, bf=True)
    Yd, _ = mo.forward(p, X)
    assert rel_l2(Y, Yb) <= 1e-3, rel_l2(Y, Yb)
    assert rel_l2(Y, Yd) <= 3e-2, rel_l2(Y, Yd)
    mlp.close()


def test_mlp_backward_matches_oracle():
    mlp = Mlp(seed=2)
    # larger last layer so that every layer's gradient is well above rounding
    p = mlp.params
    p[-(128 * 2 + 2):] *= 10.0
    mlp.params = p
    p = mlp.params.astype(np.float64)
    rng = np.random.default_rng(4)
    n = 1000
    X = rng.standard_normal((n, 40)).astype(np.float32)
    dY = rng.standard_normal((n, 2)).astype(np.float32)
    print("mlp backward n=1000", flush=True)
    g, dX = mlp.backward(X, dY)
    gb, dXb = mo.backward(p, X, dY, bf=True)
    gd, dXd = mo.backward(p, X, dY)
    G, Gb, Gd = mo.unpack(g), mo.unpack(gb), mo.unpack(gd)
    for k in G:
        assert rel_l2(G[k], Gb[k]) <= 1e-2, (k, rel_l2(G[k], Gb[k]))
        assert rel_l2(G[k], Gd[k]) <= 3e-2, (k, rel_l2(G[k], Gd[k]))
    assert rel_l2(dX, dXb) <= 1e-2
    assert rel_l2(dX, dXd) <= 3e-2
    mlp.close()


def test_mlp_adam_step_and_param_round_trip():
    mlp = Mlp(seed=3)
    p0 = mlp.params
    rng = np.random.default_rng(8)
    g = (rng.standard_normal(NPARAMS) * 10.0 ** rng.integers(-3, 2, NPARAMS)).astype(np.float32)
    lr = 1e-3
    mlp.adam_step(g, lr)
    # first Adam step with bias correction: p - lr * g / (|g| + eps)
    ref = p0.astype(np.float64) - lr * g / (np.abs(g.astype(np.float64)) + 1e-8)
    np.testing.assert_allclose(mlp.params, ref, rtol=1e-6, atol=1e-7)
    q = rng.standard_normal(NPARAMS).astype(np.float32)
    mlp.params = q
    np.testing.assert_array_equal(mlp.params, q)
    q[5] = np.nan
    with pytest.raises(Exception, match="non-finite"):
        mlp.params = q
    with pytest.raises(Exception):
        mlp.set_scales(np.zeros(40), np.ones(40), np.array([1.0, 0.0]))
    mlp.close()


def scales_for(m, gd, u):
    """Input normalisation used by the tests (the training pipeline derives
    its own from data)."""
    sh = np.zeros(40)
    sc = np.ones(40)
    sh[10:20] = np.mean(gd)
    sc[0:10] = 1.0 / 20.0
    sc[10:20] = 1.0 / (9.80616 * 500.0)
    sc[20:30] = 1.0 / 1e-5
    sc[30:40] = 1.0 / 1e-6
    return sh, sc


def setup_closure(level, precision=64, seed=2, dt=600.0, mlp_seed=5, s_out=(1e-6, 1e-3)):
    m = mesh(level)
    gd, u, ghs, _ = seeded_state(m, seed)
    g = SwmGpu(m, m.fit(), dt=dt, precision=precision)
    g.upload(gd, ghs, u=u)
    mlp = Mlp(seed=mlp_seed)
    sh, sc = scales_for(m, gd, u)
    # s_out[1]: thickness tendency (m^2 s^-3) of a unit flux output on one edge of a mean cell
    s1 = s_out[1] * np.mean(m.areaCell * RADIUS ** 2) / np.mean(m.dvEdge * RADIUS)
    out = np.array([s_out[0], s1])
    mlp.set_scales(sh, sc, out)
    return m, g, mlp, (gd, u, ghs), (sh, sc, out)


@pytest.mark.parametrize("precision", [64, 32])
def test_closure_inputs_match_oracle(precision):
    m, g, mlp, (gd, u, ghs), (sh, sc, out) = setup_closure(3, precision)
    # the seeded wind is non-divergent by construction: perturb it so that
    # the divergence input is more than round-off
    u = u + 0.5 * np.random.default_rng(1).standard_normal(m.nEdges)
    g.upload(gd, ghs, u=u)
    print(f"closure inputs L3 precision={precision}", flush=True)
    x = Closure(g, mlp).features()
    o = Oracle(m, m.fit(), dt=600.0)
    iap = o.initial_iap(gd, u)
    tg, _, d = o.stage(gd, iap, ghs)
    xr = mo.features(m, d["u"], d["gd_edge"], d["vor"], mo.divergence(m, d["u"], RADIUS), sh, sc)
    tol = 1e-6 if precision == 64 else 1e-3
    for f in range(4):
        blk = slice(10 * f, 10 * f + 10)
        assert rel_l2(x[:, blk], xr[:, blk]) <= tol, (f, rel_l2(x[:, blk], xr[:, blk]))
    g.close()
    mlp.close()


def test_zero_output_closure_equals_plain_step_bitwise():
    m, g, mlp, (gd, u, ghs), _ = setup_closure(4)
    p = mlp.params
    p[-(128 * 2 + 2):] = 0.0  # W4 and b4
    mlp.params = p
    g0 = SwmGpu(m, m.fit(), dt=600.0)
    g0.upload(gd, ghs, u=u)
    Closure(g, mlp).attach()
    print("zero-output closure L4, 3 steps", flush=True)
    i1, b1 = g.step(3)
    i0, b0 = g0.step(3)
    s1, s0 = g.download(), g0.download()
    np.testing.assert_array_equal(s1["gd"], s0["gd"])
    np.testing.assert_array_equal(s1["iap_u"], s0["iap_u"])
    np.testing.assert_array_equal(b1, b0)
    g.close()
    g0.close()
    mlp.close()


def test_closure_step_matches_oracle_and_keeps_mass():
    m, g, mlp, (gd, u, ghs), (sh, sc, out) = setup_closure(4, s_out=(2e-5, 1e-1))
    p = mlp.params.astype(np.float64)
    dt = 600.0
    o = Oracle(m, m.fit(), dt=dt)
    iap = o.initial_iap(gd, u)
    tg, _, d = o.stage(gd, iap, ghs)
    gd_p, iap_p, _, inv_p, _ = o.run(1, gd, iap, ghs)
    xr = mo.features(m, d["u"], d["gd_edge"], d["vor"], mo.divergence(m, d["u"], RADIUS), sh, sc)
    Yr, _ = mo.forward(p, xr, bf=True)
    gd_c, iap_c = mo.apply(m, gd_p, iap_p, d["gd_edge"], Yr, out, dt, RADIUS)
    Closure(g, mlp).attach()
    print("closure step L4 vs oracle", flush=True)
    g.step(1)
    st = g.download()
    assert rel_l2(gd_c - gd_p, gd_p - gd) > 1e-3  # the correction is not negligible
    assert rel_l2(st["gd"] - gd_p, gd_c - gd_p) <= 5e-3
    assert rel_l2(st["iap_u"] - iap_p, iap_c - iap_p) <= 5e-3
    A = m.areaCell
    assert abs(np.sum(A * st["gd"]) / np.sum(A * gd) - 1) <= 1e-14
    # the closure's invariants come from the corrected state
    inv = g.invariants()
    assert abs(inv[0] / inv_p[0, 0] - 1) <= 1e-13
    g.close()
    mlp.close()


def test_closure_interval_holds_the_correction():
    """set_interval(3): the MLP runs on step 0 and its correction (with the
    sqrt(gd_edge) of step 0) is added after each of the 3 steps, or (lumped)
    once after the third step with 3 dt."""
    m, g, mlp, (gd, u, ghs), (sh, sc, out) = setup_closure(3, s_out=(2e-5, 1e-1))
    p = mlp.params.astype(np.float64)
    dt = 600.0
    o = Oracle(m, m.fit(), dt=dt)
    iap = o.initial_iap(gd, u)
    tg, _, d = o.stage(gd, iap, ghs)
    xr = mo.features(m, d["u"], d["gd_edge"], d["vor"], mo.divergence(m, d["u"], RADIUS), sh, sc)
    Yr, _ = mo.forward(p, xr, bf=True)
    gd_c, iap_c = gd, iap
    for _ in range(3):
        gp, ip, _, _, _ = o.run(1, gd_c, iap_c, ghs)
        gd_c, iap_c = mo.apply(m, gp, ip, d["gd_edge"], Yr, out, dt, RADIUS)
    gd_p, iap_p, _, _, _ = o.run(3, gd, iap, ghs)
    clo = Closure(g, mlp)
    clo.attach()
    clo.set_interval(3)
    print("closure interval 3, L3", flush=True)
    g.step(3)
    st = g.download()
    assert rel_l2(st["gd"] - gd_p, gd_c - gd_p) <= 5e-3
    assert rel_l2(st["iap_u"] - iap_p, iap_c - iap_p) <= 5e-3
    # lumped: 3 plain steps, then the correction once with 3 dt
    gd_l, iap_l = mo.apply(m, gd_p, iap_p, d["gd_edge"], Yr, out, 3 * dt, RADIUS)
    g.upload(gd, ghs, u=u)
    clo.set_interval(3, lumped=True)
    g.step(2)
    st2 = g.download()
    gd_p2, iap_p2, _, _, _ = o.run(2, gd, iap, ghs)
    assert rel_l2(st2["gd"], gd_p2) <= 1e-13  # no correction before the interval ends
    g.step(1)
    st = g.download()
    assert rel_l2(st["gd"] - gd_p, gd_l - gd_p) <= 5e-3
    assert rel_l2(st["iap_u"] - iap_p, iap_l - iap_p) <= 5e-3
    with pytest.raises(Exception):
        clo.set_interval(0)
    g.close()
    mlp.close()


def test_closure_gain_scales_the_correction():
    """set_gain(0) is the plain step bit for bit; set_gain(0.5) applies half
    the oracle's correction; a negative gain is refused."""
    m, g, mlp, (gd, u, ghs), (sh, sc, out) = setup_closure(3, s_out=(2e-5, 1e-1))
    p = mlp.params.astype(np.float64)
    dt = 600.0
    o = Oracle(m, m.fit(), dt=dt)
    iap = o.initial_iap(gd, u)
    tg, _, d = o.stage(gd, iap, ghs)
    xr = mo.features(m, d["u"], d["gd_edge"], d["vor"], mo.divergence(m, d["u"], RADIUS), sh, sc)
    Yr, _ = mo.forward(p, xr, bf=True)
    gd_p, iap_p, _, _, _ = o.run(1, gd, iap, ghs)
    gd_h, iap_h = mo.apply(m, gd_p, iap_p, d["gd_edge"], Yr, out, 0.5 * dt, RADIUS)
    clo = Closure(g, mlp)
    clo.attach()
    clo.set_gain(0.0)
    print("closure gain 0 / 0.5, L3", flush=True)
    g.step(1)
    g0 = SwmGpu(m, m.fit(), dt=dt)
    g0.upload(gd, ghs, u=u)
    g0.step(1)
    np.testing.assert_array_equal(g.download()["gd"], g0.download()["gd"])
    g.upload(gd, ghs, u=u)
    clo.set_gain(0.5)
    g.step(1)
    st = g.download()
    assert rel_l2(st["gd"] - gd_p, gd_h - gd_p) <= 5e-3
    assert rel_l2(st["iap_u"] - iap_p, iap_h - iap_p) <= 5e-3
    with pytest.raises(Exception):
        clo.set_gain(-1.0)
    g.close()
    g0.close()
    mlp.close()


def test_closure_training_reduces_loss_and_eval_matches_oracle():
    """Targets from a fixed 'teacher' network on the same inputs; the student
    trained by tm_closure_train on the device must fit them."""
    m, g, student, (gd, u, ghs), (sh, sc, out) = setup_closure(3, mlp_seed=11, s_out=(2e-5, 1e-3))
    teacher = Mlp(seed=12)
    teacher.set_scales(sh, sc, out)
    clo, clo_t = Closure(g, student), Closure(g, teacher)
    kap = mo.kappa(m)
    tgs = 1e-3
    xs, targets = [], []
    for k in range(6):
        gdk, uk, ghsk, _ = seeded_state(m, 40 + k)
        g.upload(gdk, ghsk, u=uk)
        x = clo_t.features()
        yt = teacher.forward(x).astype(np.float64)
        t_u = kap * out[0] * yt[:, 0]
        t_gd = mo.flux_div(m, yt[:, 1], out[1], RADIUS)
        clo.add_sample(t_u, t_gd, tgs)
        xs.append(x)
        targets.append((t_u, t_gd))
    assert clo.size() == 6
    # device loss of the untrained student == oracle loss of its outputs
    l0 = clo.eval(0, 6)
    lr_ = np.zeros(2)
    for x, (t_u, t_gd) in zip(xs, targets):
        lu, lg, _ = mo.loss_and_dy(m, student.forward(x), t_u, t_gd, out, tgs, RADIUS)
        lr_ += np.array([lu, lg]) / 6
    np.testing.assert_allclose(l0, lr_, rtol=1e-4)
    print("closure training L3, 6 samples x 60 epochs", flush=True)
    hist = clo.train(5, 60, 2e-3, seed=1)
    assert np.all(np.isfinite(hist))
    assert hist[-1].sum() < 0.2 * hist[0].sum(), hist[[0, -1]]
    l1 = clo.eval(0, 5)
    assert l1.sum() < 0.2 * l0.sum()
    clo.clear()
    assert clo.size() == 0
    g.close()
    student.close()
    teacher.close()


def test_training_is_bitwise_deterministic():
    """The weight gradients are summed per fixed row range and the partials
    added in a fixed order (no atomics): two trainings from the same start,
    samples and shuffle seed give bit-identical parameters and losses."""
    m = Mesh(3)
    out = []
    for _ in range(2):
        g = SwmGpu(m, m.fit(), dt=600.0)
        gd, u, ghs, _ = seeded_state(m, 9)
        g.upload(gd, ghs, u=u)
        mlp = Mlp(seed=3)
        clo = Closure(g, mlp)
        rng = np.random.default_rng(0)
        for _ in range(3):
            clo.add_sample(rng.standard_normal(m.nEdges) * 1e-5, rng.standard_normal(m.nCells) * 1e-3, 1e-3)
        hist = clo.train(3, 4, 1e-3, seed=5)
        out.append((np.array(mlp.params), np.array(hist)))
        mlp.close()
        g.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
