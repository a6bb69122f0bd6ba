"""Batched ensemble: every member must match its own single-model run.

FP64 members agree with tm_step to <= 1e-12 relative L2 (the per-element
arithmetic is the same code; only the order of the global beta sums differs).
FP32 members agree with the FP32 single-model path to <= 1e-5.
"""
import numpy as np
import pytest

from tmcore_amd.ensemble import Ensemble
from tmcore_amd.mesh import Mesh, seeded_state
from tmcore_amd.swm import SwmGpu

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def _run_case(precision, tol, ntracers=0, tuning=None, specs=((3, 5, 600.0, 6), (4, 3, 300.0, 4)), slices=None):
    # specs: (level, members, dt, steps): uneven sizes and counts
    meshes = {L: Mesh(L) for L, *_ in specs}
    ens = Ensemble([(meshes[L], meshes[L].fit(), M, dt) for L, M, dt, _ in specs], precision=precision,
                   ntracers=ntracers, slices=slices)
    if slices is not None:
        assert ens.slices == list(slices)
    if tuning:
        ens.set_tuning(*tuning)
    states = {}
    for li, (L, M, dt, _) in enumerate(specs):
        for m in range(M):
            gd, u, ghs, q = seeded_state(meshes[L], 1000 + 10 * L + m, [5000 + m] * ntracers)
            states[li, m] = (gd, u, ghs, q)
            ens.upload(li, m, gd, u, ghs, q if ntracers else None)
    ens.run([s for *_, s in specs])
    for li, (L, M, dt, steps) in enumerate(specs):
        for m in range(M):
            gd, u, ghs, q = states[li, m]
            g = SwmGpu(meshes[L], meshes[L].fit(), dt=dt, precision=precision, ntracers=ntracers)
            g.upload(gd, ghs, u=u, tracers=q if ntracers else None)
            inv, _ = g.step(steps)
            ref = g.download()
            got = ens.download(li, m)
            assert rel_l2(got["gd"], ref["gd"]) <= tol, (L, m)
            assert rel_l2(got["iap_u"], ref["iap_u"]) <= tol, (L, m)
            if ntracers:
                assert rel_l2(got["tracers"], ref["tracers"]) <= tol
            ei = ens.invariants(li, m)
            np.testing.assert_allclose(ei[:3], inv[-1, :3], rtol=max(tol, 1e-12))
            g.close()
    ens.close()


def test_ensemble_fp64_members_match_single_runs():
    _run_case(64, 1e-12, ntracers=1)


def test_ensemble_fp32_members_match_single_runs():
    _run_case(32, 1e-5)


def test_ensemble_tuning_does_not_change_results():
    _run_case(64, 1e-12, tuning=(64, 1, 1))


@pytest.mark.parametrize("precision,width", [(32, 1), (32, 2), (32, 4), (64, 2)])
def test_ensemble_members_per_thread(precision, width):
    """W members per thread (vector loads of member rows) gives the same members."""
    print(f"ensemble precision={precision} members_per_thread={width}", flush=True)
    _run_case(precision, 1e-12 if precision == 64 else 1e-5, ntracers=1, tuning=(128, 1, width),
              specs=((3, 8, 600.0, 6), (4, 4, 300.0, 4)))


@pytest.mark.parametrize("precision,slices", [(32, (8, 2)), (32, (4, 8)), (64, (2, 4))])
def test_ensemble_member_slices_match_single_runs(precision, slices):
    """Members cut into slices (each slice its own member-fastest arrays,
    the slices dealt to the XCDs) still match their single-model runs."""
    print(f"ensemble precision={precision} slices={slices}", flush=True)
    _run_case(precision, 1e-12 if precision == 64 else 1e-5, ntracers=1, slices=slices,
              specs=((3, 16, 600.0, 6), (4, 8, 300.0, 4)))


def test_ensemble_slice_rules():
    """Slices divide the level's members, number 1..8, and are fixed once
    state is uploaded; the default keeps 16 members in a slice."""
    m = Mesh(2)
    ens = Ensemble([(m, m.fit(), 16, 600.0), (m, m.fit(), 6, 600.0), (m, m.fit(), 128, 600.0)], precision=32)
    assert ens.slices == [1, 1, 8]
    for bad in (0, 3, 9, 16):
        with pytest.raises(Exception):
            ens.set_slices(0, bad)
    ens.set_slices(0, 2)
    assert ens.slices == [2, 1, 8]
    gd, u, ghs, _ = seeded_state(m, 3)
    ens.upload(0, 0, gd, u, ghs)
    with pytest.raises(Exception, match="before the first upload"):
        ens.set_slices(0, 4)
    ens.close()


def test_ensemble_rejects_bad_member_width():
    m = Mesh(2)
    ens = Ensemble([(m, m.fit(), 6, 600.0)], precision=32)
    with pytest.raises(Exception):
        ens.set_tuning(128, 1, 4)  # 4 does not divide 6
    ens.set_tuning(128, 1, 2)
    ens.close()


def test_ensemble_autotune_restores_state():
    m = Mesh(3)
    ens = Ensemble([(m, m.fit(), 4, 600.0)], precision=32)
    gd, u, ghs, _ = seeded_state(m, 11)
    for k in range(4):
        ens.upload(0, k, gd, u, ghs)
    before = ens.download(0, 2)
    (chunk, bpc, width), ms = ens.autotune(ticks=1)
    after = ens.download(0, 2)
    np.testing.assert_array_equal(before["gd"], after["gd"])
    assert chunk in (64, 128, 256, 512, 1024) and bpc >= 1 and width in (1, 2, 4) and ms > 0
    ens.close()
