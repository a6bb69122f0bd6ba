"""tm_invariants_noalloc (no device allocation; static per-GPU scratch)
against tm_invariants, through the C-ABI."""
import ctypes as C

import numpy as np
import pytest

from tmcore_amd.mesh import Mesh, seeded_state
from tmcore_amd.swm import SwmGpu, gpu_lib

pytestmark = pytest.mark.gpu


def _model(level, precision, steps):
    m = Mesh(level)
    g = SwmGpu(m, m.fit(), dt=600.0, ntracers=1, precision=precision)
    gd, u, ghs, q = seeded_state(m, 7, [301])
    g.upload(gd, ghs, u=u, tracers=q)
    if steps:
        g.step(steps)
    return g


@pytest.mark.parametrize("precision", [64, 32])
def test_invariants_noalloc_matches_tm_invariants(precision):
    """Two calls in a row on one GPU, the second on a smaller mesh (level 5,
    then level 4, both contexts alive): each equals tm_invariants bit for
    bit (same kernels, same partial layout), and the larger context's call
    still agrees after the smaller one has used the shared scratch."""
    L = gpu_lib()
    big = _model(5, precision, 3)
    small = _model(4, precision, 2)
    out = np.zeros(4)
    assert L.tm_invariants_noalloc(big.h, out.ctypes.data_as(C.POINTER(C.c_double))) == 0
    np.testing.assert_array_equal(out, big.invariants())
    out_small = small.invariants_noalloc()
    np.testing.assert_array_equal(out_small, small.invariants())
    assert not np.array_equal(out_small, out)
    big.step(1)
    np.testing.assert_array_equal(big.invariants_noalloc(), big.invariants())
    big.close()
    small.close()


def test_invariants_noalloc_rejects_meshes_beyond_the_scratch():
    """Level 9 needs 30,720 edge workgroups, more than the 16,384 the static
    scratch holds: a clear error, and tm_invariants still serves the mesh."""
    m = Mesh(9)
    g = SwmGpu(m, None, dt=300.0, pv_scheme=1, precision=32)
    gd, u, ghs, _ = seeded_state(m, 7)
    g.upload(gd, ghs, u=u)
    out = np.zeros(4)
    rc = gpu_lib().tm_invariants_noalloc(g.h, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == -1  # TM_EINVAL
    assert b"4,194,304" in gpu_lib().tm_last_error(g.h)
    assert np.isfinite(g.invariants()).all()
    g.close()
