"""The Fortran FP64 CPU path (libtmcpu.so, tmcore_amd/fortran/tm_swm_cpu_mod.F90):
the reference's operators, stage, CRK4 step with beta, tracer advection and
invariants in Fortran with OpenMP element loops, for the benchmark
configuration (pv_scheme 3 at interp_pv_order 2-4, adv_order 2).  bench.py
times it as the CPU baseline beside the GPU path.

    gd, iap, q, inv, beta, seconds = cpu_run(mesh, fit, gd, iap, ghs, q, dt, nsteps, threads)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from .mesh import RADIUS, Fit, Mesh

_lib = None


def cpu_lib() -> C.CDLL:
    global _lib
    if _lib is None:
        lib = _ffi._load("libtmcpu.so")
        dp = C.POINTER(C.c_double)
        lib.tm_cpu_run.argtypes = [C.POINTER(_ffi.MeshDesc), C.POINTER(_ffi.FitDesc), C.c_double, C.c_double,
                                   C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, dp, dp, dp, dp, dp]
        lib.tm_cpu_run.restype = C.c_int
        _lib = lib
    return _lib


def cpu_run(mesh: Mesh, fit: Fit | None, gd, iap, ghs, q, dt: float, nsteps: int, threads: int = 0,
            interp_pv_order: int = 3, radius: float = RADIUS):
    """nsteps CRK4 steps of the Fortran CPU path from (gd, iap_u, tracers q
    [T][nCells]) in reference order.  Returns the new gd, iap_u, q, the
    invariants [nsteps][4] and beta [nsteps] of every step, and the wall time
    of the step loop in seconds.  threads 0 keeps OpenMP's default."""
    gd = np.array(gd, dtype=np.float64, copy=True)
    iap = np.array(iap, dtype=np.float64, copy=True)
    ghs = np.ascontiguousarray(ghs, dtype=np.float64)
    q = np.zeros((0, mesh.nCells)) if q is None else np.array(q, dtype=np.float64, copy=True)
    nT = q.shape[0]
    qbuf = q if nT else np.zeros(1)
    inv = np.zeros((max(nsteps, 1), 4))
    beta = np.zeros(max(nsteps, 1))
    el = C.c_double(0.0)
    md = mesh.desc()
    fd = fit.desc() if fit is not None else _ffi.FitDesc()
    rc = cpu_lib().tm_cpu_run(C.byref(md), C.byref(fd), radius, dt, interp_pv_order, nT, nsteps, threads,
                              _ffi.dptr(gd), _ffi.dptr(iap), _ffi.dptr(ghs), _ffi.dptr(qbuf), _ffi.dptr(inv),
                              _ffi.dptr(beta), C.cast(C.pointer(el), C.POINTER(C.c_double)))
    _ffi.check(rc, "tm_cpu_run", "interp_pv_order must be 2..4, with the vertex fits at 3 and 4")
    return gd, iap, q, inv[:nsteps], beta[:nsteps], el.value
