"""Batched ensemble over meshes of uneven size (tm_ens_* in include/tmcore_gpu.h).

All members of all level groups advance in one persistent kernel launch per
``run``; members of a level share its mesh constants on the device.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from .mesh import RADIUS
from .swm import Opts, _f64, _p, gpu_lib

_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


class EnsLevel(C.Structure):
    _fields_ = [("mesh", C.POINTER(_ffi.MeshDesc)), ("fit", C.POINTER(_ffi.FitDesc)), ("members", C.c_int),
                ("dt", C.c_double)]


def _bind(L):
    if getattr(L, "_ens_bound", False):
        return L
    L.tm_ens_create.argtypes = [C.POINTER(_vp), C.c_int, C.POINTER(EnsLevel), C.POINTER(Opts)]
    L.tm_ens_destroy.argtypes = [_vp]
    L.tm_ens_last_error.argtypes = [_vp]
    L.tm_ens_last_error.restype = C.c_char_p
    L.tm_ens_upload.argtypes = [_vp, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
    L.tm_ens_download.argtypes = [_vp, C.c_int, C.c_int, _dp, _dp, _dp]
    L.tm_ens_run.argtypes = [_vp, C.POINTER(C.c_int), _dp]
    L.tm_ens_invariants.argtypes = [_vp, C.c_int, C.c_int, _dp]
    L.tm_ens_set_tuning.argtypes = [_vp, C.c_int, C.c_int, C.c_int]
    L.tm_ens_autotune.argtypes = [_vp, C.c_int, C.POINTER(C.c_int), _dp]
    L.tm_ens_set_slices.argtypes = [_vp, C.c_int, C.c_int]
    L.tm_ens_get_slices.argtypes = [_vp, C.c_int, C.POINTER(C.c_int)]
    for f in ("tm_ens_create", "tm_ens_destroy", "tm_ens_upload", "tm_ens_download", "tm_ens_run",
              "tm_ens_invariants", "tm_ens_set_tuning", "tm_ens_autotune", "tm_ens_set_slices",
              "tm_ens_get_slices"):
        getattr(L, f).restype = C.c_int
    L._ens_bound = True
    return L


class Ensemble:
    """levels: list of (mesh, fit, members, dt).  slices: member slices per
    level (tm_ens_set_slices), None for the library's default."""

    def __init__(self, levels, *, precision: int = 32, ntracers: int = 0, radius: float = RADIUS,
                 interp_pv_order: int = 3, device: int = 0, slices=None):
        L = _bind(gpu_lib())
        o = Opts()
        L.tm_opts_default(C.byref(o))
        o.precision, o.ntracers, o.radius, o.interp_pv_order, o.device = precision, ntracers, radius, \
            interp_pv_order, device
        self.levels = levels
        self._keep = []
        arr = (EnsLevel * len(levels))()
        for i, (mesh, fit, members, dt) in enumerate(levels):
            md, fd = mesh.desc(), fit.desc()
            self._keep += [md, fd]
            arr[i] = EnsLevel(C.pointer(md), C.pointer(fd), members, dt)
        h = _vp()
        rc = L.tm_ens_create(C.byref(h), len(levels), arr, C.byref(o))
        if rc != 0:
            raise _ffi.TmError(rc, "tm_ens_create")
        self.h = h
        self.nT = ntracers
        if slices is not None:
            for lev, s in enumerate(slices):
                self.set_slices(lev, s)

    def set_slices(self, level, slices):
        """tm_ens_set_slices: before the first upload."""
        self._check(gpu_lib().tm_ens_set_slices(self.h, level, int(slices)), "tm_ens_set_slices")

    @property
    def slices(self):
        out = []
        for lev in range(len(self.levels)):
            s = C.c_int()
            self._check(gpu_lib().tm_ens_get_slices(self.h, lev, C.byref(s)), "tm_ens_get_slices")
            out.append(s.value)
        return out

    def _check(self, rc, where):
        if rc != 0:
            msg = gpu_lib().tm_ens_last_error(self.h)
            raise _ffi.TmError(rc, where, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            gpu_lib().tm_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def upload(self, level, member, gd, u, ghs, tracers=None):
        gd, u, ghs, tr = (_f64(x) for x in (gd, u, ghs, tracers))
        self._check(gpu_lib().tm_ens_upload(self.h, level, member, _p(gd), _p(u), _p(ghs), _p(tr)),
                    "tm_ens_upload")

    def download(self, level, member):
        mesh = self.levels[level][0]
        gd = np.zeros(mesh.nCells)
        iap = np.zeros(mesh.nEdges)
        tr = np.zeros((self.nT, mesh.nCells))
        self._check(gpu_lib().tm_ens_download(self.h, level, member, _p(gd), _p(iap), _p(tr) if self.nT else None),
                    "tm_ens_download")
        return {"gd": gd, "iap_u": iap, "tracers": tr}

    def run(self, nsteps) -> float:
        """Advance level l by nsteps[l] steps in one launch; returns device ms."""
        ns = (C.c_int * len(self.levels))(*nsteps)
        ms = C.c_double()
        self._check(gpu_lib().tm_ens_run(self.h, ns, C.byref(ms)), "tm_ens_run")
        return ms.value

    def invariants(self, level, member):
        out = np.zeros(4)
        self._check(gpu_lib().tm_ens_invariants(self.h, level, member, _p(out)), "tm_ens_invariants")
        return out

    def set_tuning(self, chunk, blocks_per_cu, members_per_thread=1):
        self._check(gpu_lib().tm_ens_set_tuning(self.h, chunk, blocks_per_cu, members_per_thread),
                    "tm_ens_set_tuning")

    def autotune(self, ticks=2):
        """Returns ((chunk, blocks_per_cu, members_per_thread), ms per tick)."""
        best = (C.c_int * 3)()
        ms = C.c_double()
        self._check(gpu_lib().tm_ens_autotune(self.h, ticks, best, C.byref(ms)), "tm_ens_autotune")
        return (best[0], best[1], best[2]), ms.value
