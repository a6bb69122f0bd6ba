"""TEST INFRASTRUCTURE: ctypes wrapper of the C oracle (build/liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker of the HIP path.  The product package tmcore_amd
never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from tmcore_amd import _ffi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "liboracle.so")

_dp = C.POINTER(C.c_double)


class OmParams(C.Structure):
    _fields_ = [("time_scheme", C.c_int), ("pv_scheme", C.c_int), ("interp_pv_order", C.c_int),
                ("adv_order", C.c_int), ("adv_mono", C.c_int), ("conserve_energy", C.c_int),
                ("apvm_weight", C.c_double), ("dt", C.c_double)]


class OmDiag(C.Structure):
    _fields_ = [(k, _dp) for k in ("gd_edge", "u", "tend_gd", "tend_gd_edge", "tend_gd_vertex", "gd_vertex",
                                   "vor", "v", "ke", "pv_vertex", "pv_cell", "pv_edge", "pv_flx", "u_tend",
                                   "iap_tend")]


DIAG_WHERE = {"gd_edge": "e", "u": "e", "tend_gd": "c", "tend_gd_edge": "e", "tend_gd_vertex": "v",
              "gd_vertex": "v", "vor": "v", "v": "e", "ke": "c", "pv_vertex": "v", "pv_cell": "c",
              "pv_edge": "e", "pv_flx": "e", "u_tend": "e", "iap_tend": "e"}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp = C.c_void_p
        L.om_create.restype = vp
        L.om_create.argtypes = [C.POINTER(_ffi.MeshDesc), C.POINTER(_ffi.FitDesc), C.c_double, C.POINTER(OmParams)]
        L.om_destroy.argtypes = [vp]
        L.om_set_params.argtypes = [vp, C.POINTER(OmParams)]
        L.om_poly_fit.argtypes = [C.POINTER(_ffi.MeshDesc), C.c_double, C.POINTER(_ffi.FitDesc)]
        L.om_spatial_operators.argtypes = [vp, _dp, _dp, _dp, C.c_int, _dp, _dp, _dp, _dp, C.POINTER(OmDiag)]
        L.om_initial_iap.argtypes = [vp, _dp, _dp, _dp]
        L.om_invariants.argtypes = [vp, _dp, _dp, _dp, _dp]
        L.om_run.argtypes = [vp, C.c_int, _dp, _dp, _dp, C.c_int, _dp, _dp, _dp]
        for name in ("om_div", "om_curl", "om_c2e", "om_c2v", "om_v2c", "om_v2e", "om_ke", "om_tangent_wind"):
            getattr(L, name).argtypes = [vp, _dp, _dp]
        for name in ("om_iap", "om_inverse_iap", "om_pv_vertex", "om_upwind_cell"):
            getattr(L, name).argtypes = [vp, _dp, _dp, _dp]
        L.om_tangent_vor_flux.argtypes = [vp, _dp, _dp, _dp, _dp]
        L.om_pv_edge.argtypes = [vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.om_gd_tend.argtypes = [vp, _dp, _dp, _dp]
        L.om_u_tend.argtypes = [vp] + [_dp] * 9
        L.om_adv.argtypes = [vp, _dp, _dp, _dp, _dp]
        L.om_reconstruct_coeffs.argtypes = [C.POINTER(_ffi.MeshDesc), C.c_double, _dp]
        L.om_reconstruct.argtypes = [C.POINTER(_ffi.MeshDesc), _dp, _dp, _dp]
        L.om_set_threads.argtypes = [C.c_int]
        L.om_set_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_dp)


class Oracle:
    """FP64 CPU restatement of the reference path on one mesh."""

    def __init__(self, mesh, fit=None, radius=6371220.0, dt=600.0, time_scheme=2, pv_scheme=3,
                 interp_pv_order=3, adv_order=2, adv_mono=True, conserve_energy=True, apvm_weight=0.5):
        self.mesh = mesh
        self.params = OmParams(time_scheme, pv_scheme, interp_pv_order, adv_order, int(adv_mono),
                               int(conserve_energy), apvm_weight, dt)
        self._md = mesh.desc()
        self._fd = fit.desc() if fit is not None else None
        self.h = lib().om_create(C.byref(self._md), C.byref(self._fd) if fit is not None else None, radius,
                                 C.byref(self.params))
        self.nC, self.nE, self.nV = mesh.nCells, mesh.nEdges, mesh.nVertices

    def __del__(self):
        if getattr(self, "h", None):
            lib().om_destroy(self.h)
            self.h = None

    def n(self, where):
        return {"c": self.nC, "e": self.nE, "v": self.nV}[where]

    def op(self, name, *args, out_where):
        """Call a single-output operator om_<name>(ctx, in..., out)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in args]
        out = np.zeros(self.n(out_where))
        getattr(lib(), "om_" + name)(self.h, *[a.ctypes.data_as(_dp) for a in arrs], out.ctypes.data_as(_dp))
        return out

    def initial_iap(self, gd, u):
        iap = np.zeros(self.nE)
        lib().om_initial_iap(self.h, _p(gd), _p(u), iap.ctypes.data_as(_dp))
        return iap

    def stage(self, gd, iap, ghs, q=None):
        q = np.zeros((0, self.nC)) if q is None else np.ascontiguousarray(q, dtype=np.float64)
        nT = q.shape[0]
        tg, ti, tq = np.zeros(self.nC), np.zeros(self.nE), np.zeros((nT, self.nC))
        diag = {k: np.zeros(self.n(w)) for k, w in DIAG_WHERE.items()}
        d = OmDiag(**{k: v.ctypes.data_as(_dp) for k, v in diag.items()})
        gd, iap, ghs = (np.ascontiguousarray(x, dtype=np.float64) for x in (gd, iap, ghs))
        lib().om_spatial_operators(self.h, _p(gd), _p(iap), _p(ghs), nT, q.ctypes.data_as(_dp) if nT else None,
                                   tg.ctypes.data_as(_dp), ti.ctypes.data_as(_dp),
                                   tq.ctypes.data_as(_dp) if nT else None, C.byref(d))
        diag["tend_q"] = tq
        return tg, ti, diag

    def invariants(self, gd, iap, ghs):
        out = np.zeros(4)
        lib().om_invariants(self.h, _p(gd), _p(iap), _p(ghs), out.ctypes.data_as(_dp))
        return out

    def run(self, nsteps, gd, iap, ghs, q=None):
        """Returns (gd, iap, q, invariants[nsteps,4], beta[nsteps]) after nsteps."""
        gd = np.array(gd, dtype=np.float64)
        iap = np.array(iap, dtype=np.float64)
        q = np.zeros((0, self.nC)) if q is None else np.array(q, dtype=np.float64)
        ghs = np.ascontiguousarray(ghs, dtype=np.float64)
        nT = q.shape[0]
        inv = np.zeros((nsteps, 4))
        beta = np.zeros(nsteps)
        rc = lib().om_run(self.h, nsteps, gd.ctypes.data_as(_dp), iap.ctypes.data_as(_dp), ghs.ctypes.data_as(_dp),
                          nT, q.ctypes.data_as(_dp) if nT else None, inv.ctypes.data_as(_dp),
                          beta.ctypes.data_as(_dp))
        if rc != 0:
            raise RuntimeError(f"om_run failed ({rc})")
        return gd, iap, q, inv, beta


def oracle_poly_fit(mesh, radius=6371220.0, max_fit=31, order4=True):
    """om_poly_fit into numpy arrays (same layout as tmcore_amd.mesh.Fit)."""
    nC, nE, nV = mesh.nCells, mesh.nEdges, mesh.nVertices
    r = {}
    f = _ffi.FitDesc()
    f.maxFit = max_fit
    fams = [("Cells", nC, "Cell"), ("Vertices", nV, "Vertex")]
    for order in ((2, 4) if order4 else (2,)):
        for fam, n, dname in fams:
            r[f"nFit{order}{fam}"] = np.zeros(n, np.int32)
            r[f"fit{order}{fam}"] = np.zeros((n, max_fit), np.int32)
            r[f"deriv{order}On{dname}"] = np.zeros((nE, 2, max_fit))
            setattr(f, f"nFit{order}{fam}", _ffi.iptr(r[f"nFit{order}{fam}"]))
            setattr(f, f"fit{order}{fam}", _ffi.iptr(r[f"fit{order}{fam}"]))
            setattr(f, f"deriv{order}On{dname}", _ffi.dptr(r[f"deriv{order}On{dname}"]))
    md = mesh.desc()
    rc = lib().om_poly_fit(C.byref(md), radius, C.byref(f))
    if rc != 0:
        raise RuntimeError(f"om_poly_fit failed ({rc})")
    return r


def oracle_reconstruct_coeffs(mesh, radius=6371220.0):
    """om_reconstruct_coeffs: coeffs_reconstruct as [nCells, maxEdges, 3]."""
    co = np.zeros((mesh.nCells, mesh.maxEdges, 3))
    md = mesh.desc()
    lib().om_reconstruct_coeffs(C.byref(md), radius, co.ctypes.data_as(_dp))
    return co


def oracle_reconstruct(mesh, coeffs, u):
    """om_reconstruct: (uX, uY, uZ, zonal, meridional), each [nCells]."""
    out = np.zeros((5, mesh.nCells))
    md = mesh.desc()
    lib().om_reconstruct(C.byref(md), _p(coeffs), _p(u), out.ctypes.data_as(_dp))
    return tuple(out)


def set_threads(n: int) -> int:
    """Threads of the oracle's element loops (n <= 0 keeps the current
    setting); returns the number in effect."""
    return int(lib().om_set_threads(int(n)))
