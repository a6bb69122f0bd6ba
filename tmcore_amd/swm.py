"""Python host mirror of libtmgpu (include/tmcore_gpu.h).

``SwmGpu`` follows the reference driver's shape (tmcore_swm_mod.F90:29-90):
create = mesh_init + poly_fit_init + state_init, ``upload`` = the initial
condition + the run prologue, ``step`` = the time loop with per-step
invariants and beta.  The single operators keep the names and argument order
of operators_mod.F90.  Errors raise TmError with the library's message, the
counterpart of the reference's log_error + stop.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from .mesh import RADIUS, Fit, Mesh

_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


class Opts(C.Structure):
    _fields_ = [("time_scheme", C.c_int), ("pv_scheme", C.c_int), ("interp_pv_order", C.c_int),
                ("adv_order", C.c_int), ("adv_mono", C.c_int), ("conserve_energy", C.c_int),
                ("apvm_weight", C.c_double), ("dt", C.c_double), ("radius", C.c_double),
                ("ntracers", C.c_int), ("precision", C.c_int), ("device", C.c_int), ("renumber", C.c_int),
                ("geometry_scaled", C.c_int)]


class StageOut(C.Structure):
    _fields_ = [(k, _dp) for k in ("tend_gd", "tend_iap", "tend_q", "gd_edge", "u", "ke", "vor", "gd_vertex",
                                   "pv_vertex", "v", "pv_edge", "pv_flx", "u_tend")]


STAGE_WHERE = {"tend_gd": "c", "tend_iap": "e", "gd_edge": "e", "u": "e", "ke": "c", "vor": "v",
               "gd_vertex": "v", "pv_vertex": "v", "v": "e", "pv_edge": "e", "pv_flx": "e", "u_tend": "e"}

_gpu = None


def gpu_lib() -> C.CDLL:
    """libtmgpu.so; raises if it is not built (there is no fallback path)."""
    global _gpu
    if _gpu is None:
        L = _ffi._load("libtmgpu.so")
        L.tm_opts_default.argtypes = [C.POINTER(Opts)]
        L.tm_opts_default.restype = None
        L.tm_ctx_create.argtypes = [C.POINTER(_vp), C.POINTER(_ffi.MeshDesc), C.POINTER(_ffi.FitDesc),
                                    C.POINTER(Opts)]
        L.tm_ctx_destroy.argtypes = [_vp]
        L.tm_last_error.argtypes = [_vp]
        L.tm_last_error.restype = C.c_char_p
        L.tm_state_upload.argtypes = [_vp, _dp, _dp, _dp, _dp, _dp]
        L.tm_state_download.argtypes = [_vp, _dp, _dp, _dp, _dp]
        L.tm_step.argtypes = [_vp, C.c_int, _dp, _dp]
        L.tm_step_async.argtypes = [_vp, C.c_int]
        L.tm_sync.argtypes = [_vp]
        L.tm_invariants.argtypes = [_vp, _dp]
        L.tm_ctx_set_layout.argtypes = [_vp, C.c_int]
        L.tm_ctx_get_layout.argtypes = [_vp, C.POINTER(C.c_int)]
        L.tm_ctx_set_layout.restype = C.c_int
        L.tm_ctx_get_layout.restype = C.c_int
        L.tm_invariants_noalloc.argtypes = [_vp, _dp]
        L.tm_invariants_noalloc.restype = C.c_int
        L.tm_stage.argtypes = [_vp, C.POINTER(StageOut)]
        L.tm_last_elapsed_ms.argtypes = [_vp, _dp]
        L.tm_step_bytes.argtypes = [_vp, _dp]
        L.tm_profile_steps.argtypes = [_vp, C.c_int, _dp, C.POINTER(C.c_int)]
        L.tm_kernel_bytes.argtypes = [_vp, _dp]
        L.tm_profile_steps.restype = C.c_int
        L.tm_kernel_bytes.restype = C.c_int
        L.tm_hbm_copy_probe.argtypes = [C.c_int, C.c_double, C.c_int, _dp]
        L.tm_hbm_copy_probe.restype = C.c_int
        for n in ("div", "curl", "c2e", "c2v", "v2c", "v2e", "ke", "tangent_wind"):
            getattr(L, "tm_op_" + n).argtypes = [_vp, _dp, _dp]
        L.tm_op_tangent_vor_flux.argtypes = [_vp, _dp, _dp, _dp, _dp]
        L.tm_op_upwind_cell.argtypes = [_vp, _dp, _dp, _dp]
        L.tm_history_setup.argtypes = [_vp, _dp, _dp, _dp]
        L.tm_history_fields.argtypes = [_vp] + [_dp] * 7
        L.tm_op_pv_vertex.argtypes = [_vp, _dp, _dp, _dp]
        L.tm_op_pv_edge.argtypes = [_vp] + [_dp] * 7
        L.tm_op_u_tend.argtypes = [_vp] + [_dp] * 9
        for f in ("tm_ctx_create", "tm_ctx_destroy", "tm_state_upload", "tm_state_download", "tm_step",
                  "tm_step_async", "tm_sync", "tm_invariants", "tm_stage", "tm_last_elapsed_ms", "tm_step_bytes"):
            getattr(L, f).restype = C.c_int
        _gpu = L
    return _gpu


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def hbm_copy_gbps(nbytes: float = 2.0 * 2**30, iters: int = 20, device: int = 0) -> float:
    """Device-to-device copy bandwidth (read + write GB/s), tm_hbm_copy_probe."""
    g = C.c_double()
    rc = gpu_lib().tm_hbm_copy_probe(device, nbytes, iters, C.byref(g))
    if rc != 0:
        raise _ffi.TmError(rc, "tm_hbm_copy_probe")
    return g.value


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class SwmGpu:
    """One shallow-water model on the device (CRK4, pv_scheme 3 by default)."""

    def __init__(self, mesh: Mesh, fit: Fit | None = None, *, dt: float = 600.0, radius: float = RADIUS,
                 ntracers: int = 0, precision: int = 64, interp_pv_order: int = 3, conserve_energy: bool = True,
                 renumber: bool = True, device: int = 0, **scheme):
        L = gpu_lib()
        o = Opts()
        L.tm_opts_default(C.byref(o))
        o.dt, o.radius, o.ntracers, o.precision = dt, radius, ntracers, precision
        o.interp_pv_order, o.conserve_energy, o.renumber, o.device = interp_pv_order, int(conserve_energy), \
            int(renumber), device
        for k, v in scheme.items():
            setattr(o, k, v)
        self.opts = o
        self.mesh = mesh
        needs_fit = (o.pv_scheme == 3 and interp_pv_order >= 3) or o.adv_order >= 3
        if fit is None and needs_fit:
            fit = mesh.fit(radius)
        self.fit = fit
        self.nC, self.nE, self.nV, self.nT = mesh.nCells, mesh.nEdges, mesh.nVertices, ntracers
        h = _vp()
        md = mesh.desc()
        fd = fit.desc() if fit is not None else None
        rc = L.tm_ctx_create(C.byref(h), C.byref(md), C.byref(fd) if fd is not None else None, C.byref(o))
        if rc != 0:
            raise _ffi.TmError(rc, "tm_ctx_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            gpu_lib().tm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, where):
        if rc != 0:
            msg = gpu_lib().tm_last_error(self.h)
            raise _ffi.TmError(rc, where, msg.decode() if msg else "")

    # -- state
    def upload(self, gd, ghs, u=None, iap_u=None, tracers=None):
        gd, ghs, u, iap_u, tracers = (_f64(x) for x in (gd, ghs, u, iap_u, tracers))
        self._check(gpu_lib().tm_state_upload(self.h, _p(gd), _p(u), _p(iap_u), _p(ghs), _p(tracers)),
                    "tm_state_upload")

    def download(self, out=None):
        """Returns dict gd, u, iap_u, tracers (reference order); with `out` (a
        dict an earlier download returned) the arrays are filled in place."""
        if out is None:
            out = {"gd": np.zeros(self.nC), "u": np.zeros(self.nE), "iap_u": np.zeros(self.nE),
                   "tracers": np.zeros((self.nT, self.nC))}
        gd, u, iap, tr = out["gd"], out["u"], out["iap_u"], out["tracers"]
        for a, n in ((gd, (self.nC,)), (u, (self.nE,)), (iap, (self.nE,)), (tr, (self.nT, self.nC))):
            if a.shape != n or a.dtype != np.float64 or not a.flags.c_contiguous:
                raise ValueError(f"download: out array of shape {a.shape} / {a.dtype}, expected {n} float64")
        self._check(gpu_lib().tm_state_download(self.h, _p(gd), _p(u), _p(iap), _p(tr) if self.nT else None),
                    "tm_state_download")
        return {"gd": gd, "u": u, "iap_u": iap, "tracers": tr}

    # -- time loop
    def step(self, nsteps: int):
        """nsteps CRK4 steps; returns (invariants[nsteps, 4], beta[nsteps])."""
        inv = np.zeros((nsteps, 4))
        beta = np.zeros(nsteps)
        self._check(gpu_lib().tm_step(self.h, nsteps, _p(inv), _p(beta)), "tm_step")
        return inv, beta

    def step_async(self, nsteps: int):
        self._check(gpu_lib().tm_step_async(self.h, nsteps), "tm_step_async")

    def sync(self):
        self._check(gpu_lib().tm_sync(self.h), "tm_sync")

    def last_elapsed_ms(self) -> float:
        ms = C.c_double()
        self._check(gpu_lib().tm_last_elapsed_ms(self.h, C.byref(ms)), "tm_last_elapsed_ms")
        return ms.value

    def step_bytes(self):
        b = np.zeros(3)
        self._check(gpu_lib().tm_step_bytes(self.h, _p(b)), "tm_step_bytes")
        return {"stage": b[0], "finalize": b[1], "step": b[2]}

    KERNELS = ("k_edge_prep", "k_cell_vertex", "k_edge_pv", "k_edge_tend", "k_reduce", "k_finalize")

    def profile(self, nsteps: int):
        """Per-kernel device time over nsteps eager steps (HIP events on the
        context's stream): {kernel: (avg_ms, launches, bytes_per_launch)}."""
        ms = np.zeros(6)
        n = (C.c_int * 6)()
        self._check(gpu_lib().tm_profile_steps(self.h, nsteps, _p(ms), n), "tm_profile_steps")
        by = np.zeros(6)
        self._check(gpu_lib().tm_kernel_bytes(self.h, _p(by)), "tm_kernel_bytes")
        return {k: (ms[i] / max(n[i], 1), n[i], by[i]) for i, k in enumerate(self.KERNELS)}

    def invariants(self):
        out = np.zeros(4)
        self._check(gpu_lib().tm_invariants(self.h, _p(out)), "tm_invariants")
        return out

    @property
    def layout(self) -> int:
        """Kernel layout (tm_ctx_set_layout): bit 1 and bit 2, edgesOnEdge
        derived from the cells' edge rows in the PV and tendency kernels;
        bit 4, the tendency kernel forms the next RK stage's edge values;
        bit 8, the tendency kernel gathers from an LDS tile; bit 16, the PV
        kernel does."""
        m = C.c_int()
        self._check(gpu_lib().tm_ctx_get_layout(self.h, C.byref(m)), "tm_ctx_get_layout")
        return m.value

    @layout.setter
    def layout(self, mode: int):
        self._check(gpu_lib().tm_ctx_set_layout(self.h, int(mode)), "tm_ctx_set_layout")

    def invariants_noalloc(self):
        """tm_invariants_noalloc: the invariants with no device allocation
        (static per-GPU scratch; meshes up to 4,194,304 elements)."""
        out = np.zeros(4)
        self._check(gpu_lib().tm_invariants_noalloc(self.h, _p(out)), "tm_invariants_noalloc")
        return out

    def stage(self):
        """One spatial_operators call on the current state: dict of outputs."""
        n = {"c": self.nC, "e": self.nE, "v": self.nV}
        res = {k: np.zeros(n[w]) for k, w in STAGE_WHERE.items()}
        res["tend_q"] = np.zeros((self.nT, self.nC))
        so = StageOut(**{k: _p(v) for k, v in res.items() if k != "tend_q"})
        so.tend_q = _p(res["tend_q"]) if self.nT else None
        self._check(gpu_lib().tm_stage(self.h, C.byref(so)), "tm_stage")
        return res

    # -- single operators (operators_mod.F90 names)
    def _op(self, name, out_where, *args):
        arrs = [_f64(a) for a in args]
        out = np.zeros({"c": self.nC, "e": self.nE, "v": self.nV}[out_where])
        self._check(getattr(gpu_lib(), "tm_op_" + name)(self.h, *[_p(a) for a in arrs], _p(out)), "tm_op_" + name)
        return out

    def div_operator(self, f_edge):
        return self._op("div", "c", f_edge)

    def curl_operator(self, f_edge):
        return self._op("curl", "v", f_edge)

    def scalar_c2e_interp_operator(self, f_cell):
        return self._op("c2e", "e", f_cell)

    def scalar_c2v_interp_operator(self, f_cell):
        return self._op("c2v", "v", f_cell)

    def scalar_v2c_interp_operator(self, f_vertex):
        return self._op("v2c", "c", f_vertex)

    def scalar_v2e_interp_operator(self, f_vertex):
        return self._op("v2e", "e", f_vertex)

    def calc_kinetic_energy(self, u_edge):
        return self._op("ke", "c", u_edge)

    def calc_tangent_wind(self, u_edge):
        return self._op("tangent_wind", "e", u_edge)

    def calc_tangent_vor_flux(self, u_edge, gd_edge, pv_edge):
        return self._op("tangent_vor_flux", "e", u_edge, gd_edge, pv_edge)

    # -- history output (history_mod.F90:125-174)
    def history_setup(self, radius: float | None = None):
        """Upload the MPAS reconstruction coefficients of this mesh."""
        m = self.mesh
        co = m.reconstruct_coeffs(self.opts.radius if radius is None else radius)
        self._check(gpu_lib().tm_history_setup(self.h, _p(co), _p(_f64(m.latCell)), _p(_f64(m.lonCell))),
                    "tm_history_setup")

    def history_fields(self) -> dict:
        """u, zonal, meridional, h, pv, vor, div of the current state (reference order)."""
        out = {k: np.zeros(n) for k, n in (("u", self.nE), ("zonal", self.nC), ("meridional", self.nC),
                                          ("h", self.nC), ("pv", self.nV), ("vor", self.nV), ("div", self.nC))}
        self._check(gpu_lib().tm_history_fields(self.h, *[_p(out[k]) for k in ("u", "zonal", "meridional", "h",
                                                                               "pv", "vor", "div")]),
                    "tm_history_fields")
        return out

    def calc_pv_on_vertex(self, vor_vertex, gd_vertex):
        return self._op("pv_vertex", "v", vor_vertex, gd_vertex)

    def calc_pv_on_edge(self, u_edge, v_edge, gd_edge, gd_tend_vertex, pv_vertex, pv_cell):
        return self._op("pv_edge", "e", u_edge, v_edge, gd_edge, gd_tend_vertex, pv_vertex, pv_cell)

    def calc_u_tend_on_edge(self, u_edge, ke_cell, gd_cell, gd_edge, ghs_cell, pv_flx_edge, gd_tend_edge):
        """(u_tend_edge, iap_u_tend_edge)"""
        ins = [_f64(a) for a in (u_edge, ke_cell, gd_cell, gd_edge, ghs_cell, pv_flx_edge, gd_tend_edge)]
        ut, it = np.zeros(self.nE), np.zeros(self.nE)
        self._check(gpu_lib().tm_op_u_tend(self.h, *[_p(a) for a in ins], _p(ut), _p(it)), "tm_op_u_tend")
        return ut, it

    def add_upwind_correction_on_cell(self, f_cell, u_edge, f_edge):
        """f_edge plus the upwind correction of f_cell at the context's adv_order
        (a new array; f_edge itself is left unchanged)."""
        out = np.array(f_edge, dtype=np.float64)
        if out.shape != (self.nE,):
            raise ValueError("f_edge must be [nEdges]")
        self._check(gpu_lib().tm_op_upwind_cell(self.h, _p(_f64(f_cell)), _p(_f64(u_edge)), _p(out)),
                    "tm_op_upwind_cell")
        return out
