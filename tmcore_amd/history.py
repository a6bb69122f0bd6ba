"""History files in the reference's history_mod layout (libtmhost
``tm_history_*``, include/tmcore_host.h), filled from a device context
(``SwmGpu.history_fields``)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from .mesh import RADIUS, Mesh


class History:
    """One NetCDF (CDF-2) history file; ``write`` appends one output time."""

    def __init__(self, path, mesh: Mesh, radius: float = RADIUS, dt: float = 0.0, time_scheme: int = 2,
                 case_name: str = ""):
        self.mesh = mesh
        h = C.c_void_p()
        d = mesh.desc()
        _ffi.check(_ffi.host().tm_history_create(C.byref(h), str(path).encode(), C.byref(d), radius, dt,
                                                 time_scheme, case_name.encode()), "tm_history_create")
        self.h = h

    def write(self, time_seconds: float, fields: dict, total_mass: float, total_energy: float):
        """fields: u [nEdges], zonal, meridional, h, div [nCells], pv, vor [nVertices]."""
        m = self.mesh
        n = {"u": m.nEdges, "zonal": m.nCells, "meridional": m.nCells, "h": m.nCells, "pv": m.nVertices,
             "vor": m.nVertices, "div": m.nCells}
        arrs = {}
        for k, size in n.items():
            a = np.ascontiguousarray(fields[k], dtype=np.float64)
            if a.shape != (size,):
                raise ValueError(f"{k} must have shape ({size},)")
            arrs[k] = a
        _ffi.check(_ffi.host().tm_history_write(self.h, time_seconds, *[_ffi.dptr(arrs[k]) for k in n],
                                                total_mass, total_energy), "tm_history_write")

    def close(self):
        if getattr(self, "h", None):
            _ffi.check(_ffi.host().tm_history_close(self.h), "tm_history_close")
            self.h = None

    def __del__(self):
        if getattr(self, "h", None):
            _ffi.host().tm_history_close(self.h)
            self.h = None
