"""Mesh, fit weights and initial states in the reference's array layout.

Arrays follow the reference's Fortran layout (reference src/core/mesh_mod.F90
:10-71): connectivity is 1-based int32 and 2-D arrays are (n, slots) in C
order, i.e. Fortran (slots, n) column-major.  Geometry is on the unit sphere
as a grid file holds it; consumers scale by the radius (mesh_mod.F90:205-224).
"""
from __future__ import annotations

import ctypes as C
import functools

import numpy as np

from . import _ffi

RADIUS = 6371220.0   # reference src/utils/const_mod.F90:11
G = 9.80616          # const_mod.F90:12
OMEGA = 7.292e-5     # const_mod.F90:13

_F64_CELL = ("latCell", "lonCell", "xCell", "yCell", "zCell", "areaCell")
_F64_EDGE = ("latEdge", "lonEdge", "xEdge", "yEdge", "zEdge",
             "dvEdge", "dv1Edge", "dv2Edge", "dcEdge", "angleEdge")
_F64_VERT = ("latVertex", "lonVertex", "xVertex", "yVertex", "zVertex", "areaTriangle")


_I32_2D = {"cellsOnCell": "maxEdges", "edgesOnCell": "maxEdges", "verticesOnCell": "maxEdges",
           "edgesOnEdge": "maxEdges2", "cellsOnVertex": "vertexDegree", "edgesOnVertex": "vertexDegree"}


class Mesh:
    """An icosahedral Voronoi mesh generated by libtmhost (level 0..11), or
    an MPAS grid file read by it (``Mesh.from_mpas``)."""

    def __init__(self, level: int):
        lib = _ffi.host()
        nc, ne, nv = C.c_int32(), C.c_int32(), C.c_int32()
        _ffi.check(lib.tm_mesh_icos_dims(level, C.byref(nc), C.byref(ne), C.byref(nv)),
                   "tm_mesh_icos_dims")
        self.level = level
        self._alloc(nc.value, ne.value, nv.value, 6, 12, 3)
        d = self.desc()
        _ffi.check(lib.tm_mesh_icos_generate(level, C.byref(d)), "tm_mesh_icos_generate")

    @classmethod
    def from_mpas(cls, path: str) -> "Mesh":
        """An MPAS grid file (NetCDF classic) through tm_mesh_mpas_read; the
        file's sphere_radius is kept as ``sphere_radius`` (geometry is used as
        stored: unit-sphere files are scaled by the model radius)."""
        lib = _ffi.host()
        dims = np.zeros(6, np.int32)
        r = C.c_double()
        bpath = str(path).encode()
        rc = lib.tm_mesh_mpas_dims(bpath, _ffi.iptr(dims), C.byref(r))
        _ffi.check(rc, "tm_mesh_mpas_dims", (lib.tm_mesh_mpas_error() or b"").decode() if rc else "")
        self = cls.__new__(cls)
        self.level = None
        self.sphere_radius = r.value
        self._alloc(*[int(x) for x in dims])
        d = self.desc()
        rc = lib.tm_mesh_mpas_read(bpath, C.byref(d))
        _ffi.check(rc, "tm_mesh_mpas_read", (lib.tm_mesh_mpas_error() or b"").decode() if rc else "")
        return self

    def to_mpas(self, path: str, version: int = 2):
        """Write the mesh arrays mesh_init reads (mesh_mod.F90:79-164) as an
        MPAS grid file, NetCDF classic (version 1) or 64-bit offset (2), with
        scipy's NetCDF writer; indexToCellID/EdgeID/VertexID are 1..n."""
        from scipy.io import netcdf_file
        with netcdf_file(str(path), "w", version=version) as f:
            f.on_a_sphere = "YES"
            f.sphere_radius = 1.0
            dims = {"nCells": self.nCells, "nEdges": self.nEdges, "nVertices": self.nVertices,
                    "maxEdges": self.maxEdges, "maxEdges2": self.maxEdges2, "TWO": 2,
                    "vertexDegree": self.vertexDegree}
            for k, v in dims.items():
                f.createDimension(k, v)
            where = {"Cell": "nCells", "Edge": "nEdges", "Vertex": "nVertices", "Triangle": "nVertices"}

            def put(name, arr, shape):
                v = f.createVariable(name, "d" if arr.dtype == np.float64 else "i", shape)
                v[:] = arr

            for k in _F64_CELL + _F64_EDGE + _F64_VERT:
                put(k, getattr(self, k), (next(d for s, d in where.items() if k.endswith(s)),))
            for k, n in (("indexToCellID", "nCells"), ("indexToEdgeID", "nEdges"), ("indexToVertexID", "nVertices")):
                put(k, np.arange(1, dims[n] + 1, dtype=np.int32), (n,))
            put("nEdgesOnCell", self.nEdgesOnCell, ("nCells",))
            put("nEdgesOnEdge", self.nEdgesOnEdge, ("nEdges",))
            for k, slot in _I32_2D.items():
                put(k, getattr(self, k), ("nCells" if k.endswith("Cell") else
                                          ("nEdges" if k.endswith("Edge") else "nVertices"), slot))
            put("cellsOnEdge", self.cellsOnEdge, ("nEdges", "TWO"))
            put("verticesOnEdge", self.verticesOnEdge, ("nEdges", "TWO"))
            put("weightsOnEdge", self.weightsOnEdge, ("nEdges", "maxEdges2"))
            put("kiteAreasOnVertex", self.kiteAreasOnVertex, ("nVertices", "vertexDegree"))

    def _alloc(self, nC, nE, nV, maxEdges, maxEdges2, vertexDegree):
        self.nCells, self.nEdges, self.nVertices = nC, nE, nV
        self.maxEdges, self.maxEdges2, self.vertexDegree = maxEdges, maxEdges2, vertexDegree
        for k in _F64_CELL:
            setattr(self, k, np.zeros(nC))
        for k in _F64_EDGE:
            setattr(self, k, np.zeros(nE))
        for k in _F64_VERT:
            setattr(self, k, np.zeros(nV))
        i32 = np.int32
        self.nEdgesOnCell = np.zeros(nC, i32)
        self.nEdgesOnEdge = np.zeros(nE, i32)
        self.cellsOnCell = np.zeros((nC, maxEdges), i32)
        self.edgesOnCell = np.zeros((nC, maxEdges), i32)
        self.verticesOnCell = np.zeros((nC, maxEdges), i32)
        self.cellsOnEdge = np.zeros((nE, 2), i32)
        self.verticesOnEdge = np.zeros((nE, 2), i32)
        self.edgesOnEdge = np.zeros((nE, maxEdges2), i32)
        self.weightsOnEdge = np.zeros((nE, maxEdges2))
        self.cellsOnVertex = np.zeros((nV, vertexDegree), i32)
        self.edgesOnVertex = np.zeros((nV, vertexDegree), i32)
        self.kiteAreasOnVertex = np.zeros((nV, vertexDegree))

    def desc(self) -> _ffi.MeshDesc:
        d = _ffi.MeshDesc()
        d.nCells, d.nEdges, d.nVertices = self.nCells, self.nEdges, self.nVertices
        d.maxEdges, d.maxEdges2, d.vertexDegree = self.maxEdges, self.maxEdges2, self.vertexDegree
        for name, typ in _ffi.MeshDesc._fields_[6:]:
            arr = getattr(self, name)
            assert arr.flags.c_contiguous
            setattr(d, name, _ffi.dptr(arr) if arr.dtype == np.float64 else _ffi.iptr(arr))
        self._desc_keepalive = d
        return d

    @functools.lru_cache(maxsize=2)
    def reconstruct_coeffs(self, radius: float = RADIUS) -> np.ndarray:
        """coeffs_reconstruct of the MPAS vector reconstruction, [nCells, maxEdges, 3]
        (tm_reconstruct_coeffs)."""
        co = np.zeros((self.nCells, self.maxEdges, 3))
        d = self.desc()
        _ffi.check(_ffi.host().tm_reconstruct_coeffs(C.byref(d), radius, _ffi.dptr(co)), "tm_reconstruct_coeffs")
        return co

    @functools.lru_cache(maxsize=4)
    def fit(self, radius: float = RADIUS, max_fit: int = 13, order4: bool = False) -> "Fit":
        return Fit(self, radius, max_fit, order4)


class Fit:
    """poly_fit_run weights (reference src/utils/poly_fit_mod.F90:119-609)."""

    def __init__(self, mesh: Mesh, radius: float, max_fit: int, order4: bool):
        nC, nE, nV = mesh.nCells, mesh.nEdges, mesh.nVertices
        self.maxFit = max_fit
        i32 = np.int32
        self.nFit2Cells = np.zeros(nC, i32)
        self.fit2Cells = np.zeros((nC, max_fit), i32)
        self.deriv2OnCell = np.zeros((nE, 2, max_fit))
        self.nFit2Vertices = np.zeros(nV, i32)
        self.fit2Vertices = np.zeros((nV, max_fit), i32)
        self.deriv2OnVertex = np.zeros((nE, 2, max_fit))
        f = _ffi.FitDesc()
        f.maxFit = max_fit
        f.nFit2Cells, f.fit2Cells = _ffi.iptr(self.nFit2Cells), _ffi.iptr(self.fit2Cells)
        f.deriv2OnCell = _ffi.dptr(self.deriv2OnCell)
        f.nFit2Vertices, f.fit2Vertices = _ffi.iptr(self.nFit2Vertices), _ffi.iptr(self.fit2Vertices)
        f.deriv2OnVertex = _ffi.dptr(self.deriv2OnVertex)
        if order4:
            self.nFit4Cells = np.zeros(nC, i32)
            self.fit4Cells = np.zeros((nC, max_fit), i32)
            self.deriv4OnCell = np.zeros((nE, 2, max_fit))
            self.nFit4Vertices = np.zeros(nV, i32)
            self.fit4Vertices = np.zeros((nV, max_fit), i32)
            self.deriv4OnVertex = np.zeros((nE, 2, max_fit))
            f.nFit4Cells, f.fit4Cells = _ffi.iptr(self.nFit4Cells), _ffi.iptr(self.fit4Cells)
            f.deriv4OnCell = _ffi.dptr(self.deriv4OnCell)
            f.nFit4Vertices, f.fit4Vertices = _ffi.iptr(self.nFit4Vertices), _ffi.iptr(self.fit4Vertices)
            f.deriv4OnVertex = _ffi.dptr(self.deriv4OnVertex)
        self.order4 = order4
        md = mesh.desc()
        _ffi.check(_ffi.host().tm_poly_fit(C.byref(md), radius, C.byref(f)), "tm_poly_fit")
        self._f = f

    def desc(self) -> _ffi.FitDesc:
        return self._f


def seeded_state(mesh: Mesh, seed: int, tracer_seeds=(), radius: float = RADIUS):
    """Seeded solid-body + bumps + mountain state (see tm_case_seeded).

    Returns (gd[nC], u[nE], ghs[nC], tracers[T, nC]) in reference order."""
    nT = len(tracer_seeds)
    gd = np.zeros(mesh.nCells)
    u = np.zeros(mesh.nEdges)
    ghs = np.zeros(mesh.nCells)
    tr = np.zeros((nT, mesh.nCells))
    ts = (C.c_uint64 * max(nT, 1))(*tracer_seeds)
    md = mesh.desc()
    _ffi.check(_ffi.host().tm_case_seeded(C.byref(md), radius, seed, nT, ts, _ffi.dptr(gd),
                                          _ffi.dptr(u), _ffi.dptr(ghs),
                                          _ffi.dptr(tr) if nT else None), "tm_case_seeded")
    return gd, u, ghs, tr


def steady_geostrophic_state(mesh: Mesh, radius: float = RADIUS):
    """Williamson case 2 (steady_geostrophic_flow_test_mod.F90:20-54)."""
    gd = np.zeros(mesh.nCells)
    u = np.zeros(mesh.nEdges)
    ghs = np.zeros(mesh.nCells)
    md = mesh.desc()
    _ffi.check(_ffi.host().tm_case_steady_geostrophic(C.byref(md), radius, _ffi.dptr(gd),
                                                      _ffi.dptr(u), _ffi.dptr(ghs)),
               "tm_case_steady_geostrophic")
    return gd, u, ghs
