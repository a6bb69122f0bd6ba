"""ctypes bindings of the C-ABI declared in include/tmcore_host.h and
include/tmcore_gpu.h.

The libraries are built in-tree (``tmcore_amd/lib``) by ``tmcore_amd.build``.
Nothing here falls back to Python arithmetic: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)

TM_ERRORS = {
    0: "ok",
    -1: "invalid argument",
    -2: "mesh index out of range",
    -3: "fit stencil exceeds stride",
    -4: "singular least-squares matrix",
    -5: "HIP runtime error",
    -6: "allocation failed",
    -7: "non-finite value",
    -8: "call out of order",
    -9: "grid file error",
}


class TmError(RuntimeError):
    def __init__(self, code: int, where: str, detail: str = ""):
        msg = f"{where}: {TM_ERRORS.get(code, 'error')} ({code})"
        if detail:
            msg += f": {detail}"
        super().__init__(msg)
        self.code = code


class MeshDesc(C.Structure):
    _fields_ = [
        ("nCells", C.c_int32), ("nEdges", C.c_int32), ("nVertices", C.c_int32),
        ("maxEdges", C.c_int32), ("maxEdges2", C.c_int32), ("vertexDegree", C.c_int32),
        ("latCell", _dp), ("lonCell", _dp), ("xCell", _dp), ("yCell", _dp), ("zCell", _dp),
        ("latEdge", _dp), ("lonEdge", _dp), ("xEdge", _dp), ("yEdge", _dp), ("zEdge", _dp),
        ("latVertex", _dp), ("lonVertex", _dp), ("xVertex", _dp), ("yVertex", _dp), ("zVertex", _dp),
        ("nEdgesOnCell", _ip), ("nEdgesOnEdge", _ip),
        ("cellsOnCell", _ip), ("edgesOnCell", _ip), ("verticesOnCell", _ip),
        ("cellsOnEdge", _ip), ("verticesOnEdge", _ip),
        ("edgesOnEdge", _ip), ("weightsOnEdge", _dp),
        ("cellsOnVertex", _ip), ("edgesOnVertex", _ip), ("kiteAreasOnVertex", _dp),
        ("dvEdge", _dp), ("dv1Edge", _dp), ("dv2Edge", _dp), ("dcEdge", _dp), ("angleEdge", _dp),
        ("areaCell", _dp), ("areaTriangle", _dp),
    ]


class FitDesc(C.Structure):
    _fields_ = [
        ("maxFit", C.c_int32),
        ("nFit2Cells", _ip), ("fit2Cells", _ip), ("deriv2OnCell", _dp),
        ("nFit4Cells", _ip), ("fit4Cells", _ip), ("deriv4OnCell", _dp),
        ("nFit2Vertices", _ip), ("fit2Vertices", _ip), ("deriv2OnVertex", _dp),
        ("nFit4Vertices", _ip), ("fit4Vertices", _ip), ("deriv4OnVertex", _dp),
    ]


def _load(name: str) -> C.CDLL:
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise OSError(
            f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python -m tmcore_amd.build`"
        )
    return C.CDLL(path)


_host = None


def host() -> C.CDLL:
    """libtmhost.so: mesh generator, poly_fit weights, initial states."""
    global _host
    if _host is None:
        lib = _load("libtmhost.so")
        lib.tm_mesh_icos_dims.argtypes = [C.c_int, _ip, _ip, _ip]
        lib.tm_mesh_icos_generate.argtypes = [C.c_int, C.POINTER(MeshDesc)]
        lib.tm_poly_fit.argtypes = [C.POINTER(MeshDesc), C.c_double, C.POINTER(FitDesc)]
        lib.tm_case_seeded.argtypes = [C.POINTER(MeshDesc), C.c_double, C.c_uint64, C.c_int,
                                       C.POINTER(C.c_uint64), _dp, _dp, _dp, _dp]
        lib.tm_case_steady_geostrophic.argtypes = [C.POINTER(MeshDesc), C.c_double, _dp, _dp, _dp]
        lib.tm_mesh_mpas_dims.argtypes = [C.c_char_p, _ip, _dp]
        lib.tm_mesh_mpas_read.argtypes = [C.c_char_p, C.POINTER(MeshDesc)]
        lib.tm_mesh_mpas_error.restype = C.c_char_p
        lib.tm_reconstruct_coeffs.argtypes = [C.POINTER(MeshDesc), C.c_double, _dp]
        lib.tm_history_create.argtypes = [C.POINTER(C.c_void_p), C.c_char_p, C.POINTER(MeshDesc), C.c_double,
                                          C.c_double, C.c_int, C.c_char_p]
        lib.tm_history_write.argtypes = [C.c_void_p, C.c_double] + [_dp] * 7 + [C.c_double, C.c_double]
        lib.tm_history_close.argtypes = [C.c_void_p]
        for f in (lib.tm_mesh_icos_dims, lib.tm_mesh_icos_generate, lib.tm_poly_fit,
                  lib.tm_case_seeded, lib.tm_case_steady_geostrophic, lib.tm_mesh_mpas_dims,
                  lib.tm_mesh_mpas_read, lib.tm_reconstruct_coeffs):
            f.restype = C.c_int
        _host = lib
    return _host


def check(code: int, where: str, detail: str = "") -> None:
    if code != 0:
        raise TmError(code, where, detail)


def dptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(_ip) if a is not None else None
