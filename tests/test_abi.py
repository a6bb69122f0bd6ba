"""The C-ABI libraries load and export every function include/*.h declares;
argument validation that happens before any device call (CPU-only)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from tmcore_amd import _ffi
from tmcore_amd.mesh import Mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tm_[a-z0-9_]+)\s*\(", txt)))


@pytest.mark.parametrize("header,libname", [("tmcore_host.h", "libtmhost.so"), ("tmcore_gpu.h", "libtmgpu.so"),
                                            ("tmcore_cpu.h", "libtmcpu.so")])
def test_every_declared_symbol_is_exported(header, libname):
    path = os.path.join(_ffi.LIB_DIR, libname)
    if not os.path.exists(path):
        from tmcore_amd import build
        {"libtmhost.so": build.build_host, "libtmgpu.so": build.build_gpu, "libtmcpu.so": build.build_fortran}[libname]()
    lib = C.CDLL(path)
    names = declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ctx_create_rejects_bad_mesh_before_touching_the_device():
    from tmcore_amd.swm import Opts, gpu_lib
    m = Mesh(2)
    m.cellsOnEdge[5, 1] = m.nCells + 7  # out of range
    L = gpu_lib()
    o = Opts()
    L.tm_opts_default(C.byref(o))
    h = C.c_void_p()
    md = m.desc()
    fd = m.fit().desc()
    assert L.tm_ctx_create(C.byref(h), C.byref(md), C.byref(fd), C.byref(o)) == -2  # TM_ERANGE
    assert not h.value


@pytest.mark.parametrize("field,value", [("time_scheme", 9), ("pv_scheme", 4), ("interp_pv_order", 7), ("adv_order", 7),
                                         ("adv_order", 1),
                                         ("precision", 16), ("ntracers", -1), ("dt", -1.0)])
def test_ctx_create_rejects_unsupported_options(field, value):
    from tmcore_amd.swm import Opts, gpu_lib
    m = Mesh(1)
    L = gpu_lib()
    o = Opts()
    L.tm_opts_default(C.byref(o))
    setattr(o, field, value)
    h = C.c_void_p()
    md = m.desc()
    fd = m.fit().desc()
    assert L.tm_ctx_create(C.byref(h), C.byref(md), C.byref(fd), C.byref(o)) == -1  # TM_EINVAL
    assert not h.value


def test_ctx_create_rejects_missing_order4_fits():
    """adv_order 5/6 and interp_pv_order 5/6 read the order-4 fits; a fit table
    without them is refused."""
    from tmcore_amd.swm import Opts, gpu_lib
    m = Mesh(1)
    L = gpu_lib()
    o = Opts()
    L.tm_opts_default(C.byref(o))
    md = m.desc()
    fd = m.fit().desc()
    for field, value in (("adv_order", 5), ("interp_pv_order", 5)):
        L.tm_opts_default(C.byref(o))
        setattr(o, field, value)
        h = C.c_void_p()
        assert L.tm_ctx_create(C.byref(h), C.byref(md), C.byref(fd), C.byref(o)) == -1  # TM_EINVAL
        assert not h.value


def test_ctx_create_rejects_tracers_outside_crk4():
    """The reference advances tracers with CRK4 only; CRK3/PC2 with tracers is refused."""
    from tmcore_amd.swm import Opts, gpu_lib
    m = Mesh(1)
    L = gpu_lib()
    o = Opts()
    L.tm_opts_default(C.byref(o))
    o.time_scheme, o.ntracers = 1, 1
    h = C.c_void_p()
    md = m.desc()
    fd = m.fit().desc()
    assert L.tm_ctx_create(C.byref(h), C.byref(md), C.byref(fd), C.byref(o)) == -1  # TM_EINVAL
    assert not h.value


def test_host_library_errors():
    lib = _ffi.host()
    n = C.c_int32()
    assert lib.tm_mesh_icos_dims(12, C.byref(n), C.byref(n), C.byref(n)) == -1
    assert lib.tm_mesh_icos_generate(-1, None) == -1
    m = Mesh(1)
    f = _ffi.FitDesc()
    f.maxFit = 5  # too small for the 13-point vertex stencil
    arr = np.zeros((m.nEdges, 2, 5))
    f.deriv2OnVertex = _ffi.dptr(arr)
    md = m.desc()
    assert lib.tm_poly_fit(C.byref(md), 6371220.0, C.byref(f)) == -3  # TM_ESTENCIL
