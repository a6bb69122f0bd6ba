"""MPAS grid files (tm_mesh_mpas_dims / tm_mesh_mpas_read, the NetCDF reads of
mesh_init, reference src/core/mesh_mod.F90:79-164) read by libtmhost's own
NetCDF classic parser.  Files are written by scipy's NetCDF writer (CDF-1 and
CDF-2) and, for CDF-5, by the minimal writer below; every array read back must
equal the written one bit for bit and equal what scipy's own reader returns."""
import ctypes as C
import struct

import numpy as np
import pytest

from tmcore_amd import _ffi
from tmcore_amd.mesh import Mesh

ARRAYS = ("latCell", "lonCell", "xCell", "yCell", "zCell", "latEdge", "lonEdge", "xEdge", "yEdge", "zEdge",
          "latVertex", "lonVertex", "xVertex", "yVertex", "zVertex", "nEdgesOnCell", "nEdgesOnEdge",
          "cellsOnCell", "edgesOnCell", "verticesOnCell", "cellsOnEdge", "verticesOnEdge", "edgesOnEdge",
          "weightsOnEdge", "cellsOnVertex", "edgesOnVertex", "kiteAreasOnVertex", "dvEdge", "dv1Edge",
          "dv2Edge", "dcEdge", "angleEdge", "areaCell", "areaTriangle")


@pytest.fixture(scope="module")
def m3():
    return Mesh(3)


def _same(a, b):
    for k in ARRAYS:
        x, y = getattr(a, k), getattr(b, k)
        assert x.dtype == y.dtype and x.shape == y.shape, k
        np.testing.assert_array_equal(x, y, err_msg=k)


@pytest.mark.parametrize("version", [1, 2])
def test_scipy_written_grid_reads_back_exactly(tmp_path, m3, version):
    path = tmp_path / f"x3_v{version}.grid.nc"
    m3.to_mpas(path, version=version)
    r = Mesh.from_mpas(path)
    assert (r.nCells, r.nEdges, r.nVertices, r.maxEdges, r.maxEdges2, r.vertexDegree) == (
        m3.nCells, m3.nEdges, m3.nVertices, 6, 12, 3)
    assert r.sphere_radius == 1.0
    _same(r, m3)
    # and equal to scipy's reading of the same file
    from scipy.io import netcdf_file
    with netcdf_file(str(path), "r", mmap=False) as f:
        for k in ARRAYS:
            np.testing.assert_array_equal(getattr(r, k), np.asarray(f.variables[k][:]), err_msg=k)


def _write_cdf5(path, dims, variables, gatts):
    """Minimal NetCDF CDF-5 writer (64-bit counts and offsets): dims
    {name: len}, variables [(name, (dimnames...), array)], global double
    attributes {name: value}."""
    def name(s):
        b = s.encode()
        return struct.pack(">Q", len(b)) + b + b"\0" * ((4 - len(b) % 4) % 4)

    dnames = list(dims)
    hdr = b"CDF\x05" + struct.pack(">Q", 0)
    hdr += struct.pack(">IQ", 0x0A, len(dims)) + b"".join(name(k) + struct.pack(">Q", v) for k, v in dims.items())
    hdr += struct.pack(">IQ", 0x0C, len(gatts)) + b"".join(
        name(k) + struct.pack(">IQ", 6, 1) + struct.pack(">d", v) for k, v in gatts.items())
    blobs = [np.ascontiguousarray(a).astype(">f8" if a.dtype == np.float64 else ">i4").tobytes()
             for _, _, a in variables]

    def var_list(begins):
        out = struct.pack(">IQ", 0x0B, len(variables))
        for (vn, vd, a), blob, b0 in zip(variables, blobs, begins):
            out += name(vn) + struct.pack(">Q", len(vd)) + b"".join(struct.pack(">Q", dnames.index(d)) for d in vd)
            out += struct.pack(">IQ", 0, 0) + struct.pack(">IQQ", 6 if a.dtype == np.float64 else 4,
                                                          len(blob) + (-len(blob)) % 4, b0)
        return out

    size = len(hdr) + len(var_list([0] * len(variables)))
    begins, off = [], size
    for blob in blobs:
        begins.append(off)
        off += len(blob) + (-len(blob)) % 4
    with open(path, "wb") as f:
        f.write(hdr + var_list(begins))
        for blob in blobs:
            f.write(blob + b"\0" * ((-len(blob)) % 4))


def _mesh_vars(m, me=None):
    """(name, dims, array) of every mesh_init array; maxEdges padded to me."""
    me = me or m.maxEdges
    pad = lambda a: np.pad(a, ((0, 0), (0, me - a.shape[1])))  # noqa: E731
    where = lambda k: ("nCells" if k.endswith("Cell") else "nEdges" if k.endswith("Edge") else "nVertices")  # noqa: E731
    out = []
    for k in ARRAYS:
        a = getattr(m, k)
        if a.ndim == 1:
            out.append((k, (where(k.replace("Triangle", "Vertex")),), a))
        elif k in ("cellsOnCell", "edgesOnCell", "verticesOnCell"):
            out.append((k, ("nCells", "maxEdges"), pad(a)))
        elif k in ("cellsOnEdge", "verticesOnEdge"):
            out.append((k, ("nEdges", "TWO"), a))
        elif k in ("edgesOnEdge", "weightsOnEdge"):
            out.append((k, ("nEdges", "maxEdges2"), a))
        else:
            out.append((k, ("nVertices", "vertexDegree"), a))
    return out


def test_cdf5_grid_with_wider_maxedges_reads_back(tmp_path):
    m = Mesh(2)
    path = tmp_path / "x2.cdf5.nc"
    dims = {"nCells": m.nCells, "nEdges": m.nEdges, "nVertices": m.nVertices, "maxEdges": 7, "maxEdges2": 12,
            "TWO": 2, "vertexDegree": 3}
    _write_cdf5(path, dims, _mesh_vars(m, me=7), {"sphere_radius": 6371229.0})
    r = Mesh.from_mpas(path)
    assert r.maxEdges == 7 and r.sphere_radius == 6371229.0
    for k in ARRAYS:
        a = getattr(m, k)
        b = getattr(r, k)
        if k in ("cellsOnCell", "edgesOnCell", "verticesOnCell"):
            np.testing.assert_array_equal(b[:, :6], a, err_msg=k)
            assert not b[:, 6].any()
        else:
            np.testing.assert_array_equal(b, a, err_msg=k)
    # the least-squares fit reads cellsOnCell with the file's stride
    f7, f6 = r.fit(max_fit=31, order4=True), m.fit(max_fit=31, order4=True)
    for k in ("deriv2OnCell", "deriv2OnVertex", "deriv4OnCell", "deriv4OnVertex"):
        np.testing.assert_array_equal(getattr(f7, k), getattr(f6, k), err_msg=k)


def test_reader_errors(tmp_path, m3):
    lib = _ffi.host()
    dims = np.zeros(6, np.int32)
    r = C.c_double()
    missing = str(tmp_path / "none.nc").encode()
    assert lib.tm_mesh_mpas_dims(missing, _ffi.iptr(dims), C.byref(r)) == -9
    assert b"cannot open" in lib.tm_mesh_mpas_error()
    h5 = tmp_path / "h5.nc"
    h5.write_bytes(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)
    assert lib.tm_mesh_mpas_dims(str(h5).encode(), _ffi.iptr(dims), C.byref(r)) == -9
    assert b"NetCDF-4" in lib.tm_mesh_mpas_error()
    # a file without areaTriangle
    m = Mesh(1)
    path = tmp_path / "cut.nc"
    dims5 = {"nCells": m.nCells, "nEdges": m.nEdges, "nVertices": m.nVertices, "maxEdges": 6, "maxEdges2": 12,
             "TWO": 2, "vertexDegree": 3}
    _write_cdf5(path, dims5, [v for v in _mesh_vars(m) if v[0] != "areaTriangle"], {})
    with pytest.raises(_ffi.TmError, match="areaTriangle missing"):
        Mesh.from_mpas(path)
    # a truncated file
    good = tmp_path / "good.nc"
    m.to_mpas(good)
    raw = good.read_bytes()
    (tmp_path / "trunc.nc").write_bytes(raw[: len(raw) // 2])
    with pytest.raises(_ffi.TmError):
        Mesh.from_mpas(tmp_path / "trunc.nc")
    # descriptor dimensions that differ from the file's
    d = m3.desc()
    assert lib.tm_mesh_mpas_read(str(good).encode(), C.byref(d)) == -1
