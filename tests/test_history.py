"""History output: the MPAS vector reconstruction coefficients
(tm_reconstruct_coeffs against the oracle's restatement, bit for bit) and the
history file writer (tm_history_*, read back by scipy's independent NetCDF
reader)."""
import numpy as np
import pytest

from oracle.oracle import oracle_reconstruct_coeffs
from tmcore_amd.history import History
from tmcore_amd.mesh import RADIUS, Mesh


@pytest.mark.parametrize("level", [2, 4])
def test_reconstruct_coeffs_equal_oracle(level):
    m = Mesh(level)
    np.testing.assert_array_equal(m.reconstruct_coeffs(), oracle_reconstruct_coeffs(m))


def test_history_file_reads_back(tmp_path):
    from scipy.io import netcdf_file
    m = Mesh(2)
    rng = np.random.default_rng(0)
    path = tmp_path / "h0.nc"
    h = History(path, m, dt=600.0, time_scheme=2, case_name="seeded")
    recs = []
    for t in (0.0, 3600.0, 7200.0):
        f = {"u": rng.standard_normal(m.nEdges), "zonal": rng.standard_normal(m.nCells),
             "meridional": rng.standard_normal(m.nCells), "h": rng.standard_normal(m.nCells),
             "pv": rng.standard_normal(m.nVertices), "vor": rng.standard_normal(m.nVertices),
             "div": rng.standard_normal(m.nCells)}
        tm, te = rng.random(), rng.random()
        h.write(t, f, tm, te)
        recs.append((t, f, tm, te))
    h.close()
    with netcdf_file(str(path), "r", mmap=False) as nc:
        assert nc.dimensions["nCells"] == m.nCells and nc.dimensions["Time"] is None
        assert nc.variables["u"].dimensions == ("Time", "nEdges")
        assert nc.variables["cellsOnEdge"].dimensions == ("nEdges", "TWO")
        assert nc.variables["zonalWind"].long_name == b"Zonal Wind"
        assert nc.variables["h"].units == b"m"
        assert nc.sphere_radius == RADIUS and nc.time_scheme == 2 and nc.on_a_sphere == b"YES"
        np.testing.assert_array_equal(nc.variables["Time"][:], [r[0] for r in recs])
        for i, (t, f, tm, te) in enumerate(recs):
            for k, v in (("u", "u"), ("zonal", "zonalWind"), ("meridional", "merdionalWind"), ("h", "h"),
                         ("pv", "pv"), ("vor", "vor"), ("div", "div")):
                np.testing.assert_array_equal(nc.variables[v][i], f[k], err_msg=v)
            assert nc.variables["tm"][i] == tm and nc.variables["te"][i] == te
        np.testing.assert_array_equal(nc.variables["xCell"][:], m.xCell * RADIUS)
        np.testing.assert_array_equal(nc.variables["areaEdge"][:], (m.dvEdge * RADIUS) * (m.dcEdge * RADIUS))
        np.testing.assert_array_equal(nc.variables["edgesOnEdge"][:], m.edgesOnEdge)
        np.testing.assert_array_equal(nc.variables["indexToVertexID"][:], np.arange(1, m.nVertices + 1))
