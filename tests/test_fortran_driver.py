"""The Fortran host (test_tmcore_sw over iso_c_binding) against the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

from oracle.oracle import Oracle
from tmcore_amd import _ffi
from tmcore_amd.mesh import Mesh, seeded_state

EXE = os.path.join(_ffi.LIB_DIR, "test_tmcore_sw")

NAMELIST = """&time_control
run_hours = {hours}
dt        = {dt}
/
&io_control
output_file_prefix = 'test'
{io}/
&tmcore
conserve_energy = .true.
time_scheme     = 2
pv_scheme       = 3
interp_pv_order = 3
{tmcore}/
&test_case
case_name = 'seeded'
/
&gpu_control
mesh_level   = {level}
precision    = 64
ntracers     = 1
seed         = 7
tracer_seed0 = 77
print_every  = 1
{extra}/
"""

LINE = re.compile(r"step=\s*(\d+) total_mass=\s*(\S+) total_energy=\s*(\S+) total_potential_enstropy=\s*(\S+) "
                  r"total_absolute_vorticity=\s*(\S+)(?: beta=\s*(\S+))?")


def parse(out):
    rows = {}
    for line in out.splitlines():
        mt = LINE.search(line)
        if mt:
            rows[int(mt.group(1))] = [float(x) for x in mt.groups()[1:] if x is not None]
    return rows


@pytest.mark.gpu
def test_fortran_driver_matches_oracle(tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (amdflang absent when building)")
    level, dt, hours = 4, 600.0, 2
    nml = tmp_path / "namelist.tmcore_sw"
    nml.write_text(NAMELIST.format(level=level, dt=int(dt), hours=hours, extra="", tmcore="", io=""))
    r = subprocess.run([EXE, str(nml)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = parse(r.stdout)
    nsteps = int(hours * 3600 / dt)
    assert sorted(rows) == list(range(0, nsteps + 1))
    m = Mesh(level)
    gd, u, ghs, q = seeded_state(m, 7, [77])
    o = Oracle(m, m.fit(), dt=dt)
    iap = o.initial_iap(gd, u)
    inv0 = o.invariants(gd, iap, ghs)
    _, _, _, inv, beta = o.run(nsteps, gd, iap, ghs, q)
    np.testing.assert_allclose(rows[0][:3], inv0[:3], rtol=1e-12)
    for s in range(1, nsteps + 1):
        np.testing.assert_allclose(rows[s][:3], inv[s - 1, :3], rtol=1e-12)
        assert abs(rows[s][4] - beta[s - 1]) < 1e-12


def test_fortran_driver_rejects_missing_namelist(tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built")
    r = subprocess.run([EXE, str(tmp_path / "nope.nml")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "cannot open namelist" in r.stdout


def test_fortran_driver_cpu_path_matches_oracle_bitwise(tmp_path):
    """use_gpu = .false.: the same namelist run takes its steps on the host in
    the Fortran FP64 CPU path; every printed invariant and beta (17
    significant digits) equals the oracle's."""
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (amdflang absent when building)")
    level, dt, hours = 4, 600.0, 2
    nml = tmp_path / "namelist.tmcore_sw"
    nml.write_text(NAMELIST.format(level=level, dt=int(dt), hours=hours, extra="use_gpu = .false.\n", tmcore="",
                                   io=""))
    r = subprocess.run([EXE, str(nml)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "on the host CPU" in r.stdout
    rows = parse(r.stdout)
    nsteps = int(hours * 3600 / dt)
    assert sorted(rows) == list(range(0, nsteps + 1))
    m = Mesh(level)
    gd, u, ghs, q = seeded_state(m, 7, [77])
    o = Oracle(m, m.fit(), dt=dt)
    iap = o.initial_iap(gd, u)
    inv0 = o.invariants(gd, iap, ghs)
    _, _, _, inv, beta = o.run(nsteps, gd, iap, ghs, q)
    np.testing.assert_array_equal(rows[0], inv0)
    for s in range(1, nsteps + 1):
        np.testing.assert_array_equal(rows[s][:4], inv[s - 1])
        assert rows[s][4] == beta[s - 1]


@pytest.mark.parametrize("tmcore,extra", [("", "precision = 32\n"), ("adv_order = 3\n", ""),
                                          ("time_scheme = 1\n", ""), ("interp_pv_order = 5\n", "")])
def test_fortran_driver_cpu_path_refuses_what_it_does_not_implement(tmp_path, tmcore, extra):
    """A later namelist assignment overrides the template's: each case asks the
    CPU path for a scheme or form it does not have."""
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built")
    nml = tmp_path / "nml"
    nml.write_text(NAMELIST.format(level=2, dt=600, hours=1, extra="use_gpu = .false.\n" + extra, tmcore=tmcore,
                                   io=""))
    r = subprocess.run([EXE, str(nml)], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "use_gpu = .false. runs" in r.stdout


@pytest.mark.gpu
def test_fortran_driver_cpu_and_gpu_paths_agree(tmp_path):
    """The same namelist with use_gpu .true. and .false.: the device run and
    the host CPU run print the same invariants and beta to 1e-12."""
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (amdflang absent when building)")
    outs = {}
    for flag in (".true.", ".false."):
        nml = tmp_path / f"nml{flag}"
        nml.write_text(NAMELIST.format(level=5, dt=600, hours=3, extra=f"use_gpu = {flag}\n", tmcore="", io=""))
        r = subprocess.run([EXE, str(nml)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        outs[flag] = parse(r.stdout)
    assert sorted(outs[".true."]) == sorted(outs[".false."]) == list(range(0, 19))
    for s in range(19):
        np.testing.assert_allclose(outs[".true."][s][:3], outs[".false."][s][:3], rtol=1e-12)
        if s:
            assert abs(outs[".true."][s][4] - outs[".false."][s][4]) < 1e-12


@pytest.mark.gpu
def test_fortran_driver_attaches_closure_file(tmp_path):
    """closure_file: a zero-output closure leaves every diagnostic line of the
    plain run unchanged; a non-zero one changes them and keeps the mass."""
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (amdflang absent when building)")
    from tmcore_amd.closure import NPARAMS
    level, dt, hours = 3, 600.0, 1
    rng = np.random.default_rng(3)
    params = (rng.standard_normal(NPARAMS) * 0.05).astype("<f4")
    # inputs (u, gd_edge, vorticity, divergence; 10 slots each) normalised to O(1),
    # outputs small: a velocity tendency of ~1e-6 m s^-2 and an edge flux of ~1
    shift = np.concatenate([np.zeros(10), np.full(10, 4.9e4), np.zeros(20)])
    scale = np.repeat([1 / 20.0, 1 / 5e3, 1e5, 1e6], 10)
    scales = np.concatenate([shift, scale, [1e-6, 1.0]]).astype("<f4")
    zero = params.copy()
    zero[-(128 * 2 + 2):] = 0.0
    outs = {}
    for name, p in (("plain", None), ("zero", zero), ("random", params)):
        extra = ""
        if p is not None:
            f = tmp_path / f"{name}.bin"
            np.concatenate([p, scales]).astype("<f4").tofile(f)
            extra = f"closure_file = '{f}'\n"
        nml = tmp_path / f"nml_{name}"
        nml.write_text(NAMELIST.format(level=level, dt=int(dt), hours=hours, extra=extra, tmcore="", io=""))
        r = subprocess.run([EXE, str(nml)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        outs[name] = parse(r.stdout)
    assert outs["zero"] == outs["plain"]
    last = max(outs["plain"])
    assert outs["random"][last][1] != outs["plain"][last][1]  # energy changed
    assert abs(outs["random"][last][0] / outs["plain"][last][0] - 1) < 1e-13  # mass kept


def test_fortran_driver_reports_an_unreadable_grid_file(tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built")
    nml = tmp_path / "nml"
    nml.write_text(NAMELIST.format(level=3, dt=600, hours=1, extra="", io="",
                                   tmcore=f"mesh_file_path = '{tmp_path / 'absent.grid.nc'}'\n"))
    r = subprocess.run([EXE, str(nml)], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "tm_mesh_mpas_dims" in r.stdout and "cannot open" in r.stdout


@pytest.mark.gpu
def test_fortran_driver_reads_mpas_grid_file_and_upwind_order(tmp_path):
    """mesh_file_path (mesh_init, mesh_mod.F90:79-164): a level-4 grid written
    as an MPAS grid file gives the same invariant lines as the generated mesh;
    with adv_order = 4 the run matches the oracle at that order."""
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (amdflang absent when building)")
    level, dt, hours = 4, 600.0, 1
    m = Mesh(level)
    grid = tmp_path / "x4.grid.nc"
    m.to_mpas(grid)
    outs = {}
    for name, tm in (("generated", "adv_order = 4\n"),
                     ("file", f"adv_order = 4\nmesh_file_path = '{grid}'\n")):
        nml = tmp_path / f"nml_{name}"
        nml.write_text(NAMELIST.format(level=level, dt=int(dt), hours=hours, extra="", tmcore=tm, io=""))
        r = subprocess.run([EXE, str(nml)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        outs[name] = parse(r.stdout)
    assert outs["file"] == outs["generated"]
    nsteps = int(hours * 3600 / dt)
    gd, u, ghs, q = seeded_state(m, 7, [77])
    o = Oracle(m, m.fit(), dt=dt, adv_order=4)
    iap = o.initial_iap(gd, u)
    _, _, _, inv, beta = o.run(nsteps, gd, iap, ghs, q)
    for s in range(1, nsteps + 1):
        np.testing.assert_allclose(outs["file"][s][:3], inv[s - 1, :3], rtol=1e-12)


@pytest.mark.gpu
def test_fortran_driver_writes_history(tmp_path):
    """history_interval = '1 hours' (history_mod.F90): records at 0, 1 and 2 h
    whose tm equals the driver's total_mass lines and whose fields equal
    tm_history_fields of the same run through the Python binding."""
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (amdflang absent when building)")
    from scipy.io import netcdf_file
    from tmcore_amd.swm import SwmGpu
    level, dt, hours = 3, 600.0, 2
    nml = tmp_path / "nml"
    nml.write_text(NAMELIST.format(level=level, dt=int(dt), hours=hours, extra="", tmcore="",
                                   io="history_interval = '1 hours'\n"))
    r = subprocess.run([EXE, str(nml)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = parse(r.stdout)
    m = Mesh(level)
    gd, u, ghs, q = seeded_state(m, 7, [77])
    g = SwmGpu(m, m.fit(), dt=dt, ntracers=1)
    g.upload(gd, ghs, u=u, tracers=q)
    g.history_setup()
    g.step(12)
    f = g.history_fields()
    g.close()
    with netcdf_file(str(tmp_path / "test.h0.nc"), "r", mmap=False) as nc:
        np.testing.assert_array_equal(nc.variables["Time"][:], [0.0, 3600.0, 7200.0])
        for i, step in enumerate((0, 6, 12)):
            assert nc.variables["tm"][i] == rows[step][0]
        for k, v in (("u", "u"), ("zonal", "zonalWind"), ("meridional", "merdionalWind"), ("h", "h"),
                     ("pv", "pv"), ("vor", "vor"), ("div", "div")):
            np.testing.assert_array_equal(nc.variables[v][2], f[k], err_msg=v)


@pytest.mark.gpu
def test_fortran_operator_wrappers_match_oracle(tmp_path):
    """The operators of operators_mod.F90 called from Fortran under their
    reference names and dummy arguments (tm_swm_gpu_mod: div_operator ...
    calc_u_tend_on_edge, through iso_c_binding) on random arrays, against the
    oracle's operators; adv_order 3 so that add_upwind_correction_on_cell
    acts."""
    exe = os.path.join(_ffi.LIB_DIR, "test_tmcore_ops")
    if not os.path.exists(exe):
        pytest.skip("Fortran operator test not built (amdflang absent when building)")
    level = 4
    nml = tmp_path / "namelist.ops"
    nml.write_text(NAMELIST.format(level=level, dt=600, hours=1, extra="", tmcore="adv_order = 3\nadv_mono = .false.\n",
                                   io=""))
    m = Mesh(level)
    nC, nE, nV = m.nCells, m.nEdges, m.nVertices
    rng = np.random.default_rng(21)
    fc, fe, fv, ue, ve = (rng.standard_normal(n) for n in (nC, nE, nV, nE, nE))
    gde = 1.0 + rng.random(nE)
    tgv, pvv, pvc = rng.standard_normal(nV), rng.standard_normal(nV), rng.standard_normal(nC)
    kec, gdc, ghs = rng.standard_normal(nC), 1.0 + rng.random(nC), rng.standard_normal(nC)
    pvf, tge = rng.standard_normal(nE), rng.standard_normal(nE)
    fin, fout = tmp_path / "ops_in.bin", tmp_path / "ops_out.bin"
    np.concatenate([fc, fe, fv, ue, ve, gde, tgv, pvv, pvc, kec, gdc, ghs, pvf, tge]).astype("=f8").tofile(fin)
    r = subprocess.run([exe, str(nml), str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(fout, dtype="=f8")
    sizes = [nC, nV, nE, nV, nC, nE, nC, nE, nE, nE, nV, nE, nE, nE]
    assert out.size == sum(sizes)
    got = np.split(out, np.cumsum(sizes)[:-1])
    o = Oracle(m, m.fit(), dt=600.0, adv_order=3, adv_mono=False)
    from oracle.oracle import _p, lib
    upw = fe.copy()
    lib().om_upwind_cell(o.h, _p(fc), _p(ue), _p(upw))
    ut, it = np.zeros(nE), np.zeros(nE)
    lib().om_u_tend(o.h, *[_p(a) for a in (ue, kec, gdc, gde, ghs, pvf, tge)], _p(ut), _p(it))
    ref = [o.op("div", fe, out_where="c"), o.op("curl", fe, out_where="v"), o.op("c2e", fc, out_where="e"),
           o.op("c2v", fc, out_where="v"), o.op("v2c", fv, out_where="c"), o.op("v2e", fv, out_where="e"),
           o.op("ke", ue, out_where="c"), o.op("tangent_wind", ue, out_where="e"),
           o.op("tangent_vor_flux", ue, gde, fe, out_where="e"), upw,
           o.op("pv_vertex", fv, np.abs(pvv) + 1, out_where="v"),
           o.op("pv_edge", ue, ve, gde, tgv, pvv, pvc, out_where="e"), ut, it]
    names = ["div", "curl", "c2e", "c2v", "v2c", "v2e", "ke", "tangent_wind", "tangent_vor_flux", "upwind",
             "pv_vertex", "pv_edge", "u_tend", "iap_u_tend"]
    assert np.abs(upw - fe).max() > 0  # the correction acts at adv_order 3
    for name, g, rf in zip(names, got, ref):
        err = np.linalg.norm(g - rf) / np.linalg.norm(rf)
        assert err <= 1e-13, (name, err)
