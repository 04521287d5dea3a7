"""CPU tests of the on-disk formats (force.dat, airfoil_motion.dat,
motion_params.dat, marker files, Tecplot snapshot) and of the file-mode run
loop, driven with the FP64 oracle as the solver."""
import os

import numpy as np
import pytest

from lbm_mutiwing_amd import SolverConfig, naca0012_markers, preset, seeded_motions, sinusoidal_motion_table
from lbm_mutiwing_amd.formats import (MotionFileError, f15_8_record, fortran_e, fortran_f, fortran_i,
                                      load_motion_params, read_marker_file, run_file_mode, vorticity,
                                      write_motion_params, write_tecplot)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_edit_descriptors():
    assert fortran_f(1.5, 15, 8) == "     1.50000000"
    assert fortran_f(-0.0, 15, 8) == "    -0.00000000"
    assert fortran_f(-12345.123456789, 15, 8) == "-12345.12345679"
    assert fortran_f(1.0e7, 15, 8) == "*" * 15
    assert fortran_e(0.123456789, 16, 8) == "  0.12345679E+00"
    assert fortran_e(-1234.5, 16, 8) == " -0.12345000E+04"
    assert fortran_e(0.0, 16, 8) == "  0.00000000E+00"
    assert fortran_e(3.0e-120, 16, 8) == "  0.30000000-119"
    assert fortran_i(7, 5, 3) == "  007"
    assert fortran_i(1234, 5, 3) == " 1234"
    assert f15_8_record([0.5, -2.0]) == "     0.50000000    -2.00000000"


def test_motion_params_file(tmp_path):
    rows = load_motion_params(os.path.join(GOLDEN, "motion_params_head.dat"), 2)
    assert rows.shape == (3, 19)
    p = tmp_path / "m.dat"
    write_motion_params(str(p), rows)
    with open(p, "a") as fh:
        fh.write("\n\n")                       # blank lines are skipped
    assert np.array_equal(load_motion_params(str(p), 2), rows)
    bad = tmp_path / "bad.dat"
    bad.write_text("0.0 1 2 3\n")
    with pytest.raises(MotionFileError):
        load_motion_params(str(bad), 2)
    dec = tmp_path / "dec.dat"
    write_motion_params(str(dec), rows[::-1])
    with pytest.raises(MotionFileError):
        load_motion_params(str(dec), 2)
    with pytest.raises(MotionFileError):
        load_motion_params(str(tmp_path / "missing.dat"), 2)


def test_marker_file_matches_generator():
    m = read_marker_file(os.path.join(GOLDEN, "boundary.dat"))
    assert m.shape == (205, 2)
    assert np.abs(m - naca0012_markers(101)).max() <= 1e-8 + 1e-12


def test_vorticity_of_solid_rotation():
    ny, nx, sc, uu, om = 9, 11, 0.1, 0.05, 0.3
    y, x = np.meshgrid(np.arange(ny) * sc, np.arange(nx) * sc, indexing="ij")
    u, v = -om * y, om * x
    assert np.allclose(vorticity(u, v, uu, sc), 2 * om / uu, rtol=0, atol=1e-11)


def test_tecplot_layout(tmp_path):
    cfg = SolverConfig(nx=9, ny=7, scale=0.5, wing_count=1, markers=naca0012_markers(3))
    P = cfg.points_per_wing
    u = np.full((7, 9), cfg.uu)
    v = np.zeros((7, 9))
    pre = np.full((7, 9), 1.0 / 3.0)
    xb, yb = cfg.marker_array()[:, 0], cfg.marker_array()[:, 1]
    path = tmp_path / "F12.dat"
    write_tecplot(str(path), cfg, 12, u, v, pre, xb, yb)
    lines = path.read_text().splitlines()
    out_n = ((9 - 1) // 2 + 1) * ((7 - 1) // 2 + 1)
    assert lines[0].startswith('VARIABLES = "X" "Y" "U" "V" "Presure" "Vorticity" "body" "dvordt"')
    assert 'ZONE T="Flow feild"' in lines[1]
    assert len(lines) == 2 + out_n + 1 + (P + 1) + (P - 1)
    first = [float(t) for t in lines[2].split()]
    ox, oy = cfg.origin()
    assert np.allclose(first, [ox, oy, 1.0, 0.0, 1.0 / 3.0, 0.0, 0.0, 0.0], atol=1e-8)
    assert lines[2 + out_n].lstrip().startswith('ZONE T="airfoil"')
    assert lines[-1] == (
        f"  {P - 1:03d}  {P:03d}  {P + 1:03d}")


def test_file_mode_run_writes_rows(tmp_path, oracle_mod):
    cfg = preset("small")
    s = oracle_mod.OracleSolver(cfg)
    n = 30
    table = sinusoidal_motion_table(cfg, seeded_motions(cfg, 3), n)
    run_file_mode(s, table, str(tmp_path), chunk=7)
    f = np.loadtxt(tmp_path / "force.dat")
    m = np.loadtxt(tmp_path / "airfoil_motion.dat")
    assert f.shape == (n, cfg.row_len) and m.shape == (n, 1 + 3 * cfg.wing_count)
    ref = oracle_mod.OracleSolver(cfg)
    ref.set_motion_table(table)
    r = ref.advance(n, want_rows=True)
    assert np.allclose(f, r.force_rows, rtol=0, atol=5.01e-9)    # F15.8 rounding
