"""GPU tests of the two ways the flow pass solves the IBM system S y = c:
the dense mat-vec with the explicit inverse S^-1 and the band substitution
with the factors S = L D L^T (one wave per wing and component).  Steps with
rigid-body coupling take the substitution (their geometry is on the critical
path); LBM_DIRECT_SOLVE selects it for no step (0), those steps (1, the
default) or every step (2).  Steps whose wings' support rectangles meet keep
the inverse, which the coupling of wings reads.

The two agree to rounding: fields within 1e-12 relative L2 and forces within
1e-11 over 100 steps, and both stay within the oracle bars of the other
parity tests (fields 1e-10, window-averaged forces 1e-9).
"""
import numpy as np
import pytest

from lbm_mutiwing_amd import Solver, WingMotion, preset, seeded_motions, sinusoidal_motion_table
from lbm_mutiwing_amd.solver import BODY_REL_FIELDS, Dynamics

pytestmark = pytest.mark.gpu

N = 100


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from lbm_mutiwing_amd._native import load_library
    return load_library()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def _run(monkeypatch, cfg, table, mode, dyn=None, rel0=None, steps=N):
    monkeypatch.setenv("LBM_DIRECT_SOLVE", str(mode))
    s = Solver(cfg)
    s.set_motion_table(table)
    if dyn is not None:
        s.set_dynamics(dyn)
    if rel0 is not None:
        s.set_body_rel(rel0)
    r = s.advance(steps, want_rows=True)
    out = (r, s.fields()[:3], s.body_rel(), s.state()["movenum"])
    s.close()
    return out


@pytest.mark.parametrize("name", ["small", "c1"])
def test_substitution_matches_inverse_prescribed(lib, oracle_mod, monkeypatch, name):
    cfg = preset(name)
    table = sinusoidal_motion_table(cfg, seeded_motions(cfg, 11), N + 2)
    print(f"{name}: inverse vs substitution, {N} steps", flush=True)
    ri, fi, _, _ = _run(monkeypatch, cfg, table, 0)
    rd, fd, _, _ = _run(monkeypatch, cfg, table, 2)
    for a, b in zip(fd, fi):
        assert rel(a, b) < 1e-12
    assert rel(rd.force_rows, ri.force_rows) < 1e-11
    o = oracle_mod.OracleSolver(cfg)
    o.set_motion_table(table)
    ro = o.advance(N, want_rows=True)
    for a, b in zip(fd, o.fields()[:3]):
        assert rel(a, b) < 1e-10
    W = cfg.wing_count
    c = slice(1, 1 + 3 * W + 3)
    ag, ao = rd.force_rows[N // 2:, c].mean(0), ro.force_rows[N // 2:, c].mean(0)
    assert np.all(np.abs(ag - ao) <= 1e-9 * np.maximum(np.abs(ao), 1e-3 * np.abs(ao).max()))
    o.close()


def test_wings_that_share_nodes_keep_the_inverse(lib, monkeypatch):
    """Support rectangles that meet: every step keeps the explicit inverse,
    so the results are bitwise those of the inverse-only run."""
    cfg = preset("small")
    cfg.wing_dx = 1.05
    motions = [WingMotion(pitch=0.05, heave=0.02, freq=0.25, phase_p=0.0, phase_h=0.5),
               WingMotion(pitch=0.08, heave=0.03, freq=0.25, phase_p=1.0, phase_h=2.0)]
    table = sinusoidal_motion_table(cfg, motions, N + 2)
    ri, fi, _, _ = _run(monkeypatch, cfg, table, 0)
    rd, fd, _, _ = _run(monkeypatch, cfg, table, 2)
    assert np.array_equal(rd.force_rows, ri.force_rows)
    for a, b in zip(fd, fi):
        assert np.array_equal(a, b)


def test_substitution_matches_inverse_with_coupling(lib, monkeypatch):
    """Wing 1 coupled to the flow and crossing a lattice node (move_mesh):
    the substitution (default) against the inverse, same moves, body state
    within 1e-11."""
    cfg = preset("small")
    table = sinusoidal_motion_table(cfg, seeded_motions(cfg, 5), N + 2)
    dyn = Dynamics(rk4=True, solid_den=50.0, par1=-2.0, par2=50.0)
    rel0 = np.zeros((11, cfg.wing_count))
    rel0[BODY_REL_FIELDS.index("v"), 0] = 0.02
    rel0[BODY_REL_FIELDS.index("pos"), 0] = 0.5
    ri, fi, bi, mi = _run(monkeypatch, cfg, table, 0, dyn, rel0)
    rd, fd, bd, md = _run(monkeypatch, cfg, table, 1, dyn, rel0)
    assert md == mi >= 1
    for a, b in zip(fd, fi):
        assert rel(a, b) < 1e-12
    assert rel(rd.force_rows, ri.force_rows) < 1e-11
    assert np.allclose(bd, bi, rtol=1e-11, atol=1e-16)


def _run_mode(cfg, table, mode, dyn=None, rel0=None, steps=N):
    """As _run, with the solve mode set through lbm_set_ibm_solve."""
    s = Solver(cfg)
    s.set_ibm_solve(mode)
    s.set_motion_table(table)
    if dyn is not None:
        s.set_dynamics(dyn)
    if rel0 is not None:
        s.set_body_rel(rel0)
    r = s.advance(steps, want_rows=True)
    out = (r, s.fields()[:3], s.body_rel(), s.state()["movenum"])
    s.close()
    return out


def test_solve_mode_entry_point(lib):
    """lbm_set_ibm_solve: the twisted substitutions of coupled steps at C1
    (126 segments per wing) and the plain ones of prescribed steps against
    the inverse, to rounding; a mode outside 0..2 is an error."""
    from lbm_mutiwing_amd.solver import (IBM_SOLVE_COUPLED, IBM_SOLVE_INVERSE, IBM_SOLVE_SUBSTITUTION,
                                         LbmError)
    cfg = preset("c1")
    table = sinusoidal_motion_table(cfg, seeded_motions(cfg, 7), N + 2)
    print("c1 coupled: twisted substitution vs inverse", flush=True)
    dyn = Dynamics(rk4=True, solid_den=50.0, par1=-2.0, par2=50.0)
    ri, fi, bi, mi = _run_mode(cfg, table, IBM_SOLVE_INVERSE, dyn)
    rt, ft, bt, mt = _run_mode(cfg, table, IBM_SOLVE_COUPLED, dyn)
    assert mt == mi
    for a, b in zip(ft, fi):
        assert rel(a, b) < 1e-12
    assert rel(rt.force_rows, ri.force_rows) < 1e-11
    assert np.allclose(bt, bi, rtol=1e-11, atol=1e-16)
    print("c1 prescribed: plain substitution vs inverse", flush=True)
    ri, fi, _, _ = _run_mode(cfg, table, IBM_SOLVE_INVERSE)
    rp, fp, _, _ = _run_mode(cfg, table, IBM_SOLVE_SUBSTITUTION)
    for a, b in zip(fp, fi):
        assert rel(a, b) < 1e-12
    assert rel(rp.force_rows, ri.force_rows) < 1e-11
    s = Solver(cfg)
    with pytest.raises(LbmError):
        s.set_ibm_solve(3)
    s.close()
