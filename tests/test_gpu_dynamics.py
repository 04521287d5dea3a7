"""GPU parity of the rigid-body coupling of wing 1 (boundary_condition.f90:
1694-1719, rk4 :3168-3202) against the FP64 oracle, through the C ABI.

The case: the 256x128 two-wing preset with seeded pitch/heave motion (file
mode), wing 1 starting half a node from a move and drifting downstream at
0.02 lattice units per step, so that move_mesh runs during the run; the
solid density 50 keeps the explicit coupling stable.  Tolerances as the
other parity tests: fields 1e-10 relative L2, window-averaged forces 1e-9
relative, body state 1e-9 relative.
"""
import numpy as np
import pytest

from lbm_mutiwing_amd import Solver, preset, seeded_motions, sinusoidal_motion_table
from lbm_mutiwing_amd.solver import BODY_REL_FIELDS, Dynamics, LbmError

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-10
FORCE_TOL = 1e-9
N = 100


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from lbm_mutiwing_amd._native import load_library
    return load_library()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def _start(s, cfg, rk4):
    s.set_motion_table(sinusoidal_motion_table(cfg, seeded_motions(cfg, 5), N + 2))
    s.set_dynamics(Dynamics(rk4=rk4, solid_den=50.0, par1=-2.0, par2=50.0))
    rel0 = np.zeros((11, cfg.wing_count))
    rel0[BODY_REL_FIELDS.index("v"), 0] = 0.02
    rel0[BODY_REL_FIELDS.index("pos"), 0] = 0.5
    s.set_body_rel(rel0)


@pytest.mark.parametrize("rk4", [True, False])
def test_dynamics_match_oracle(lib, oracle_mod, rk4):
    cfg = preset("small")
    g, o = Solver(cfg), oracle_mod.OracleSolver(cfg)
    for s in (g, o):
        _start(s, cfg, rk4)
    print(f"dynamics rk4={rk4}: {N} steps", flush=True)
    rg, ro = g.advance(N, want_rows=True, want_motion=True), o.advance(N, want_rows=True, want_motion=True)
    assert g.state()["movenum"] == o.state()["movenum"] >= 2
    assert g.state()["flag2"] == o.state()["flag2"]
    bg, bo = g.body_rel(), o.body_rel()
    assert np.allclose(bg, bo, rtol=1e-9, atol=1e-14), (bg[:, 0], bo[:, 0])
    for a, b in zip(g.fields()[:3], o.fields()[:3]):
        assert rel(a, b) < FIELD_TOL
    avg_g = rg.force_rows[20:, 1:10].mean(0)
    avg_o = ro.force_rows[20:, 1:10].mean(0)
    assert np.all(np.abs(avg_g - avg_o) <= FORCE_TOL * np.maximum(np.abs(avg_o), 1e-3))
    assert rel(rg.row_sum, ro.row_sum) < FORCE_TOL
    # motion columns follow the body: x and body_extra_v in every row
    assert np.allclose(rg.force_rows[:, 10:], ro.force_rows[:, 10:], rtol=1e-10, atol=1e-13)
    assert np.allclose(rg.motion_rows, ro.motion_rows, rtol=1e-10, atol=1e-13)
    ig, io = g.ibm(), o.ibm()
    assert np.array_equal(ig[0], io[0]) and np.array_equal(ig[1], io[1])
    for a, b in zip(g.markers(), o.markers()):
        assert np.allclose(a, b, rtol=0, atol=1e-11)


def test_dynamics_split_calls_bitwise(lib):
    """Steps run one at a time: splitting the advance changes nothing."""
    cfg = preset("small")
    a, b = Solver(cfg), Solver(cfg)
    for s in (a, b):
        _start(s, cfg, True)
    print("dynamics split calls", flush=True)
    ra = a.advance(60)
    b.advance(25)
    rb = b.advance(60)
    assert np.array_equal(a.populations()[0], b.populations()[0])
    assert np.array_equal(a.body_rel(), b.body_rel())
    assert np.array_equal(ra.last_row, rb.last_row)


def test_dynamics_off_is_the_prescribed_path(lib):
    """Switching the coupling on and off again leaves the default path."""
    cfg = preset("small")
    a, b = Solver(cfg), Solver(cfg)
    rows = sinusoidal_motion_table(cfg, seeded_motions(cfg, 5), 42)
    for s in (a, b):
        s.set_motion_table(rows)
    b.set_dynamics(Dynamics(solid_den=50.0))
    b.set_dynamics(None)
    print("dynamics off", flush=True)
    a.advance(40)
    b.advance(40)
    assert np.array_equal(a.populations()[0], b.populations()[0])
    assert np.array_equal(b.body_rel(), np.zeros((11, 2)))


def test_batch_rejects_dynamics(lib):
    from lbm_mutiwing_amd import Batch
    cfg = preset("small")
    s = Solver(cfg)
    _start(s, cfg, True)
    bt = Batch(lib)
    with pytest.raises(LbmError, match="rigid-body coupling"):
        bt.advance([s], 5)
    bt.close()


def test_dynamics_fp32_storage_within_fp32_bar(lib, oracle_mod):
    """FP32-shifted populations with the coupling on: the body update itself
    runs in FP64 on the FP32 run's forces; fields within 1e-4 and the body
    state within 1e-3 of the FP64 oracle (the FP32 bars), same moves."""
    from lbm_mutiwing_amd import PRECISION_FP32_SHIFTED
    cfg = preset("small")
    o = oracle_mod.OracleSolver(cfg)
    cfg32 = preset("small", precision=PRECISION_FP32_SHIFTED)
    g = Solver(cfg32)
    for s, c in ((g, cfg32), (o, cfg)):
        _start(s, c, True)
    print("dynamics fp32-shifted", flush=True)
    rg, ro = g.advance(N, want_rows=True), o.advance(N, want_rows=True)
    assert g.state()["movenum"] == o.state()["movenum"] >= 2
    bg, bo = g.body_rel()[:, 0], o.body_rel()[:, 0]
    scale = np.maximum(np.abs(bo), 1e-12)
    for name in ("v", "s", "pos"):
        i = BODY_REL_FIELDS.index(name)
        assert abs(bg[i] - bo[i]) <= 1e-3 * scale[i], (name, bg[i], bo[i])
    for a, b in zip(g.fields()[:3], o.fields()[:3]):
        assert rel(a, b) < 1e-4
    avg_g = rg.force_rows[20:, 1:10].mean(0)
    avg_o = ro.force_rows[20:, 1:10].mean(0)
    assert np.all(np.abs(avg_g - avg_o) <= 1e-3 * np.maximum(np.abs(avg_o).max(), 1e-3))


def test_dynamics_c1_match_oracle(lib, oracle_mod):
    """The 1024x512 two-wing workload (chord 64) with wing 1 coupled, 200
    steps from a drift of 0.02 nodes per step half a node from a move: the
    same bars as the small case."""
    cfg = preset("c1")
    rows = sinusoidal_motion_table(cfg, seeded_motions(cfg, 12), 204)
    g, o = Solver(cfg), oracle_mod.OracleSolver(cfg)
    for s in (g, o):
        s.set_motion_table(rows)
        s.set_dynamics(Dynamics(solid_den=40.0, par1=-2.0, par2=50.0))
        rel0 = np.zeros((11, cfg.wing_count))
        rel0[BODY_REL_FIELDS.index("v"), 0] = 0.02
        rel0[BODY_REL_FIELDS.index("pos"), 0] = 0.5
        s.set_body_rel(rel0)
    print("dynamics c1: 200 steps", flush=True)
    rg, ro = g.advance(200, want_rows=True), o.advance(200, want_rows=True)
    assert g.state()["movenum"] == o.state()["movenum"] >= 1
    assert np.allclose(g.body_rel(), o.body_rel(), rtol=1e-9, atol=1e-14)
    for a, b in zip(g.fields()[:3], o.fields()[:3]):
        assert rel(a, b) < FIELD_TOL
    avg_g = rg.force_rows[50:, 1:10].mean(0)
    avg_o = ro.force_rows[50:, 1:10].mean(0)
    assert np.all(np.abs(avg_g - avg_o) <= FORCE_TOL * np.maximum(np.abs(avg_o), 1e-3))


def test_dynamics_with_coupled_wings_match_oracle(lib, oracle_mod):
    """Wing 1 coupled to the flow while the supports of wings 1 and 2 share
    nodes: the body moves with the force of the coupled system."""
    cfg = preset("small")
    cfg.wing_dx = 1.05
    g, o = Solver(cfg), oracle_mod.OracleSolver(cfg)
    for s in (g, o):
        _start(s, cfg, True)
    print("dynamics with coupled wings", flush=True)
    rg, ro = g.advance(N, want_rows=True), o.advance(N, want_rows=True)
    ig, io = g.ibm(), o.ibm()
    assert np.array_equal(ig[0], io[0]) and np.array_equal(ig[1], io[1])
    assert g.state()["movenum"] == o.state()["movenum"] >= 1
    assert np.allclose(g.body_rel(), o.body_rel(), rtol=1e-9, atol=1e-14)
    for a, b in zip(g.fields()[:3], o.fields()[:3]):
        assert rel(a, b) < FIELD_TOL
    assert rel(rg.row_sum, ro.row_sum) < FORCE_TOL
