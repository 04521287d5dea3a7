"""GPU parity tests: the HIP library (through its C ABI) against the FP64 CPU
oracle on the same seeded inputs.

Tolerances: the fused lattice arithmetic is bit-identical to the oracle
(same operation order, no FMA contraction), so runs without wings must match
bit for bit.  With wings the IBM system is solved by a banded Cholesky of
S = D D^T instead of the reference's LU of A = S diag(arc) (same solution,
different rounding): fields must agree within 1e-10 relative L2 and
window-averaged forces/moments within 1e-9 relative.
"""
import numpy as np
import pytest

from lbm_mutiwing_amd import Solver, SolverConfig, WingMotion, preset, seeded_motions, sinusoidal_motion_table

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-10
FORCE_TOL = 1e-9


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from lbm_mutiwing_amd._native import load_library
    return load_library()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def pair(cfg, oracle_mod, rows=None):
    g, o = Solver(cfg), oracle_mod.OracleSolver(cfg)
    if rows is not None:
        g.set_motion_table(rows)
        o.set_motion_table(rows)
    return g, o


def test_create_reset_destroy(lib):
    s = Solver(preset("small"))
    st = s.state()
    assert st["n_next"] == 1 and st["row_len"] == 28
    s.reset()
    s.close()


def test_no_wings_bit_exact(lib, oracle_mod):
    cfg = preset("small")
    cfg.wing_count = 0
    g, o = pair(cfg, oracle_mod)
    rng = np.random.default_rng(11)
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    f = w[:, None, None] * (1 + 0.02 * rng.normal(size=(9, cfg.ny, cfg.nx)))
    g.set_populations(f, f)
    o.set_populations(f, f)
    rg, ro = g.advance(60, want_rows=True), o.advance(60, want_rows=True)
    fg, f0g = g.populations()
    fo, f0o = o.populations()
    assert np.array_equal(fg, fo)
    for a, b in zip(g.fields()[:3], o.fields()[:3]):
        assert np.array_equal(a, b)
    assert rg.err == pytest.approx(ro.err, rel=1e-12)


def test_wings_fields_and_forces(lib, oracle_mod):
    cfg = preset("small")
    n = 600
    rows = sinusoidal_motion_table(cfg, seeded_motions(cfg, 5), n + 2)
    g, o = pair(cfg, oracle_mod, rows)
    rg, ro = g.advance(n, want_rows=True), o.advance(n, want_rows=True)
    for a, b in zip(g.fields()[:3], o.fields()[:3]):
        assert rel(a, b) < FIELD_TOL
    # window-averaged per-wing forces and moments (columns 2..1+3W+3)
    avg_g = rg.force_rows[100:, 1:10].mean(0)
    avg_o = ro.force_rows[100:, 1:10].mean(0)
    assert np.all(np.abs(avg_g - avg_o) <= FORCE_TOL * np.maximum(np.abs(avg_o), 1e-3))
    assert rel(rg.row_sum, ro.row_sum) < FORCE_TOL
    assert rg.err == pytest.approx(ro.err, rel=1e-8)
    ig, io = g.ibm(), o.ibm()
    assert np.array_equal(ig[0], io[0]) and np.array_equal(ig[1], io[1])  # adjoint list and owners
    xg, xo = g.markers(), o.markers()
    for a, b in zip(xg, xo):
        assert np.allclose(a, b, rtol=0, atol=1e-14)


def test_stream_mode_segments(lib, oracle_mod):
    cfg = preset("small")
    cfg.motion_source = 2
    g, o = pair(cfg, oracle_mod)
    W = cfg.wing_count
    start = np.zeros(cfg.motion_cols)
    for s_ in (g, o):
        s_.set_stream_segment(start, start)
    rg, ro = g.advance(1), o.advance(1)          # RESET runs one step
    assert rel(rg.row_sum, ro.row_sum) < FORCE_TOL
    rng = np.random.default_rng(3)
    t = 0.0
    for seg in range(3):
        t_new = t + 100 * cfg.dt
        delta = np.concatenate([[t_new], rng.normal(0, 0.1, 3 * W)])
        tgt = g.build_stream_target_row(start, t_new, delta)
        assert np.array_equal(tgt, o.build_stream_target_row(start, t_new, delta))
        for s_ in (g, o):
            s_.set_stream_segment(start, tgt)
        n_last = int(round(t_new / cfg.dt)) + 1
        rg, ro = g.advance(n_last), o.advance(n_last)
        assert rel(rg.row_sum[1:10], ro.row_sum[1:10]) < FORCE_TOL
        assert np.allclose(rg.last_row[10:], ro.last_row[10:], rtol=1e-12, atol=1e-14)
        start, t = tgt, t_new
    for a, b in zip(g.fields()[:3], o.fields()[:3]):
        assert rel(a, b) < FIELD_TOL


def test_move_mesh_shift(lib, oracle_mod):
    cfg = preset("small")
    rows = sinusoidal_motion_table(cfg, seeded_motions(cfg, 9), 120)
    g, o = pair(cfg, oracle_mod, rows)
    g.advance(40)
    o.advance(40)
    for s_ in (g, o):
        s_.set_body_rel_pos(1, 1.0)
    rg, ro = g.advance(41), o.advance(41)
    assert g.state()["movenum"] == o.state()["movenum"] == 1
    assert g.state()["flag2"] == o.state()["flag2"] == 2
    for a, b in zip(g.fields()[:3], o.fields()[:3]):
        assert rel(a, b) < FIELD_TOL
    rg, ro = g.advance(42), o.advance(42)
    assert rg.err == ro.err == 0.0
    rg, ro = g.advance(80, want_rows=True), o.advance(80, want_rows=True)
    for a, b in zip(g.fields()[:3], o.fields()[:3]):
        assert rel(a, b) < FIELD_TOL
    assert rel(rg.row_sum, ro.row_sum) < FORCE_TOL
    fg, f0g = g.populations()
    fo, f0o = o.populations()
    assert np.allclose(f0g[:, :, 0], f0o[:, :, 0], rtol=0, atol=1e-15)   # shifted f0 ring
    assert rel(fg, fo) < FIELD_TOL

    def ring_matches():
        """f0 ring columns and rows (the rows are shifted too,
        move_mesh.f90:30-34, and saved by the checkpoint)."""
        f0g, f0o = g.populations()[1], o.populations()[1]
        for sl in ((slice(None), slice(None), 0), (slice(None), slice(None), -1),
                   (slice(None), 0, slice(None)), (slice(None), -1, slice(None))):
            assert rel(f0g[sl], f0o[sl]) < FIELD_TOL, sl
    ring_matches()
    for s_ in (g, o):   # and a move back (dr = -1, move_mesh.f90:171-214)
        s_.set_body_rel_pos(1, -1.0)
    g.advance(81)
    o.advance(81)
    assert g.state()["movenum"] == o.state()["movenum"] == 0
    ring_matches()


def test_run_to_run_bitwise_deterministic(lib):
    cfg = preset("small")
    rows = sinusoidal_motion_table(cfg, seeded_motions(cfg, 21), 300)
    out = []
    for _ in range(2):
        s = Solver(cfg)
        s.set_motion_table(rows)
        r = s.advance(250, want_rows=True)
        out.append((s.populations()[0], r.force_rows))
        s.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("tile,waves", [(128, 16), (256, 24), (512, 0)])
def test_tuning_choices_bitwise_equal(lib, tile, waves, monkeypatch):
    """Every tile width and wave cap the autotuner may pick computes the same
    bits (the launch shape changes, the arithmetic per node does not)."""
    cfg = preset("small")
    rows = sinusoidal_motion_table(cfg, seeded_motions(cfg, 21), 100)
    out = []
    for t, w in ((tile, waves), (256, 0)):
        monkeypatch.setenv("LBM_LATTICE_TILE", str(t))
        monkeypatch.setenv("LBM_LATTICE_WAVES", str(w))
        s = Solver(cfg)
        s.set_motion_table(rows)
        r = s.advance(60, want_rows=True)
        assert s.tuning_full() == (t, w)
        out.append((s.populations()[0], r.force_rows))
        s.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_reset_restores_initial_state(lib):
    cfg = preset("small")
    rows = sinusoidal_motion_table(cfg, seeded_motions(cfg, 2), 200)
    s = Solver(cfg)
    s.set_motion_table(rows)
    a = s.advance(100, want_rows=True).force_rows
    s.reset()
    b = s.advance(100, want_rows=True).force_rows
    assert np.array_equal(a, b)


def test_wing_at_edge_is_an_error(lib):
    from lbm_mutiwing_amd import LbmError
    cfg = preset("small")
    s = Solver(cfg)
    rows = sinusoidal_motion_table(cfg, [WingMotion(), WingMotion()], 10)
    rows[:, 4] = -3.9      # wing 1 pivot 0.1 chord from the left edge
    s.set_motion_table(rows)
    with pytest.raises(LbmError):
        s.advance(2)


def _coupled_case(dx=1.05):
    cfg = preset("small")
    cfg.wing_dx = dx       # 0.05 chord (1.6 nodes) between trailing and leading edge
    return cfg


def test_coupled_wings_match_oracle(lib, oracle_mod):
    """Two wings whose support boxes share lattice nodes: one IBM system over
    both wings' segments, the shared nodes owned by the first wing, as the
    reference solves it (boundary_condition.f90:1558-1640, airfoil.f90:288-333).
    Fields within 1e-10, forces within 1e-9, identical adjoint list and owners."""
    cfg = _coupled_case()
    n = 200
    motions = [WingMotion(pitch=0.05, heave=0.02, freq=0.25, phase_p=0.0, phase_h=0.5),
               WingMotion(pitch=0.08, heave=0.03, freq=0.25, phase_p=1.0, phase_h=2.0)]
    rows = sinusoidal_motion_table(cfg, motions, n + 2)
    g, o = pair(cfg, oracle_mod, rows)
    rg, ro = g.advance(n, want_rows=True), o.advance(n, want_rows=True)
    ig, io = g.ibm(), o.ibm()
    assert np.array_equal(ig[0], io[0]) and np.array_equal(ig[1], io[1])  # adjoint list and owners
    assert (io[1] == 2).any() and (io[1] == 1).any()
    errs = [rel(a, b) for a, b in zip(g.fields()[:3], o.fields()[:3])]
    print(f"coupled wings: rho/u/v relative L2 {errs}", flush=True)
    assert max(errs) < FIELD_TOL
    W = cfg.wing_count
    fr = slice(1, 1 + 3 * W + 3)
    assert rel(rg.force_rows[:, fr], ro.force_rows[:, fr]) < FORCE_TOL
    assert rel(rg.row_sum, ro.row_sum) < FORCE_TOL


def test_coupled_wings_batch_equals_alone(lib):
    """Coupled instances (a pair, and a chain of three) advanced in a batch
    end bitwise as alone."""
    from lbm_mutiwing_amd import Batch
    chain = _coupled_case()
    chain.wing_count = 3
    cfgs = [_coupled_case(), preset("small"), chain]
    motions = [WingMotion(pitch=0.05), WingMotion(pitch=0.08, heave=0.03), WingMotion(pitch=0.06)]
    alone, mixed = [], []
    for c in cfgs:
        for lst in (alone, mixed):
            s = Solver(c)
            s.set_motion_table(sinusoidal_motion_table(c, motions[:c.wing_count], 80))
            lst.append(s)
    ra = [s.advance(40) for s in alone]
    b = Batch(lib)
    rb = b.advance(mixed, 40)
    for e in range(len(cfgs)):
        assert np.array_equal(alone[e].populations()[0], mixed[e].populations()[0]), e
        assert np.array_equal(ra[e].row_sum, rb[e].row_sum), e
    b.close()


def test_three_coupled_wings_match_oracle(lib, oracle_mod):
    """Three wings in a chain (the supports of wings 1-2 and 2-3 share
    nodes): one system over all the segments, each shared node owned by the
    lowest of its wings, as the oracle's dense LU over all three."""
    cfg = _coupled_case()
    cfg.wing_count = 3
    n = 120
    motions = [WingMotion(pitch=0.05, heave=0.02, freq=0.25, phase_p=0.0, phase_h=0.5),
               WingMotion(pitch=0.08, heave=0.03, freq=0.25, phase_p=1.0, phase_h=2.0),
               WingMotion(pitch=0.06, heave=0.02, freq=0.25, phase_p=2.0, phase_h=1.0)]
    rows = sinusoidal_motion_table(cfg, motions, n + 2)
    g, o = pair(cfg, oracle_mod, rows)
    print("three coupled wings: 120 steps", flush=True)
    rg, ro = g.advance(n, want_rows=True), o.advance(n, want_rows=True)
    ig, io = g.ibm(), o.ibm()
    assert np.array_equal(ig[0], io[0]) and np.array_equal(ig[1], io[1])
    assert all((io[1] == w).any() for w in (1, 2, 3))
    errs = [rel(a, b) for a, b in zip(g.fields()[:3], o.fields()[:3])]
    print(f"three coupled wings: rho/u/v relative L2 {errs}", flush=True)
    assert max(errs) < FIELD_TOL
    fr = slice(1, 1 + 3 * cfg.wing_count + 3)
    assert rel(rg.force_rows[:, fr], ro.force_rows[:, fr]) < FORCE_TOL
    assert rel(rg.row_sum, ro.row_sum) < FORCE_TOL


def test_too_many_coupling_segments_are_an_error(lib):
    """Wings overlapping over most of their chord couple more than 32
    segments of a wing: reported as a failed step."""
    from lbm_mutiwing_amd import LbmError
    cfg = _coupled_case(dx=0.2)
    s = Solver(cfg)
    s.set_motion_table(sinusoidal_motion_table(cfg, [WingMotion(), WingMotion()], 10))
    with pytest.raises(LbmError, match="more than 32 segments"):
        s.advance(2)


def test_c1_2000_steps(lib, oracle_mod):
    """The 1024x512 two-wing case (chord 64, Re 200, seeded pitch/heave),
    2000 steps: rho/u/v within 1e-10 relative L2 of the FP64 oracle after
    1000 and after 2000 steps, and the per-wing forces and moments averaged
    over steps 1001-2000 within 1e-9 relative."""
    cfg = preset("c1")
    n = 2000
    rows = sinusoidal_motion_table(cfg, seeded_motions(cfg, 17), n + 2)
    g, o = pair(cfg, oracle_mod, rows)
    for n_last in (1000, n):
        print(f"c1 parity: advancing to step {n_last}", flush=True)
        rg, ro = g.advance(n_last, want_rows=True), o.advance(n_last, want_rows=True)
        errs = [rel(a, b) for a, b in zip(g.fields()[:3], o.fields()[:3])]
        print(f"c1 parity: step {n_last}, rho/u/v relative L2 {errs}", flush=True)
        assert max(errs) < FIELD_TOL, n_last
    W = cfg.wing_count
    avg_g = rg.force_rows[:, 1:1 + 3 * W + 3].mean(0)
    avg_o = ro.force_rows[:, 1:1 + 3 * W + 3].mean(0)
    ferr = np.abs(avg_g - avg_o) / np.maximum(np.abs(avg_o), 1e-3 * np.max(np.abs(avg_o)))
    print(f"c1 parity: averaged forces and moments, largest relative difference {ferr.max():.3e}", flush=True)
    assert np.all(ferr <= FORCE_TOL)


@pytest.mark.parametrize("name", ["c1"])
def test_c1_short_parity(lib, oracle_mod, name):
    cfg = preset(name)
    rows = sinusoidal_motion_table(cfg, seeded_motions(cfg, 13), 40)
    g, o = pair(cfg, oracle_mod, rows)
    rg, ro = g.advance(30, want_rows=True), o.advance(30, want_rows=True)
    for a, b in zip(g.fields()[:3], o.fields()[:3]):
        assert rel(a, b) < FIELD_TOL
    assert rel(rg.row_sum, ro.row_sum) < FORCE_TOL
