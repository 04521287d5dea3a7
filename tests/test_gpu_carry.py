"""GPU tests of the geometry carried from one lbm_advance call to the next
(the rest of a call's last geometry batch, computed after its own steps and
used by the next call when nothing it depends on has changed).

Every result must be bitwise the one of an instance with the carry switched
off (LBM_GEO_CARRY=0) or of a single call over the same steps, whatever the
host does between the calls; a geometry error in a step after the call's
last step must not fail the call.
"""
import numpy as np
import pytest

from lbm_mutiwing_amd import LbmError, Solver, WingMotion, preset, seeded_motions, sinusoidal_motion_table

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from lbm_mutiwing_amd._native import load_library
    return load_library()


def _solver(cfg, carry, monkeypatch):
    monkeypatch.setenv("LBM_GEO_CARRY", "1" if carry else "0")
    return Solver(cfg)


def test_split_calls_equal_one_call(lib, monkeypatch):
    cfg = preset("small")
    rows = sinusoidal_motion_table(cfg, seeded_motions(cfg, 31), 160)
    a = _solver(cfg, True, monkeypatch)
    b = _solver(cfg, True, monkeypatch)
    for s in (a, b):
        s.set_motion_table(rows)
    parts = [a.advance(n, want_rows=True, want_motion=True) for n in (5, 21, 22, 37, 48, 70, 117)]
    whole = b.advance(117, want_rows=True, want_motion=True)
    assert np.array_equal(np.concatenate([p.force_rows for p in parts]), whole.force_rows)
    assert np.array_equal(np.concatenate([p.motion_rows for p in parts]), whole.motion_rows)
    assert parts[-1].err == whole.err
    assert np.array_equal(a.populations()[0], b.populations()[0])
    for x, y in zip(a.markers(), b.markers()):
        assert np.array_equal(x, y)
    ia, ib = a.ibm(), b.ibm()
    assert np.array_equal(ia[0], ib[0]) and np.array_equal(ia[1], ib[1])


def test_carry_on_and_off_bitwise_through_host_changes(lib, monkeypatch):
    """Motion table replaced, a move_mesh requested, reset and a restart
    between calls: the carry must be dropped or used exactly when valid."""
    cfg = preset("small")
    rows1 = sinusoidal_motion_table(cfg, seeded_motions(cfg, 3), 200)
    rows2 = sinusoidal_motion_table(cfg, seeded_motions(cfg, 4), 200)
    out = []
    for carry in (True, False):
        s = _solver(cfg, carry, monkeypatch)
        s.set_motion_table(rows1)
        rec = [s.advance(7, want_rows=True).force_rows]
        s.set_motion_table(rows2)
        rec.append(s.advance(30, want_rows=True).force_rows)
        s.set_body_rel_pos(1, 1.0)                  # move_mesh after the next step
        rec.append(s.advance(45, want_rows=True).force_rows)
        rec.append(s.advance(50, want_rows=True).force_rows)
        f, f0 = s.populations()
        s.reset()
        s.set_motion_table(rows1)
        rec.append(s.advance(12, want_rows=True).force_rows)
        s.set_step(40, 0)                           # restart at another step
        s.set_populations(f, f0)
        rec.append(s.advance(61, want_rows=True).force_rows)
        out.append((rec, s.populations()[0], s.state()))
        s.close()
    (ra, pa, sa), (rb, pb, sb) = out
    assert sa == sb and sa["movenum"] == 0
    for x, y in zip(ra, rb):
        assert np.array_equal(x, y)
    assert np.array_equal(pa, pb)


def test_stream_chunks_equal_one_call(lib, monkeypatch):
    """Stream mode as the episode server drives it: STEPs of 300 steps in
    chunks of 100, against one call per STEP."""
    cfg = preset("small")
    cfg.motion_source = 2
    W = cfg.wing_count
    a = _solver(cfg, True, monkeypatch)
    b = _solver(cfg, True, monkeypatch)
    start = np.zeros(cfg.motion_cols)
    for s in (a, b):
        s.set_stream_segment(start, start)
        s.advance(1)
    rng = np.random.default_rng(7)
    t, n = 0.0, 1
    for seg in range(3):
        t_new = t + 300 * cfg.dt
        delta = np.concatenate([[t_new], rng.normal(0, 0.1, 3 * W)])
        tgt = a.build_stream_target_row(start, t_new, delta)
        for s in (a, b):
            s.set_stream_segment(start, tgt)
        n_last = int(round(t_new / cfg.dt)) + 1
        ra = [a.advance(min(n + c, n_last), want_rows=True).force_rows for c in range(100, 301, 100)]
        rb = b.advance(n_last, want_rows=True).force_rows
        assert np.array_equal(np.concatenate(ra), rb)
        start, t, n = tgt, t_new, n_last
    assert np.array_equal(a.populations()[0], b.populations()[0])


def test_geometry_error_after_the_call_is_raised_at_its_step(lib, oracle_mod, monkeypatch):
    """A wing that reaches the lattice edge at step n_fail: calls that end
    before it succeed (the carry past them has the error), the call that
    runs it fails, as the oracle's step loop does."""
    cfg = preset("small")
    n_tab = 120
    rows = sinusoidal_motion_table(cfg, [WingMotion(), WingMotion()], n_tab)
    ox, _ = cfg.origin()
    rows[:, 4] = np.linspace(0.0, ox + 0.1, n_tab)   # wing 1 drifts into the left edge
    o = oracle_mod.OracleSolver(cfg)
    o.set_motion_table(rows)
    n_fail = None
    for n in range(1, n_tab):
        try:
            o.advance(n)
        except LbmError:
            n_fail = n
            break
    assert n_fail is not None and n_fail > 20
    s = _solver(cfg, True, monkeypatch)
    s.set_motion_table(rows)
    s.advance(n_fail - 15)
    s.advance(n_fail - 1)          # its carry covers n_fail
    with pytest.raises(LbmError):
        s.advance(n_fail)
