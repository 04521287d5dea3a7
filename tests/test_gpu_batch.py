"""Batched advance (lbm_batch_advance) against the same instances advanced
alone: every instance must end bitwise in the state lbm_advance leaves it in
(same kernels, same arithmetic, independent per instance)."""
import numpy as np
import pytest

from lbm_mutiwing_amd import (PRECISION_FP32_SHIFTED, PRECISION_FP64, Batch, LbmError, Solver, episode_config,
                              seeded_motions, sinusoidal_motion_table)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from lbm_mutiwing_amd._native import load_library
    return load_library()


SPECS = [  # (nx, wings, chord, Re, seed)
    (256, 2, 32, 100.0, 1),
    (384, 1, 40, 250.0, 2),
    (320, 0, 32, 150.0, 3),
    (512, 3, 32, 400.0, 4),
]


def make(spec, precision, n_table):
    nx, wings, chord, re, seed = spec
    cfg = episode_config(nx, wings, chord, re, precision)
    s = Solver(cfg)
    if wings:
        s.set_motion_table(sinusoidal_motion_table(cfg, seeded_motions(cfg, seed), n_table))
    return s


def snapshot(s):
    f, _ = s.populations()
    return f, s.fields()[:3]


@pytest.mark.parametrize("precision", [PRECISION_FP64, PRECISION_FP32_SHIFTED])
def test_batch_equals_alone(lib, precision):
    pre = [0, 7, 3, 12]          # instances start at different steps
    steps = 45                   # crosses several geometry batches and a tail
    alone = [make(sp, precision, 200) for sp in SPECS]
    mixed = [make(sp, precision, 200) for sp in SPECS]
    for s, p in zip(alone + mixed, pre + pre):
        if p:
            s.advance(p)
    ra = [s.advance(s.state()["n_next"] + steps - 1) for s in alone]
    b = Batch(lib)
    rb = b.advance(mixed, steps)
    for e, (sa, sb) in enumerate(zip(alone, mixed)):
        assert sa.state() == sb.state(), e
        fa, qa = snapshot(sa)
        fb, qb = snapshot(sb)
        assert np.array_equal(fa, fb), e
        for x, y in zip(qa, qb):
            assert np.array_equal(x, y), e
        assert np.array_equal(ra[e].row_sum, rb[e].row_sum), e
        assert np.array_equal(ra[e].last_row, rb[e].last_row), e
        assert ra[e].err == rb[e].err, e
        if SPECS[e][1]:
            ia, ib = sa.ibm(), sb.ibm()
            assert np.array_equal(ia[0], ib[0]) and np.array_equal(ia[2], ib[2])
    # a second batched call continues from where the first left off
    ra2 = [s.advance(s.state()["n_next"] + 9) for s in alone]
    rb2 = b.advance(mixed, 10)
    for e in range(len(SPECS)):
        assert np.array_equal(ra2[e].last_row, rb2[e].last_row), e
    b.close()


@pytest.mark.parametrize("precision", [PRECISION_FP64, PRECISION_FP32_SHIFTED])
def test_batch_equals_alone_band_substitution(lib, precision, monkeypatch):
    """Instances that solve every step's IBM system by band substitution
    (LBM_DIRECT_SOLVE=2, no explicit inverse) end bitwise where they end
    alone when advanced in a batch: the batched flow solves the same way."""
    monkeypatch.setenv("LBM_DIRECT_SOLVE", "2")
    steps = 40
    alone = [make(sp, precision, 200) for sp in SPECS]
    mixed = [make(sp, precision, 200) for sp in SPECS]
    ra = [s.advance(steps) for s in alone]
    b = Batch(lib)
    rb = b.advance(mixed, steps)
    for e, (sa, sb) in enumerate(zip(alone, mixed)):
        fa, qa = snapshot(sa)
        fb, qb = snapshot(sb)
        assert np.array_equal(fa, fb), e
        assert np.array_equal(ra[e].row_sum, rb[e].row_sum), e
    b.close()


@pytest.mark.parametrize("tile,rows,waves", [(128, 4, 16), (256, 2, 0), (512, 4, 24)])
def test_batch_tiles_per_workgroup(lib, tile, rows, waves, monkeypatch):
    """Work items of several lattice tiles (crossing row ends: 2-4 tiles per
    row here), with and without a cap on lattice waves per CU, give the same
    bits as the instances run alone."""
    monkeypatch.setenv("LBM_LATTICE_TILE", str(tile))
    monkeypatch.setenv("LBM_BATCH_ROWS", str(rows))
    monkeypatch.setenv("LBM_LATTICE_WAVES", str(waves))
    steps = 20
    alone = [make(sp, PRECISION_FP32_SHIFTED, 100) for sp in SPECS]
    mixed = [make(sp, PRECISION_FP32_SHIFTED, 100) for sp in SPECS]
    ra = [s.advance(steps) for s in alone]
    b = Batch(lib)
    rb = b.advance(mixed, steps)
    assert b.tuning_full() == (tile, rows, waves)
    for e, (sa, sb) in enumerate(zip(alone, mixed)):
        assert np.array_equal(sa.populations()[0], sb.populations()[0]), e
        assert np.array_equal(ra[e].last_row, rb[e].last_row), e
    b.close()


def test_batch_rejects_mixed_precision(lib):
    a = make(SPECS[0], PRECISION_FP64, 50)
    c = make(SPECS[1], PRECISION_FP32_SHIFTED, 50)
    b = Batch(lib)
    with pytest.raises(LbmError):
        b.advance([a, c], 5)
    with pytest.raises(LbmError):
        b.advance([a, a], 5)
    b.close()


def test_batch_pending_move_is_rejected(lib):
    a = make(SPECS[0], PRECISION_FP64, 50)
    a.set_body_rel_pos(1, 1.0)
    b = Batch(lib)
    with pytest.raises(LbmError):
        b.advance([a], 5)
    b.close()


@pytest.mark.parametrize("precision,steps,specs", [
    (PRECISION_FP64, 12, SPECS), (PRECISION_FP32_SHIFTED, 12, SPECS),   # captured whole
    (PRECISION_FP64, 100, SPECS), (PRECISION_FP32_SHIFTED, 100, SPECS),  # batches 1-4 from one graph
    (PRECISION_FP64, 70, SPECS[2:3]),                                   # no wings: 16-step periods
])
def test_batch_graph_replays(lib, precision, steps, specs, monkeypatch):
    """Same composition and step count four times, graph mode 2.  A short
    advance runs eagerly first, is captured on the second call and replayed on
    the third and fourth; a long one replays the graph of one geometry batch
    for its middle batches from the first call on.  Each call is checked against the
    instances run alone, with eager work and synchronisations of the host in
    between."""
    n_table = 4 * steps + 50
    alone = [make(sp, precision, n_table) for sp in specs]
    mixed = [make(sp, precision, n_table) for sp in specs]
    monkeypatch.setenv("LBM_BATCH_GRAPHS", "2")   # read when the batch object is created
    b = Batch(lib)
    for call in range(4):
        print(f"graph replay: precision {precision} steps {steps} instances {len(specs)} call {call}", flush=True)
        ra = [s.advance(s.state()["n_next"] + steps - 1) for s in alone]
        rb = b.advance(mixed, steps)
        for e, (sa, sb) in enumerate(zip(alone, mixed)):
            assert sa.state() == sb.state(), (call, e)
            assert np.array_equal(ra[e].last_row, rb[e].last_row), (call, e)
            assert np.array_equal(ra[e].row_sum, rb[e].row_sum), (call, e)
            assert ra[e].err == rb[e].err, (call, e)
            assert np.array_equal(sa.populations()[0], sb.populations()[0]), (call, e)
            for x, y in zip(sa.fields()[:3], sb.fields()[:3]):
                assert np.array_equal(x, y), (call, e)
    b.close()
