"""CPU tests of the online protocol mirror (zmq_comm.f90 parsing and reply
layout, driver.f90 RESET/STEP/SAVE/STOP state machine and FORCE averaging),
driven with the FP64 CPU oracle as the solver."""
import numpy as np
import pytest

from lbm_mutiwing_amd import MOTION_STREAM, preset
from lbm_mutiwing_amd.protocol import (CMD_INVALID, CMD_OUTPUT, CMD_RESET, CMD_STEP, CMD_STOP, ERR_ALIGN, ERR_BEHIND,
                                       ERR_RESET_NONZERO, ERR_SEGMENT_START, ERR_STEP_ARGS, ERR_UNKNOWN, Session, fortran_es24_16,
                                       parse_command, run_job_alone)


def test_es24_16_layout():
    assert fortran_es24_16(1.0) == "1.0000000000000000E+00"
    assert fortran_es24_16(-0.375) == "-3.7500000000000000E-01"
    assert fortran_es24_16(-2.5e-7) == "-2.4999999999999999E-07"   # correctly rounded binary value
    assert fortran_es24_16(0.0) == "0.0000000000000000E+00"
    assert fortran_es24_16(123456.789) == "1.2345678900000000E+05"
    assert fortran_es24_16(1e100) == "1.0000000000000000+100"
    assert fortran_es24_16(-2.0 ** -500) == "-3.0549363634996047-151"
    # 24-character field in Fortran: sign + 18 mantissa characters + 4 exponent
    assert len(" " + fortran_es24_16(1.0)) <= 24


def test_parse_commands():
    W = 2
    c, r, e = parse_command("RESET", W)
    assert c == CMD_RESET and r[0] == 0.0
    c, r, e = parse_command("start 0.0", W)
    assert c == CMD_RESET and r[0] == 0.0
    c, r, e = parse_command("reset 1.5d0", W)
    assert c == CMD_RESET and r[0] == 1.5
    c, r, e = parse_command("STEP 0.05 0.1 0 0, 0 0 0", W)
    assert c == CMD_STEP and r[0] == 0.05 and r[1] == 0.1 and r[6] == 0.0
    c, r, e = parse_command("STEP 0.05 0.1 0 0", W)
    assert c == CMD_INVALID and e == ERR_STEP_ARGS
    assert parse_command("save", W)[0] == CMD_OUTPUT
    assert parse_command("OUTPUT", W)[0] == CMD_OUTPUT
    assert parse_command("exit", W)[0] == CMD_STOP
    c, r, e = parse_command("JUMP 1", W)
    assert c == CMD_INVALID and e == ERR_UNKNOWN + "JUMP"


@pytest.fixture()
def session(oracle_mod):
    cfg = preset("small")
    cfg.motion_source = MOTION_STREAM
    return Session(oracle_mod.OracleSolver(cfg))


def force_values(reply):
    toks = reply.split()
    assert toks[0] == "FORCE"
    n = int(toks[1])
    assert len(toks) == 2 + n
    return np.array([float(t) for t in toks[2:]])


def test_session_state_machine(session, oracle_mod):
    s = session
    cfg = s.cfg
    dt = cfg.dt
    assert s.receive("STEP 0.1 0 0 0 0 0 0")[0] == "ERROR online mode requires RESET before STEP"
    assert s.receive("SAVE")[0] == "ERROR online mode requires RESET before SAVE"
    assert s.receive("RESET 1.0")[0] == "ERROR " + ERR_RESET_NONZERO
    reply, job = s.receive("RESET 0.0")
    assert reply is None and job.target_cycle == 1
    f0 = force_values(run_job_alone(job))
    assert f0.size == cfg.row_len and f0[0] == 0.0
    # a misaligned STEP is refused and leaves the segment unchanged
    assert s.receive(f"STEP {dt * 10.5!r} 0 0 0 0 0 0")[0] == "ERROR " + ERR_ALIGN
    t1 = 40 * dt
    reply, job = s.receive(f"STEP {t1!r} 0.02 0.01 -0.01 0 0.005 0")
    assert reply is None and job.target_cycle == 41
    vals = force_values(run_job_alone(job, chunk=7))
    # same segment on a second oracle instance, force rows kept
    ref = oracle_mod.OracleSolver(cfg)
    ref.reset()
    z = np.zeros(cfg.motion_cols)
    ref.set_stream_segment(z, z)
    ref.advance(1)
    tgt = ref.build_stream_target_row(z, t1, np.array([t1, 0.02, 0.01, -0.01, 0, 0.005, 0]))
    ref.set_stream_segment(z, tgt)
    rr = ref.advance(41, want_rows=True)
    W = cfg.wing_count
    expect = rr.force_rows[-1].copy()
    expect[1:1 + 3 * W + 3] = rr.force_rows[:, 1:1 + 3 * W + 3].mean(0)
    expect[0] = t1
    assert np.allclose(vals, expect, rtol=1e-14, atol=1e-15)
    # a STEP before the segment start, or not past the last step run, is refused
    assert s.receive(f"STEP {20 * dt!r} 0 0 0 0 0 0")[0] == "ERROR " + ERR_SEGMENT_START
    assert s.receive(f"STEP {t1!r} 0 0 0 0 0 0")[0] == "ERROR " + ERR_BEHIND
    assert s.receive("SAVE")[0] == "STATUS FLOW_SAVED F41.dat"
    assert s.receive("STOP")[0] == "STATUS STOPPING"
