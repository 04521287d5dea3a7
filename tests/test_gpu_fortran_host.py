"""GPU tests of the Fortran host program (lbm_mutiwing_amd/lbm_host) over
liblbm_hip.so.

File mode: the host reads its namelist, the marker files and
motion_params.dat, runs the step loop and writes force.dat and
airfoil_motion.dat (main.f90:235-441, solution.f90:386-402).  Every record is
checked against the FP64 CPU oracle on the same inputs (tolerance: the
'(f15.8)' rounding of 5e-9 plus 1e-9 relative, the force parity bar), and
against the same run through the Python host and its formatter (identical
text: same library, same GPU, same edit descriptors).

Online mode: the host answers RESET / STEP / SAVE / STOP on a ZeroMQ REP
socket (driver.f90:59-196).  The replies are checked against the protocol
mirror (lbm_mutiwing_amd.protocol.Session) driven over the oracle (FORCE
values within 1e-9 relative, all other replies identical) and over the GPU
library (identical text), and the SAVE snapshot against the Python Tecplot
writer (identical file).
"""
import os
import subprocess
import uuid

import numpy as np
import pytest

from lbm_mutiwing_amd import MOTION_STREAM, Solver, build, preset, seeded_motions, sinusoidal_motion_table
from lbm_mutiwing_amd import formats
from lbm_mutiwing_amd.protocol import Session, run_job_alone

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-9
F15_8_HALF_ULP = 5.0e-9


@pytest.fixture(scope="module")
def host():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    if not os.path.exists(build.HOST):
        build.build_host()
    return build.HOST


def _namelist(cfg, out_dir="out", extra=""):
    ox, oy = cfg.origin()
    vals = dict(nx=cfg.nx, ny=cfg.ny, scale=cfg.scale, ox=ox, oy=oy, uu=cfg.uu, re=cfg.re,
                wing_count=cfg.wing_count, points_per_wing=cfg.points_per_wing, wing_x0=cfg.wing_x0,
                wing_dx=cfg.wing_dx, wing_y0=cfg.wing_y0, xc=cfg.xc, yc=cfg.yc, xp=cfg.xp, yp=cfg.yp,
                precision=cfg.precision)
    body = ", ".join(f"{k}={v!r}" for k, v in vals.items())
    s = ", ".join(repr(float(x)) for x in cfg.s)
    return f"&lbm_case {body}, s={s}, boundary_dir='.', motion_file='motion_params.dat', " \
           f"output_dir='{out_dir}' {extra} /\n"


def _write_case(tmp, cfg, rows=None):
    P = cfg.points_per_wing
    m = cfg.marker_array()
    for w in range(cfg.wing_count):
        name = "boundary.dat" if w == 0 else f"boundary{w}.dat"
        formats.write_marker_file(str(tmp / name), m[w * P:(w + 1) * P])
    if rows is not None:
        formats.write_motion_params(str(tmp / "motion_params.dat"), rows)
    (tmp / "case.nml").write_text(_namelist(cfg))


def _read_case(tmp, cfg):
    """The configuration as the host sees it: markers read back from the
    files ('(2f14.8)', exact for these 8-decimal markers)."""
    P = cfg.points_per_wing
    wm = [formats.read_marker_file(str(tmp / ("boundary.dat" if w == 0 else f"boundary{w}.dat")), P)
          for w in range(cfg.wing_count)]
    cfg.wing_markers = wm
    cfg.markers = None
    return cfg


def _read_f15(path, ncols):
    out = []
    for line in open(path):
        line = line.rstrip("\n")
        assert len(line) == 15 * ncols, line
        out.append([float(line[15 * k:15 * (k + 1)]) for k in range(ncols)])
    return np.array(out)


def test_file_mode_matches_oracle(host, oracle_mod, tmp_path):
    cfg = preset("small")
    n = 150
    rows = sinusoidal_motion_table(cfg, seeded_motions(cfg, 21), n)
    _write_case(tmp_path, cfg, rows)
    print(f"fortran host: file mode, small 256x128, {n} steps", flush=True)
    r = subprocess.run([host, "case.nml"], cwd=tmp_path, capture_output=True, timeout=300)
    assert r.returncode == 0, r.stdout.decode() + r.stderr.decode()
    cfg = _read_case(tmp_path, cfg)
    table = formats.load_motion_params(str(tmp_path / "motion_params.dat"), cfg.wing_count)
    fr = _read_f15(tmp_path / "out" / "force.dat", cfg.row_len)
    mr = _read_f15(tmp_path / "out" / "airfoil_motion.dat", 1 + 3 * cfg.wing_count)
    assert fr.shape[0] == n and mr.shape[0] == n

    o = oracle_mod.OracleSolver(cfg)
    o.set_motion_table(table)
    ro = o.advance(n, want_rows=True, want_motion=True)
    tol_f = F15_8_HALF_ULP * 1.0001 + FORCE_TOL * np.abs(ro.force_rows)
    assert np.all(np.abs(fr - ro.force_rows) <= tol_f)
    assert np.all(np.abs(mr - ro.motion_rows) <= F15_8_HALF_ULP * 1.0001 + 1e-12 * np.abs(ro.motion_rows))

    # the same run through the Python host writes the same text
    g = Solver(cfg)
    formats.run_file_mode(g, table, str(tmp_path / "py"), chunk=100)
    for name in ("force.dat", "airfoil_motion.dat"):
        assert (tmp_path / "out" / name).read_text() == (tmp_path / "py" / name).read_text(), name


def _commands(dt):
    t = lambda k: repr(k * dt)  # noqa: E731
    return ["STEP 0.1 0 0 0 0 0 0", "SAVE", "JUMP", "RESET 0.5", "RESET 0",
            f"STEP {t(40)} 0.01 0.002 -0.003 0.0 0.001 0.002",
            f"STEP {repr(40.5 * dt)} 0 0 0 0 0 0",
            f"STEP {t(40)} 0 0 0 0 0 0",
            f"STEP {t(20)} 0 0 0 0 0 0",
            "SAVE",
            f"STEP {t(90)} -0.01 0.003 0.001 0.002 0 0.001",
            "RESET",
            f"STEP {t(25)} 0.02 0 0 0 0 0",
            "STOP"]


def _mirror_replies(solver, commands, save_dir=None):
    hook = formats.tecplot_save_hook(save_dir) if save_dir else None
    sess = Session(solver, save_hook=hook)
    out = []
    for msg in commands:
        reply, job = sess.receive(msg)
        out.append(reply if job is None else run_job_alone(job))
    return out


def _force(reply):
    toks = reply.split()
    assert toks[0] == "FORCE" and len(toks) == 2 + int(toks[1])
    return np.array([float(x) for x in toks[2:]])


def test_online_mode_session(host, oracle_mod, tmp_path):
    from lbm_mutiwing_amd.build import zmq_library
    from zmq_client import ReqClient
    cfg = preset("small")
    cfg.motion_source = MOTION_STREAM
    _write_case(tmp_path, cfg)
    cfg = _read_case(tmp_path, cfg)
    cmds = _commands(cfg.dt)
    endpoint = f"ipc:///tmp/lbm_host_{uuid.uuid4().hex[:12]}.ipc"
    env = dict(os.environ, LBM_MOTION_SOURCE="ZMQ", LBM_ZMQ_ENDPOINT=endpoint)
    print(f"fortran host: online mode on {endpoint}", flush=True)
    log = open(tmp_path / "host.log", "w")
    proc = subprocess.Popen([host, "case.nml"], cwd=tmp_path, env=env, stdout=log, stderr=subprocess.STDOUT)
    client = ReqClient(zmq_library(), endpoint, timeout_ms=120000)
    try:
        got = []
        for msg in cmds:
            print(f"  -> {msg}", flush=True)
            got.append(client.request(msg))
        rc = proc.wait(timeout=60)
    finally:
        client.close()
        if proc.poll() is None:
            proc.kill()
            proc.wait()
        log.close()
    assert rc == 0, (tmp_path / "host.log").read_text()

    want_o = _mirror_replies(oracle_mod.OracleSolver(cfg), cmds)
    want_g = _mirror_replies(Solver(cfg), cmds, save_dir=str(tmp_path / "py_save"))
    assert [r.split()[0] for r in got] == [r.split()[0] for r in want_o]
    for msg, a, b, c in zip(cmds, got, want_o, want_g):
        if a.startswith("FORCE"):
            fa, fb = _force(a), _force(b)
            assert fa.shape == fb.shape, msg
            # relative per value, floored at 1e-3 of the row's largest value
            # (the first step's moments are rounding noise beside its forces)
            floor = 1e-3 * np.max(np.abs(fb))
            assert np.all(np.abs(fa - fb) <= FORCE_TOL * np.maximum(np.abs(fb), floor)), msg
        else:
            assert a == b, (msg, a, b)
        assert a == c, (msg, a, c)

    saved = [r for r in got if r.startswith("STATUS FLOW_SAVED")]
    assert saved == ["STATUS FLOW_SAVED F41.dat"]
    host_file = tmp_path / "out" / "episode_1" / "F41.dat"
    py_file = tmp_path / "py_save" / "F41.dat"
    assert host_file.read_text() == py_file.read_text()
    for ep in (1, 2):
        assert (tmp_path / "out" / f"force_episode_{ep}.dat").exists()


def test_file_mode_checkpoint_and_restart(host, tmp_path):
    """saveflowfield / readflowfield (solution.f90:1616-1851, main.f90:343-347,
    436-439): the checkpoint written at step 100 holds the state the library
    has after 100 steps, and a restarted run writes for steps 101..150 the
    same force and motion records as the uninterrupted run; with the main
    checkpoint damaged the restart falls back to the copy."""
    cfg = preset("small")
    n = 150
    rows = sinusoidal_motion_table(cfg, seeded_motions(cfg, 23), n)
    _write_case(tmp_path, cfg, rows)
    (tmp_path / "full.nml").write_text(_namelist(cfg, out_dir="full", extra=", checkpoint_every=100"))
    print("fortran host: file mode with a checkpoint at step 100", flush=True)
    r = subprocess.run([host, "full.nml"], cwd=tmp_path, capture_output=True, timeout=300)
    assert r.returncode == 0, r.stdout.decode() + r.stderr.decode()
    ckp = tmp_path / "full" / "saved_flow_feild.dat"
    assert ckp.read_bytes() == (tmp_path / "full" / "saved_flow_feild_copy.dat").read_bytes()

    # the checkpoint against the library's own state after 100 steps
    cfg = _read_case(tmp_path, cfg)
    table = formats.load_motion_params(str(tmp_path / "motion_params.dat"), cfg.wing_count)
    g = Solver(cfg)
    g.set_motion_table(table)
    f_init, _ = g.populations()
    g.advance(100)
    ck = formats.read_checkpoint(str(ckp), cfg.wing_count, cfg.nx, cfg.ny)
    _, u, v, _ = g.fields()
    f, f0 = g.populations()
    assert (ck.cycle, ck.movenum) == (100, 0)
    assert np.array_equal(ck.u, u) and np.array_equal(ck.v, v) and np.array_equal(ck.f, f)
    ring0 = formats.ring_from_populations(f_init)
    for name in ("left2", "right2"):
        assert np.array_equal(ck.ring[name][:, 1:-1], ring0[name][:, 1:-1]), name
    for name in ("up2", "down2", "f0_top", "f0_bottom"):
        assert np.array_equal(ck.ring[name], ring0[name]), name
    assert np.array_equal(ck.ring["f0_left"][:, 1:-1], f0[:, 1:-1, 0])
    assert np.array_equal(ck.ring["f0_right"][:, 1:-1], f0[:, 1:-1, -1])

    full_f = (tmp_path / "full" / "force.dat").read_text().splitlines()
    full_m = (tmp_path / "full" / "airfoil_motion.dat").read_text().splitlines()
    assert len(full_f) == n
    for name, damage in (("rs", False), ("rs_copy", True)):
        if damage:  # a truncated main checkpoint: the copy is read instead
            ckp.write_bytes(ckp.read_bytes()[:1000])
        (tmp_path / f"{name}.nml").write_text(
            _namelist(cfg, out_dir=name, extra=", checkpoint_every=100, restart=.true., restart_dir='full'"))
        print(f"fortran host: restart {name}", flush=True)
        r = subprocess.run([host, f"{name}.nml"], cwd=tmp_path, capture_output=True, timeout=300)
        out = r.stdout.decode()
        assert r.returncode == 0, out + r.stderr.decode()
        assert ("启动备用存档" in out) == damage, out
        assert (tmp_path / name / "force.dat").read_text().splitlines() == full_f[100:], name
        assert (tmp_path / name / "airfoil_motion.dat").read_text().splitlines() == full_m[100:], name


def test_file_mode_dynamics_checkpoint_and_restart(host, tmp_path):
    """Rigid-body coupling switched on from the namelist (the reference's
    solve_dynamic_equations and constants, main.f90:98, 122-151): the host's
    force.dat equals the library's rows with the same coupling, the
    checkpoint at step 60 holds the library's full body state
    (solution.f90:1663-1665), and a restart from it continues with the same
    records as the uninterrupted run."""
    from lbm_mutiwing_amd.solver import Dynamics
    cfg = preset("small")
    n = 100                                   # one checkpoint, at step 60
    rows = sinusoidal_motion_table(cfg, seeded_motions(cfg, 23), n)
    _write_case(tmp_path, cfg, rows)
    dyn_nml = ", solve_dynamic_equations=.true., solid_den=50.0, par1=-2.0, par2=50.0"
    (tmp_path / "full.nml").write_text(_namelist(cfg, out_dir="full", extra=dyn_nml + ", checkpoint_every=60"))
    print("fortran host: file mode with rigid-body coupling, checkpoint at step 60", flush=True)
    r = subprocess.run([host, "full.nml"], cwd=tmp_path, capture_output=True, timeout=300)
    assert r.returncode == 0, r.stdout.decode() + r.stderr.decode()

    cfg = _read_case(tmp_path, cfg)
    table = formats.load_motion_params(str(tmp_path / "motion_params.dat"), cfg.wing_count)
    g = Solver(cfg)
    g.set_motion_table(table)
    g.set_dynamics(Dynamics(solid_den=50.0, par1=-2.0, par2=50.0))
    r60 = g.advance(60, want_rows=True)
    ck = formats.read_checkpoint(str(tmp_path / "full" / "saved_flow_feild.dat"), cfg.wing_count, cfg.nx, cfg.ny)
    assert ck.cycle == 60
    assert np.array_equal(ck.body_rel, g.body_rel())
    assert ck.body_rel[2, 0] != 0.0 and ck.body_rel[4, 0] != 0.0     # v and s of wing 1 moved
    r_end = g.advance(n, want_rows=True)
    lib_rows = np.concatenate([r60.force_rows, r_end.force_rows])
    fr = _read_f15(tmp_path / "full" / "force.dat", cfg.row_len)
    assert np.all(np.abs(fr - lib_rows) <= F15_8_HALF_ULP * 1.01)

    full_f = (tmp_path / "full" / "force.dat").read_text().splitlines()
    (tmp_path / "rs.nml").write_text(
        _namelist(cfg, out_dir="rs", extra=dyn_nml + ", checkpoint_every=60, restart=.true., restart_dir='full'"))
    print("fortran host: restart with rigid-body coupling", flush=True)
    r = subprocess.run([host, "rs.nml"], cwd=tmp_path, capture_output=True, timeout=300)
    assert r.returncode == 0, r.stdout.decode() + r.stderr.decode()
    assert (tmp_path / "rs" / "force.dat").read_text().splitlines() == full_f[60:]


def test_python_file_mode_restarts_from_fortran_checkpoint(host, tmp_path):
    """The Python host's file mode (lbm_mutiwing_amd/file_mode.py) on the same
    case file: a full run writes the Fortran host's force.dat as text, and a
    restart from the checkpoint the Fortran host wrote at step 60 continues
    with the uninterrupted run's records (coupling on, so the body state
    travels through the file too)."""
    from lbm_mutiwing_amd import file_mode
    from lbm_mutiwing_amd.case import read_case
    cfg = preset("small")
    n = 100
    rows = sinusoidal_motion_table(cfg, seeded_motions(cfg, 31), n)
    _write_case(tmp_path, cfg, rows)
    dyn_nml = ", solve_dynamic_equations=.true., solid_den=50.0, par1=-2.0, par2=50.0, checkpoint_every=60"
    (tmp_path / "full.nml").write_text(_namelist(cfg, out_dir="full", extra=dyn_nml))
    print("fortran host: file mode with coupling, checkpoint at step 60", flush=True)
    r = subprocess.run([host, "full.nml"], cwd=tmp_path, capture_output=True, timeout=300)
    assert r.returncode == 0, r.stdout.decode() + r.stderr.decode()
    full_f = (tmp_path / "full" / "force.dat").read_text().splitlines()
    (tmp_path / "py.nml").write_text(_namelist(cfg, out_dir="py", extra=dyn_nml))
    (tmp_path / "pyrs.nml").write_text(
        _namelist(cfg, out_dir="pyrs", extra=dyn_nml + ", restart=.true., restart_dir='full'"))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        print("python host: the same case, then a restart from the Fortran checkpoint", flush=True)
        assert file_mode.run_case(read_case("py.nml")) == n
        assert file_mode.run_case(read_case("pyrs.nml")) == n
    finally:
        os.chdir(cwd)
    assert (tmp_path / "py" / "force.dat").read_text().splitlines() == full_f
    assert (tmp_path / "pyrs" / "force.dat").read_text().splitlines() == full_f[60:]
    assert (tmp_path / "py" / "saved_flow_feild.dat").read_bytes() == \
        (tmp_path / "full" / "saved_flow_feild.dat").read_bytes()
