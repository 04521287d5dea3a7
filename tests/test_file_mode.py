"""CPU tests of the Python host's file mode (lbm_mutiwing_amd/file_mode.py),
driven over the oracle through the same host interface: the checkpoint
cadence, the restart from the checkpoint and from its copy, and the rigid-
body state carried through the file."""
import os

import numpy as np

from lbm_mutiwing_amd import SolverConfig, formats, naca0012_markers
from lbm_mutiwing_amd.case import read_case
from lbm_mutiwing_amd.config import sinusoidal_motion_table, WingMotion
from lbm_mutiwing_amd import file_mode


def _case(tmp_path, extra=""):
    s = (0.0, 1.4, 1.4, 0.0, 1.2, 0.0, 1.2, 0.0, 0.0)
    cfg = SolverConfig(nx=160, ny=80, scale=1 / 20, uu=0.05, re=100.0, s=s, wing_count=1,
                       markers=naca0012_markers(20), ox=-1.5, oy=0.5 - 0.5 * 79 / 20, name="fm")
    P = cfg.points_per_wing
    formats.write_marker_file(str(tmp_path / "boundary.dat"), cfg.marker_array()[:P])
    rows = sinusoidal_motion_table(cfg, [WingMotion(pitch=0.1, heave=0.05)], 40)
    formats.write_motion_params(str(tmp_path / "motion_params.dat"), rows)
    ox, oy = cfg.origin()
    body = (f"&lbm_case nx=160, ny=80, scale=0.05, ox={ox!r}, oy={oy!r}, uu=0.05, re=100.0, "
            f"s=0.0, 1.4, 1.4, 0.0, 1.2, 0.0, 1.2, 0.0, 0.0, wing_count=1, points_per_wing={P}, "
            f"wing_x0=0.0, wing_y0=0.5, rows_per_call=7, print_every=0")
    return body, extra


def _run(tmp_path, oracle_mod, name, text):
    (tmp_path / f"{name}.nml").write_text(text + f", output_dir='{name}' /\n")
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        return file_mode.run_case(read_case(f"{name}.nml"), solver_cls=oracle_mod.OracleSolver)
    finally:
        os.chdir(cwd)


def test_file_mode_checkpoint_and_restart_over_the_oracle(tmp_path, oracle_mod):
    body, _ = _case(tmp_path)
    dyn = ", solve_dynamic_equations=.true., solid_den=40.0, checkpoint_every=25"
    n = _run(tmp_path, oracle_mod, "full", body + dyn)
    full = (tmp_path / "full" / "force.dat").read_text().splitlines()
    assert len(full) == n and n > 25
    ck = formats.read_checkpoint(str(tmp_path / "full" / "saved_flow_feild.dat"), 1, 160, 80)
    assert ck.cycle == 25 and ck.body_rel[2, 0] != 0.0          # the body moved by then
    assert (tmp_path / "full" / "saved_flow_feild.dat").read_bytes() == \
        (tmp_path / "full" / "saved_flow_feild_copy.dat").read_bytes()
    _run(tmp_path, oracle_mod, "rs", body + dyn + ", restart=.true., restart_dir='full'")
    assert (tmp_path / "rs" / "force.dat").read_text().splitlines() == full[25:]
    # a damaged main checkpoint: the copy is read
    p = tmp_path / "full" / "saved_flow_feild.dat"
    p.write_bytes(p.read_bytes()[:100])
    _run(tmp_path, oracle_mod, "rs2", body + dyn + ", restart=.true., restart_dir='full'")
    assert (tmp_path / "rs2" / "force.dat").read_text().splitlines() == full[25:]
    m = np.loadtxt(tmp_path / "full" / "airfoil_motion.dat")
    assert m.shape == (n, 4)
