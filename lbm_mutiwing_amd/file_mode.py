"""File mode of the Python host: the offline run of a case file, as the
Fortran host (lbm_mutiwing_amd/fortran/lbm_host.f90) runs it.

    python -m lbm_mutiwing_amd.file_mode case.nml

The reference's offline program (main.f90:235-441) reads its compile-time
case, the marker files and motion_params.dat, runs the step loop while
(n-1)*dt <= motion_end_time and writes force.dat and airfoil_motion.dat
('(1000f15.8)', solution.f90:386-402); every checkpoint_every steps it writes
the restart checkpoint saved_flow_feild.dat and its copy
(solution.f90:1616-1719), and with restart on it continues from them
(solution.f90:1721-1851, main.f90:284-347).  This module does the same from
the case namelist /lbm_case/ (lbm_mutiwing_amd/case.py), over the HIP
library, with the rigid-body coupling when the case switches it on.  Its
checkpoints are the Fortran host's, byte for byte, so either host restarts
from the other's.
"""
from __future__ import annotations

import argparse
import os
import sys
from typing import Optional, Tuple

import numpy as np

from . import formats
from .case import case_config, case_dynamics, read_case
from .config import MOTION_FILE
from .solver import Solver

CHECKPOINT_NAMES = ("saved_flow_feild", "saved_flow_feild_copy")


def save_checkpoint(solver: Solver, out_dir: str, n: int, ring: dict) -> None:
    """saveflowfield: the state after step n, written twice."""
    st = solver.state()
    _, u, v, _ = solver.fields()
    f, f0 = solver.populations()
    ring["f0_left"] = f0[:, :, 0].copy()      # the f0 ring as the library holds it (move_mesh shifts it)
    ring["f0_right"] = f0[:, :, -1].copy()
    ring["f0_bottom"] = f0[:, 0, :].copy()
    ring["f0_top"] = f0[:, -1, :].copy()
    W = solver.cfg.wing_count
    rel = solver.body_rel() if W else np.zeros((11, 0))
    ck = formats.Checkpoint(cycle=n, movenum=st["movenum"], body_rel=rel, u=u, v=v, f=f, ring=ring)
    for name in CHECKPOINT_NAMES:
        formats.write_checkpoint(os.path.join(out_dir, name + ".dat"), ck)


def restore_checkpoint(solver: Solver, ck_dir: str) -> Tuple[dict, int]:
    """readflowfield: the main checkpoint, its copy if that fails; returns
    the ring records and the next step."""
    cfg = solver.cfg
    try:
        ck = formats.read_checkpoint(os.path.join(ck_dir, CHECKPOINT_NAMES[0] + ".dat"), cfg.wing_count, cfg.nx,
                                     cfg.ny)
        print("sucess!", flush=True)
    except (OSError, ValueError) as e:
        print(str(e), flush=True)
        print("启动备用存档！", flush=True)
        ck = formats.read_checkpoint(os.path.join(ck_dir, CHECKPOINT_NAMES[1] + ".dat"), cfg.wing_count, cfg.nx,
                                     cfg.ny)
        print("读取保存的流场数据成功!", flush=True)
    f0 = np.zeros_like(ck.f)
    f0[:, :, 0] = ck.ring["f0_left"]
    f0[:, :, -1] = ck.ring["f0_right"]
    f0[:, 0, :] = ck.ring["f0_bottom"]
    f0[:, -1, :] = ck.ring["f0_top"]
    solver.set_step(ck.cycle + 1, ck.movenum)
    if cfg.wing_count:
        solver.set_body_rel(ck.body_rel)
    solver.set_populations(ck.f, f0)
    return ck.ring, ck.cycle + 1


def run_case(case: dict, lib=None, solver_cls=Solver) -> int:
    """Run the case to the end of its motion table; returns the last step.
    ``solver_cls`` is the host class bound to the library (the HIP one
    unless a test drives another implementation of the same interface)."""
    cfg = case_config(case, MOTION_FILE)
    solver = solver_cls(cfg) if lib is None else solver_cls(cfg, lib)
    try:
        dyn = case_dynamics(case)
        if dyn is not None:
            solver.set_dynamics(dyn)
        table = formats.load_motion_params(str(case["motion_file"]), cfg.wing_count)
        end_time = solver.set_motion_table(table)
        n_end = cfg.n_steps_for(end_time)
        out_dir = str(case["output_dir"])
        os.makedirs(out_dir, exist_ok=True)
        every = int(case["checkpoint_every"])
        if case["restart"]:
            ring, n_next = restore_checkpoint(solver, str(case["restart_dir"]) or out_dir)
        else:
            ring, n_next = formats.ring_from_populations(solver.populations()[0]), 1
        ff = formats.RowFile(os.path.join(out_dir, "force.dat"))
        fm = formats.RowFile(os.path.join(out_dir, "airfoil_motion.dat"))
        try:
            while n_next <= n_end:
                n_last = min(n_end, n_next + int(case["rows_per_call"]) - 1)
                if every > 0:
                    n_last = min(n_last, ((n_next - 1) // every + 1) * every)
                r = solver.advance(n_last, want_rows=True, want_motion=True)
                ff.write_rows(r.force_rows)
                fm.write_rows(r.motion_rows)
                pe = int(case["print_every"])
                if pe > 0 and n_last // pe > (n_next - 1) // pe:
                    print(f"step {n_last}  err {r.err:12.5E}", flush=True)
                if every > 0 and n_last % every == 0:
                    save_checkpoint(solver, out_dir, n_last, ring)
                n_next = n_last + 1
        finally:
            ff.close()
            fm.close()
        print(f"file mode done: {n_end} steps", flush=True)
        return n_end
    finally:
        solver.close()


def main(argv: Optional[list] = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("case", nargs="?", default="lbm_case.nml", help="namelist file with the group /lbm_case/")
    args = ap.parse_args(argv)
    run_case(read_case(args.case))
    return 0


if __name__ == "__main__":
    sys.exit(main())
