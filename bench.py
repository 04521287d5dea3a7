#!/usr/bin/env python3
"""Benchmark of the D2Q9 MRT + IBM multi-airfoil time step on one MI355X.

Workload (default): single instance, FP64, 4096 x 2048 lattice, chord 256
nodes, Re 1000 (tau 0.5384), U 0.05, two NACA0012 wings with 511 markers
each, seeded sinusoidal pitch (5-15 deg) and heave (0.1-0.3 c) at
f = U/(4c), file-mode motion table.  One step = erranalysis + motion + marker
update + adjoint build + IBM solve + fused stream/collide + boundary ring +
force row, i.e. the reference's whole run-loop body.

Prints one JSON line: value = whole-job MLUPS (Nx*Ny*steps / s / 1e6) with
all data resident in HBM, the roofline of the fused lattice kernel measured
with HIP events on the library's stream during the timed steps, and a CPU
baseline (the FP64 oracle on the host cores, a bounded sample of the same
workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_NODE = {"fp64": 144.0,   # 9 population reads + 9 writes of 8 B
                  "fp32": 72.0}    # the same in FP32 deviations


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c5", "small"])
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"],
                    help="population storage: FP64, or FP32 deviations from the weights")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="after the timed steps, copy a device buffer of this size 3 times "
                         "(calibration of rocprofv3 FETCH_SIZE/WRITE_SIZE against known bytes)")
    return ap.parse_args()


def workload(name, seed, precision="fp64"):
    from lbm_mutiwing_amd import PRECISION_FP32_SHIFTED, PRECISION_FP64, preset, seeded_motions
    cfg = preset(name, precision=PRECISION_FP32_SHIFTED if precision == "fp32" else PRECISION_FP64)
    if name == "c2":
        mot = seeded_motions(cfg, seed, pitch_deg=(5.0, 15.0), heave=(0.1, 0.3))
    else:
        mot = seeded_motions(cfg, seed, pitch_deg=(10.0, 10.0), heave=(0.25, 0.25))
    return cfg, mot


def cpu_baseline(name, seed, budget_s):
    """The FP64 oracle (C restatement of the reference CPU path, OpenMP over
    the host cores) on the first steps of the same workload."""
    from lbm_mutiwing_amd import sinusoidal_motion_table
    from oracle import oracle
    cores = min(16, os.cpu_count() or 1)
    oracle.set_threads(cores)
    cfg, mot = workload(name, seed)
    o = oracle.OracleSolver(cfg)
    o.set_motion_table(sinusoidal_motion_table(cfg, mot, 64))
    o.advance(1)  # first step outside the timing (page faults of fresh buffers)
    steps, t = 0, 0.0
    while t < budget_s and steps < 50:
        t0 = time.perf_counter()
        o.advance(2 + steps)
        t += time.perf_counter() - t0
        steps += 1
    o.close()
    return {"value": cfg.nx * cfg.ny * steps / t / 1e6, "unit": "MLUPS", "cores": cores,
            "kind": "port", "sample": f"{name} steps 2..{1 + steps} ({steps} steps, {t:.1f} s)"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(episodes are independent and shard one instance per GPU)"}))
        sys.exit(2)
    import torch  # noqa: F401  (device visibility check only)
    from lbm_mutiwing_amd import sinusoidal_motion_table
    from lbm_mutiwing_amd.solver import TimedSolver

    cfg, mot = workload(args.config, args.seed, args.precision)
    bpn = BYTES_PER_NODE[args.precision]
    n_total = args.warmup + args.steps + 1
    s = TimedSolver(cfg)
    s.set_motion_table(sinusoidal_motion_table(cfg, mot, n_total + 1))
    s.advance(args.warmup)
    s.set_timing(True)
    t0 = time.perf_counter()
    r = s.advance(args.warmup + args.steps)
    dt = time.perf_counter() - t0
    lat_ms, launches = s.timing()
    st = s.state()
    nodes = cfg.nx * cfg.ny
    mlups = nodes * args.steps / dt / 1e6
    ms_step = dt * 1e3 / args.steps
    lat_avg_ms = lat_ms / max(launches, 1)
    interior = (cfg.nx - 2) * (cfg.ny - 2)
    achieved = bpn * interior / (lat_avg_ms * 1e-3) / 1e9
    traffic = None
    tp = os.path.join(ROOT, "profiles", f"traffic_{args.config}_{args.precision}.json")
    if os.path.exists(tp):
        with open(tp) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    out = {
        "metric": f"MLUPS (whole time step incl. IBM), {args.precision.upper()} single instance",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded sinusoidal pitch/heave motion table, generated NACA0012 markers)",
        "config": {"workload": f"{args.config}: {cfg.nx}x{cfg.ny} D2Q9 MRT + IBM, chord {1 / cfg.scale:.0f} nodes, "
                               f"Re {cfg.re:g}, {cfg.wing_count} wings x {cfg.points_per_wing} markers",
                   "nx": cfg.nx, "ny": cfg.ny, "wings": cfg.wing_count, "seed": args.seed,
                   "adjoint_nodes": st["adjoint_n"]},
        "hbm_roofline_frac_whole_step": round(mlups * 1e6 * bpn / (HBM_PEAK_GBS * 1e9), 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_lattice (fused pull-stream + moments + MRT collide)",
                     "kernel_avg_us": round(lat_avg_ms * 1e3, 2),
                     "algorithmic_bytes_per_launch": bpn * interior},
        "last_force_row": [round(x, 8) for x in r.last_row[:10]],
    }
    s.close()
    if args.copy_mb > 0:
        n = args.copy_mb * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        out["copy_calibration_bytes"] = 2 * n * 8
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.config, args.seed, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
