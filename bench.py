#!/usr/bin/env python3
"""Benchmark of the D2Q9 MRT + IBM multi-airfoil time step on one MI355X.

Workload (default): single instance, FP64, 4096 x 2048 lattice, chord 256
nodes, Re 1000 (tau 0.5384), U 0.05, two NACA0012 wings with 511 markers
each, seeded sinusoidal pitch (5-15 deg) and heave (0.1-0.3 c) at
f = U/(4c), file-mode motion table.  One step = motion + marker update +
adjoint build + IBM solve + fused stream/collide + boundary ring + force row,
i.e. the reference's run-loop body; the timed call of K steps is one
lbm_advance, which also computes erranalysis of its last step (the value the
reference prints, main.f90:418) and the move_mesh checks.  The geometry of a
call's first steps may have been computed by the previous call (the rest of
its last geometry batch); each call computes the rest of its own last batch
for the next one.

Prints one JSON line: value = whole-job MLUPS (Nx*Ny*steps / s / 1e6) with
all data resident in HBM, the roofline of the fused lattice kernel measured
with HIP events on the library's stream during the timed steps, the ceiling
of that kernel's memory pattern, the FP64 mass before and after the timed
steps, a CPU baseline (the FP64 oracle on the host cores, a bounded sample
of the same workload; beside it the same on one core, and the 1024x512
workload on all cores and on one) with the GPU path's error against it, and
two sub-results of the same run: "batched" (the --mode batch workload) and
"serving" (the --mode serve workload with 32 episodes of 4 STEPs), each with
its error against the FP64 oracle on a short window (--no-extras skips them).

More workloads, each printing its own JSON line:
  --mode batch  64 uneven episodes (lattices log-uniform 512x256 .. 2048x1024,
                1-3 wings of chord 32-128, Re 100-500, seeded sinusoidal
                motions), FP32-shifted, advanced together by lbm_batch_advance;
                value = env-steps/s (1000-step STEPs completed per second over
                all episodes), beside the plain FP64 and FP32 one-episode-at-a-
                time GPU paths timed on a sample of the same episodes.
  --mode serve  the episode server with continuous batching: Poisson arrivals,
                each episode RESET 0 then STEPs of 1000 steps with per-wing
                (domega, dxv, dyv) increments N(0, 0.1); p50/p99 STEP latency
                (submission to FORCE reply) and aggregate env-steps/s.
  --mode mix    the default-scale served mix: 16 episodes on the reference's
                default 40x20-chord domain (4001 x 2001, chord 100, Re 1000,
                two wings), all connected at once, RESET 0 then STEPs of 1000
                steps with seeded per-wing increments N(0, 0.1), served in
                FP32-shifted storage; the same commands are then served in
                FP64 and every FORCE value of the FP32 run is compared with
                the FP64 one.
  --mode dynamics  the single-instance workload with wing 1 coupled to the
                flow (the reference's rigid-body solver switched on with its
                own constants: solid density 10, closed boundary, RK4); each
                step's geometry waits for the previous step's force, so the
                steps run one at a time; the same instance's prescribed-motion
                step is timed beside it.
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
    ap.add_argument("--mode", default="single", choices=["single", "batch", "serve", "mix", "dynamics"])
    ap.add_argument("--episodes", type=int, default=None, help="batch: 64, serve: 32, mix: 16")
    ap.add_argument("--substeps", type=int, default=1000, help="steps per STEP (env-step)")
    ap.add_argument("--env-steps", type=int, default=4, help="serve: STEPs per episode")
    ap.add_argument("--rate", type=float, default=40.0, help="serve: episode arrivals per second")
    ap.add_argument("--chunk", type=int, default=100, help="serve: steps per scheduler iteration")
    ap.add_argument("--hbm-gb", type=float, default=200.0, help="serve: HBM budget for episodes")
    ap.add_argument("--spacing", type=float, default=1.5, help="serve: wing spacing in chords")
    ap.add_argument("--coupled", type=int, default=2,
                    help="mix: episodes whose wing 1 is coupled to the flow (drives move_mesh)")
    ap.add_argument("--base-sample", type=int, default=8, help="batch: episodes timed one at a time")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c5", "small"])
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"],
                    help="population storage: FP64, or FP32 deviations from the weights")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="single mode: skip the batched and served sub-results")
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


def cpu_baseline(name, seed, budget_s, precision):
    """The FP64 oracle (C restatement of the reference CPU path, OpenMP over
    the host cores) on the first steps of the same workload, timed; then the
    GPU path over the same steps from a fresh instance, with its error against
    the oracle's fields and force rows."""
    import numpy as np
    from lbm_mutiwing_amd import Solver, sinusoidal_motion_table
    from oracle import oracle
    cores = min(16, os.cpu_count() or 1)
    oracle.set_threads(cores)
    cfg, mot = workload(name, seed)
    table = sinusoidal_motion_table(cfg, mot, 64)
    o = oracle.OracleSolver(cfg)
    o.set_motion_table(table)
    r1 = o.advance(1, want_rows=True)  # first step outside the timing (page faults of fresh buffers)
    rows = [r1.force_rows]
    steps, t = 0, 0.0
    while t < budget_s and steps < 50:
        t0 = time.perf_counter()
        r = o.advance(2 + steps, want_rows=True)
        t += time.perf_counter() - t0
        rows.append(r.force_rows)
        steps += 1
    fo, mo = o.fields()[:3], o.mass()
    o.close()
    ro = np.concatenate(rows)
    gcfg, _ = workload(name, seed, precision)
    g = Solver(gcfg)
    g.set_motion_table(table)
    rg = g.advance(1 + steps, want_rows=True)
    fg, mg = g.fields()[:3], g.mass()
    g.close()

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(float(np.linalg.norm(b)), 1e-300))
    W = cfg.wing_count
    fr = slice(1, 1 + 3 * W + 3)
    err = {"steps": 1 + steps, "rho": rel(fg[0], fo[0]), "u": rel(fg[1], fo[1]), "v": rel(fg[2], fo[2]),
           "force_rows": rel(rg.force_rows[:, fr], ro[:, fr]), "mass": abs(mg - mo) / abs(mo),
           "bar": "1e-10 fields, 1e-9 forces, 1e-12 mass (FP64); 1e-4 fields, 1e-3 forces, 2e-7 mass "
                  "(FP32-shifted)"}
    base = {"value": cfg.nx * cfg.ny * steps / t / 1e6, "unit": "MLUPS", "cores": cores,
            "kind": "port", "sample": f"{name} steps 2..{1 + steps} ({steps} steps, {t:.1f} s)"}
    base["single_core"] = _oracle_rate(name, seed, 1, min(8.0, budget_s), 4)
    other = "c1" if name != "c1" else "c2"
    base[f"{other}_all_cores"] = _oracle_rate(other, seed, cores, min(5.0, budget_s), 50)
    base[f"{other}_single_core"] = _oracle_rate(other, seed, 1, min(5.0, budget_s), 10)
    oracle.set_threads(cores)
    return base, err


def _oracle_rate(name, seed, threads, budget_s, max_steps):
    """MLUPS of the FP64 oracle on `threads` host cores over the first steps
    of a workload (step 1 untimed), stopping after budget_s or max_steps."""
    from lbm_mutiwing_amd import sinusoidal_motion_table
    from oracle import oracle
    oracle.set_threads(threads)
    cfg, mot = workload(name, seed)
    o = oracle.OracleSolver(cfg)
    o.set_motion_table(sinusoidal_motion_table(cfg, mot, max_steps + 4))
    o.advance(1)
    steps, t = 0, 0.0
    while t < budget_s and steps < max_steps:
        t0 = time.perf_counter()
        o.advance(2 + steps)
        t += time.perf_counter() - t0
        steps += 1
    o.close()
    return {"value": round(cfg.nx * cfg.ny * steps / t / 1e6, 3), "cores": threads,
            "sample": f"{name} steps 2..{1 + steps} ({steps} steps, {t:.1f} s)"}


def note(msg):
    """Progress line on stderr (stdout carries only the JSON result)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_batch(args):
    import numpy as np
    from lbm_mutiwing_amd import (PRECISION_FP32_SHIFTED, PRECISION_FP64, Batch, Solver, random_episode,
                                  seeded_motions, sinusoidal_motion_table)
    from lbm_mutiwing_amd.solver import IBM_SOLVE_SUBSTITUTION
    n_ep = args.episodes or 64
    K = args.substeps
    rng = np.random.default_rng(args.seed)
    cfgs = [random_episode(rng, PRECISION_FP32_SHIFTED) for _ in range(n_ep)]
    warm = args.warmup
    tables = [sinusoidal_motion_table(c, seeded_motions(c, args.seed + i), warm + 3 * K + 2)
              for i, c in enumerate(cfgs)]

    def build(precision, idx, batched=False):
        out = []
        for i in idx:
            c = cfgs[i]
            c.precision = precision
            s = Solver(c)
            if batched:   # no inverse formed: the geometry of 64 episodes weighs less beside the lattice kernel
                s.set_ibm_solve(IBM_SOLVE_SUBSTITUTION)
            s.set_motion_table(tables[i])
            out.append(s)
        return out

    solvers = build(PRECISION_FP32_SHIFTED, range(n_ep), batched=True)
    b = Batch()
    note(f"batch: {n_ep} episodes built, warm-up of {warm} steps")
    b.advance(solvers, warm)
    note(f"batch: untimed STEP of {K} steps")
    b.advance(solvers, K)
    note(f"batch: timed STEP of {K} steps")
    t0 = time.perf_counter()
    b.advance(solvers, K)
    t_batch = time.perf_counter() - t0
    tile, rows, waves = b.tuning_full()
    cpol = b.cache_policy()
    nodes = sum(c.nx * c.ny for c in cfgs)
    for s in solvers:
        s.close()
    b.close()
    sample = list(range(min(args.base_sample, n_ep)))
    ksample = max(50, K // 5)

    def one_at_a_time(precision):
        note(f"batch: one episode at a time, precision {precision}")
        ss = build(precision, sample)
        tot = 0.0
        for s in ss:
            s.advance(warm)
            t = time.perf_counter()
            s.advance(warm + ksample)
            tot += (time.perf_counter() - t) * K / ksample
            s.close()
        return len(ss) / tot

    base64 = one_at_a_time(PRECISION_FP64)
    base32 = one_at_a_time(PRECISION_FP32_SHIFTED)
    value = n_ep * K / 1000.0 / t_batch
    nodes_of = [c.nx * c.ny for c in cfgs]
    err = None
    if not args.no_cpu_baseline:
        err = _batch_oracle_error(cfgs, tables, [int(np.argmin(nodes_of)), int(np.argmax(nodes_of))], 40)
    return {
        "metric": "env-steps/s (1000-step STEPs), 64 uneven episodes batched, FP32-shifted",
        "value": round(value, 2), "unit": "env-steps/s", "n_gpus": 1, "steps": K, "warmup": warm,
        "ms_per_step": round(t_batch * 1e3 / K, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded random episodes and sinusoidal motions)",
        "config": {"workload": f"{n_ep} episodes, lattices 512x256..2048x1024 log-uniform, 1-3 wings, "
                               f"chord 32-128, Re 100-500", "episodes": n_ep, "seed": args.seed,
                   "total_nodes": nodes},
        "aggregate_mlups": round(nodes * K / t_batch / 1e6, 1),
        "lattice_tile": tile, "lattice_tiles_per_workgroup": rows, "lattice_waves_per_cu_cap": waves,
        "lattice_cache_policy": ["plain", "non-temporal stores", "non-temporal loads and stores"][cpol],
        "timed": "second STEP of the same episodes",
        "one_at_a_time": {"fp64_env_steps_per_s": round(base64, 3), "fp32_env_steps_per_s": round(base32, 3),
                          "sample": f"episodes 0..{len(sample) - 1}, {ksample} steps each, scaled to {K}"},
        "speedup_vs_fp64_one_at_a_time": round(value / base64, 3),
        "speedup_vs_fp32_one_at_a_time": round(value / base32, 3),
        "error_vs_fp64_oracle": err,
    }


def _rel(a, b):
    import numpy as np
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(float(np.linalg.norm(b)), 1e-300))


def _batch_oracle_error(cfgs, tables, idx, steps):
    """The FP32-shifted batched path against the FP64 oracle on a short
    window: the given episodes advanced together by lbm_batch_advance (each
    is bitwise what it is in any composition), their fields, averaged forces
    and mass compared with the oracle's."""
    import dataclasses
    from lbm_mutiwing_amd import PRECISION_FP32_SHIFTED, PRECISION_FP64, Batch, Solver
    from lbm_mutiwing_amd.solver import IBM_SOLVE_SUBSTITUTION
    from oracle import oracle
    oracle.set_threads(min(16, os.cpu_count() or 1))
    ss = []
    for i in idx:
        g = Solver(dataclasses.replace(cfgs[i], precision=PRECISION_FP32_SHIFTED))
        g.set_ibm_solve(IBM_SOLVE_SUBSTITUTION)   # as the batched episodes are solved
        g.set_motion_table(tables[i])
        ss.append(g)
    b = Batch()
    rb = b.advance(ss, steps)
    out = {"episodes": [cfgs[i].name for i in idx], "steps": steps}
    worst = {"rho": 0.0, "u": 0.0, "v": 0.0, "forces": 0.0, "mass": 0.0}
    for g, r, i in zip(ss, rb, idx):
        note(f"batch: oracle check of episode {cfgs[i].name}, {steps} steps")
        o = oracle.OracleSolver(dataclasses.replace(cfgs[i], precision=PRECISION_FP64))
        o.set_motion_table(tables[i])
        ro = o.advance(steps)
        fc = slice(1, 1 + 3 * cfgs[i].wing_count + 3)
        for k, a, bb in zip(("rho", "u", "v"), g.fields()[:3], o.fields()[:3]):
            worst[k] = max(worst[k], _rel(a, bb))
        worst["forces"] = max(worst["forces"], _rel(r.row_sum[fc], ro.row_sum[fc]))
        worst["mass"] = max(worst["mass"], abs(g.mass() - o.mass()) / abs(o.mass()))
        o.close()
        g.close()
    b.close()
    out.update(worst)
    out["bar"] = "1e-4 fields, 1e-3 forces (averaged over the window), 2e-7 mass"
    return out


def run_serve(args):
    import numpy as np
    from lbm_mutiwing_amd import MOTION_STREAM, PRECISION_FP32_SHIFTED, random_episode
    from lbm_mutiwing_amd.server import EpisodeServer
    n_ep = args.episodes or 32
    K, S = args.substeps, args.env_steps
    rng = np.random.default_rng(args.seed)
    cfgs = [random_episode(rng, PRECISION_FP32_SHIFTED, MOTION_STREAM, spacing=args.spacing) for _ in range(n_ep)]
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, n_ep))
    noise = [rng.normal(0.0, 0.1, (S, 3 * c.wing_count)) for c in cfgs]

    def next_command(i, k, reply):
        """Client of episode i after its k-th reply: STEP k+1 (k < S) or STOP,
        with N(0, 0.1) velocity targets around a return to the start pose."""
        c = cfgs[i]
        if k >= S:
            return "STOP"
        d = _return_to_start(c, noise[i][k], reply, K * c.dt)
        return "STEP " + repr((k + 1) * K * c.dt) + "".join(" " + repr(float(v)) for v in d)
    srv = EpisodeServer(int(args.hbm_gb * (1 << 30)), chunk=args.chunk)
    for i, c in enumerate(cfgs):
        srv.connect(i, c)
    nxt = [0] * n_ep            # replies received so far per episode
    pending = ["RESET 0.0"] * n_ep
    sent = [[] for _ in range(n_ep)]
    got = [[] for _ in range(n_ep)]
    busy = [False] * n_ep
    finished = [False] * n_ep
    lat = []
    active_max = 0
    t_start = time.perf_counter()
    done_steps = 0
    while True:
        now = time.perf_counter() - t_start
        for i in range(n_ep):
            if not busy[i] and not finished[i] and pending[i] and arrivals[i] <= now:
                srv.submit(i, pending[i])
                sent[i].append(pending[i])
                busy[i] = True
        active_max = max(active_max, len(srv.jobs) + len(srv.waiting))
        for r in srv.poll():
            i = r.client
            busy[i] = False
            got[i].append(r.text)
            cmd = pending[i].split(" ", 1)[0]
            if cmd == "STOP":
                finished[i] = True
                note(f"serve: episode {i} stopped ({sum(finished)}/{n_ep})")
                continue
            if not r.text.startswith("FORCE "):
                raise RuntimeError(f"episode {i} ({cfgs[i].name}): {r.text}")
            if cmd == "STEP":
                lat.append(r.latency)
                done_steps += 1
            nxt[i] += 1
            pending[i] = next_command(i, nxt[i] - 1, r.text)
        if all(finished):
            break
        if srv.idle() and not any(busy):
            time.sleep(0.0005)
    wall = time.perf_counter() - t_start - float(arrivals[0])
    srv.close()
    lat_ms = np.array(lat) * 1e3
    err = None
    if not args.no_cpu_baseline:
        k = int(np.argmin([c.nx * c.ny for c in cfgs]))
        err = _serve_oracle_error(cfgs[k], sent[k][:2], got[k][:2])
    return {
        "metric": "p50/p99 STEP latency and env-steps/s, episode server with continuous batching, FP32-shifted",
        "value": round(done_steps / wall, 2), "unit": "env-steps/s", "n_gpus": 1, "steps": K, "warmup": 0,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded random episodes, Poisson arrivals, N(0, 0.1) velocity targets around a "
                "return-to-start pose per STEP)",
        "config": {"workload": f"{n_ep} episodes arriving at {args.rate}/s, RESET + {S} STEPs of {K} steps, "
                               f"chunk {args.chunk}", "episodes": n_ep, "seed": args.seed},
        "step_latency_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 2),
                            "p99": round(float(np.percentile(lat_ms, 99)), 2),
                            "mean": round(float(lat_ms.mean()), 2), "n": int(lat_ms.size)},
        "max_concurrent_jobs": active_max,
        "error_vs_fp64_oracle": err,
    }


def _reply_errors(got, ref):
    """Per FORCE value, |got - ref| / max(|ref|, 1e-3 max|ref|) over the
    replies; and per averaged force/moment column, |got - ref| over the
    column's largest |ref|."""
    import numpy as np
    per_value, per_col = [], []
    for a, b in zip(got, ref):
        va, vb = np.array(_force_values(a)), np.array(_force_values(b))
        per_value.append(float(np.max(np.abs(va - vb) / np.maximum(np.abs(vb), 1e-3 * np.max(np.abs(vb))))))
        per_col.append(np.abs(va - vb))
    return per_value, per_col


def _serve_oracle_error(cfg, cmds, replies):
    """The served replies of one episode (FP32-shifted) against the same
    commands run through the protocol over the FP64 oracle."""
    import dataclasses
    from lbm_mutiwing_amd import PRECISION_FP64
    from lbm_mutiwing_amd.protocol import Session, run_job_alone
    from oracle import oracle
    oracle.set_threads(min(16, os.cpu_count() or 1))
    note(f"serve: oracle check of episode {cfg.name}: {len(cmds)} commands")
    o = oracle.OracleSolver(dataclasses.replace(cfg, precision=PRECISION_FP64))
    sess = Session(o)
    ref = []
    for cmd in cmds:
        reply, job = sess.receive(cmd, 0.0)
        ref.append(run_job_alone(job) if job is not None else reply)
    o.close()
    per_value, _ = _reply_errors(replies, ref)
    return {"episode": cfg.name, "commands": [c.split(" ", 1)[0] for c in cmds],
            "force_rel_err_max": max(per_value),
            "bar": "1e-3 per value, floored at 1e-3 of the reply's largest |value|"}


def _force_values(reply):
    toks = reply.split()
    if not toks or toks[0] != "FORCE":
        raise RuntimeError(f"expected a FORCE reply, got: {reply[:200]}")
    return [float(x) for x in toks[2:]]


def _return_to_start(cfg, noise_k, reply, T):
    """STEP values for the next segment of one episode: per wing the target
    velocities of (rotation, x, y) are noise_k around the velocity that brings
    the wing back to its start pose within one STEP, read from the motion
    columns of the FORCE reply; the STEP carries target minus current."""
    import numpy as np
    W = cfg.wing_count
    row = np.array([float(t) for t in reply.split()[2:]])
    d = []
    for w in range(W):
        m = row[1 + 3 * W + 3 + 9 * w:]
        x, xv, y, yv, a, om = m[0], m[1], m[3], m[4], m[6], m[7]
        x0 = cfg.wing_x0 + w * cfg.wing_dx
        tgt = noise_k[3 * w:3 * w + 3] - np.array([a, x - x0, y - cfg.wing_y0]) / T
        d += [tgt[0] - om, tgt[1] - xv, tgt[2] - yv]
    return d


def _reply_value_errors(got, ref):
    """Per FORCE reply, the largest per-value relative error of ``got``
    against ``ref``, floored at 1e-3 of the reply's largest |value|."""
    import numpy as np
    out = []
    for a, b in zip(got, ref):
        ra, rb = np.array(_force_values(a)), np.array(_force_values(b))
        out.append(float(np.max(np.abs(ra - rb) / np.maximum(np.abs(rb), 1e-3 * np.max(np.abs(rb))))))
    return out


def _mix_sensitivity(args, episodes, cmds, r64, n_cpl):
    """How far FP64 itself moves the mix's FORCE values when its populations
    carry the rounding of FP32-shifted storage: each listed episode's
    commands are replayed alone in FP64 (same chunking as the server, so the
    unperturbed replay equals the served FP64 replies byte for byte), once as
    served and once with every population deviation d = f - w scaled by
    (1 + 2^-24 u), u uniform in [-1, 1] and seeded, after the RESET and after
    each STEP: the relative rounding of an FP32 deviation, applied four or
    five times instead of at every step."""
    import numpy as np
    from lbm_mutiwing_amd import MOTION_STREAM, PRECISION_FP64, preset
    from lbm_mutiwing_amd.protocol import Session, run_job_alone
    from lbm_mutiwing_amd.solver import IBM_SOLVE_SUBSTITUTION, Dynamics, Solver
    w = np.array([4.0 / 9.0] + [1.0 / 9.0] * 4 + [1.0 / 36.0] * 4, dtype=np.float32).astype(np.float64)
    rng = np.random.default_rng(args.seed + 7)
    out = {"episodes": [], "replay_equals_served": True, "perturbed_max": 0.0, "perturbed_max_by_step": None}
    for i in sorted(set(episodes)):
        for perturb in (False, True):
            note(f"mix: FP64 replay of episode {i}{' with FP32-level perturbations' if perturb else ''}")
            cfg = preset("c5")
            cfg.precision, cfg.motion_source = PRECISION_FP64, MOTION_STREAM
            s = Solver(cfg)
            s.set_ibm_solve(IBM_SOLVE_SUBSTITUTION)   # as the server solves its episodes
            if i < n_cpl:
                s.set_dynamics(Dynamics())
            sess = Session(s)
            replies = []
            for c in cmds[i]:
                if c == "STOP":
                    break
                text, job = sess.receive(c)
                if job is not None:
                    text = run_job_alone(job, args.chunk)
                replies.append(text)
                if perturb:
                    f, _ = s.populations()
                    d = f - w[:, None, None]
                    s.set_populations(w[:, None, None] + d * (1.0 + 2.0 ** -24 * rng.uniform(-1.0, 1.0, f.shape)))
            s.close()
            ref = r64[i][:len(replies)]
            if not perturb:
                out["replay_equals_served"] &= replies == ref
            else:
                e = _reply_value_errors(replies[1:], ref[1:])
                out["episodes"].append({"episode": i, "coupled": i < n_cpl, "max_by_step": e})
                out["perturbed_max"] = max(out["perturbed_max"], max(e))
    steps = [ep["max_by_step"] for ep in out["episodes"]]
    out["perturbed_max_by_step"] = [max(v) for v in zip(*steps)] if steps else None
    out["what"] = ("FP64 against FP64 with the populations' deviations perturbed at FP32 rounding after the RESET "
                   "and after each STEP; per FORCE value, relative, floored at 1e-3 of the reply's largest |value|")
    return out


def _mix_oracle_error(args, noise0, steps):
    """One episode of the mix, FP32-shifted and solved as the server solves
    it, against the FP64 oracle over its RESET and the first steps of a STEP
    with the same velocity increments: fields, averaged forces and mass."""
    import dataclasses
    import numpy as np
    from lbm_mutiwing_amd import MOTION_STREAM, PRECISION_FP32_SHIFTED, PRECISION_FP64, preset
    from lbm_mutiwing_amd.solver import IBM_SOLVE_SUBSTITUTION, Solver
    from oracle import oracle
    oracle.set_threads(min(16, os.cpu_count() or 1))
    base = preset("c5")
    base.motion_source = MOTION_STREAM
    g = Solver(dataclasses.replace(base, precision=PRECISION_FP32_SHIFTED))
    g.set_ibm_solve(IBM_SOLVE_SUBSTITUTION)
    o = oracle.OracleSolver(dataclasses.replace(base, precision=PRECISION_FP64))
    T = args.substeps * base.dt
    res = []
    for x in (g, o):
        x.reset()
        start = np.zeros(base.motion_cols)
        target = x.build_stream_target_row(start, T, np.concatenate([[T], noise0]))
        x.set_stream_segment(start, target)
        res.append(x.advance(steps))
    note(f"mix: oracle check over {steps} steps done")
    out = {"episode": "c5 (episode 0's first STEP increments)", "steps": steps}
    for k, a, b in zip(("rho", "u", "v"), g.fields()[:3], o.fields()[:3]):
        out[k] = _rel(a, b)
    fc = slice(1, 1 + 3 * base.wing_count + 3)
    out["forces"] = _rel(res[0].row_sum[fc], res[1].row_sum[fc])
    out["mass"] = abs(g.mass() - o.mass()) / abs(o.mass())
    out["bar"] = "1e-4 fields, 1e-3 forces (averaged over the window), 2e-7 mass"
    g.close()
    o.close()
    return out


def run_mix(args):
    """Default-scale served mix: seeded command streams served in FP32-shifted
    storage, then the same command strings served in FP64; every FORCE value
    of the FP32 run is compared with the FP64 one (relative, floored at 1e-3
    of the reply's largest value: the first steps' moments are rounding noise
    beside the forces)."""
    import numpy as np
    from lbm_mutiwing_amd import MOTION_STREAM, PRECISION_FP32_SHIFTED, PRECISION_FP64, preset
    from lbm_mutiwing_amd.server import EpisodeServer
    n_ep = args.episodes or 16
    K, S = args.substeps, args.env_steps
    base = preset("c5")
    rng = np.random.default_rng(args.seed)
    noise = [rng.normal(0.0, 0.1, (S, 3 * base.wing_count)) for _ in range(n_ep)]
    T = K * base.dt

    n_cpl = min(args.coupled, n_ep)

    def serve(precision, script=None):
        """Serve every episode's commands; without a script the STEPs are
        built from the replies (return to start plus noise) and recorded.
        The first n_cpl episodes couple wing 1 to the flow (the reference's
        rigid-body solver and constants), which drives move_mesh."""
        from lbm_mutiwing_amd.solver import Dynamics
        srv = EpisodeServer(int(args.hbm_gb * (1 << 30)), chunk=args.chunk)
        for i in range(n_ep):
            cfg = preset("c5")
            cfg.precision = precision
            cfg.motion_source = MOTION_STREAM
            srv.connect(i, cfg, Dynamics() if i < n_cpl else None)
        sent = [["RESET 0.0"] for _ in range(n_ep)]
        replies = [[] for _ in range(n_ep)]
        lat = []
        t0 = time.perf_counter()
        for i in range(n_ep):
            srv.submit(i, "RESET 0.0")
        done = 0
        while done < n_ep:
            for r in srv.poll():
                i = r.client
                if not r.text.startswith(("FORCE ", "STATUS STOPPING")):
                    raise RuntimeError(f"episode {i}: {r.text}")
                if sent[i][-1].startswith("STEP"):
                    lat.append(r.latency)
                replies[i].append(r.text)
                k = len(replies[i]) - 1          # STEPs answered so far
                if sent[i][-1] == "STOP":
                    done += 1
                    continue
                if script is not None:
                    nxt = script[i][len(sent[i])]
                elif k >= S:
                    nxt = "STOP"
                else:
                    d = _return_to_start(base, noise[i][k], r.text, T)
                    nxt = "STEP " + repr((k + 1) * T) + "".join(" " + repr(float(v)) for v in d)
                sent[i].append(nxt)
                srv.submit(i, nxt)
            if srv.idle():
                time.sleep(0.0005)
        wall = time.perf_counter() - t0
        final = dict(srv.final)
        srv.close()
        note(f"mix: precision {precision} served in {wall:.1f} s")
        return sent, replies, np.array(lat) * 1e3, wall, final

    cmds, r32, lat32, wall32, fin32 = serve(PRECISION_FP32_SHIFTED)
    _, r64, _, wall64, fin64 = serve(PRECISION_FP64, script=cmds)
    # per value, relative with a floor of 1e-3 of the reply's largest value;
    # and per force/moment column, against that column's largest magnitude
    # over the episode's STEPs
    W = base.wing_count
    nf = 3 * W + 3                      # averaged (Fx, Fy, M) per wing and total
    errs, col_errs, worst = [], [], None
    for i in range(n_ep):
        fa = np.array([_force_values(a) for a in r32[i][1:S + 1]])
        fb = np.array([_force_values(b) for b in r64[i][1:S + 1]])
        for k, (ra, rb) in enumerate(zip(fa, fb)):
            floor = 1e-3 * np.max(np.abs(rb))
            e = np.abs(ra - rb) / np.maximum(np.abs(rb), floor)
            errs.append(float(e.max()))
            if worst is None or e.max() > worst["err"]:
                c = int(e.argmax())
                worst = {"err": float(e.max()), "episode": i, "step": k + 1, "value_index": c,
                         "fp32": float(ra[c]), "fp64": float(rb[c]), "floor": float(floor)}
        scale = np.max(np.abs(fb[:, 1:1 + nf]), axis=0)
        col_errs.append(np.max(np.abs(fa[:, 1:1 + nf] - fb[:, 1:1 + nf]), axis=0) / np.maximum(scale, 1e-300))
    errs = np.array(errs)
    col_errs = np.array(col_errs)
    by_step = errs.reshape(n_ep, S).max(axis=0)
    sens = _mix_sensitivity(args, [worst["episode"], n_cpl % n_ep, (n_cpl + 1) % n_ep], cmds, r64, n_cpl)
    oerr = None if args.no_cpu_baseline else _mix_oracle_error(args, noise[0][0], 20)
    # mass: every episode starts from the same uniform state; the drift from
    # it (mass leaves and enters through the open far field, and move_mesh
    # drops a column and repeats another) is compared between the precisions
    m0 = {}
    for prec in (PRECISION_FP32_SHIFTED, PRECISION_FP64):
        c0 = preset("c5")
        c0.precision, c0.motion_source = prec, MOTION_STREAM
        from lbm_mutiwing_amd import Solver
        s0 = Solver(c0)
        m0[prec] = s0.mass()
        s0.close()
    ids = [i for i in range(n_ep) if i in fin32 and i in fin64]
    mass = [abs(fin32[i][0] - fin64[i][0]) / abs(fin64[i][0]) for i in ids]
    drift = [abs((fin32[i][0] - m0[PRECISION_FP32_SHIFTED]) - (fin64[i][0] - m0[PRECISION_FP64])) /
             abs(fin64[i][0] - m0[PRECISION_FP64]) for i in ids]
    moves = [fin64[i][1] for i in range(n_cpl) if i in fin64]
    env_steps = n_ep * S
    return {
        "metric": "env-steps/s and p50/p99 STEP latency, default-scale served mix, FP32-shifted; FORCE error vs FP64",
        "value": round(env_steps / wall32, 3), "unit": "env-steps/s", "n_gpus": 1, "steps": K, "warmup": 0,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded N(0, 0.1) velocity targets around a return-to-start pose per STEP)",
        "config": {"workload": f"{n_ep} episodes of 4001x2001 (chord 100, Re 1000, 2 wings x {base.points_per_wing} "
                               f"markers), {n_cpl} of them with wing 1 coupled to the flow, all connected at once, "
                               f"RESET + {S} STEPs of {K} steps, chunk {args.chunk}",
                   "episodes": n_ep, "seed": args.seed},
        "step_latency_ms": {"p50": round(float(np.percentile(lat32, 50)), 2),
                            "p99": round(float(np.percentile(lat32, 99)), 2), "n": int(lat32.size)},
        "fp64_env_steps_per_s": round(env_steps / wall64, 3),
        "force_rel_err_vs_fp64": {"max": float(errs.max()), "median": float(np.median(errs)), "n": int(errs.size),
                                  "p99": float(np.percentile(errs, 99)), "bar": 1e-3,
                                  "bar_met": bool(errs.max() < 1e-3), "worst": worst,
                                  "what": "per FORCE value, relative, floored at 1e-3 of the reply's largest "
                                          "|value|"},
        "error_vs_fp64_oracle": oerr,
        "force_rel_err_vs_fp64_max_by_step": [float(v) for v in by_step],
        "fp64_sensitivity": sens,
        "mass_vs_fp64": {"rel_err_max": float(max(mass)) if mass else None,
                         "drift_rel_err_max": float(max(drift)) if drift else None, "drift_bar": 1e-3,
                         "what": "final mass of each episode (lbm_get_mass) FP32 against FP64, and its drift "
                                 "from the initial mass against the FP64 drift"},
        "coupled_episodes": n_cpl, "move_mesh_per_coupled_episode": moves,
        "force_err_vs_fp64_per_column": {"max": float(col_errs.max()), "median": float(np.median(col_errs)),
                                         "bar": 1e-3, "bar_met": bool(col_errs.max() < 1e-3),
                                         "max_by_column": [float(v) for v in col_errs.max(axis=0)],
                                         "columns": "Fx, Fy, M per wing, then the totals; error over the "
                                                    "column's largest |value| in the episode"},
        "note": "STEP commands carry velocity targets only; as in the reference, only the rigid-body coupling "
                "moves body_rel_pos, so move_mesh fires in the coupled episodes (wing 1 drifts downstream under "
                "its drag and the lattice re-centres each time it crosses a node)",
    }


def run_dynamics(args):
    from lbm_mutiwing_amd import sinusoidal_motion_table
    from lbm_mutiwing_amd.solver import BODY_REL_FIELDS, Dynamics, Solver
    cfg, mot = workload(args.config, args.seed, args.precision)
    steps, warmup = args.steps, args.warmup
    table = sinusoidal_motion_table(cfg, mot, warmup + 2 * steps + 2)
    out_ms = {}
    for name, dyn in (("prescribed", None), ("coupled", Dynamics())):
        s = Solver(cfg)
        s.set_motion_table(table)
        s.set_dynamics(dyn)
        s.advance(warmup)
        t0 = time.perf_counter()
        r = s.advance(warmup + steps)
        out_ms[name] = (time.perf_counter() - t0) * 1e3 / steps
        note(f"dynamics: {name} {out_ms[name]:.4f} ms/step")
        if dyn is not None:
            rel = s.body_rel()[:, 0]
            body = {k: float(rel[i]) for i, k in enumerate(BODY_REL_FIELDS) if k in ("a", "v", "s", "pos")}
            movenum = s.state()["movenum"]
        s.close()
    nodes = cfg.nx * cfg.ny
    return {
        "metric": f"MLUPS (whole time step incl. IBM and rigid-body coupling), {args.precision.upper()} single instance",
        "value": round(nodes / (out_ms["coupled"] * 1e-3) / 1e6, 1),
        "unit": "MLUPS", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(out_ms["coupled"], 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded sinusoidal pitch/heave motion table, generated NACA0012 markers)",
        "config": {"workload": f"{args.config}: {cfg.nx}x{cfg.ny}, {cfg.wing_count} wings x {cfg.points_per_wing} "
                               "markers, wing 1 coupled (solid density 10, closed boundary, RK4)",
                   "seed": args.seed},
        "prescribed_ms_per_step": round(out_ms["prescribed"], 4),
        "body_after_warmup_plus_steps": body, "movenum": movenum,
        "error_vs_fp64_oracle": None if args.no_cpu_baseline else _dynamics_oracle_error(cfg, table, 20),
    }


def _dynamics_oracle_error(cfg, table, steps):
    """The coupled instance against the FP64 oracle (with the same rigid-body
    coupling) over the first steps of the workload: fields, averaged forces
    and the body state."""
    import numpy as np
    from lbm_mutiwing_amd.solver import Dynamics, Solver
    from oracle import oracle
    oracle.set_threads(min(16, os.cpu_count() or 1))
    out = {"steps": steps}
    g, o = Solver(cfg), oracle.OracleSolver(cfg)
    res = []
    for x in (g, o):
        x.set_motion_table(table)
        x.set_dynamics(Dynamics())
        res.append(x.advance(steps))
    note(f"dynamics: oracle check over {steps} steps done")
    for k, a, b in zip(("rho", "u", "v"), g.fields()[:3], o.fields()[:3]):
        out[k] = _rel(a, b)
    fc = slice(1, 1 + 3 * cfg.wing_count + 3)
    out["forces"] = _rel(res[0].row_sum[fc], res[1].row_sum[fc])
    bg, bo = g.body_rel(), o.body_rel()
    out["body_rel"] = float(np.max(np.abs(bg - bo)) / max(float(np.max(np.abs(bo))), 1e-300))
    out["bar"] = "1e-10 fields, 1e-9 forces, 1e-9 body state (FP64)"
    g.close()
    o.close()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(episodes are independent and shard one instance per GPU)"}))
        sys.exit(2)
    import torch  # noqa: F401  (device visibility check only)
    if args.mode == "batch":
        print(json.dumps(run_batch(args)), flush=True)
        return
    if args.mode == "serve":
        print(json.dumps(run_serve(args)), flush=True)
        return
    if args.mode == "mix":
        print(json.dumps(run_mix(args)), flush=True)
        return
    if args.mode == "dynamics":
        print(json.dumps(run_dynamics(args)), flush=True)
        return
    from lbm_mutiwing_amd import sinusoidal_motion_table
    from lbm_mutiwing_amd.solver import TimedSolver

    cfg, mot = workload(args.config, args.seed, args.precision)
    bpn = BYTES_PER_NODE[args.precision]
    n_total = args.warmup + 2 * args.steps + 1
    s = TimedSolver(cfg)
    s.set_motion_table(sinusoidal_motion_table(cfg, mot, n_total + 1))
    s.advance(args.warmup)
    mass0 = s.mass()
    t0 = time.perf_counter()
    r = s.advance(args.warmup + args.steps)
    dt = time.perf_counter() - t0
    mass1 = s.mass()
    # the lattice-kernel events of a second pass of the same length
    s.set_timing(True)
    s.advance(args.warmup + 2 * args.steps)
    lat_ms, launches = s.timing()
    tile, waves = s.tuning_full()
    cpol = s.cache_policy()
    st = s.state()
    # the lattice kernel's memory pattern alone (pull of nine planes, no
    # arithmetic), timed on the same device: the ceiling of that pattern
    import ctypes
    ceil_us = ctypes.c_double()
    fn = s.lib.lbm_measure_pull_copy
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
    if fn(cfg.nx, cfg.ny, cfg.precision, 20, ctypes.byref(ceil_us)) != 0:
        raise RuntimeError("lbm_measure_pull_copy failed")
    nodes = cfg.nx * cfg.ny
    mlups = nodes * args.steps / dt / 1e6
    ms_step = dt * 1e3 / args.steps
    lat_avg_ms = lat_ms / max(launches, 1)
    interior = (cfg.nx - 2) * (cfg.ny - 2)
    achieved = bpn * interior / (lat_avg_ms * 1e-3) / 1e9
    traffic = None
    tp = os.path.join(ROOT, "profiles", f"traffic_{args.config}_{args.precision}.json")
    if os.path.exists(tp):  # the lattice kernel's measured HBM bytes at the tile width and cache policy in use
        with open(tp) as fh:
            traffic = json.load(fh).get("by_tile_and_policy", {}).get(f"{tile}/{cpol}", {}).get("hbm_bytes_per_launch")
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
                     "kernel_avg_us": round(lat_avg_ms * 1e3, 2), "tile": tile, "waves_per_cu_cap": waves,
                     "cache_policy": ["plain", "non-temporal stores", "non-temporal loads and stores"][cpol],
                     "algorithmic_bytes_per_launch": bpn * interior},
        "pull_copy_ceiling": {"avg_us": round(ceil_us.value, 2),
                              "achieved": round(bpn * interior / (ceil_us.value * 1e-6) / 1e9, 1), "unit": "GB/s",
                              "frac": round(bpn * interior / (ceil_us.value * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                              "kernel_over_ceiling": round(lat_avg_ms * 1e3 / ceil_us.value, 4),
                              "what": "the lattice kernel's loads and stores alone (nine pulled planes in, nine out, "
                                      "no arithmetic), same lattice and storage, fastest of its cache policies and "
                                      "wave caps, 20 launches each"},
        "last_force_row": [round(x, 8) for x in r.last_row[:10]],
        "mass_check": {"before": mass0, "after": mass1, "drift_rel": (mass1 - mass0) / mass0,
                       "what": "FP64 sum of every population of the lattice before and after the timed steps "
                               "(lbm_get_mass); the open far field exchanges mass with the outside"},
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
        out["cpu_baseline"], out["error_vs_fp64_oracle"] = cpu_baseline(args.config, args.seed, args.cpu_seconds,
                                                                        args.precision)
    if not args.no_extras:
        import copy
        keep = ("metric", "value", "unit", "ms_per_step", "config", "error_vs_fp64_oracle")
        bargs = copy.copy(args)
        bargs.episodes, bargs.substeps = 64, 1000
        note("extras: batched uneven episodes")
        b = run_batch(bargs)
        out["batched"] = {k: b[k] for k in keep + ("aggregate_mlups", "one_at_a_time",
                                                   "speedup_vs_fp64_one_at_a_time", "speedup_vs_fp32_one_at_a_time")}
        sargs = copy.copy(args)
        sargs.episodes, sargs.substeps, sargs.env_steps = 32, 1000, 4
        note("extras: episode server")
        v = run_serve(sargs)
        out["serving"] = {k: v[k] for k in keep + ("step_latency_ms", "max_concurrent_jobs")}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
