"""Device-resident episode server with continuous batching.

Each connected client is a protocol :class:`~lbm_mutiwing_amd.protocol.Session`
(RESET / STEP / SAVE / STOP, driver.f90:59-196) bound to its own solver
instance, whose lattice, IBM geometry and motion state stay in HBM for the
life of the episode.  The scheduler:

* admits a client's first RESET when its estimated device footprint fits the
  HBM budget (otherwise the RESET waits until STOPs free memory);
* turns every STEP / RESET into an advance job and, at each iteration, runs
  all active jobs together with one batched advance (lbm_batch_advance) of up
  to ``chunk`` steps; jobs that arrive while others run join at the next
  iteration, and each job advances by min(chunk, remaining) so its chunking
  depends only on its own segment: a reply is bitwise the same alone and in
  any mix;
* answers a job with the segment-averaged FORCE reply as soon as it reaches
  its target step, recording the latency from submission to reply;
* advances the episodes that couple wing 1 to the flow (one step at a time,
  lbm_batch_advance rejects them) on a worker thread while the batched
  launches run, so their steps overlap on the device; each instance is still
  used by one thread at a time.

The transport is not part of this module: ``submit`` takes the command text
and ``poll`` returns reply texts, so a ZeroMQ REP loop, a test or a benchmark
drives it the same way.
"""
from __future__ import annotations

import math
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional, Tuple

from .config import PRECISION_FP32_SHIFTED, SolverConfig
from .protocol import CMD_RESET, CMD_STOP, Session, parse_command
from .solver import IBM_SOLVE_SUBSTITUTION, Batch, LbmError, Solver


# lbm_batch_advance captures advances of at most this many steps as HIP
# graphs (include/lbm_hip.h, "Launch graphs")
GRAPH_MAX_STEPS = 64


def device_bytes(cfg: SolverConfig) -> int:
    """Device memory of one instance (the allocations of lbm_create)."""
    nodes = cfg.nx * cfg.ny
    fsize = 4 if cfg.precision == PRECISION_FP32_SHIFTED else 8
    total = nodes * (2 * 9 * fsize + 10 * 8)
    W = cfg.wing_count
    if W:
        m = cfg.marker_array()
        spw = cfg.points_per_wing - 1
        r = float(max((m[:, 0] + cfg.xc - cfg.xp) ** 2 + (m[:, 1] + cfg.yc - cfg.yp) ** 2)) ** 0.5
        side = math.ceil(2.0 * r / cfg.scale) + 12
        per_slot = W * spw * spw * 8 + W * 2 * side * side * 4 + 36 * W * spw * 12 + W * spw * 160
        per_slot += 64 * 64 * 8 if W > 1 else 0         # coupling matrix M
        per_slot += W * ((spw + 24) * 25 + spw + 1) * 8  # factors of S
        nslot = max(1, min(16, (256 << 20) // per_slot))
        total += 2 * nslot * per_slot
    return int(total)


@dataclass
class Reply:
    client: object
    text: str
    submitted: float
    replied: float
    kind: str          # "FORCE", "STATUS" or "ERROR"

    @property
    def latency(self) -> float:
        return self.replied - self.submitted


class EpisodeServer:
    def __init__(self, hbm_budget: int, chunk: int = 100, lib=None, clock: Callable[[], float] = time.perf_counter,
                 save_hook=None):
        self.lib = lib
        self.batch = Batch(lib)
        self.hbm_budget = int(hbm_budget)
        self.used = 0
        self.chunk = int(chunk)
        self.clock = clock
        self.save_hook = save_hook
        self.configs: Dict[object, SolverConfig] = {}
        self.dynamics: Dict[object, object] = {}   # client -> solver.Dynamics (rigid-body coupling) or None
        self.sessions: Dict[object, Session] = {}
        self.waiting: deque = deque()          # (client, message, submitted) first RESETs
        self.jobs: List = []
        self.replies: List[Reply] = []
        # per client, when it stopped: its episode's mass (lbm_get_mass) and movenum
        self.final: Dict[object, Tuple[float, int]] = {}
        self._pool: Optional[ThreadPoolExecutor] = None

    # ---------------------------------------------------------------- clients
    def connect(self, client, cfg: SolverConfig, dynamics=None) -> None:
        """Register a client and the configuration of its episodes; with
        ``dynamics`` (a solver.Dynamics) its episodes couple wing 1 to the
        flow.  Such an episode steps one step at a time, so the scheduler
        advances it on its own, beside the batched launches."""
        self.configs[client] = cfg
        self.dynamics[client] = dynamics

    def submit(self, client, message: str, submitted: Optional[float] = None) -> None:
        t = self.clock() if submitted is None else submitted
        sess = self.sessions.get(client)
        if sess is None:
            cfg = self.configs[client]
            cmd, _, err = parse_command(message, cfg.wing_count)
            if cmd == CMD_RESET:
                self.waiting.append((client, message, t))
                return
            if cmd == CMD_STOP:
                self._reply(client, "STATUS STOPPING", t)
                return
            if err:
                self._reply(client, "ERROR " + err, t)
                return
            self._reply(client, "ERROR online mode requires RESET before " +
                        ("SAVE" if message.strip().upper().startswith(("SAVE", "OUTPUT")) else "STEP"), t)
            return
        reply, job = sess.receive(message, t)
        if job is not None:
            job.tag = client
            self.jobs.append(job)
        else:
            self._reply(client, reply, t)
            if sess.stopped:
                self._release(client)

    def _reply(self, client, text: str, submitted: float) -> None:
        self.replies.append(Reply(client, text, submitted, self.clock(), text.split(" ", 1)[0]))

    def _release(self, client) -> None:
        sess = self.sessions.pop(client, None)
        if sess is not None:
            self.used -= device_bytes(sess.cfg)
            try:
                self.final[client] = (sess.solver.mass(), sess.solver.state()["movenum"])
            except LbmError:
                pass
            sess.solver.close()

    # ------------------------------------------------------------- scheduling
    def _admit(self) -> None:
        while self.waiting:
            client, message, t = self.waiting[0]
            cfg = self.configs[client]
            need = device_bytes(cfg)
            if self.sessions and self.used + need > self.hbm_budget:
                return
            self.waiting.popleft()
            solver = Solver(cfg, self.lib)
            # batched episodes solve the IBM system by band substitution:
            # no inverse is formed, so the geometry of all episodes weighs
            # less beside the batched lattice kernel
            solver.set_ibm_solve(IBM_SOLVE_SUBSTITUTION)
            if self.dynamics.get(client) is not None:
                solver.set_dynamics(self.dynamics[client])
            sess = Session(solver, self.save_hook)
            self.sessions[client] = sess
            self.used += need
            reply, job = sess.receive(message, t)
            if job is not None:
                job.tag = client
                self.jobs.append(job)
            else:
                self._reply(client, reply, t)

    def poll(self) -> List[Reply]:
        """One scheduler iteration: admit, advance every active job by up to
        ``chunk`` steps in batched calls, answer the jobs that are done."""
        self._admit()
        if self.jobs:
            groups: Dict[Tuple[int, int], list] = {}
            alone = []
            for job in self.jobs:
                c = min(self.chunk, job.remaining())
                if self.dynamics.get(job.tag) is not None:
                    alone.append((job, c))
                else:
                    groups.setdefault((job.session.cfg.precision, c), []).append(job)
            failed = []

            def step_alone():
                out = []
                for j, c in alone:
                    s = j.session.solver
                    try:
                        r = s.advance(s.state()["n_next"] + c - 1)
                    except LbmError as e:
                        out.append((j, str(e)))
                    else:
                        j.add(r)
                return out

            pending = None
            # a batched advance of at most GRAPH_MAX_STEPS steps may be
            # captured as a HIP graph, and a capture fails when another
            # thread synchronises meanwhile: overlap only longer chunks
            if alone and groups and min(c for _, c in groups) > GRAPH_MAX_STEPS:
                if self._pool is None:
                    self._pool = ThreadPoolExecutor(max_workers=1)
                pending = self._pool.submit(step_alone)
            else:
                failed += step_alone()
            for (_, c), jobs in groups.items():
                results = self.batch.advance([j.session.solver for j in jobs], c, partial=True)
                for j, r in zip(jobs, results):
                    if isinstance(r, LbmError):
                        failed.append((j, str(r)))
                    else:
                        j.add(r)
            if pending is not None:
                failed += pending.result()
            for j, msg in failed:
                self.jobs.remove(j)
                self._reply(j.tag, "ERROR " + msg, j.arrival)
            done = [j for j in self.jobs if j.remaining() <= 0]
            self.jobs = [j for j in self.jobs if j.remaining() > 0]
            for j in done:
                self._reply(j.tag, j.session.finish(j), j.arrival)
        out, self.replies = self.replies, []
        return out

    def idle(self) -> bool:
        return not self.jobs and not self.waiting

    def close(self) -> None:
        if self._pool is not None:
            self._pool.shutdown(wait=True)
            self._pool = None
        for c in list(self.sessions):
            self._release(c)
        self.batch.close()
