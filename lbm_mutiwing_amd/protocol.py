"""Host side of the online episode protocol, transport-agnostic.

Mirrors the reference's ZeroMQ front end: command parsing and reply layout
(zmq_comm.f90:183-281) and the per-episode RESET / STEP / SAVE / STOP state
machine with segment-averaged FORCE replies (driver.f90:59-196, 270-408).
The time stepping itself is done by the HIP library; a :class:`Session`
turns a command into either an immediate reply or an :class:`AdvanceJob`
(run the solver up to a target step), so that a server can run the jobs of
many sessions together and complete them with :meth:`Session.finish`.

Reply strings are byte-compatible with the reference: ``FORCE <n> v1 .. vn``
with every value in Fortran ``es24.16`` form (left-adjusted), ``STATUS <text>``
and ``ERROR <text>``; the error texts are the reference's own.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

CMD_INVALID, CMD_RESET, CMD_STEP, CMD_OUTPUT, CMD_STOP = 0, 1, 2, 3, 4

# error texts of the reference (zmq_comm.f90:240-276, driver.f90:103-113, 293-317, mode.f90:303)
ERR_PARSE = "无法解析 ZeroMQ 指令"
ERR_RESET_TIME_MISSING = "RESET 指令缺少初始时间"
ERR_STEP_ARGS = "STEP 指令缺少目标时刻或速度增量"
ERR_UNKNOWN = "未知 ZeroMQ 指令: "
ERR_SAVE_BEFORE_RESET = "online mode requires RESET before SAVE"
ERR_STEP_BEFORE_RESET = "online mode requires RESET before STEP"
ERR_RESET_NONZERO = "RESET command must provide initial time 0.0"
ERR_EXPECTED = "expected STEP, SAVE, RESET or STOP"
ERR_DT = "求解器时间步无效"
ERR_ALIGN = "STEP 时间必须与 LBM 时间步严格对齐"
ERR_BEHIND = "STEP 时间必须大于等于当前待推进时刻"
ERR_SEGMENT_START = "STEP 时间必须大于等于当前段起始时刻"
ERR_SAVE_TRANSITION = "当前时刻处于移网格后的过渡阶段，请在下一次 STEP 后再保存流场"


def fortran_es24_16(x: float) -> str:
    """Fortran ``es24.16`` edit of one value, left-adjusted (adjustl + trim):
    one digit before the point, 16 after, exponent E+dd (three-digit
    exponents drop the E, as Fortran does when the field has no room)."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Infinity" if x > 0 else "-Infinity"
    s = f"{x:.16E}"
    mant, exp = s.split("E")
    e = int(exp)
    if abs(e) <= 99:
        return f"{mant}E{'+' if e >= 0 else '-'}{abs(e):02d}"
    return f"{mant}{'+' if e >= 0 else '-'}{abs(e):03d}"


def format_force_reply(row: np.ndarray) -> str:
    """zmq_send_force_reply (zmq_comm.f90:183-203)."""
    return "FORCE " + str(len(row)) + "".join(" " + fortran_es24_16(v) for v in row)


_REAL = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([eEdDqQ][+-]?\d+|[+-]\d+)?$")
_REPEAT = re.compile(r"^(\d+)\*(.*)$", re.S)
_SEPARATORS = " \t,;/"


def list_items(record: str) -> Tuple[List[Optional[str]], bool]:
    """Items of a list-directed input record, as Fortran reads them: values
    are separated by blanks (or tabs) around at most one comma or semicolon;
    a comma that follows a separator gives a null item (None); ``r*c`` is r
    copies of c and ``r*`` r null items; a quoted string is one item; a slash
    ends the record.  Returns (items, ended_by_slash)."""
    items: List[Optional[str]] = []
    pos, n = 0, len(record)

    def skip_blanks(p):
        while p < n and record[p] in " \t":
            p += 1
        return p

    while True:
        pos = skip_blanks(pos)
        if pos >= n:
            return items, False
        c = record[pos]
        if c == "/":
            return items, True
        if c in ",;":                     # separator after a separator: null
            items.append(None)
            pos += 1
            continue
        if c in "'\"":
            end = record.find(c, pos + 1)
            end = n if end < 0 else end
            tok, pos = record[pos + 1:end], end + 1
            items.append(tok)
        else:
            end = pos
            while end < n and record[end] not in _SEPARATORS:
                end += 1
            tok, pos = record[pos:end], end
            m = _REPEAT.match(tok)
            if m:
                items.extend([m.group(2) if m.group(2) else None] * int(m.group(1)))
            else:
                items.append(tok)
        pos = skip_blanks(pos)
        if pos < n and record[pos] in ",;":
            pos += 1


def fortran_real(tok: str) -> float:
    """A real in list-directed input: optional sign, digits with an optional
    point, exponent letter E, D or Q, or a signed exponent without letter."""
    m = _REAL.match(tok.strip())
    if not m:
        raise ValueError(tok)
    mant, ex = m.group(1), m.group(2) or ""
    sign = "-" if tok.strip().startswith("-") else ""
    if ex and ex[0] in "eEdDqQ":
        ex = ex[1:]
    return float(f"{sign}{mant}e{ex}" if ex else f"{sign}{mant}")


def read_reals(items: List[Optional[str]], slash: bool, row: np.ndarray, count: int) -> bool:
    """Fill row[:count] from items as ``read(...) row(1:count)`` does: null
    items keep the old value, a slash keeps the rest; False when the record
    runs out or an item is not a real."""
    for k in range(count):
        if k >= len(items):
            return slash
        if items[k] is None:
            continue
        try:
            row[k] = fortran_real(items[k])
        except ValueError:
            return False
    return True


def parse_command(message: str, wing_count: int) -> Tuple[int, np.ndarray, str]:
    """parse_command_message (zmq_comm.f90:225-281): (command, row, error).
    The keyword is the first list-directed item (at most 32 characters,
    ASCII-uppercased).  RESET reads its time only when the message is longer
    than the keyword (len_trim comparison, so leading blanks make the time
    mandatory); STEP reads exactly 1 + 3*wing_count reals."""
    row = np.zeros(1 + 9 * 8)
    items, slash = list_items(message)
    if not items and not slash:
        return CMD_INVALID, row, ERR_PARSE
    key = (items[0] or "") if items else ""
    key = "".join(ch.upper() if "a" <= ch <= "z" else ch for ch in key[:32])
    if key in ("RESET", "START"):
        if len(message.rstrip(" ")) != len(key.rstrip(" ")):
            if not read_reals(items[1:], slash, row, 1):
                return CMD_INVALID, row, ERR_RESET_TIME_MISSING
        return CMD_RESET, row, ""
    if key == "STEP":
        if not read_reals(items[1:], slash, row, 1 + 3 * wing_count):
            return CMD_INVALID, row, ERR_STEP_ARGS
        return CMD_STEP, row, ""
    if key in ("OUTPUT", "SAVE"):
        return CMD_OUTPUT, row, ""
    if key in ("STOP", "EXIT"):
        return CMD_STOP, row, ""
    return CMD_INVALID, row, (ERR_UNKNOWN + key).rstrip(" ")


def fortran_nint(x: float) -> int:
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


@dataclass
class AdvanceJob:
    """Run a session's solver through step ``target_cycle`` (inclusive);
    row_sum / samples / last_row accumulate the force rows of the segment."""
    session: "Session"
    kind: int
    target_cycle: int
    end_time: float
    arrival: float = 0.0
    row_sum: Optional[np.ndarray] = None
    samples: int = 0
    last_row: Optional[np.ndarray] = None
    err: float = 0.0
    tag: object = None

    def remaining(self) -> int:
        return self.target_cycle - self.session.solver.state()["n_next"] + 1

    def add(self, result) -> None:
        self.row_sum = result.row_sum.copy() if self.row_sum is None else self.row_sum + result.row_sum
        self.samples += result.n_steps
        self.last_row = result.last_row.copy()
        self.err = result.err


class Session:
    """One client of the online protocol bound to one solver instance in
    stream-motion mode (driver.f90:59-196)."""

    def __init__(self, solver, save_hook=None):
        self.solver = solver
        self.cfg = solver.cfg
        self.W = self.cfg.wing_count
        self.ncols = self.cfg.motion_cols
        self.active = False           # an episode is running (RESET done)
        self.start_row = np.zeros(self.ncols)
        self.target_row = np.zeros(self.ncols)
        self.stopped = False
        self.save_hook = save_hook    # callable(session, cycle) -> file name, for SAVE
        self.episode_index = 0

    # --------------------------------------------------------------- commands
    def receive(self, message: str, arrival: float = 0.0):
        """Returns (reply, None) for an immediate reply or (None, job)."""
        cmd, row, err = parse_command(message, self.W)
        if cmd == CMD_INVALID:
            return "ERROR " + err, None
        if cmd == CMD_STOP:
            self.stopped = True
            return "STATUS STOPPING", None
        if not self.active:
            if cmd == CMD_OUTPUT:
                return "ERROR " + ERR_SAVE_BEFORE_RESET, None
            if cmd != CMD_RESET:
                return "ERROR " + ERR_STEP_BEFORE_RESET, None
        if cmd == CMD_RESET:
            return self._reset(row, arrival)
        if cmd == CMD_OUTPUT:
            return self._save(), None
        if cmd == CMD_STEP:
            return self._step(row, arrival)
        return "ERROR " + ERR_EXPECTED, None

    def _reset(self, row, arrival):
        if abs(row[0]) > 1.0e-12:
            return "ERROR " + ERR_RESET_NONZERO, None
        self.episode_index += 1
        self.solver.reset()
        self.start_row = np.zeros(self.ncols)
        self.target_row = np.zeros(self.ncols)
        self.start_row[0] = self.target_row[0] = row[0]
        self.solver.set_stream_segment(self.start_row, self.target_row)
        self.active = True
        return self._segment(CMD_RESET, row[0], arrival, restore=None)

    def _step(self, row, arrival):
        start = self.target_row.copy()
        try:
            tgt = self.solver.build_stream_target_row(start, row[0], row[:1 + 3 * self.W])
        except Exception:
            return "ERROR " + ERR_SEGMENT_START, None
        self.start_row, self.target_row = start, tgt
        self.solver.set_stream_segment(self.start_row, self.target_row)
        return self._segment(CMD_STEP, row[0], arrival, restore=start)

    def _segment(self, kind, target_time, arrival, restore):
        """advance_zmq_segment (driver.f90:322-367) checks, then a job."""
        dt = self.cfg.dt
        if dt <= 0.0:
            return self._abort(ERR_DT, restore)
        target_cycle = fortran_nint(target_time / dt) + 1
        aligned = (target_cycle - 1) * dt
        if abs(aligned - target_time) > 1.0e-10:
            return self._abort(ERR_ALIGN, restore)
        if target_cycle < self.solver.state()["n_next"]:
            return self._abort(ERR_BEHIND, restore)
        return None, AdvanceJob(self, kind, target_cycle, float(target_time), arrival)

    def _abort(self, text, restore):
        if restore is not None:  # the target row of the failed STEP is dropped
            self.target_row = restore
            self.solver.set_stream_segment(self.start_row, self.target_row)
        return "ERROR " + text, None

    def _save(self) -> str:
        """save_current_flow_output (driver.f90:291-320)."""
        st = self.solver.state()
        if st["flag2"] != 0:
            return "ERROR " + ERR_SAVE_TRANSITION
        cycle = max(1, st["n_next"] - 1)
        name = f"F{cycle}.dat"
        if self.save_hook is not None:
            name = self.save_hook(self, cycle) or name
        return "STATUS FLOW_SAVED " + name

    # ---------------------------------------------------------------- replies
    def finish(self, job: AdvanceJob) -> str:
        """finalize_force_reply_average (driver.f90:270-289) + FORCE reply."""
        L = self.cfg.row_len
        if job.samples <= 0 or job.row_sum is None:
            return format_force_reply(np.zeros(0))
        row = job.last_row.copy()
        end = min(L, 1 + 3 * self.W + 3)
        row[1:end] = job.row_sum[1:end] / float(job.samples)
        row[0] = job.end_time
        return format_force_reply(row)


def run_job_alone(job: AdvanceJob, chunk: int = 0) -> str:
    """Run one job on its own solver (chunks of ``chunk`` steps, 0 = one call)
    and return the FORCE reply."""
    s = job.session.solver
    while job.remaining() > 0:
        c = job.remaining() if chunk <= 0 else min(chunk, job.remaining())
        job.add(s.advance(s.state()["n_next"] + c - 1))
    return job.session.finish(job)
