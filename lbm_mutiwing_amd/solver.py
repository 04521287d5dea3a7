"""Host-side handle of one solver instance behind the C ABI (include/lbm_hip.h).

The reference's host program calls evolution/erranalysis/move_mesh/init
directly (main.f90:315-411, driver.f90:205-408).  :class:`Solver` exposes the
same operations over the C ABI: ``reset`` (init + read_boundary),
``set_motion_table`` / ``set_stream_segment`` (mode.f90 motion sources),
``advance`` (the step loop up to a target step) and field / marker / IBM
readback for SAVE and tests.

The class is bound to a function table ``lib`` with a symbol prefix; the
product binds it to liblbm_hip.so (prefix ``lbm_``).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .config import SolverConfig

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)

_PROTOS = {
    "create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "reset": (ctypes.c_int, [ctypes.c_void_p]),
    "set_motion_table": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_int32, ctypes.c_int32, _dp]),
    "set_stream_segment": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, ctypes.c_int32]),
    "build_stream_target_row": (ctypes.c_int, [ctypes.c_int32, _dp, ctypes.c_double, _dp, _dp]),
    "advance": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _dp, _dp, _dp, _dp, _dp]),
    "get_state": (ctypes.c_int, [ctypes.c_void_p, _ip, _ip, _ip, _ip, _ip]),
    "get_fields": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _dp]),
    "get_populations": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "set_populations": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "get_markers": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _dp]),
    "get_ibm": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _ip, _ip, _dp, _dp]),
    "set_body_rel_pos": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double]),
    "get_body_rel": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "set_body_rel": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "set_dynamics": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "set_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "get_mass": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "last_error": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]),
}


IBM_SOLVE_INVERSE, IBM_SOLVE_COUPLED, IBM_SOLVE_SUBSTITUTION = 0, 1, 2   # lbm_set_ibm_solve

BODY_REL_FIELDS = ("a", "a_prev", "v", "v_prev", "s", "s_prev", "pos", "theta_prev", "theta", "temp_theta_prev",
                   "temp_theta")


class Dynamics(ctypes.Structure):
    """lbm_dynamics: rigid-body coupling of wing 1 (boundary_condition.f90:
    1694-1719; constants main.f90:98, 122-151, whose values are the
    defaults here, with the solver switched on)."""
    _fields_ = [("enable", ctypes.c_int32), ("closed_boundary", ctypes.c_int32), ("rk4", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("solid_den", ctypes.c_double), ("area", ctypes.c_double),
                ("par1", ctypes.c_double), ("par2", ctypes.c_double), ("alphad", ctypes.c_double)]

    def __init__(self, enable=True, closed_boundary=True, rk4=True, solid_den=10.0, area=6.807946476488794e-2,
                 par1=0.0, par2=0.0, alphad=10.0 * 3.1415926535897932384626433832795 / 180.0):
        super().__init__(int(bool(enable)), int(bool(closed_boundary)), int(bool(rk4)), 0, float(solid_den),
                         float(area), float(par1), float(par2), float(alphad))


def bind(lib: ctypes.CDLL, prefix: str) -> None:
    """Attach argument/result types to every C-ABI entry point of ``lib``."""
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, prefix + name)
        fn.restype = res
        fn.argtypes = args


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_dp)


class LbmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


@dataclass
class AdvanceResult:
    force_rows: Optional[np.ndarray]
    motion_rows: Optional[np.ndarray]
    row_sum: np.ndarray
    last_row: np.ndarray
    err: float
    n_steps: int


class Solver:
    """One solver instance (one episode's lattice, IBM and motion state)."""

    prefix = "lbm_"

    def __init__(self, cfg: SolverConfig, lib: Optional[ctypes.CDLL] = None):
        if lib is None:
            from ._native import load_library
            lib = load_library()
        self.lib = lib
        self.cfg = cfg
        self._h = ctypes.c_void_p()
        c = cfg.to_c()
        rc = self._fn("create")(ctypes.byref(c), ctypes.byref(self._h))
        if rc != 0:
            raise LbmError(rc, self._msg(None))
        self.nx, self.ny = cfg.nx, cfg.ny
        self.row_len = cfg.row_len

    # ------------------------------------------------------------ plumbing
    def _fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def _msg(self, h) -> str:
        buf = ctypes.create_string_buffer(512)
        self._fn("last_error")(h, buf, 512)
        return buf.value.decode(errors="replace")

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise LbmError(rc, self._msg(self._h))

    def close(self) -> None:
        if self._h:
            self._fn("destroy")(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ----------------------------------------------------------- operations
    def reset(self) -> None:
        self._check(self._fn("reset")(self._h))

    def set_motion_table(self, rows: np.ndarray) -> float:
        rows = np.ascontiguousarray(rows, dtype=np.float64)
        end = ctypes.c_double()
        self._check(self._fn("set_motion_table")(self._h, _p(rows), rows.shape[0], rows.shape[1],
                                                 ctypes.byref(end)))
        return end.value

    def set_stream_segment(self, start_row: np.ndarray, target_row: np.ndarray) -> None:
        s = np.ascontiguousarray(start_row, dtype=np.float64)
        t = np.ascontiguousarray(target_row, dtype=np.float64)
        self._check(self._fn("set_stream_segment")(self._h, _p(s), _p(t), s.shape[0]))

    def build_stream_target_row(self, start_row: np.ndarray, target_time: float,
                                delta_row: np.ndarray) -> np.ndarray:
        s = np.ascontiguousarray(start_row, dtype=np.float64)
        d = np.zeros(1 + 3 * self.cfg.wing_count, dtype=np.float64)
        dd = np.asarray(delta_row, dtype=np.float64)
        d[: min(d.size, dd.size)] = dd[: d.size]
        out = np.zeros_like(s)
        rc = self._fn("build_stream_target_row")(self.cfg.wing_count, _p(s), float(target_time),
                                                 _p(d), _p(out))
        if rc != 0:
            raise LbmError(rc, "STEP time is before the segment start")
        return out

    def state(self):
        vals = [ctypes.c_int32() for _ in range(5)]
        self._check(self._fn("get_state")(self._h, *[ctypes.byref(v) for v in vals]))
        keys = ("n_next", "movenum", "flag2", "row_len", "adjoint_n")
        return dict(zip(keys, (v.value for v in vals)))

    def _n_next(self) -> int:
        n = ctypes.c_int32()
        self._check(self._fn("get_state")(self._h, ctypes.byref(n), None, None, None, None))
        return n.value

    def advance(self, n_last: int, want_rows: bool = False, want_motion: bool = False) -> AdvanceResult:
        nsteps = max(0, n_last - self._n_next() + 1)
        rows = np.zeros((max(nsteps, 1), self.row_len)) if want_rows else None
        mrows = np.zeros((max(nsteps, 1), 1 + 3 * self.cfg.wing_count)) if want_motion else None
        rsum = np.zeros(self.row_len)
        last = np.zeros(self.row_len)
        err = ctypes.c_double()
        self._check(self._fn("advance")(self._h, int(n_last), _p(rows), _p(mrows), _p(rsum),
                                        _p(last), ctypes.byref(err)))
        if rows is not None:
            rows = rows[:nsteps]
        if mrows is not None:
            mrows = mrows[:nsteps]
        return AdvanceResult(rows, mrows, rsum, last, err.value, nsteps)

    def fields(self):
        n = self.nx * self.ny
        rho, u, v, pre = (np.zeros(n) for _ in range(4))
        self._check(self._fn("get_fields")(self._h, _p(rho), _p(u), _p(v), _p(pre)))
        shp = (self.ny, self.nx)  # [j, i]: i fastest, as the C layout
        return rho.reshape(shp), u.reshape(shp), v.reshape(shp), pre.reshape(shp)

    def populations(self):
        n = self.nx * self.ny
        f = np.zeros(9 * n)
        f0 = np.zeros(9 * n)
        self._check(self._fn("get_populations")(self._h, _p(f), _p(f0)))
        return f.reshape(9, self.ny, self.nx), f0.reshape(9, self.ny, self.nx)

    def set_populations(self, f: np.ndarray, f0: Optional[np.ndarray] = None) -> None:
        fa = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        f0a = None if f0 is None else np.ascontiguousarray(f0, dtype=np.float64).reshape(-1)
        self._check(self._fn("set_populations")(self._h, _p(fa), _p(f0a)))

    def mass(self) -> float:
        """Sum over every node of its nine populations after the last step,
        reduced in FP64 (lbm_get_mass)."""
        m = ctypes.c_double()
        self._check(self._fn("get_mass")(self._h, ctypes.byref(m)))
        return m.value

    def markers(self):
        npts = self.cfg.wing_count * self.cfg.points_per_wing
        nseg = self.cfg.wing_count * max(self.cfg.points_per_wing - 1, 0)
        xb1, yb1 = np.zeros(max(npts, 1)), np.zeros(max(npts, 1))
        xb2, yb2 = np.zeros(max(nseg, 1)), np.zeros(max(nseg, 1))
        self._check(self._fn("get_markers")(self._h, _p(xb1), _p(yb1), _p(xb2), _p(yb2)))
        return xb1[:npts], yb1[:npts], xb2[:nseg], yb2[:nseg]

    def ibm(self):
        n = self.state()["adjoint_n"]
        ij = np.zeros(2 * max(n, 1), dtype=np.int32)
        wid = np.zeros(max(n, 1), dtype=np.int32)
        wf = np.zeros(16)
        wm = np.zeros(8)
        self._check(self._fn("get_ibm")(self._h, n, ij.ctypes.data_as(_ip), wid.ctypes.data_as(_ip),
                                        _p(wf), _p(wm)))
        return ij[: 2 * n].reshape(n, 2), wid[:n], wf.reshape(8, 2), wm

    def set_step(self, n_next: int, movenum: int = 0) -> None:
        """Restart: the next step is ``n_next``, the frame shift ``movenum``
        (readflowfield, solution.f90:1721-1851).  Call before
        :meth:`set_populations`."""
        self._check(self._fn("set_step")(self._h, int(n_next), int(movenum)))

    def set_body_rel_pos(self, wing: int, pos: float) -> None:
        self._check(self._fn("set_body_rel_pos")(self._h, int(wing), float(pos)))

    def body_rel(self) -> np.ndarray:
        """body_rel_* as an (11, wing_count) array, rows in restart-file
        order (BODY_REL_FIELDS)."""
        out = np.zeros(11 * max(self.cfg.wing_count, 1))
        self._check(self._fn("get_body_rel")(self._h, _p(out)))
        return out[: 11 * self.cfg.wing_count].reshape(11, self.cfg.wing_count)

    def set_body_rel(self, rel) -> None:
        a = np.ascontiguousarray(np.asarray(rel, dtype=np.float64).reshape(11, self.cfg.wing_count))
        self._check(self._fn("set_body_rel")(self._h, _p(a)))

    def set_dynamics(self, dyn: Optional[Dynamics]) -> None:
        """Switch the rigid-body coupling on (a :class:`Dynamics`) or off
        (None)."""
        d = dyn if dyn is not None else Dynamics(enable=False)
        self._check(self._fn("set_dynamics")(self._h, ctypes.byref(d)))

    def set_ibm_solve(self, mode: int) -> None:
        """How the IBM system is solved at flow time (lbm_set_ibm_solve):
        IBM_SOLVE_INVERSE, IBM_SOLVE_COUPLED (the default) or
        IBM_SOLVE_SUBSTITUTION.  A library without the entry point (the CPU
        oracle, which solves its dense system directly) has no modes."""
        fn = getattr(self.lib, "lbm_set_ibm_solve", None)
        if fn is None:
            return
        fn.restype, fn.argtypes = ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]
        self._check(fn(self._h, int(mode)))

    def tuning(self) -> int:
        """Lattice tile width chosen by the autotuner (HIP library only)."""
        return self.tuning_full()[0]

    def tuning_full(self):
        """(tile width, cap on lattice waves per CU, 0 = none) chosen by the
        autotuner (HIP library only)."""
        fn = self.lib.lbm_get_tuning
        fn.restype, fn.argtypes = ctypes.c_int, [ctypes.c_void_p, _ip, _ip]
        v, w = ctypes.c_int32(), ctypes.c_int32()
        self._check(fn(self._h, ctypes.byref(v), ctypes.byref(w)))
        return v.value, w.value

    def cache_policy(self) -> int:
        """Cache policy of the lattice kernel's population accesses chosen by
        the autotuner: 0 ordinary, 1 non-temporal stores, 2 non-temporal
        loads and stores (HIP library only)."""
        fn = self.lib.lbm_get_lattice_policy
        fn.restype, fn.argtypes = ctypes.c_int, [ctypes.c_void_p, _ip]
        c = ctypes.c_int32()
        self._check(fn(self._h, ctypes.byref(c)))
        return c.value


class TimedSolver(Solver):
    """Solver with the lattice-kernel event timing of the HIP library."""

    def set_timing(self, enable: bool) -> None:
        fn = self.lib.lbm_set_timing
        fn.restype, fn.argtypes = ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]
        self._check(fn(self._h, 1 if enable else 0))

    def timing(self):
        fn = self.lib.lbm_get_timing
        fn.restype = ctypes.c_int
        fn.argtypes = [ctypes.c_void_p, _dp, _ip]
        ms, n = ctypes.c_double(), ctypes.c_int32()
        self._check(fn(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


class Batch:
    """Batched advance of several instances of the HIP library
    (lbm_batch_advance): one launch per kernel phase over all instances."""

    MAX_ROW = 4 + 12 * 8

    def __init__(self, lib: Optional[ctypes.CDLL] = None):
        if lib is None:
            from ._native import load_library
            lib = load_library()
        self.lib = lib
        lib.lbm_batch_create.restype = ctypes.c_int
        lib.lbm_batch_create.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        lib.lbm_batch_destroy.restype = ctypes.c_int
        lib.lbm_batch_destroy.argtypes = [ctypes.c_void_p]
        lib.lbm_batch_advance.restype = ctypes.c_int
        lib.lbm_batch_advance.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32,
                                          ctypes.c_int32, _dp, _dp, _dp, _ip]
        lib.lbm_batch_last_error.restype = ctypes.c_int
        lib.lbm_batch_last_error.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
        self._b = ctypes.c_void_p()
        rc = lib.lbm_batch_create(ctypes.byref(self._b))
        if rc != 0:
            raise LbmError(rc, "lbm_batch_create failed")

    def close(self) -> None:
        if self._b:
            self.lib.lbm_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def tuning(self) -> int:
        """Lattice tile width chosen by the batch's autotuner."""
        return self.tuning_full()[0]

    def tuning_full(self):
        """(tile width, tiles per workgroup, cap on lattice waves per CU)
        chosen by the batch's autotuner."""
        fn = self.lib.lbm_batch_get_tuning
        fn.restype, fn.argtypes = ctypes.c_int, [ctypes.c_void_p, _ip, _ip, _ip]
        v, r, w = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        fn(self._b, ctypes.byref(v), ctypes.byref(r), ctypes.byref(w))
        return v.value, r.value, w.value

    def cache_policy(self) -> int:
        """Cache policy of the batched lattice kernel (0 ordinary, 2
        non-temporal loads and stores)."""
        fn = self.lib.lbm_batch_get_lattice_policy
        fn.restype, fn.argtypes = ctypes.c_int, [ctypes.c_void_p, _ip]
        c = ctypes.c_int32()
        fn(self._b, ctypes.byref(c))
        return c.value

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def advance(self, solvers, steps: int, partial: bool = False):
        """Advance every solver by `steps` steps; returns one AdvanceResult per
        solver (row_sum, last_row, err; no per-step rows).  A step error of any
        instance raises LbmError, unless ``partial``: then the failing
        instances get an LbmError object in their place."""
        n = len(solvers)
        hs = (ctypes.c_void_p * n)(*[s._h.value for s in solvers])
        rs = np.zeros((n, self.MAX_ROW))
        lr = np.zeros((n, self.MAX_ROW))
        er = np.zeros(n)
        st = np.zeros(n, dtype=np.int32)
        rc = self.lib.lbm_batch_advance(self._b, hs, n, int(steps), _p(rs), _p(lr), _p(er),
                                        st.ctypes.data_as(_ip))
        if rc != 0 and not (partial and np.any(st != 0)):
            buf = ctypes.create_string_buffer(512)
            self.lib.lbm_batch_last_error(self._b, buf, 512)
            raise LbmError(rc, buf.value.decode(errors="replace"))
        out = []
        for e, s in enumerate(solvers):
            if st[e] != 0:
                out.append(LbmError(int(st[e]), s._msg(s._h)))
                continue
            L = s.row_len
            out.append(AdvanceResult(None, None, rs[e, :L].copy(), lr[e, :L].copy(), float(er[e]), int(steps)))
        return out

