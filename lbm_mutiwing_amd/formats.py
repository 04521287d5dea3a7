"""On-disk formats of the reference run, written and read by the host.

* ``force.dat`` / ``airfoil_motion.dat``: one ``(1000f15.8)`` record per step
  (solution.f90:386-402);
* ``motion_params.dat``: list-directed rows of 1 + 9*W columns, blank lines
  skipped, time column non-decreasing (load_motion_table, mode.f90:104-178);
* marker files ``boundary.dat``: ``(2f14.8)`` records, one per marker
  (read_boundary, airfoil.f90:207-242);
* the Tecplot flow snapshot ``F<n>.dat`` of SAVE / output (solution.f90:132-241):
  the flow zone sampled every second node with velocity, pressure and
  vorticity in ``(25e16.8)`` records, then the airfoil FE-triangle zone;
* the restart checkpoint ``saved_flow_feild.dat`` (saveflowfield /
  readflowfield, solution.f90:1616-1851): :func:`write_checkpoint`.

The offline (file-mode) run loop of the reference (main.f90:374-441) is
:func:`run_file_mode`: it drives a solver over a motion table and writes the
two per-step files.
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass
from typing import Iterable, List, Optional, TextIO

import numpy as np

from .protocol import list_items, read_reals

OUTPUT_STRIDE_X = 2     # main.f90:38
OUTPUT_STRIDE_Y = 2


# ------------------------------------------------------------ edit descriptors

def fortran_f(x: float, w: int, d: int) -> str:
    """Fortran Fw.d: right-justified, w asterisks when the value does not fit."""
    s = f"{float(x):{w}.{d}f}"
    if len(s) > w:
        return "*" * w
    return s


def fortran_e(x: float, w: int, d: int) -> str:
    """Fortran Ew.d (0.ddddE+xx form), right-justified."""
    x = float(x)
    if x == 0.0:
        body = "0." + "0" * d + "E+00"
        body = ("-" if np.signbit(x) else "") + body
    else:
        m, e = f"{abs(x):.{d - 1}e}".split("e")
        digits = m.replace(".", "")
        ex = int(e) + 1
        body = ("-" if x < 0 else "") + "0." + digits + (f"E{ex:+03d}" if abs(ex) < 100 else f"{ex:+04d}")
    if len(body) > w:
        return "*" * w
    return body.rjust(w)


def fortran_i(v: int, w: int, m: int = 1) -> str:
    """Fortran Iw.m: at least m digits, right-justified in w columns."""
    s = f"{abs(int(v)):0{m}d}"
    s = ("-" if v < 0 else "") + s
    return "*" * w if len(s) > w else s.rjust(w)


def f15_8_record(values: Iterable[float]) -> str:
    """One record of format '(1000f15.8)'."""
    return "".join(fortran_f(v, 15, 8) for v in values)


def e16_8_records(values: Iterable[float], per_line: int = 25) -> List[str]:
    vals = list(values)
    return ["".join(fortran_e(v, 16, 8) for v in vals[i:i + per_line]) for i in range(0, len(vals), per_line)]


def f14_8_record(values: Iterable[float]) -> str:
    return "".join(fortran_f(v, 14, 8) for v in values)


def list_directed(*items) -> str:
    """List-directed output in the usual style: a leading blank, strings as
    they are, integers right-justified in 12 columns, reals with 17
    significant digits."""
    out = []
    for it in items:
        if isinstance(it, str):
            out.append(it)
        elif isinstance(it, (int, np.integer)):
            out.append(f"{int(it):12d}")
        else:
            out.append(f"{float(it):25.17f}" if 1e-3 <= abs(float(it)) < 1e8 or float(it) == 0.0
                       else f"{float(it):25.16E}")
    return " " + "".join(out)


# --------------------------------------------------------------- motion table

class MotionFileError(ValueError):
    pass


def load_motion_params(path: str, wing_count: int) -> np.ndarray:
    """load_motion_table (mode.f90:104-178): rows of 1 + 9*W columns."""
    ncols = 1 + 9 * wing_count
    if not os.path.exists(path):
        raise MotionFileError(f"无法打开运动参数文件: {path}")
    rows = []
    with open(path) as fh:
        for line in fh:
            if not line.strip():
                continue
            items, slash = list_items(line.rstrip("\n"))
            vals = np.zeros(ncols)
            if not read_reals(items, slash, vals, ncols):
                raise MotionFileError(f"运动参数文件格式错误，列数应为: {ncols} {path}")
            rows.append(vals)
    if not rows:
        raise MotionFileError(f"运动参数文件为空: {path}")
    tab = np.array(rows, dtype=np.float64)
    if np.any(np.diff(tab[:, 0]) < 0.0):
        raise MotionFileError(f"运动参数文件时间列必须非递减: {path}")
    return tab


def write_motion_params(path: str, rows: np.ndarray) -> None:
    with open(path, "w") as fh:
        for r in rows:
            fh.write(" ".join(repr(float(v)) for v in r) + "\n")


# -------------------------------------------------------------------- markers

def read_marker_file(path: str, points: Optional[int] = None) -> np.ndarray:
    """read_boundary (airfoil.f90:229): (x, y) per record, chord units."""
    out = []
    with open(path) as fh:
        for line in fh:
            if not line.strip():
                continue
            x, y = line.split()[:2]
            out.append((float(x), float(y)))
            if points is not None and len(out) == points:
                break
    return np.array(out, dtype=np.float64)


def write_marker_file(path: str, markers: np.ndarray) -> None:
    with open(path, "w") as fh:
        for x, y in markers:
            fh.write(f14_8_record((x, y)) + "\n")


# -------------------------------------------------------- per-step row files

class RowFile:
    """force.dat / airfoil_motion.dat: one '(1000f15.8)' record per row."""

    def __init__(self, path: str, mode: str = "w"):
        self.fh: TextIO = open(path, mode)

    def write_rows(self, rows: np.ndarray) -> None:
        for r in np.atleast_2d(rows):
            self.fh.write(f15_8_record(r) + "\n")

    def close(self) -> None:
        self.fh.close()


def run_file_mode(solver, table: np.ndarray, out_dir: str, chunk: int = 100) -> float:
    """The offline run loop (main.f90:374-441): advance while
    (n-1)*dt <= motion_end_time, writing force.dat and airfoil_motion.dat.
    Returns the last erranalysis value."""
    os.makedirs(out_dir, exist_ok=True)
    end_time = solver.set_motion_table(table)
    n_end = solver.cfg.n_steps_for(end_time)
    ff = RowFile(os.path.join(out_dir, "force.dat"))
    fm = RowFile(os.path.join(out_dir, "airfoil_motion.dat"))
    err = 0.0
    try:
        while solver.state()["n_next"] <= n_end:
            last = min(n_end, solver.state()["n_next"] + chunk - 1)
            r = solver.advance(last, want_rows=True, want_motion=True)
            ff.write_rows(r.force_rows)
            fm.write_rows(r.motion_rows)
            err = r.err
    finally:
        ff.close()
        fm.close()
    return err


# ---------------------------------------------------------------- Tecplot

def vorticity(u: np.ndarray, v: np.ndarray, uu: float, scale: float) -> np.ndarray:
    """dv/dx - du/dy in units of U per chord (velocities in lattice units,
    scale = chord per node), central inside and second-order one-sided on
    the edges (solution.f90:183-207).  Arrays are [j, i]."""
    ny, nx = u.shape
    dvx = np.empty_like(v)
    dvx[:, 1:-1] = (v[:, 2:] - v[:, :-2]) / (uu * 2.0 * scale)
    dvx[:, -1] = (1.5 * v[:, -1] - 2.0 * v[:, -2] + 0.5 * v[:, -3]) / (uu * scale)
    dvx[:, 0] = (-1.5 * v[:, 0] + 2.0 * v[:, 1] - 0.5 * v[:, 2]) / (uu * scale)
    duy = np.empty_like(u)
    duy[1:-1, :] = (u[2:, :] - u[:-2, :]) / (uu * 2.0 * scale)
    duy[-1, :] = (1.5 * u[-1, :] - 2.0 * u[-2, :] + 0.5 * u[-3, :]) / (uu * scale)
    duy[0, :] = (-1.5 * u[0, :] + 2.0 * u[1, :] - 0.5 * u[2, :]) / (uu * scale)
    return dvx - duy


def body_map(nx: int, ny: int, adjoint_ij: np.ndarray) -> np.ndarray:
    """The reference's body flags, [j, i]: 1 at the adjoint nodes of the last
    step (1-based (i, j) pairs), 0 elsewhere (airfoil.f90:276-329)."""
    body = np.zeros((ny, nx), dtype=np.int32)
    ij = np.asarray(adjoint_ij, dtype=np.int64).reshape(-1, 2)
    if ij.size:
        body[ij[:, 1] - 1, ij[:, 0] - 1] = 1
    return body


def write_tecplot(path: str, cfg, cycle: int, u: np.ndarray, v: np.ndarray, pre: np.ndarray,
                  xb1: np.ndarray, yb1: np.ndarray, movenum: int = 0,
                  body: Optional[np.ndarray] = None) -> None:
    """output (solution.f90:132-241) for one lattice: fields are [j, i];
    body is the flag map of :func:`body_map` (zeros when None)."""
    ny, nx = u.shape
    ox, oy = cfg.origin()
    sc, uu = cfg.scale, cfg.uu
    vor = vorticity(u, v, uu, sc)
    out_nx = (nx - 1) // OUTPUT_STRIDE_X + 1
    out_ny = (ny - 1) // OUTPUT_STRIDE_Y + 1
    tsol = cycle * uu * sc
    W, P = cfg.wing_count, cfg.points_per_wing
    with open(path, "w") as fh:
        fh.write('VARIABLES = "X" "Y" "U" "V" "Presure" "Vorticity" "body" "dvordt" \n')
        fh.write(list_directed('ZONE T="Flow feild", I=', out_nx, ',J=', out_ny, ',SOLUTIONTIME=', tsol) + "\n")
        jj = np.arange(0, ny, OUTPUT_STRIDE_Y)
        ii = np.arange(0, nx, OUTPUT_STRIDE_X)
        for j in jj:
            y = oy + j * sc
            for i in ii:
                x = ox + i * sc + movenum * sc
                b = 0.0 if body is None else float(body[j, i])
                fh.write(e16_8_records((x, y, u[j, i] / uu, v[j, i] / uu, pre[j, i], vor[j, i], b, 0.0))[0] + "\n")
        npts = W * P
        fh.write(list_directed('ZONE T="airfoil", NODES=', npts + W, ' ELEMENTS,=', W * (P - 1),
                               ', DATAPACKING=POINT, ZONETYPE=FETRIANGLE', ',SOLUTIONTIME=', tsol) + "\n")
        for q in range(npts):
            fh.write(f14_8_record((xb1[q] + movenum * sc, yb1[q], 255.0, 0, 0, 0, 0, 0)) + "\n")
        for w in range(W):
            j = w * P
            k = j + (P - 1) // 2
            fh.write(f14_8_record((0.5 * ((xb1[j] + movenum * sc) + (xb1[k] + movenum * sc)),
                                   0.5 * (yb1[j] + yb1[k]), 255.0, 0, 0, 0, 0, 0)) + "\n")
        for w in range(W):
            j = w * P + 1
            for k in range(j, j + P - 1):
                fh.write("".join(fortran_i(val, 5, 3) for val in (k, k + 1, npts + w + 1)) + "\n")


def tecplot_save_hook(out_dir: str):
    """SAVE hook for protocol sessions: writes F<cycle>.dat into out_dir."""
    os.makedirs(out_dir, exist_ok=True)

    def hook(session, cycle):
        s = session.solver
        rho, u, v, pre = s.fields()
        xb1, yb1, _, _ = s.markers()
        body = body_map(s.cfg.nx, s.cfg.ny, s.ibm()[0])
        name = f"F{cycle}.dat"
        write_tecplot(os.path.join(out_dir, name), s.cfg, cycle, u, v, pre, xb1, yb1, s.state()["movenum"], body)
        return name

    return hook


# ------------------------------------------------------------ checkpoint

RING_NAMES = ("left2", "f0_left", "right2", "f0_right", "up2", "f0_top", "down2", "f0_bottom")


@dataclass
class Checkpoint:
    """Contents of a restart checkpoint.  u, v: (ny, nx); f: (9, ny, nx);
    body_rel: (11, W), the rigid-body arrays in file order (body_rel_a, _a_prev,
    _v, _v_prev, _s, _s_prev, _pos, _theta_prev, _theta, _temp_theta_prev,
    _temp_theta); ring: the previous-step boundary populations and the f0
    ring, ``left2``/``f0_left``/``right2``/``f0_right`` (9, ny) (rows 2..ny-1
    are stored) and ``up2``/``f0_top``/``down2``/``f0_bottom`` (9, nx)."""
    cycle: int
    movenum: int
    body_rel: np.ndarray
    u: np.ndarray
    v: np.ndarray
    f: np.ndarray
    ring: dict


def checkpoint_bytes(wing_count: int, nx: int, ny: int) -> int:
    return 12 + 88 * wing_count + 88 * nx * ny + 288 * (ny - 2) + 288 * nx


def ring_from_populations(f: np.ndarray) -> dict:
    """The ring records as init leaves them, from the populations f
    (9, ny, nx) just after init (solution.f90:55-63)."""
    return {"left2": f[:, :, 0].copy(), "f0_left": f[:, :, 0].copy(), "right2": f[:, :, -1].copy(),
            "f0_right": f[:, :, -1].copy(), "up2": f[:, -1, :].copy(), "f0_top": f[:, -1, :].copy(),
            "down2": f[:, 0, :].copy(), "f0_bottom": f[:, 0, :].copy()}


def write_checkpoint(path: str, ck: Checkpoint) -> None:
    """saveflowfield (solution.f90:1660-1686): int32 step, int32 wing count,
    11*W doubles, int32 movenum, then (u, v, f1..f9) per node with i fastest,
    then per row j = 2..ny-1 (leftf2, f0(1,j), rightf2, f0(nx,j)), then per
    column i = 1..nx (upf2, f0(i,ny), downf2, f0(i,1)); 9 doubles each, no
    record markers (form='binary')."""
    ny, nx = ck.u.shape
    W = ck.body_rel.shape[1]
    node = np.empty((ny, nx, 11))
    node[:, :, 0] = ck.u
    node[:, :, 1] = ck.v
    node[:, :, 2:] = np.moveaxis(ck.f, 0, -1)
    r = ck.ring
    rows = np.stack([r["left2"].T, r["f0_left"].T, r["right2"].T, r["f0_right"].T], axis=1)[1:ny - 1]
    cols = np.stack([r["up2"].T, r["f0_top"].T, r["down2"].T, r["f0_bottom"].T], axis=1)
    with open(path, "wb") as fh:
        fh.write(struct.pack("<ii", ck.cycle, W))
        fh.write(np.ascontiguousarray(ck.body_rel, dtype="<f8").tobytes())
        fh.write(struct.pack("<i", ck.movenum))
        fh.write(node.astype("<f8").tobytes())
        fh.write(np.ascontiguousarray(rows, dtype="<f8").tobytes())
        fh.write(np.ascontiguousarray(cols, dtype="<f8").tobytes())


def read_checkpoint(path: str, wing_count: int, nx: int, ny: int) -> Checkpoint:
    """readflowfield (solution.f90:1721-1787) for a case of wing_count wings
    on an nx x ny lattice; raises ValueError on a size or wing-count
    mismatch."""
    data = open(path, "rb").read()
    if len(data) != checkpoint_bytes(wing_count, nx, ny):
        raise ValueError(f"checkpoint size {len(data)} does not fit {nx}x{ny} with {wing_count} wings")
    cycle, W = struct.unpack_from("<ii", data, 0)
    if W != wing_count:
        raise ValueError("wing count of the checkpoint differs from the case")
    o = 8
    body_rel = np.frombuffer(data, "<f8", 11 * W, o).reshape(11, W).copy()
    o += 88 * W
    (movenum,) = struct.unpack_from("<i", data, o)
    o += 4
    node = np.frombuffer(data, "<f8", 11 * nx * ny, o).reshape(ny, nx, 11)
    o += 88 * nx * ny
    rows = np.frombuffer(data, "<f8", 36 * (ny - 2), o).reshape(ny - 2, 4, 9)
    o += 288 * (ny - 2)
    cols = np.frombuffer(data, "<f8", 36 * nx, o).reshape(nx, 4, 9)
    ring = {}
    for q, name in enumerate(("left2", "f0_left", "right2", "f0_right")):
        a = np.zeros((9, ny))
        a[:, 1:ny - 1] = rows[:, q, :].T
        ring[name] = a
    for q, name in enumerate(("up2", "f0_top", "down2", "f0_bottom")):
        ring[name] = cols[:, q, :].T.copy()
    # the f0 ring's corners are the end columns' rows 1 and ny
    ring["f0_left"][:, 0], ring["f0_left"][:, -1] = ring["f0_bottom"][:, 0], ring["f0_top"][:, 0]
    ring["f0_right"][:, 0], ring["f0_right"][:, -1] = ring["f0_bottom"][:, -1], ring["f0_top"][:, -1]
    return Checkpoint(cycle, movenum, body_rel, node[:, :, 0].copy(), node[:, :, 1].copy(),
                      np.moveaxis(node[:, :, 2:], -1, 0).copy(), ring)

