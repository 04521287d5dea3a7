"""Run-time configuration of one solver instance.

The reference keeps every physical and geometric knob as a compile-time
Fortran parameter (main.f90:4-159, ini_mesh.f90:15-67).  Here they are fields
of :class:`SolverConfig`, converted to the C struct ``lbm_config`` of
``include/lbm_hip.h`` by :meth:`SolverConfig.to_c`.

Also here: the NACA0012 marker generator (layout of the reference's
boundary.dat: leading edge, upper surface, trailing edge, lower surface, back
to the leading edge, pivot at the quarter chord) and the sinusoidal
pitch/heave motion table in the 19-column motion_params layout
(README.md section 5, mode.f90:180-218).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

MAX_WINGS = 8
MOTION_COLS_PER_WING = 9
MOTION_FILE = 1
MOTION_STREAM = 2
PRECISION_FP64 = 0
PRECISION_FP32_SHIFTED = 1


class LbmConfigC(ctypes.Structure):
    """ctypes mirror of ``lbm_config`` (include/lbm_hip.h)."""

    _fields_ = [
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("wing_count", ctypes.c_int32),
        ("points_per_wing", ctypes.c_int32),
        ("scale", ctypes.c_double),
        ("ox", ctypes.c_double),
        ("oy", ctypes.c_double),
        ("uu", ctypes.c_double),
        ("re", ctypes.c_double),
        ("s", ctypes.c_double * 9),
        ("wing_x0", ctypes.c_double),
        ("wing_dx", ctypes.c_double),
        ("wing_y0", ctypes.c_double),
        ("xc", ctypes.c_double),
        ("yc", ctypes.c_double),
        ("xp", ctypes.c_double),
        ("yp", ctypes.c_double),
        ("motion_source", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("xb0", ctypes.POINTER(ctypes.c_double)),
        ("yb0", ctypes.POINTER(ctypes.c_double)),
    ]


def naca0012_markers(n_intervals: int = 101, spacing: str = "reference",
                     thickness: float = 0.12) -> np.ndarray:
    """Closed NACA 00xx contour as (2*n_intervals+3, 2) chord-unit points.

    ``spacing="reference"`` reproduces the x stations of the reference's
    boundary.dat ({0, 0.25, 1, 2, ..., N}/N of the chord, N=101 gives its 205
    points); ``"cosine"`` uses x = (1-cos(pi*t))/2 on N+2 stations.
    The half-thickness uses the closed-trailing-edge coefficient -0.1036
    (boundary.dat ends at y=0).  The pivot (x=0) is at the quarter chord.
    """
    N = int(n_intervals)
    if N < 3:
        raise ValueError("n_intervals must be >= 3")
    if spacing == "reference":
        xs = np.concatenate([[0.0, 0.25], np.arange(1, N + 1, dtype=float)]) / N
    elif spacing == "cosine":
        t = np.arange(N + 3, dtype=float) / (N + 2)
        xs = 0.5 * (1.0 - np.cos(np.pi * t))
    else:
        raise ValueError(f"unknown spacing {spacing!r}")
    yt = 5.0 * thickness * (0.2969 * np.sqrt(xs) - 0.1260 * xs - 0.3516 * xs ** 2
                            + 0.2843 * xs ** 3 - 0.1036 * xs ** 4)
    yt[0] = 0.0
    yt[-1] = 0.0
    # the file stores 8 decimals ('(2f14.8)', airfoil.f90:229); the lower
    # surface is the rounded upper surface mirrored
    xs = np.round(xs - 0.25, 8)
    yt = np.round(yt, 8)
    upper = np.stack([xs, yt], axis=1)
    lower = np.stack([xs[::-1][1:], -yt[::-1][1:]], axis=1)
    return np.concatenate([upper, lower], axis=0)


@dataclass
class WingMotion:
    """Sinusoidal pitch/heave of one wing: angle = pitch*sin(w t + phase_p),
    y = y0 + heave*sin(w t + phase_h), w = 2*pi*freq (convective time)."""

    pitch: float = 0.0          # radians
    heave: float = 0.0          # chord units
    freq: float = 0.25          # cycles per convective time unit c/U
    phase_p: float = 0.0
    phase_h: float = 0.0


@dataclass
class SolverConfig:
    nx: int = 1024
    ny: int = 512
    scale: float = 1.0 / 64.0
    uu: float = 0.05
    re: float = 200.0
    s: Sequence[float] = (0.0, 1.4, 1.4, 0.0, 1.2, 0.0, 1.2, 0.0, 0.0)
    ox: Optional[float] = None
    oy: Optional[float] = None
    wing_count: int = 2
    markers: Optional[np.ndarray] = None     # (points, 2) per wing, same for all
    wing_markers: Optional[List[np.ndarray]] = None
    wing_x0: float = 0.0
    wing_dx: float = 1.5
    wing_y0: float = 0.5
    xc: float = 0.0
    yc: float = 0.0
    xp: float = 0.0
    yp: float = 0.0
    motion_source: int = MOTION_FILE
    precision: int = PRECISION_FP64
    name: str = "custom"
    _keep: list = field(default_factory=list, repr=False)

    @property
    def dt(self) -> float:
        """Solver step in convective time units (main.f90:375)."""
        return self.scale * self.uu

    @property
    def tau(self) -> float:
        niu = self.uu * (1.0 / self.scale) / self.re
        return 3.0 * niu + 0.5

    @property
    def chord_nodes(self) -> float:
        return 1.0 / self.scale

    def origin(self):
        ox = self.ox if self.ox is not None else -0.5 * (self.nx - 1) * self.scale
        oy = self.oy if self.oy is not None else -0.5 * (self.ny - 1) * self.scale
        return ox, oy

    def marker_array(self) -> np.ndarray:
        if self.wing_markers is not None:
            ms = [np.asarray(m, dtype=np.float64) for m in self.wing_markers]
        else:
            m = self.markers if self.markers is not None else naca0012_markers()
            ms = [np.asarray(m, dtype=np.float64)] * self.wing_count
        if len(ms) != self.wing_count:
            raise ValueError("one marker array per wing is required")
        if not ms:
            return np.zeros((0, 2))
        p = ms[0].shape[0]
        if any(mm.shape != (p, 2) for mm in ms):
            raise ValueError("all wings need the same number of markers")
        return np.concatenate(ms, axis=0) if ms else np.zeros((0, 2))

    @property
    def points_per_wing(self) -> int:
        if self.wing_count == 0:
            return 0
        return int(self.marker_array().shape[0] // self.wing_count)

    @property
    def motion_cols(self) -> int:
        return 1 + MOTION_COLS_PER_WING * self.wing_count

    @property
    def row_len(self) -> int:
        return 4 + 12 * self.wing_count

    def to_c(self) -> LbmConfigC:
        c = LbmConfigC()
        ox, oy = self.origin()
        m = self.marker_array()
        xb = np.ascontiguousarray(m[:, 0], dtype=np.float64) if m.size else np.zeros(1)
        yb = np.ascontiguousarray(m[:, 1], dtype=np.float64) if m.size else np.zeros(1)
        self._keep[:] = [xb, yb]
        c.nx, c.ny = int(self.nx), int(self.ny)
        c.wing_count = int(self.wing_count)
        c.points_per_wing = int(m.shape[0] // self.wing_count) if self.wing_count else 0
        c.scale, c.ox, c.oy = float(self.scale), float(ox), float(oy)
        c.uu, c.re = float(self.uu), float(self.re)
        for k in range(9):
            c.s[k] = float(self.s[k])
        c.wing_x0, c.wing_dx, c.wing_y0 = float(self.wing_x0), float(self.wing_dx), float(self.wing_y0)
        c.xc, c.yc, c.xp, c.yp = float(self.xc), float(self.yc), float(self.xp), float(self.yp)
        c.motion_source = int(self.motion_source)
        c.precision = int(self.precision)
        c.xb0 = xb.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        c.yb0 = yb.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        return c

    def n_steps_for(self, end_time: float) -> int:
        """Last step n of the file-mode loop: n runs while
        (n-1)*dt <= end_time + 1e-12 (main.f90:374-376)."""
        return int(math.floor((end_time + 1e-12) / self.dt)) + 1


def sinusoidal_motion_table(cfg: SolverConfig, motions: Sequence[WingMotion],
                            n_steps: int, every: int = 1) -> np.ndarray:
    """19-column (for 2 wings) motion_params rows sampled every ``every``
    solver steps over ``n_steps`` steps.  x/y are absolute positions as the
    file format requires (README.md section 5, mode.f90:210-215)."""
    if len(motions) != cfg.wing_count:
        raise ValueError("one WingMotion per wing")
    n_rows = (n_steps - 1) // every + 1
    t = np.arange(n_rows, dtype=np.float64) * (cfg.dt * every)
    rows = np.zeros((n_rows, cfg.motion_cols), dtype=np.float64)
    rows[:, 0] = t
    for wi, m in enumerate(motions):
        col = 1 + wi * MOTION_COLS_PER_WING
        om = 2.0 * math.pi * m.freq
        ap = om * t + m.phase_p
        ah = om * t + m.phase_h
        rows[:, col + 0] = m.pitch * np.sin(ap)
        rows[:, col + 1] = m.pitch * om * np.cos(ap)
        rows[:, col + 2] = -m.pitch * om * om * np.sin(ap)
        rows[:, col + 3] = cfg.wing_x0 + wi * cfg.wing_dx
        rows[:, col + 6] = cfg.wing_y0 + m.heave * np.sin(ah)
        rows[:, col + 7] = m.heave * om * np.cos(ah)
        rows[:, col + 8] = -m.heave * om * om * np.sin(ah)
    return rows


def seeded_motions(cfg: SolverConfig, seed: int, pitch_deg=(10.0, 10.0),
                   heave=(0.25, 0.25), freq: Optional[float] = None) -> List[WingMotion]:
    """Per-wing sinusoidal motions with seeded phases and amplitudes drawn
    uniformly from the given (low, high) ranges.  ``freq`` defaults to
    f = U/(4c) in lattice units, i.e. 0.25 per convective time unit."""
    rng = np.random.default_rng(seed)
    f = 0.25 if freq is None else freq
    out = []
    for _ in range(cfg.wing_count):
        p = math.radians(rng.uniform(*pitch_deg))
        h = rng.uniform(*heave)
        ph = rng.uniform(0.0, 2.0 * math.pi)
        out.append(WingMotion(pitch=p, heave=h, freq=f, phase_p=ph + 0.5 * math.pi, phase_h=ph))
    return out


def preset(name: str, precision: int = PRECISION_FP64) -> SolverConfig:
    """Named workloads.

    c1: 1024x512, c=64, U=0.05, Re=200, 2 wings x 127 markers.
    c2: 4096x2048, c=256, U=0.05, Re=1000, 2 wings x 511 markers.
    c5: 4001x2001 (40x20 chords at 0.01c), c=100, U=0.05, Re=1000, the
        reference's 205-marker wings.
    small: 256x128, c=32, Re=100, 2 wings x 67 markers (parity tests).
    """
    s = (0.0, 1.4, 1.4, 0.0, 1.2, 0.0, 1.2, 0.0, 0.0)
    if name == "c1":
        cfg = SolverConfig(nx=1024, ny=512, scale=1 / 64, uu=0.05, re=200.0, s=s,
                           markers=naca0012_markers(62), name="c1")
    elif name == "c2":
        cfg = SolverConfig(nx=4096, ny=2048, scale=1 / 256, uu=0.05, re=1000.0, s=s,
                           markers=naca0012_markers(254), name="c2")
    elif name == "c5":
        cfg = SolverConfig(nx=4001, ny=2001, scale=0.01, uu=0.05, re=1000.0,
                           s=(0.0, 1.1, 1.05, 0.0, 1.2, 0.0, 1.2, 0.0, 0.0),
                           ox=-20.0, oy=-10.0, markers=naca0012_markers(101), name="c5")
    elif name == "small":
        cfg = SolverConfig(nx=256, ny=128, scale=1 / 32, uu=0.05, re=100.0, s=s,
                           markers=naca0012_markers(32), name="small")
    else:
        raise ValueError(f"unknown preset {name!r}")
    cfg.precision = precision
    return cfg


def episode_config(nx: int, wings: int, chord: int, re: float, precision: int = PRECISION_FP64,
                   motion_source: int = MOTION_FILE, name: str = "episode", spacing: float = 1.5) -> SolverConfig:
    """One episode of the batched / served workloads: an nx x nx/2 lattice,
    ``wings`` NACA0012 wings of ``chord`` lattice units (markers about one
    lattice spacing apart) ``spacing`` chords apart, the first pivot 1.5
    chords from the inflow edge, centred vertically on y = 0.5."""
    ny = nx // 2
    c = float(chord)
    cfg = SolverConfig(nx=nx, ny=ny, scale=1.0 / c, uu=0.05, re=float(re),
                       s=(0.0, 1.4, 1.4, 0.0, 1.2, 0.0, 1.2, 0.0, 0.0), wing_count=wings,
                       markers=naca0012_markers(max(16, int(chord) - 2)), ox=-1.5,
                       oy=0.5 - 0.5 * (ny - 1) / c, motion_source=motion_source, precision=precision, name=name)
    cfg.wing_dx = float(spacing)
    return cfg


def random_episode(rng: np.random.Generator, precision: int = PRECISION_FP32_SHIFTED,
                   motion_source: int = MOTION_FILE, min_nx: int = 512, max_nx: int = 2048,
                   spacing: float = 1.5) -> SolverConfig:
    """Episode of the uneven batched workload: lattice log-uniform from
    min_nx x min_nx/2 to max_nx x max_nx/2 (2:1), 1-3 wings of chord 32-128
    lattice units ``spacing`` chords apart (chord capped so that the wings
    and 1.5 chords of margin fit), Re uniform in [100, 500]."""
    nx = int(round(math.exp(rng.uniform(math.log(min_nx), math.log(max_nx)))))
    nx -= nx % 2
    wings = int(rng.integers(1, 4))
    cmax = max(32, min(128, int(nx / (spacing * (wings - 1) + 4.0))))
    chord = int(rng.integers(32, cmax + 1))
    re = float(rng.uniform(100.0, 500.0))
    return episode_config(nx, wings, chord, re, precision, motion_source,
                          name=f"ep{nx}x{nx // 2}w{wings}c{chord}", spacing=spacing)

