"""The host's case file: the namelist /lbm_case/ that the Fortran host
(lbm_mutiwing_amd/fortran/lbm_host.f90) reads, for the Python hosts.

The reference fixes its case at compile time (main.f90:4-159,
ini_mesh.f90:15-67); the hosts here read it from one namelist group:

    &lbm_case nx=4001, ny=2001, scale=0.01, ox=-20.0, oy=-10.0, uu=0.05,
      re=1000.0, s=0.0, 1.1, 1.05, 0.0, 1.2, 0.0, 1.2, 0.0, 0.0,
      wing_count=2, points_per_wing=205, boundary_dir='.', output_dir='out' /

This reader takes the forms the hosts' files use: ``name=value`` items
separated by commas or blanks, values continuing an array after its first
one, integers, reals (``d`` exponents allowed), ``.true.``/``.false.`` and
quoted strings.  Unknown names are an error, as in Fortran namelist input.
"""
from __future__ import annotations

import os
import re
from typing import Dict, List

from .config import MOTION_FILE, SolverConfig

# name -> (kind, default) as declared in lbm_host.f90
FIELDS = {
    "nx": ("int", 0), "ny": ("int", 0), "wing_count": ("int", 2), "points_per_wing": ("int", 0),
    "precision": ("int", 0), "rows_per_call": ("int", 100), "print_every": ("int", 100),
    "checkpoint_every": ("int", 100000),
    "scale": ("real", 0.01), "ox": ("real", 0.0), "oy": ("real", 0.0), "uu": ("real", 0.05),
    "re": ("real", 1000.0), "wing_x0": ("real", 0.0), "wing_dx": ("real", 1.5), "wing_y0": ("real", 0.5),
    "xc": ("real", 0.0), "yc": ("real", 0.0), "xp": ("real", 0.0), "yp": ("real", 0.0),
    "s": ("real9", [0.0, 1.4, 1.4, 0.0, 1.2, 0.0, 1.2, 0.0, 0.0]),
    "restart": ("logical", False),
    "boundary_dir": ("str", "."), "motion_file": ("str", "motion_params.dat"), "output_dir": ("str", "output"),
    "restart_dir": ("str", ""),
    "solve_dynamic_equations": ("logical", False), "closed_boundary": ("logical", True), "if_rk4": ("logical", True),
    "solid_den": ("real", 10.0), "area_airfoil": ("real", 6.807946476488794e-2), "par1": ("real", 0.0),
    "par2": ("real", 0.0), "alphad": ("real", 10.0 * 3.1415926535897932384626433832795 / 180.0),
}

_TOKEN = re.compile(r"'(?:[^']|'')*'|\"(?:[^\"]|\"\")*\"|[^\s,=]+=?|=")


def _value(kind: str, tok: str):
    if kind == "str":
        if tok[:1] in "'\"":
            q = tok[0]
            return tok[1:-1].replace(q + q, q)
        return tok
    if kind == "logical":
        t = tok.lower().lstrip(".")
        if t[:1] in ("t", "f"):
            return t[0] == "t"
        raise ValueError(f"bad logical value {tok!r}")
    v = tok.lower().replace("d", "e")
    return int(v) if kind == "int" else float(v)


def read_case(path: str) -> Dict[str, object]:
    """The namelist group /lbm_case/ of the file, over the hosts' defaults."""
    text = open(path).read()
    m = re.search(r"&lbm_case\b(.*?)(?:^|\s)/", text, re.S | re.I)
    if m is None:
        raise ValueError(f"no namelist group /lbm_case/ in {path}")
    case = {k: (list(d) if isinstance(d, list) else d) for k, (_, d) in FIELDS.items()}
    name, index = None, 0
    toks: List[str] = _TOKEN.findall(m.group(1))
    i = 0
    while i < len(toks):
        t = toks[i]
        if t.endswith("=") and t != "=" or (i + 1 < len(toks) and toks[i + 1] == "="):
            name = t.rstrip("=").lower()
            if name not in FIELDS:
                raise ValueError(f"unknown name {name!r} in /lbm_case/")
            index = 0
            i += 1 if t.endswith("=") else 2
            continue
        if name is None:
            raise ValueError(f"value {t!r} before any name in /lbm_case/")
        kind = FIELDS[name][0]
        if kind == "real9":
            if index >= 9:
                raise ValueError("too many values for s")
            case["s"][index] = _value("real", t)
            index += 1
        else:
            if index:
                raise ValueError(f"too many values for {name}")
            case[name] = _value(kind, t)
            index += 1
        i += 1
    return case


def case_dynamics(case: Dict[str, object]):
    """The rigid-body coupling of the case (a solver.Dynamics), or None when
    solve_dynamic_equations is off (the reference's setting, main.f90:124-125)."""
    from .solver import Dynamics
    if not case["solve_dynamic_equations"]:
        return None
    return Dynamics(enable=True, closed_boundary=case["closed_boundary"], rk4=case["if_rk4"],
                    solid_den=case["solid_den"], area=case["area_airfoil"], par1=case["par1"], par2=case["par2"],
                    alphad=case["alphad"])


def case_config(case: Dict[str, object], motion_source: int = MOTION_FILE) -> SolverConfig:
    """The solver configuration of a case, with the markers read from its
    boundary files (read_boundary, airfoil.f90:207-242)."""
    from .formats import read_marker_file
    W, P = int(case["wing_count"]), int(case["points_per_wing"])
    markers = []
    for w in range(W):
        name = "boundary.dat" if w == 0 else f"boundary{w}.dat"
        mk = read_marker_file(os.path.join(str(case["boundary_dir"]), name), P)
        if mk.shape != (P, 2):
            raise ValueError(f"{name}: {P} markers expected, {mk.shape[0]} read")
        markers.append(mk)
    cfg = SolverConfig(nx=int(case["nx"]), ny=int(case["ny"]), scale=float(case["scale"]), uu=float(case["uu"]),
                       re=float(case["re"]), s=tuple(float(v) for v in case["s"]), ox=float(case["ox"]),
                       oy=float(case["oy"]), wing_count=W, motion_source=motion_source,
                       precision=int(case["precision"]), name="case")
    cfg.wing_markers = markers if W else None
    cfg.wing_x0, cfg.wing_dx, cfg.wing_y0 = float(case["wing_x0"]), float(case["wing_dx"]), float(case["wing_y0"])
    cfg.xc, cfg.yc, cfg.xp, cfg.yp = float(case["xc"]), float(case["yc"]), float(case["xp"]), float(case["yp"])
    return cfg
