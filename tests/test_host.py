"""CPU tests of the host side: the C-ABI library exports what the header
declares, configuration and marker geometry, and reference data formats."""
import ctypes
import os
import re

import numpy as np
import pytest

from lbm_mutiwing_amd import SolverConfig, naca0012_markers, preset
from lbm_mutiwing_amd.config import LbmConfigC

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "lbm_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(lbm_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol(hip_lib_path):
    syms = _declared_symbols()
    assert len(syms) >= 18
    lib = ctypes.CDLL(hip_lib_path)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_oracle_mirrors_the_abi(oracle_mod):
    lib = oracle_mod.load()
    skip = {"lbm_version", "lbm_set_timing", "lbm_get_timing", "lbm_batch_create", "lbm_batch_destroy",
            "lbm_batch_advance", "lbm_batch_last_error", "lbm_get_tuning", "lbm_batch_get_tuning",
            "lbm_get_lattice_policy", "lbm_batch_get_lattice_policy",
            "lbm_measure_pull_copy", "lbm_get_mass_noalloc", "lbm_set_ibm_solve"}
    missing = [s for s in _declared_symbols() if s not in skip and not hasattr(lib, "orc_" + s[4:])]
    assert not missing, missing


def test_config_struct_layout_matches_header():
    # offsets of the C struct on x86-64 / gfx950 host ABI
    assert LbmConfigC.scale.offset == 16
    assert LbmConfigC.s.offset == 56
    assert LbmConfigC.motion_source.offset == 56 + 72 + 7 * 8
    assert ctypes.sizeof(LbmConfigC) == LbmConfigC.yb0.offset + 8


def test_dynamics_struct_layout_matches_header():
    from lbm_mutiwing_amd.solver import Dynamics
    txt = open(os.path.join(ROOT, "include", "lbm_hip.h")).read()
    body = txt[txt.index("typedef struct lbm_dynamics {"):txt.index("} lbm_dynamics;")]
    fields = re.findall(r"^\s*(int32_t|double)\s+([\w, ]+);", body, re.M)
    names = [n.strip() for _, group in fields for n in group.split(",")]
    assert names == [f[0] for f in Dynamics._fields_]
    assert Dynamics.solid_den.offset == 16 and Dynamics.alphad.offset == 48 and ctypes.sizeof(Dynamics) == 56


def test_naca_generator_reproduces_reference_marker_file():
    ref = np.loadtxt(os.path.join(ROOT, "tests", "golden", "boundary.dat"))
    gen = naca0012_markers(101)
    assert gen.shape == ref.shape == (205, 2)
    assert np.abs(gen - ref).max() <= 1e-8 + 1e-12
    assert np.array_equal(gen[0], gen[-1])        # closed contour


def test_reference_motion_rows_parse():
    rows = np.loadtxt(os.path.join(ROOT, "tests", "golden", "motion_params_head.dat"))
    assert rows.shape == (3, 19)
    assert np.allclose(np.diff(rows[:, 0]), 1e-4)
    # static wings at their base offsets, as in the shipped file
    assert rows[0, 4] == 0.0 and rows[0, 7] == 0.5 and rows[0, 13] == 1.5 and rows[0, 16] == 0.5


def test_presets_match_workload_definitions():
    c1, c2, c5 = preset("c1"), preset("c2"), preset("c5")
    assert (c1.nx, c1.ny, round(1 / c1.scale)) == (1024, 512, 64)
    assert c1.tau == pytest.approx(0.548)
    assert (c2.nx, c2.ny, round(1 / c2.scale)) == (4096, 2048, 256)
    assert c2.tau == pytest.approx(0.5384)
    assert c2.points_per_wing == 511
    assert (c5.nx, c5.ny) == (4001, 2001) and c5.points_per_wing == 205
    assert c1.row_len == 28 and c1.motion_cols == 19


def test_n_steps_for_matches_file_loop():
    cfg = preset("c5")
    cfg.uu = 0.01
    # default reference run: motion_end_time 22.0, dt 1e-4 -> 220001 steps
    assert cfg.n_steps_for(22.0) == 220001


def test_kernels_hold_no_device_wide_fence():
    """gfx950 compiles an agent-scope (device-wide) fence to a write-back and
    an invalidation of the XCD's whole L2 (buffer_wbl2 / buffer_inv).  In a
    kernel that runs beside the lattice pass it throws away the rows the pass
    has cached: a marker pass with two such fences per workgroup took the
    lattice pass of a batched geometry step (64 episodes) from 1.00 to
    1.41 ms.  The kernels order their data at workgroup scope (barriers) or
    at launch boundaries instead, and no device- or system-wide fence may
    come back into their sources."""
    csrc = os.path.join(ROOT, "lbm_mutiwing_amd", "csrc")
    bad = re.compile(r'__threadfence\s*\(\s*\)|__threadfence_system|__atomic_thread_fence'
                     r'|__builtin_amdgcn_fence\s*\([^)]*"(agent|)"\s*\)')
    hits = []
    for name in sorted(os.listdir(csrc)):
        if not name.endswith((".hip", ".h")):
            continue
        with open(os.path.join(csrc, name)) as fh:
            for n, line in enumerate(fh, 1):
                if bad.search(line.split("//")[0]):
                    hits.append(f"{name}:{n}: {line.strip()}")
    assert not hits, "\n".join(hits)
