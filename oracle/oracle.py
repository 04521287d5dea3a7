"""ctypes front end of the CPU oracle (liblbm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  The oracle is
driven through the same host class as the HIP library (Solver), bound to the
``orc_`` symbols, so the two see identical inputs.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

from lbm_mutiwing_amd.solver import Solver, bind

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_oracle.so")
_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE, "-B" if force else "all"])
    return LIB_PATH


def configure(lib):
    """Attach argument/result types to every entry point of an oracle build."""
    bind(lib, "orc_")
    lib.orc_set_threads.restype = ctypes.c_int
    lib.orc_set_threads.argtypes = [ctypes.c_int]
    lib.orc_delta_kernel.restype = ctypes.c_double
    lib.orc_delta_kernel.argtypes = [ctypes.c_double]
    _dp = ctypes.POINTER(ctypes.c_double)
    lib.orc_mrt_matrices.restype = None
    lib.orc_mrt_matrices.argtypes = [_dp, _dp]
    lib.orc_weights.restype = None
    lib.orc_weights.argtypes = [_dp]
    lib.orc_collide_node.restype = ctypes.c_int
    lib.orc_collide_node.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
    lib.orc_bc_node.restype = ctypes.c_int
    lib.orc_bc_node.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 6 + [_dp, _dp]
    lib.orc_get_ibm_nodes.restype = ctypes.c_int
    lib.orc_get_ibm_nodes.argtypes = [ctypes.c_void_p, ctypes.c_int32, _dp, _dp, _dp, _dp, _dp]
    lib.orc_get_rates.restype = ctypes.c_int
    lib.orc_get_rates.argtypes = [ctypes.c_void_p, _dp]
    lib.orc_lodi_node.restype = ctypes.c_int
    lib.orc_lodi_node.argtypes = [ctypes.c_void_p, ctypes.c_int32, _dp, _dp, _dp, _dp]
    lib.orc_body_step.restype = None
    lib.orc_body_step.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 4 + [_dp]
    lib.orc_get_ibm_system.restype = ctypes.c_int
    lib.orc_get_ibm_system.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
    lib.orc_get_mass.restype = ctypes.c_int
    lib.orc_get_mass.argtypes = [ctypes.c_void_p, _dp]
    return lib


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = configure(ctypes.CDLL(LIB_PATH))
    return _lib


def build_variant(source: str, out_path: str) -> str:
    """Compile an edited copy of the oracle source (a planted mistake, for
    the tests that show a check fails on it) with the flags of the Makefile;
    the copy includes lbm_oracle.h from this directory."""
    cmd = ["gcc", "-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-w",
           "-iquote", HERE, "-shared", "-o", out_path, source, "-lm"]
    subprocess.check_call(cmd)
    return out_path


def body_step(dyn, scale: float, uu: float, fx: float, mom: float, st):
    """One rigid-body update of wing 1's 11 body_rel fields (in place)."""
    import numpy as np
    assert st.dtype == np.float64 and st.shape == (11,) and st.flags.c_contiguous
    load().orc_body_step(ctypes.byref(dyn), scale, uu, fx, mom, st.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))


def set_threads(n: int) -> int:
    return load().orc_set_threads(int(n))


class OracleSolver(Solver):
    prefix = "orc_"

    def __init__(self, cfg, lib=None):
        super().__init__(cfg, lib=lib if lib is not None else load())

    def ibm_system(self):
        import numpy as np
        n = self.cfg.wing_count * (self.cfg.points_per_wing - 1)
        a = np.zeros(n * n)
        b = np.zeros(2 * n)
        x = np.zeros(2 * n)
        dp = ctypes.POINTER(ctypes.c_double)
        rc = self.lib.orc_get_ibm_system(self._h, a.ctypes.data_as(dp), b.ctypes.data_as(dp),
                                         x.ctypes.data_as(dp))
        if rc != 0:
            raise RuntimeError("no IBM system yet")
        return a.reshape(n, n, order="F"), b.reshape(2, n).T, x.reshape(2, n).T

    # ---- single-routine hooks (for property tests)
    def _dp(self, a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    def collide(self, f, rho, u, v, gx=0.0, gy=0.0):
        import numpy as np
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = np.zeros(9)
        self.lib.orc_collide_node(self._h, self._dp(f), rho, u, v, gx, gy, self._dp(out))
        return out

    def bc_node(self, rb, ub, vb, rn, un, vn, f0n):
        import numpy as np
        f0n = np.ascontiguousarray(f0n, dtype=np.float64)
        out = np.zeros(9)
        self.lib.orc_bc_node(self._h, rb, ub, vb, rn, un, vn, self._dp(f0n), self._dp(out))
        return out

    def ibm_nodes(self):
        import numpy as np
        n = self.state()["adjoint_n"]
        arrs = [np.zeros(max(n, 1)) for _ in range(5)]
        self.lib.orc_get_ibm_nodes(self._h, n, *[self._dp(a) for a in arrs])
        return [a[:n] for a in arrs]

    def lodi(self, right, r, u, v):
        import numpy as np
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (r, u, v)]
        out = np.zeros(3)
        self.lib.orc_lodi_node(self._h, int(right), *[self._dp(x) for x in a], self._dp(out))
        return out

    def rates(self):
        import numpy as np
        s = np.zeros(9)
        self.lib.orc_get_rates(self._h, self._dp(s))
        return s
