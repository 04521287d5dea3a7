"""MI355X (gfx950) time step of the D2Q9 MRT lattice-Boltzmann + implicit
immersed-boundary multi-airfoil solver, behind the C ABI of include/lbm_hip.h."""
from .config import (PRECISION_FP32_SHIFTED, PRECISION_FP64, SolverConfig, WingMotion, naca0012_markers,
                     preset, seeded_motions, sinusoidal_motion_table)
from .solver import AdvanceResult, LbmError, Solver

__all__ = ["PRECISION_FP64", "PRECISION_FP32_SHIFTED", "SolverConfig", "WingMotion", "naca0012_markers", "preset", "seeded_motions",
           "sinusoidal_motion_table", "Solver", "LbmError", "AdvanceResult"]
