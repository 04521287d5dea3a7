"""MI355X (gfx950) time step of the D2Q9 MRT lattice-Boltzmann + implicit
immersed-boundary multi-airfoil solver, behind the C ABI of include/lbm_hip.h."""
from .config import (MOTION_FILE, MOTION_STREAM, PRECISION_FP32_SHIFTED, PRECISION_FP64, SolverConfig, WingMotion,
                     episode_config, naca0012_markers, preset, random_episode, seeded_motions,
                     sinusoidal_motion_table)
from .solver import AdvanceResult, Batch, LbmError, Solver

__all__ = ["MOTION_FILE", "MOTION_STREAM", "episode_config", "random_episode", "PRECISION_FP64", "PRECISION_FP32_SHIFTED", "SolverConfig", "WingMotion", "naca0012_markers", "preset", "seeded_motions",
           "sinusoidal_motion_table", "Solver", "Batch", "LbmError", "AdvanceResult"]
