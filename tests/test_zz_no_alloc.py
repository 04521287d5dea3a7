"""GPU test of lbm_get_mass_noalloc, the allocation-free variant of
lbm_get_mass: on caller-owned device populations it must give the value of
lbm_get_mass on the same populations bit for bit (the same kernels and
reduction order), call after call, including a second call on a smaller
lattice, and refuse a lattice larger than its fixed scratch."""
import ctypes

import numpy as np
import pytest

from lbm_mutiwing_amd import (PRECISION_FP32_SHIFTED, PRECISION_FP64, Solver, SolverConfig, preset, seeded_motions,
                              sinusoidal_motion_table)

pytestmark = pytest.mark.gpu

W = np.array([np.float32(4.0) / np.float32(9.0)] + [np.float32(1.0) / np.float32(9.0)] * 4 +
             [np.float32(1.0) / np.float32(36.0)] * 4, dtype=np.float64)


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from lbm_mutiwing_amd._native import load_library
    lib = load_library()
    lib.lbm_get_mass_noalloc.restype = ctypes.c_int
    lib.lbm_get_mass_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.c_void_p, ctypes.c_void_p]
    lib.lbm_last_error.restype = ctypes.c_int
    lib.lbm_last_error.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
    return lib


def _noalloc_mass(lib, f_host, precision):
    """Upload the populations as the library stores them and reduce them on
    the device with the allocation-free entry point."""
    import torch
    if precision == PRECISION_FP32_SHIFTED:
        dev = torch.from_numpy((f_host - W[:, None, None]).astype(np.float32)).cuda()
    else:
        dev = torch.from_numpy(np.ascontiguousarray(f_host)).cuda()
    out = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    ny, nx = f_host.shape[1], f_host.shape[2]
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.lbm_get_mass_noalloc(dev.data_ptr(), nx, ny, precision, out.data_ptr(), stream)
    assert rc == 0
    torch.cuda.synchronize()
    return float(out.item())


def _instance(cfg, steps):
    s = Solver(cfg)
    if cfg.wing_count:
        s.set_motion_table(sinusoidal_motion_table(cfg, seeded_motions(cfg, 12), steps + 2))
    s.advance(steps)
    return s


@pytest.mark.parametrize("precision", [PRECISION_FP64, PRECISION_FP32_SHIFTED])
def test_noalloc_mass_equals_get_mass(lib, precision):
    big = _instance(preset("small", precision=precision), 60)
    small_cfg = SolverConfig(nx=96, ny=40, scale=1 / 16, uu=0.05, re=80.0, wing_count=0, precision=precision)
    small = _instance(small_cfg, 25)
    for s in (big, small, big):           # two calls in a row, the second smaller, then back
        f, _ = s.populations()
        got = _noalloc_mass(lib, f, precision)
        assert got == s.mass(), (got, s.mass())
    big.close()
    small.close()


def test_noalloc_mass_rejects_more_rows_than_its_scratch(lib):
    import torch
    buf = torch.zeros(16, dtype=torch.float64, device="cuda")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    rc = lib.lbm_get_mass_noalloc(buf.data_ptr(), 1, 40000, PRECISION_FP64, out.data_ptr(), None)
    assert rc == 4                        # LBM_E_RANGE, nothing launched
    msg = ctypes.create_string_buffer(256)
    lib.lbm_last_error(None, msg, 256)
    assert b"scratch" in msg.value
    rc = lib.lbm_get_mass_noalloc(buf.data_ptr(), 1, 16, 7, out.data_ptr(), None)
    assert rc == 1                        # LBM_E_INVALID precision
