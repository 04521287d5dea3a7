"""Loader of the in-tree HIP library liblbm_hip.so.

There is no fallback: if the library is missing or cannot be loaded the
import fails with an error naming the build step.
"""
from __future__ import annotations

import ctypes
import os

from .solver import bind

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liblbm_hip.so")
_lib = None


def load_library() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -m lbm_mutiwing_amd.build` (or __graft_entry__.build())")
        lib = ctypes.CDLL(LIB_PATH)
        bind(lib, "lbm_")
        lib.lbm_version.restype = ctypes.c_char_p
        lib.lbm_version.argtypes = []
        _lib = lib
    return _lib
