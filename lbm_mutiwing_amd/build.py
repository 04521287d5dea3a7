"""Compile the gfx950 HIP library liblbm_hip.so in-tree with hipcc, and the
Fortran host program lbm_host over it with flang.

Flags: -ffp-contract=off keeps the FP64 lattice arithmetic in the reference's
evaluation order (no fused multiply-add), so the fused kernel reproduces a
sequential evaluation bit for bit; packed FP32 instructions are disabled
(-packed-fp32-ops) for every HIP source.

The host (fortran/*.f90) links liblbm_hip.so from its own directory and the
libzmq of the image for its online front end.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblbm_hip.so")
SOURCES = ["lbm_api.hip", "lbm_lattice.hip", "lbm_ibm.hip"]
ARCH = os.environ.get("LBM_OFFLOAD_ARCH", "gfx950")
FORTRAN = os.path.join(HERE, "fortran")
HOST = os.path.join(HERE, "lbm_host")
HOST_MODULES = ["lbm_hip.f90", "lbm_front.f90", "lbm_host_io.f90"]
ZMQ_DIRS = ["/opt/conda/lib", "/usr/lib/x86_64-linux-gnu", "/usr/local/lib", "/usr/lib64"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def flang() -> str:
    for cand in (os.environ.get("FLANG"), "/opt/rocm/lib/llvm/bin/flang", "/opt/rocm/llvm/bin/flang"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("flang not found")


def zmq_dir() -> str:
    for d in ZMQ_DIRS:
        if os.path.exists(os.path.join(d, "libzmq.so")) or os.path.exists(os.path.join(d, "libzmq.so.5")):
            return d
    raise RuntimeError("libzmq not found in " + ", ".join(ZMQ_DIRS))


def zmq_library() -> str:
    d = zmq_dir()
    for name in ("libzmq.so.5", "libzmq.so"):
        if os.path.exists(os.path.join(d, name)):
            return os.path.join(d, name)
    raise RuntimeError("libzmq not found")


def fortran_program(main_source: str, out: str, verbose: bool = False) -> str:
    """flang build of one program over the host modules (lbm_hip, lbm_front,
    lbm_host_io), linked with liblbm_hip.so from this directory and libzmq."""
    import tempfile
    fc = flang()
    zd = zmq_dir()
    with tempfile.TemporaryDirectory(prefix="lbm_host_") as tmp:
        objs = []
        for src in [os.path.join(FORTRAN, m) for m in HOST_MODULES] + [main_source]:
            obj = os.path.join(tmp, os.path.basename(src).replace(".f90", ".o"))
            cmd = [fc, "-O2", "-module-dir", tmp, "-c", src, "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.check_call(cmd)
            objs.append(obj)
        tmp_out = out + ".tmp"
        cmd = [fc, *objs, "-o", tmp_out, f"-L{HERE}", "-llbm_hip", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{HERE}",
               f"-L{zd}", "-lzmq", f"-Wl,-rpath,{zd}"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        os.replace(tmp_out, out)
    return out


def build_host(verbose: bool = False) -> str:
    """The Fortran host program lbm_host (needs liblbm_hip.so built first)."""
    return fortran_program(os.path.join(FORTRAN, "lbm_host.f90"), HOST, verbose)


def flags():
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def build(verbose: bool = False) -> str:
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        cmd = [hipcc(), *flags(), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_host(verbose="-v" in sys.argv))
