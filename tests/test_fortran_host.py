"""CPU tests of the Fortran host (lbm_mutiwing_amd/fortran): its command
front end against the Python mirror of the protocol (parse and FORCE reply
layout, zmq_comm.f90:183-281), and its input checks that stop the program
before any device work.  The host and the check driver are built with flang
here; no test calls into the GPU library."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from lbm_mutiwing_amd import build
from lbm_mutiwing_amd.protocol import (CMD_INVALID, CMD_OUTPUT, CMD_RESET, CMD_STEP, CMD_STOP, format_force_reply,
                                       fortran_es24_16, parse_command)

HERE = os.path.dirname(os.path.abspath(__file__))

# command codes of the Fortran front end (main.f90:91-95 numbering)
F_CODES = {0: CMD_INVALID, 1: CMD_RESET, 2: CMD_STEP, 3: CMD_STOP, 4: CMD_OUTPUT}

MESSAGES = [
    "RESET", "reset 0", "START 1.5d0", "  RESET", "RESET  ", "RESET /", "RESET ,", "RESET 0 junk", "RESET abc",
    "RESET 1.0e", ",RESET", "/", "", "   ",
    "STEP 0.05 0.1 0 0, 0 0 0", "STEP 0.05 0.1 0 0", "STEP 0.05 7*0.0", "STEP 0.05,,1,2,3,4,5",
    "STEP 0.05 0.1 0 0 0 0 0 extra", "STEP 0.05 abc 0 0 0 0 0", "step 1e-3 1 2 3 4 5 6 / 7", "STEP 0.05 /",
    "STEP 0.05 2*1.5 3* 7", "STEP 0.05 1 2 3 4 5 ,", "STEP 0.05, 1 ,2 , 3,,5, 6", "STEP\t0.05\t1 2 3 4 5 6",
    "\"STEP\" 1 2 3 4 5 6 7", "STEP 0.1d0 1.d0 .5 -3 +4 5E-1 6", "STEP 1.5+3 1 2 3 4 5 6",
    "save", "OUTPUT", "exit", "STOP now", "ST OP", "JUMP 1",
]


def _have_toolchain():
    try:
        build.flang()
        build.zmq_dir()
        return True
    except RuntimeError:
        return False


pytestmark = pytest.mark.skipif(not _have_toolchain(), reason="flang or libzmq not in this image")


@pytest.fixture(scope="module")
def front_check(tmp_path_factory, hip_lib_path):
    out = str(tmp_path_factory.mktemp("fortran") / "front_check")
    return build.fortran_program(os.path.join(HERE, "fortran", "front_check.f90"), out)


@pytest.fixture(scope="module")
def host(hip_lib_path):
    if not os.path.exists(build.HOST):
        build.build_host()
    return build.HOST


def _run_front(exe, wing_count, lines):
    r = subprocess.run([exe], input=(f"{wing_count}\n" + "\n".join(lines) + "\n").encode(),
                       capture_output=True, timeout=60)
    assert r.returncode == 0, r.stderr.decode()
    return r.stdout.decode().splitlines()


@pytest.mark.parametrize("wing_count", [1, 2])
def test_parse_matches_python_mirror(front_check, wing_count):
    out = _run_front(front_check, wing_count, MESSAGES)
    assert len(out) == 2 * len(MESSAGES)
    for k, msg in enumerate(MESSAGES):
        head, values = out[2 * k], out[2 * k + 1]
        code, ok, err = (head.split(" ", 2) + [""])[:3]
        cmd, row, perr = parse_command(msg, wing_count)
        assert F_CODES[int(code)] == cmd, (msg, head)
        assert (ok == "T") == (cmd != CMD_INVALID), msg
        assert err == perr, (msg, err, perr)
        if cmd != CMD_INVALID:
            n = 1 + 3 * wing_count
            assert values.split() == [fortran_es24_16(v) for v in row[:n]], msg


def test_force_reply_layout_matches(front_check):
    rng = np.random.default_rng(3)
    rows = [np.array([1.0, -2.5e-7, 1e100]), np.array([0.0, -0.0, 1e-310, 123456.789]),
            rng.normal(size=28) * 10.0 ** rng.integers(-12, 12, size=28)]
    lines = ["#FORCE " + " ".join([str(len(r))] + [repr(float(v)) for v in r]) for r in rows]
    out = _run_front(front_check, 2, lines)
    assert out == [format_force_reply(r) for r in rows]


def test_host_missing_case_file(host, tmp_path):
    r = subprocess.run([host, str(tmp_path / "none.nml")], capture_output=True, timeout=60, cwd=tmp_path)
    assert r.returncode == 1
    assert b"cannot open case file" in r.stdout


def test_host_missing_marker_file(host, tmp_path):
    (tmp_path / "case.nml").write_text("&lbm_case nx=64, ny=32, wing_count=1, points_per_wing=10, "
                                       "boundary_dir='" + str(tmp_path) + "' /\n")
    r = subprocess.run([host, "case.nml"], capture_output=True, timeout=60, cwd=tmp_path)
    assert r.returncode == 1
    assert "无法打开边界文件".encode() in r.stdout


def test_host_links_library_and_zmq(host):
    if shutil.which("ldd") is None:
        pytest.skip("no ldd")
    out = subprocess.run(["ldd", host], capture_output=True, timeout=60).stdout.decode()
    assert "liblbm_hip.so" in out and "libzmq" in out
    assert "not found" not in out


# ------------------------------------------------------------ checkpoint I/O

@pytest.fixture(scope="module")
def checkpoint_check(tmp_path_factory, hip_lib_path):
    out = str(tmp_path_factory.mktemp("fortran") / "checkpoint_check")
    return build.fortran_program(os.path.join(HERE, "fortran", "checkpoint_check.f90"), out)


def test_checkpoint_layout_of_fortran_writer(checkpoint_check, tmp_path):
    """The Fortran writer's bytes, read back by the Python reader, hold the
    driver's known values at the places saveflowfield writes them
    (solution.f90:1660-1686)."""
    from lbm_mutiwing_amd import formats
    nx, ny, W = 7, 5, 2
    path = tmp_path / "ck.dat"
    r = subprocess.run([checkpoint_check, "write", str(path), str(nx), str(ny), str(W), "300", "-2"],
                       capture_output=True, timeout=60)
    assert r.returncode == 0, r.stdout.decode() + r.stderr.decode()
    assert path.stat().st_size == formats.checkpoint_bytes(W, nx, ny)
    ck = formats.read_checkpoint(str(path), W, nx, ny)
    assert (ck.cycle, ck.movenum) == (300, -2)
    expect_rel = np.arange(1, 12)[:, None] + np.arange(1, W + 1)[None, :] / 8.0
    assert np.array_equal(ck.body_rel, expect_rel)
    i = np.arange(1, nx + 1)[None, :]
    j = np.arange(1, ny + 1)[:, None]
    assert np.array_equal(ck.u, i + j / 1024.0)
    assert np.array_equal(ck.v, -i - j / 2048.0)
    k = np.arange(1, 10)[:, None, None]
    assert np.array_equal(ck.f, k + i[None] / 1024.0 + j[None] / 1048576.0)
    kk = np.arange(1, 10)[:, None]
    for q, name in enumerate(("left2", "f0_left", "right2", "f0_right")):
        want = 16.0 * (q + 1) + kk + np.arange(1, ny + 1)[None, :] / 1024.0
        assert np.array_equal(ck.ring[name][:, 1:ny - 1], want[:, 1:ny - 1]), name
    for q, name in enumerate(("up2", "f0_top", "down2", "f0_bottom")):
        want = 16.0 * (q + 5) + kk + np.arange(1, nx + 1)[None, :] / 1024.0
        assert np.array_equal(ck.ring[name], want), name


def test_checkpoint_python_fortran_round_trip(checkpoint_check, tmp_path):
    """A checkpoint written by the Python mirror, read and written again by
    the Fortran host's routines, comes back byte for byte."""
    from lbm_mutiwing_amd import formats
    rng = np.random.default_rng(11)
    nx, ny, W = 9, 6, 3
    f = rng.normal(size=(9, ny, nx))
    ring = formats.ring_from_populations(rng.normal(size=(9, ny, nx)))
    ck = formats.Checkpoint(cycle=12345, movenum=3, body_rel=np.zeros((11, W)), u=rng.normal(size=(ny, nx)),
                            v=rng.normal(size=(ny, nx)), f=f, ring=ring)
    ck.body_rel[:] = rng.normal(size=(11, W))
    a, b = tmp_path / "a.dat", tmp_path / "b.dat"
    formats.write_checkpoint(str(a), ck)
    r = subprocess.run([checkpoint_check, "copy", str(a), str(b), str(nx), str(ny), str(W)],
                       capture_output=True, timeout=60)
    assert r.returncode == 0, r.stdout.decode()
    assert a.read_bytes() == b.read_bytes()
    back = formats.read_checkpoint(str(b), W, nx, ny)
    assert np.array_equal(back.f, f) and back.cycle == 12345 and back.movenum == 3


def test_checkpoint_reader_rejects_mismatch(checkpoint_check, tmp_path):
    path = tmp_path / "ck.dat"
    subprocess.run([checkpoint_check, "write", str(path), "7", "5", "2", "1", "0"], check=True, timeout=60)
    r = subprocess.run([checkpoint_check, "copy", str(path), str(tmp_path / "o.dat"), "7", "6", "2"],
                       capture_output=True, timeout=60)
    assert r.returncode == 1 and "续算存档文件大小与网格不一致".encode() in r.stdout
    # the right size, but the wing-count field says 3
    data = bytearray(path.read_bytes())
    data[4:8] = (3).to_bytes(4, "little")
    path.write_bytes(bytes(data))
    r = subprocess.run([checkpoint_check, "copy", str(path), str(tmp_path / "o.dat"), "7", "5", "2"],
                       capture_output=True, timeout=60)
    assert r.returncode == 1 and "续算文件中的翼数量与当前编译设置不一致".encode() in r.stdout
