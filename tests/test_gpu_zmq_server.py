"""GPU test of the multi-client ZeroMQ episode server
(python -m lbm_mutiwing_amd.zmq_server): three REQ clients on one ROUTER
endpoint, each running its own episodes with the unchanged protocol
(driver.f90:59-196), served together by continuous batching.  Every reply
must be the text the same commands get from a protocol session over the same
library alone (replies are bitwise the same alone and in any mix), and the
SAVE snapshot the file the Python Tecplot writer makes alone."""
import os
import subprocess
import sys
import threading
import uuid

import numpy as np
import pytest

from lbm_mutiwing_amd import MOTION_STREAM, Solver, formats, preset
from lbm_mutiwing_amd.build import zmq_library
from lbm_mutiwing_amd.case import case_config, read_case
from lbm_mutiwing_amd.protocol import Session, run_job_alone
from lbm_mutiwing_amd.solver import IBM_SOLVE_SUBSTITUTION

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _commands(dt, k):
    rng = np.random.default_rng(100 + k)
    out = ["RESET 0.0"]
    t = 0.0
    for m in range(3):
        t += (30 + 10 * k + 7 * m) * dt
        out.append(f"STEP {t!r}" + "".join(f" {v!r}" for v in rng.normal(0.0, 0.01, 6)))
        if m == 1 and k == 0:
            out.append("SAVE")
    if k == 1:
        out += ["RESET 0.0", f"STEP {25 * dt!r} 0 0 0 0 0 0"]
    return out + ["STOP"]


def test_three_clients_match_alone(tmp_path, monkeypatch):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from zmq_client import ReqClient
    cfg = preset("small")
    P = cfg.points_per_wing
    m = cfg.marker_array()
    for w in range(cfg.wing_count):
        formats.write_marker_file(str(tmp_path / ("boundary.dat" if w == 0 else f"boundary{w}.dat")),
                                  m[w * P:(w + 1) * P])
    ox, oy = cfg.origin()
    (tmp_path / "case.nml").write_text(
        f"&lbm_case nx={cfg.nx}, ny={cfg.ny}, scale={cfg.scale!r}, ox={ox!r}, oy={oy!r}, uu={cfg.uu!r}, "
        f"re={cfg.re!r}, s=" + ", ".join(repr(float(v)) for v in cfg.s) +
        f", wing_count={cfg.wing_count}, points_per_wing={P}, boundary_dir='.', output_dir='out' /\n")
    monkeypatch.chdir(tmp_path)        # the case's paths are relative to the working directory, as for the hosts
    base = case_config(read_case("case.nml"), MOTION_STREAM)
    cmds = [_commands(base.dt, k) for k in range(3)]

    endpoint = f"ipc:///tmp/lbm_srv_{uuid.uuid4().hex[:12]}.ipc"
    print(f"zmq episode server on {endpoint}", flush=True)
    log = open(tmp_path / "server.log", "w")
    env = dict(os.environ, PYTHONPATH=ROOT + os.pathsep + os.environ.get("PYTHONPATH", ""))
    proc = subprocess.Popen([sys.executable, "-m", "lbm_mutiwing_amd.zmq_server", "case.nml", "--endpoint", endpoint,
                             "--exit-after", "3", "--chunk", "16"], cwd=tmp_path, env=env, stdout=log,
                            stderr=subprocess.STDOUT)
    got, errors = {}, []

    def client(k):
        c = ReqClient(zmq_library(), endpoint, timeout_ms=120000)
        try:
            got[k] = [c.request(msg) for msg in cmds[k]]
        except Exception as exc:  # noqa: BLE001 (reported below)
            errors.append((k, repr(exc)))
        finally:
            c.close()

    threads = [threading.Thread(target=client, args=(k,)) for k in range(3)]
    try:
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=300)
        rc = proc.wait(timeout=120)
    finally:
        if proc.poll() is None:
            proc.kill()
            proc.wait()
        log.close()
    assert not errors, (errors, (tmp_path / "server.log").read_text())
    assert rc == 0, (tmp_path / "server.log").read_text()

    for k in range(3):
        s = Solver(base)
        s.set_ibm_solve(IBM_SOLVE_SUBSTITUTION)   # as the server solves its episodes
        sess = Session(s, save_hook=formats.tecplot_save_hook(str(tmp_path / f"alone_{k}")))
        want = []
        for msg in cmds[k]:
            reply, job = sess.receive(msg)
            want.append(reply if job is None else run_job_alone(job, chunk=16))   # the server's chunking
        assert got[k] == want, (k, got[k], want)
    saved = [r for r in got[0] if r.startswith("STATUS FLOW_SAVED")]
    assert len(saved) == 1
    name = saved[0].split()[-1]
    srv_file = tmp_path / "out" / "client_1" / name
    if not srv_file.exists():      # client order of first contact is not fixed
        srv_file = next((tmp_path / "out").glob(f"client_*/{name}"))
    assert srv_file.read_text() == (tmp_path / "alone_0" / name).read_text()
