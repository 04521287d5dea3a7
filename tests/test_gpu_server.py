"""Episode server with continuous batching: every client's replies are the
same (byte for byte) when it is served alone and when it shares the GPU with
others that arrive and leave while it runs; admission waits for HBM."""
import numpy as np
import pytest

from lbm_mutiwing_amd import MOTION_STREAM, PRECISION_FP32_SHIFTED, PRECISION_FP64, episode_config
from lbm_mutiwing_amd.server import EpisodeServer, device_bytes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from lbm_mutiwing_amd._native import load_library
    return load_library()


def script(cfg, seed, steps=3, seg=40):
    rng = np.random.default_rng(seed)
    cmds = ["RESET 0.0"]
    for k in range(1, steps + 1):
        t = k * seg * cfg.dt
        d = rng.normal(0.0, 0.1, 3 * cfg.wing_count)
        cmds.append("STEP " + repr(t) + "".join(" " + repr(float(x)) for x in d))
    cmds += ["SAVE", "STOP"]
    return cmds


def serve(lib, clients, budget, chunk=16, start_after=None, dynamics=None):
    """clients: {name: (cfg, cmds)}; a client submits its next command as soon
    as it has the previous reply; client c starts after start_after[c] polls."""
    srv = EpisodeServer(budget, chunk=chunk, lib=lib)
    for name, (cfg, _) in clients.items():
        srv.connect(name, cfg, (dynamics or {}).get(name))
    nxt = {name: 0 for name in clients}
    busy = {name: False for name in clients}
    replies = {name: [] for name in clients}
    order = []
    polls = 0
    while any(nxt[n] < len(c[1]) or busy[n] for n, c in clients.items()):
        for name, (_, cmds) in clients.items():
            if busy[name] or nxt[name] >= len(cmds):
                continue
            if start_after and polls < start_after.get(name, 0):
                continue
            srv.submit(name, cmds[nxt[name]])
            nxt[name] += 1
            busy[name] = True
        for r in srv.poll():
            replies[r.client].append(r.text)
            order.append((r.client, r.text.split(" ", 1)[0], cmds_of(clients, r.client, len(replies[r.client]) - 1)))
            busy[r.client] = False
        polls += 1
        assert polls < 10000
    srv.close()
    return replies, order


def cmds_of(clients, name, i):
    return clients[name][1][i].split(" ", 1)[0]


@pytest.mark.parametrize("precision", [PRECISION_FP64, PRECISION_FP32_SHIFTED])
def test_replies_same_alone_and_mixed(lib, precision):
    cfgs = {
        "a": episode_config(256, 2, 32, 100.0, precision, MOTION_STREAM),
        "b": episode_config(320, 1, 40, 300.0, precision, MOTION_STREAM),
        "c": episode_config(384, 3, 32, 200.0, precision, MOTION_STREAM),
    }
    clients = {n: (cfg, script(cfg, i)) for i, (n, cfg) in enumerate(cfgs.items())}
    big = 1 << 40
    alone = {n: serve(lib, {n: clients[n]}, big)[0][n] for n in clients}
    mixed, _ = serve(lib, clients, big, start_after={"b": 3, "c": 7})
    for n in clients:
        assert mixed[n] == alone[n], n
        assert alone[n][0].startswith("FORCE ") and alone[n][-2].startswith("STATUS FLOW_SAVED F")
        assert alone[n][-1] == "STATUS STOPPING"


def test_dynamics_client_same_alone_and_mixed(lib):
    """A client whose episodes couple wing 1 to the flow steps on its own
    beside the batch: its replies do not depend on the other clients, and
    equal those of its solver advanced directly."""
    from lbm_mutiwing_amd.protocol import Session, run_job_alone
    from lbm_mutiwing_amd.solver import Dynamics, Solver
    dyn = Dynamics(solid_den=50.0, par1=-2.0, par2=50.0)
    cfgs = {"d": episode_config(256, 2, 32, 100.0, PRECISION_FP64, MOTION_STREAM),
            "p": episode_config(320, 1, 40, 300.0, PRECISION_FP64, MOTION_STREAM)}
    clients = {n: (cfg, script(cfg, 10 + i)) for i, (n, cfg) in enumerate(cfgs.items())}
    big = 1 << 40
    alone = serve(lib, {"d": clients["d"]}, big, dynamics={"d": dyn})[0]["d"]
    mixed, _ = serve(lib, clients, big, start_after={"p": 2}, dynamics={"d": dyn})
    assert mixed["d"] == alone
    plain = serve(lib, {"d": clients["d"]}, big)[0]["d"]
    assert plain[1] != alone[1]                # the coupling changes the forces
    s = Solver(cfgs["d"], lib)
    s.set_dynamics(dyn)
    sess = Session(s)
    direct = []
    for cmd in clients["d"][1][:-2]:
        reply, job = sess.receive(cmd, 0.0)
        direct.append(run_job_alone(job, 16) if job is not None else reply)   # the chunks of serve()
    assert direct == alone[:-2]
    s.close()


def test_admission_waits_for_hbm(lib):
    cfg = episode_config(256, 1, 32, 100.0, PRECISION_FP32_SHIFTED, MOTION_STREAM)
    clients = {n: (cfg, script(cfg, i, steps=2)) for i, n in enumerate(["p", "q", "r"])}
    budget = 2 * device_bytes(cfg) + 1      # two episodes fit, the third waits
    replies, order = serve(lib, clients, budget)
    for n in clients:
        assert replies[n][-1] == "STATUS STOPPING"
    first_r = next(i for i, (c, k, _) in enumerate(order) if c == "r")
    first_stop = next(i for i, (c, k, cmd) in enumerate(order) if cmd == "STOP")
    assert first_r > first_stop


def test_errors_before_reset(lib):
    cfg = episode_config(256, 1, 32, 100.0, PRECISION_FP64, MOTION_STREAM)
    srv = EpisodeServer(1 << 40, lib=lib)
    srv.connect("x", cfg)
    srv.submit("x", "STEP 0.1 0 0 0")
    srv.submit("x", "SAVE")
    srv.submit("x", "HELLO")
    out = [r.text for r in srv.poll()]
    assert out[0] == "ERROR online mode requires RESET before STEP"
    assert out[1] == "ERROR online mode requires RESET before SAVE"
    assert out[2].startswith("ERROR ")
    srv.close()
