"""CPU tests of the multi-client ZeroMQ front end (lbm_mutiwing_amd/zmq_server.py)
and of the case-file reader (lbm_mutiwing_amd/case.py); no GPU work.

The front end's envelope handling is checked with an echo loop: REQ clients
in threads, each request answered to its own client, in any interleaving."""
import threading
import uuid

import pytest

from lbm_mutiwing_amd import build
from lbm_mutiwing_amd.case import read_case


def _have_zmq():
    try:
        build.zmq_library()
        return True
    except RuntimeError:
        return False


def test_read_case_forms(tmp_path):
    p = tmp_path / "case.nml"
    p.write_text("! comment line\n&lbm_case nx=400, ny=200, scale=1.0d-2, ox=-2.0, oy = -1.0,\n"
                 "  s=0.0, 1.1, 1.05, 0.0, 1.2, 0.0, 1.2, 0.0, 0.0, wing_count=1, points_per_wing=67,\n"
                 "  boundary_dir='wings dir', restart=.true., output_dir='it''s' /\n")
    c = read_case(str(p))
    assert (c["nx"], c["ny"], c["wing_count"], c["points_per_wing"]) == (400, 200, 1, 67)
    assert c["scale"] == 0.01 and c["ox"] == -2.0 and c["oy"] == -1.0
    assert c["s"] == [0.0, 1.1, 1.05, 0.0, 1.2, 0.0, 1.2, 0.0, 0.0]
    assert c["boundary_dir"] == "wings dir" and c["output_dir"] == "it's" and c["restart"] is True
    assert c["uu"] == 0.05 and c["checkpoint_every"] == 100000      # defaults of lbm_host.f90
    p.write_text("&lbm_case nx=10, bogus=3 /\n")
    with pytest.raises(ValueError):
        read_case(str(p))


def test_read_case_dynamics(tmp_path):
    from lbm_mutiwing_amd.case import case_dynamics
    p = tmp_path / "case.nml"
    p.write_text("&lbm_case nx=64, ny=32 /\n")
    assert case_dynamics(read_case(str(p))) is None          # the reference's setting
    p.write_text("&lbm_case nx=64, ny=32, solve_dynamic_equations=.true., if_rk4=.false.,\n"
                 "  solid_den=2.5d1, par1=-1.0 /\n")
    d = case_dynamics(read_case(str(p)))
    assert (d.enable, d.closed_boundary, d.rk4) == (1, 1, 0)
    assert d.solid_den == 25.0 and d.area == 6.807946476488794e-2 and d.par1 == -1.0 and d.par2 == 0.0
    assert d.alphad == 10.0 * 3.1415926535897932384626433832795 / 180.0


@pytest.mark.skipif(not _have_zmq(), reason="libzmq not in this image")
def test_router_front_envelopes():
    from lbm_mutiwing_amd.zmq_server import RouterFront
    from zmq_client import ReqClient
    lib = build.zmq_library()
    endpoint = f"ipc:///tmp/lbm_front_{uuid.uuid4().hex[:12]}.ipc"
    front = RouterFront(lib, endpoint)
    n_clients, n_msgs = 4, 5
    got = {}

    def client(k):
        c = ReqClient(lib, endpoint, timeout_ms=20000)
        try:
            got[k] = [c.request(f"REQ {k} {m}") for m in range(n_msgs)]
        finally:
            c.close()

    threads = [threading.Thread(target=client, args=(k,)) for k in range(n_clients)]
    for t in threads:
        t.start()
    answered = 0
    idents = set()
    try:
        while answered < n_clients * n_msgs:
            assert front.wait(20000), "no request within 20 s"
            while True:
                r = front.recv()
                if r is None:
                    break
                ident, text = r
                idents.add(ident)
                front.send(ident, "ECHO " + text)
                answered += 1
    finally:
        for t in threads:
            t.join(timeout=30)
        front.close()
    assert len(idents) == n_clients
    for k in range(n_clients):
        assert got[k] == [f"ECHO REQ {k} {m}" for m in range(n_msgs)]
