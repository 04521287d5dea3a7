"""ZeroMQ front end of the episode server: many protocol clients, one GPU.

The reference's online mode answers one client on a REP socket
(driver.f90:59-196, zmq_comm.f90:1-366): one process, one episode at a time.
Here one ROUTER socket takes any number of the same REQ clients, each
speaking the unchanged protocol (RESET / STEP / SAVE / STOP, one request, one
reply; replies laid out as zmq_comm.f90:183-281).  Every connected REQ
socket is one client of the :class:`~lbm_mutiwing_amd.server.EpisodeServer`:
its requests are submitted as they arrive, each scheduler iteration advances
every active episode together (batched, continuous), and each reply goes back
to its client in the ROUTER envelope as soon as its job is done.

    python -m lbm_mutiwing_amd.zmq_server case.nml [--endpoint tcp://*:5555]
        [--hbm-gb 200] [--chunk 100] [--exit-after N]

The case file is the Fortran host's namelist (lbm_mutiwing_amd/case.py);
every client's episodes use that case, in stream motion mode.  SAVE writes
F<cycle>.dat under <output_dir>/client_<k>/ (k = order of first contact).
``--exit-after N`` ends the server once N clients have had their STOP
answered (tests); otherwise it runs until killed.
"""
from __future__ import annotations

import argparse
import ctypes
import os
import sys
from typing import Dict, Optional, Tuple

ZMQ_ROUTER = 6
ZMQ_LINGER = 17
ZMQ_RCVMORE = 13
ZMQ_SNDMORE = 2
ZMQ_DONTWAIT = 1
ZMQ_POLLIN = 1
EAGAIN = 11


class _PollItem(ctypes.Structure):
    _fields_ = [("socket", ctypes.c_void_p), ("fd", ctypes.c_int), ("events", ctypes.c_short),
                ("revents", ctypes.c_short)]


class RouterFront:
    """A bound ROUTER socket over libzmq's C API: requests of REQ clients in,
    replies out, keyed by the client's routing identity."""

    def __init__(self, libpath: str, endpoint: str):
        z = ctypes.CDLL(libpath, use_errno=True)
        z.zmq_ctx_new.restype = ctypes.c_void_p
        z.zmq_socket.restype = ctypes.c_void_p
        z.zmq_socket.argtypes = [ctypes.c_void_p, ctypes.c_int]
        z.zmq_setsockopt.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t]
        z.zmq_getsockopt.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]
        z.zmq_bind.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        z.zmq_send.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        z.zmq_recv.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        z.zmq_poll.argtypes = [ctypes.POINTER(_PollItem), ctypes.c_int, ctypes.c_long]
        z.zmq_close.argtypes = [ctypes.c_void_p]
        z.zmq_ctx_term.argtypes = [ctypes.c_void_p]
        z.zmq_errno.restype = ctypes.c_int
        self.z = z
        self.ctx = z.zmq_ctx_new()
        self.sock = z.zmq_socket(self.ctx, ZMQ_ROUTER)
        v = ctypes.c_int(0)
        z.zmq_setsockopt(self.sock, ZMQ_LINGER, ctypes.byref(v), ctypes.sizeof(v))
        if z.zmq_bind(self.sock, endpoint.encode()) != 0:
            raise RuntimeError(f"zmq_bind({endpoint}) failed, errno {z.zmq_errno()}")
        self.buf = ctypes.create_string_buffer(65536)

    def _more(self) -> bool:
        v, n = ctypes.c_int(0), ctypes.c_size_t(ctypes.sizeof(ctypes.c_int))
        self.z.zmq_getsockopt(self.sock, ZMQ_RCVMORE, ctypes.byref(v), ctypes.byref(n))
        return v.value != 0

    def wait(self, timeout_ms: int) -> bool:
        """True when a request is waiting (or arrives within timeout_ms)."""
        item = _PollItem(self.sock, 0, ZMQ_POLLIN, 0)
        return self.z.zmq_poll(ctypes.byref(item), 1, timeout_ms) > 0 and bool(item.revents & ZMQ_POLLIN)

    def recv(self) -> Optional[Tuple[bytes, str]]:
        """(identity, request text) of the next waiting request, or None."""
        frames = []
        flags = ZMQ_DONTWAIT
        while True:
            n = self.z.zmq_recv(self.sock, self.buf, len(self.buf), flags)
            if n < 0:
                if not frames and self.z.zmq_errno() == EAGAIN:
                    return None
                raise RuntimeError(f"zmq_recv failed, errno {self.z.zmq_errno()}")
            frames.append(self.buf.raw[:min(n, len(self.buf))])
            if not self._more():
                break
            flags = 0
        if len(frames) < 3 or frames[1] != b"":
            return None                      # not a REQ envelope: drop it
        return frames[0], frames[-1].decode(errors="replace")

    def send(self, identity: bytes, text: str) -> None:
        data = text.encode()
        for part, flag in ((identity, ZMQ_SNDMORE), (b"", ZMQ_SNDMORE), (data, 0)):
            if self.z.zmq_send(self.sock, part, len(part), flag) < 0:
                raise RuntimeError(f"zmq_send failed, errno {self.z.zmq_errno()}")

    def close(self) -> None:
        if self.sock:
            self.z.zmq_close(self.sock)
            self.sock = None
        if self.ctx:
            self.z.zmq_ctx_term(self.ctx)
            self.ctx = None


def serve(front: RouterFront, cfg_for_client, hbm_budget: int, chunk: int, save_dir: Optional[str] = None,
          exit_after: int = 0, dynamics=None) -> int:
    """Run the episode server behind ``front`` until ``exit_after`` clients
    have had their STOP answered (0: forever); every client's episodes use
    ``dynamics`` (rigid-body coupling, a solver.Dynamics, or None).  Returns
    the number of replies sent."""
    from . import formats
    from .server import EpisodeServer
    hooks: Dict[bytes, object] = {}

    def save_hook(session, cycle):
        owner = next(ident for ident, s in srv.sessions.items() if s is session)
        return hooks[owner](session, cycle)

    srv = EpisodeServer(hbm_budget, chunk=chunk, save_hook=save_hook if save_dir else None)
    order: Dict[bytes, int] = {}
    stopped = 0
    sent = 0
    try:
        while exit_after <= 0 or stopped < exit_after:
            busy = not srv.idle()
            if busy or front.wait(5):
                while True:
                    got = front.recv()
                    if got is None:
                        break
                    ident, text = got
                    if ident not in order:
                        order[ident] = len(order) + 1
                        srv.connect(ident, cfg_for_client(ident), dynamics)
                        if save_dir:
                            hooks[ident] = formats.tecplot_save_hook(
                                os.path.join(save_dir, f"client_{order[ident]}"))
                    srv.submit(ident, text)
            for r in srv.poll():
                front.send(r.client, r.text)
                sent += 1
                if r.text.startswith("STATUS STOPPING"):
                    stopped += 1
    finally:
        srv.close()
    return sent


def main(argv=None) -> int:
    from .build import zmq_library
    from .case import case_config, case_dynamics, read_case
    from .config import MOTION_STREAM
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("case", help="namelist file with the group /lbm_case/")
    ap.add_argument("--endpoint", default="tcp://*:5555")
    ap.add_argument("--hbm-gb", type=float, default=200.0)
    ap.add_argument("--chunk", type=int, default=100)
    ap.add_argument("--exit-after", type=int, default=0)
    args = ap.parse_args(argv)
    case = read_case(args.case)
    base = case_config(case, MOTION_STREAM)
    front = RouterFront(zmq_library(), args.endpoint)
    print(f"episode server on {args.endpoint}: RESET / STEP / SAVE / STOP from any number of clients", flush=True)
    try:
        serve(front, lambda ident: base, int(args.hbm_gb * (1 << 30)), args.chunk,
              save_dir=str(case["output_dir"]), exit_after=args.exit_after, dynamics=case_dynamics(case))
    finally:
        front.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
