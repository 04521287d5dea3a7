"""Minimal ZeroMQ REQ client over ctypes (libzmq's C API), for driving the
Fortran host's online mode from tests.  Every receive has a time-out, so a
host that stops answering fails the test instead of hanging it."""
from __future__ import annotations

import ctypes

ZMQ_REQ = 3
ZMQ_LINGER = 17
ZMQ_RCVTIMEO = 27
ZMQ_SNDTIMEO = 28


class ReqClient:
    def __init__(self, libpath: str, endpoint: str, timeout_ms: int = 60000):
        z = ctypes.CDLL(libpath)
        z.zmq_ctx_new.restype = ctypes.c_void_p
        z.zmq_socket.restype = ctypes.c_void_p
        z.zmq_socket.argtypes = [ctypes.c_void_p, ctypes.c_int]
        z.zmq_setsockopt.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t]
        z.zmq_connect.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        z.zmq_send.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int]
        z.zmq_recv.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        z.zmq_close.argtypes = [ctypes.c_void_p]
        z.zmq_ctx_term.argtypes = [ctypes.c_void_p]
        self.z = z
        self.ctx = z.zmq_ctx_new()
        self.sock = z.zmq_socket(self.ctx, ZMQ_REQ)
        for opt, val in ((ZMQ_LINGER, 0), (ZMQ_RCVTIMEO, timeout_ms), (ZMQ_SNDTIMEO, timeout_ms)):
            v = ctypes.c_int(val)
            z.zmq_setsockopt(self.sock, opt, ctypes.byref(v), ctypes.sizeof(v))
        if z.zmq_connect(self.sock, endpoint.encode()) != 0:
            raise RuntimeError(f"zmq_connect({endpoint}) failed")
        self.buf = ctypes.create_string_buffer(65536)

    def request(self, text: str) -> str:
        data = text.encode()
        if self.z.zmq_send(self.sock, data, len(data), 0) < 0:
            raise RuntimeError(f"zmq_send failed for {text!r}")
        n = self.z.zmq_recv(self.sock, self.buf, len(self.buf), 0)
        if n < 0:
            raise TimeoutError(f"no reply to {text!r}")
        return self.buf.raw[:min(n, len(self.buf))].decode()

    def close(self) -> None:
        if self.sock:
            self.z.zmq_close(self.sock)
            self.sock = None
        if self.ctx:
            self.z.zmq_ctx_term(self.ctx)
            self.ctx = None
