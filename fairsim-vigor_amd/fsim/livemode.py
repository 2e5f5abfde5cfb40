"""Serving form of the reconstruction: the VIGOR live-mode loop.

``ReconstructionServer`` wraps the native continuous-batching server
(fsim_server_*).  Its surface follows livemode/ReconstructionRunner.java:
``queueImage`` offers a stack and returns False when the queue is full
(:140-150, the reference counts these as missed frames); ``take`` waits for
a result (the ``finalRecon`` queue, :57-58, 564).  Channels are fused
``Reconstructor`` objects whose filters stay resident on the device, as the
per-channel vectors of ``ReconstructionThread`` do (:380-413).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import FsimBusy, SERVER_CONFIG, SERVER_STATS, call, lib


class ReconstructionServer:
    def __init__(self, mem_budget=8 << 30, max_batch=32, max_queue=1024, nslots=4):
        _lib.init()
        cfg = SERVER_CONFIG(int(mem_budget), int(max_batch), int(max_queue), int(nslots))
        self._h = C.c_void_p()
        call("fsim_server_create", C.byref(cfg), C.byref(self._h))
        self._channels = []
        self._keep = {}

    def close(self):
        if self._h is not None and self._h.value and _lib._lib is not None:
            lib().fsim_server_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_channel(self, reconstructor):
        cid = C.c_int()
        call("fsim_server_add_channel", self._h, reconstructor.handle, C.byref(cid))
        self._channels.append(reconstructor)
        return cid.value

    def submit(self, channel, raw, out, device=False):
        """raw/out: device addresses (device=True) or host addresses / arrays;
        host arrays must stay alive until ``wait`` returns.  Raises FsimBusy
        when the queue is full."""
        t = C.c_uint64()
        rp = raw if isinstance(raw, int) else raw.ctypes.data
        op = out if isinstance(out, int) else out.ctypes.data
        call("fsim_server_submit", self._h, int(channel), C.c_void_p(rp), C.c_void_p(op),
             _lib.FSIM_SUBMIT_DEVICE if device else 0, C.byref(t))
        if not isinstance(raw, int):
            self._keep[t.value] = (raw, out)
        return t.value

    def queueImage(self, channel, raw, out):
        """ReconstructionRunner.queueImage: False when the request is refused."""
        try:
            return self.submit(channel, raw, out)
        except FsimBusy:
            return False

    def wait(self, ticket, timeout_ms=-1):
        """Block until the request is done; returns its latency in ms."""
        lat = C.c_double()
        try:
            call("fsim_server_wait", self._h, C.c_uint64(ticket), int(timeout_ms), C.byref(lat))
        finally:
            self._keep.pop(ticket, None)
        return lat.value

    take = wait

    def stats(self, reset=False):
        st = SERVER_STATS()
        call("fsim_server_get_stats", self._h, C.byref(st), 1 if reset else 0)
        return {k: getattr(st, k) for k, _ in SERVER_STATS._fields_}


def pinned(shape, dtype):
    """A numpy array over pinned host memory (fsim_alloc_host)."""
    nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
    p = C.c_void_p()
    call("fsim_alloc_host", nbytes, C.byref(p))
    buf = (C.c_char * nbytes).from_address(p.value)
    arr = np.frombuffer(buf, dtype=dtype).reshape(shape)
    return arr, p
