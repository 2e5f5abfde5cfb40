"""Serving form of the reconstruction: the VIGOR live-mode loop.

``ReconstructionServer`` wraps the native continuous-batching server
(fsim_server_*).  Its surface follows livemode/ReconstructionRunner.java:
``queueImage`` offers a stack and returns False when the queue is full
(:140-150, the reference counts these as missed frames); ``take`` waits for
a result (the ``finalRecon`` queue, :57-58, 564).  Channels are fused
``Reconstructor`` objects whose filters stay resident on the device, as the
per-channel vectors of ``ReconstructionThread`` do (:380-413).

``SimSequenceExtractor`` is the live stream's front end (fsim_ingest_*):
camera packets in the transport/ImageWrapper.java wire format are sorted per
channel and cut into SIM sequences at sync frames as
livemode/SimSequenceExtractor.java does, and can feed a
``ReconstructionServer`` directly.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (FsimBusy, INGEST_CONFIG, INGEST_SEQ_INFO, INGEST_STATS, SERVER_CONFIG, SERVER_STATS, call,
                   lib)


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

    def submit_params(self, channel, raw, out, count, params, device=False):
        """fsim_server_submit_params: ``count`` consecutive stacks in ``raw``
        reconstructed into ``count`` consecutive images in ``out`` with the
        SIM parameters ``params`` (an fsim_sim_params or a SimParam)."""
        from .sim_algorithm import sim_params
        sp = params if isinstance(params, _lib.SIM_PARAMS) else sim_params(params)
        t = C.c_uint64()
        rp = raw if isinstance(raw, int) else raw.ctypes.data
        op = out if isinstance(out, int) else out.ctypes.data
        call("fsim_server_submit_params", self._h, int(channel), C.c_void_p(rp), C.c_void_p(op), int(count),
             C.byref(sp), _lib.FSIM_SUBMIT_DEVICE if device else 0, C.byref(t))
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


class ReconstructionBatch:
    """Fixed-shape batch (fsim_batch_*): slot i reconstructs with
    ``reconstructors[i]``; sizes and channels may be mixed.  ``run`` takes
    one device raw address and one device output address per slot (8-byte
    aligned); the class pipelines run concurrently and, with
    ``use_graph``, as one replayed hipGraph whose job lists carry the
    addresses of the run.  With ``persistent`` the slots of sides 256 to
    2048 run through persistent work-list kernels: per stage one launch per
    group of sides (256/512, 1024/2048) walks the (slot class, tile or row
    block) items of all its slots."""

    GRAPH, PERSISTENT = 1, 2

    def __init__(self, reconstructors, use_graph=True, persistent=False):
        _lib.init()
        self._recs = list(reconstructors)
        n = len(self._recs)
        hs = (C.c_void_p * n)(*[r.handle.value for r in self._recs])
        self._h = C.c_void_p()
        flags = (self.GRAPH if use_graph else 0) | (self.PERSISTENT if persistent else 0)
        call("fsim_batch_create", hs, n, flags, C.byref(self._h))
        self.count = n

    def run(self, raw_ptrs, out_ptrs, stream=None):
        if len(raw_ptrs) != self.count or len(out_ptrs) != self.count:
            raise ValueError("one raw and one output address per slot")
        rp = (C.c_void_p * self.count)(*[int(p) for p in raw_ptrs])
        op = (C.c_void_p * self.count)(*[int(p) for p in out_ptrs])
        call("fsim_batch_run", self._h, rp, op, C.c_void_p(stream))

    def info(self):
        nc, cap, nb = C.c_int(), C.c_int(), C.c_int64()
        call("fsim_batch_info", self._h, C.byref(nc), C.byref(cap), C.byref(nb))
        return {"classes": nc.value, "captured": bool(cap.value), "device_bytes": nb.value}

    def close(self):
        if self._h is not None and self._h.value and _lib._lib is not None:
            lib().fsim_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pinned(shape, dtype):
    """A numpy array over pinned host memory (fsim_alloc_host)."""
    nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
    p = C.c_void_p()
    call("fsim_alloc_host", nbytes, C.byref(p))
    buf = (C.c_char * nbytes).from_address(p.value)
    arr = np.frombuffer(buf, dtype=dtype).reshape(shape)
    return arr, p


class SimSequenceExtractor:
    """Per-channel sequence assembly of camera packets (fsim_ingest):
    ``push`` one packet (128-byte header + pixels), ``pop`` an assembled
    sequence, or ``attach`` a ReconstructionServer and ``take`` the
    reconstructions in submission order.  The configuration names follow the
    reference's Conf entries (SyncMode, SyncFrameDelay, SyncFrameAvr,
    SyncFrameFreq; SimSequenceExtractor.java:59-90)."""

    def __init__(self, channel_numbers, ndir=3, nphase=3, sync_mode=0, sync_delay=12000, sync_avr=3000,
                 seq_count=1, sort_capacity=20, queue_capacity=None, max_size=2048):
        raws = ndir * nphase
        cfg = INGEST_CONFIG(ndir, nphase, int(sync_mode), int(sync_delay), int(sync_avr), int(seq_count),
                            int(sort_capacity), int(queue_capacity or raws * raws * 10), int(max_size),
                            int(max_size))
        self.raws = raws
        self.nch = len(channel_numbers)
        nums = (C.c_int * self.nch)(*channel_numbers)
        self._h = C.c_void_p()
        call("fsim_ingest_create", C.byref(cfg), nums, self.nch, C.byref(self._h))
        self._n = None
        self._server = None
        self._max = int(max_size)
        self._popbuf = None

    def close(self):
        if self._h is not None and self._h.value and _lib._lib is not None:
            lib().fsim_ingest_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def push(self, packet):
        """One packet as bytes / bytearray / uint8 array."""
        buf = np.frombuffer(packet, dtype=np.uint8) if not isinstance(packet, np.ndarray) else packet
        call("fsim_ingest_push", self._h, C.c_void_p(buf.ctypes.data), int(buf.size))

    def pop(self, channel):
        """(frames [raws, h, w] uint16, info) of the next sequence of channel
        index ``channel``, or None."""
        mp = self.raws * self._max * self._max
        if self._popbuf is None:
            self._popbuf = np.empty(mp, np.uint16)
        out = self._popbuf
        info = INGEST_SEQ_INFO()
        got = lib().fsim_ingest_pop(self._h, int(channel), C.c_void_p(out.ctypes.data), int(mp), C.byref(info))
        if got < 0:
            _lib.check(got)
        if got == 0:
            return None
        npx = self.raws * info.width * info.height
        return out[:npx].reshape(self.raws, info.height, info.width).copy(), info

    def stats(self, channel):
        st = INGEST_STATS()
        call("fsim_ingest_get_stats", self._h, int(channel), C.byref(st))
        return {f: getattr(st, f) for f, _ in INGEST_STATS._fields_}

    def attach(self, server, server_channels, n, max_inflight=8):
        """Feed ``server`` directly: channel index c goes to server channel
        server_channels[c]; frames must be n x n."""
        ids = (C.c_int * self.nch)(*server_channels)
        call("fsim_ingest_attach_server", self._h, server._h, ids, int(n), int(max_inflight))
        self._n = n
        self._server = server

    def take(self, timeout_ms=-1):
        """The oldest submitted sequence's reconstruction ((2n)^2 float32) and
        its info; raises FsimBusy on timeout or when nothing is in flight."""
        out = np.empty((2 * self._n, 2 * self._n), np.float32)
        info = INGEST_SEQ_INFO()
        call("fsim_ingest_result", self._h, int(timeout_ms), C.c_void_p(out.ctypes.data), C.byref(info))
        return out, info


def read_registration_file(path):
    """The registration file as Registration reads it (Registration.java:
    70-80, 334-385): ``...=width`` and ``...=height`` on the first two lines,
    then an "X Trans" and a "Y Trans" block of space-separated rows of
    source positions (tokens that do not parse are skipped), each ended by an
    empty line.  Returns (xtrans, ytrans) float32 arrays [height, width]."""
    with open(path) as f:
        lines = f.read().splitlines()
    width = int(lines[0].split("=")[1])
    height = int(lines[1].split("=")[1])
    tabs = {"x": np.zeros((height, width), np.float32), "y": np.zeros((height, width), np.float32)}
    rows = {"x": 0, "y": 0}
    cur = None
    i = 2
    while i < len(lines):
        line = lines[i]
        if line.startswith("X Trans") or line.startswith("Y Trans"):
            cur = "x" if line.startswith("X") else "y"
            i += 1
            if i >= len(lines):
                break
            line = lines[i]
        elif line == "":
            cur = None
        if cur is not None:
            col = 0
            for word in line.split(" "):
                try:
                    v = np.float32(float(word))
                except ValueError:
                    continue
                tabs[cur][rows[cur], col] = v
                col += 1
            rows[cur] += 1
        i += 1
    return tabs["x"], tabs["y"]


class Registration:
    """Registration of widefield and reconstructed images on the device
    (fsim_registration_*): ``register`` remaps ``count`` images through the
    per-pixel source positions with the reference's bilinear transPixel."""

    def __init__(self, xtrans, ytrans):
        _lib.init()
        xt = np.ascontiguousarray(xtrans, np.float32)
        yt = np.ascontiguousarray(ytrans, np.float32)
        if xt.shape != yt.shape:
            raise ValueError("missmatch in dimensions")
        self.height, self.width = xt.shape
        self._h = C.c_void_p()
        call("fsim_registration_create", C.c_void_p(xt.ctypes.data), C.c_void_p(yt.ctypes.data), self.width,
             self.height, C.byref(self._h))

    @staticmethod
    def from_file(path):
        return Registration(*read_registration_file(path))

    def close(self):
        if self._h is not None and self._h.value and _lib._lib is not None:
            lib().fsim_registration_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def register(self, src_dev, dst_dev, w, h, count=1, widefield=False, stream=None):
        """registerReconImage / registerWfImage for ``count`` images of w x h
        (device addresses)."""
        call("fsim_register_images", self._h, 1 if widefield else 0, C.c_void_p(int(src_dev)),
             C.c_void_p(int(dst_dev)), int(w), int(h), int(count), C.c_void_p(stream) if stream else None)


class PerChannel:
    """ReconstructionRunner.PerChannel (ReconstructionRunner.java:72-110): a
    channel's SIM parameters and OTF from its configuration folder
    ("channel-<label>"), the pixel size from the global settings, and the
    live filter settings setupFilters applies (:773-813): Wiener 0.05,
    attenuation 0.99 / 1.25 switched on, apodization bend 1 / cutoff 2, fade
    10 px, band 0's OTF for every component.  ``offset`` is read as the
    reference reads it (and, as there, not applied)."""

    def __init__(self, fld, image_size, microns, label):
        from .sim_algorithm import OtfProvider, SimParam
        self.param = SimParam.loadConfig(fld)
        self.param.setPxlSize(image_size, microns)
        self.param.otf(OtfProvider.loadFromConfig(fld))
        self.offset = float(np.float32(fld.getDbl("offset").val()))
        self.chNumber = fld.getInt("ChannelNumber").val()
        self.label = label
        self.wienParam, self.attStr, self.attFWHM, self.useAttenuation = 0.05, 0.99, 1.25, True

    def reconstructor(self, mode=0):
        from .sim_algorithm import Reconstructor
        otf = self.param.otf()
        otf.setAttenuation(self.attStr, self.attFWHM)
        otf.switchAttenuation(self.useAttenuation)
        return Reconstructor(self.param, wiener=self.wienParam, apo_bend=1.0, apo_cut=2.0, fade_px=10, mode=mode,
                             otf_band0=True)


def load_live_channels(cfg, labels):
    """The channel set-up of ReconstructionRunner's constructor
    (ReconstructionRunner.java:160-200): ``cfg`` is the vigor-settings
    folder, ``labels`` the channels to run (e.g. ["488", "568"])."""
    size = cfg.getInt("RawPxlCount").val()
    microns = cfg.getDbl("RawPxlSize").val()
    return [PerChannel(cfg.cd("channel-" + lb), size, microns, lb) for lb in labels]


def sequence_extractor_from_conf(cfg, channels, **kw):
    """SimSequenceExtractor with the vigor-settings' SyncMode, SyncFrameFreq,
    SyncFrameAvr, SyncFrameDelay, NrAngles and NrPhases
    (SimSequenceExtractor.java:59-90); channels: PerChannel objects."""
    return SimSequenceExtractor([c.chNumber for c in channels], ndir=cfg.getInt("NrAngles").val(),
                                nphase=cfg.getInt("NrPhases").val(), sync_mode=cfg.getInt("SyncMode").val(),
                                sync_delay=cfg.getInt("SyncFrameDelay").val(),
                                sync_avr=cfg.getInt("SyncFrameAvr").val(),
                                seq_count=cfg.getInt("SyncFrameFreq").val(), **kw)
