"""Live-stream ingest (fsim_wire_*, fsim_ingest_*) against the CPU
restatement of ImageWrapper and SimSequenceExtractor (oracle/ingest_ref.py).

No GPU: the wire format and the sequence assembly are host code.  The
header layout is pinned byte by byte against the header description's
offsets; the assembly is pinned by known answers on clean streams (a sync
frame every seq_count sequences gives exactly those sequences) and compared
with the oracle on shuffled, gapped and duplicated-free random streams in
both sync modes.
"""
import ctypes as C
import struct

import numpy as np
import pytest

from oracle import ingest_ref as ir

W = H = 4


class HDR(C.Structure):
    _fields_ = [("seq_nr", C.c_int64), ("npix", C.c_int32), ("version", C.c_int32), ("bpp", C.c_int32),
                ("width", C.c_int32), ("height", C.c_int32), ("pos_a", C.c_int32), ("pos_b", C.c_int32),
                ("pos0", C.c_int32), ("pos1", C.c_int32), ("pos2", C.c_int32), ("time_camera", C.c_int64),
                ("time_capture", C.c_int64), ("time_record", C.c_int64)]


def _lib():
    from fsim import _lib
    return _lib


def test_header_layout_and_roundtrip():
    L = _lib().lib()
    h = HDR(seq_nr=0x0102030405060708, width=640, height=480, bpp=2, pos_a=2, pos_b=1, pos0=7, pos1=488,
            pos2=123456, time_camera=1_700_000_000_000_001, time_capture=-1, time_record=0)
    buf = (C.c_uint8 * 128)()
    assert L.fsim_wire_write_header(C.byref(h), buf) == 0
    b = bytes(buf)
    # offsets of transport/header-definition.md, little endian
    assert struct.unpack_from("<q", b, 0)[0] == 0x0102030405060708
    assert struct.unpack_from("<i", b, 12)[0] == 640 * 480
    assert b[16] == 1 and b[17] == 2
    assert struct.unpack_from("<hhhhhh", b, 32) == (640, 480, 2, 1, 7, 488)
    assert struct.unpack_from("<i", b, 44)[0] == 123456
    assert struct.unpack_from("<qqq", b, 48) == (1_700_000_000_000_001, -1, 0)
    assert b[72:] == bytes(56) and b[8:12] == bytes(4) and b[18:32] == bytes(14)
    # the oracle writes the same bytes
    assert b == ir.write_header(h.seq_nr, 640, 480, 2, 2, 1, 7, 488, 123456, h.time_camera, -1, 0)
    back = HDR()
    assert L.fsim_wire_parse_header(buf, 1024, 1024, C.byref(back)) == 0
    for f, _ in HDR._fields_:
        want = 640 * 480 if f == "npix" else (1 if f == "version" else getattr(h, f))
        assert getattr(back, f) == want, f


@pytest.mark.parametrize("mutate,msg", [
    (lambda b: struct.pack_into("<i", b, 12, 17), "Width x Height != Len"),
    (lambda b: b.__setitem__(17, 3), "bpp neither 1 nor 2"),
    (lambda b: b.__setitem__(16, 2), "unsupported version"),
    (lambda b: struct.pack_into("<hh", b, 32, 2048, 4), "too large"),
])
def test_header_checks(mutate, msg):
    L = _lib().lib()
    b = bytearray(ir.write_header(5, 8, 8))
    mutate(b)
    buf = (C.c_uint8 * 128).from_buffer_copy(bytes(b))
    out = HDR()
    assert L.fsim_wire_parse_header(buf, 1024, 1024, C.byref(out)) == -1
    assert msg in L.fsim_last_error().decode()
    with pytest.raises(ir.BrokenHeader, match=msg.split()[0]):
        ir.parse_header(bytes(b), 1024, 1024)


class CFG(C.Structure):
    _fields_ = [("ndir", C.c_int), ("nphase", C.c_int), ("sync_mode", C.c_int), ("sync_delay", C.c_int64),
                ("sync_avr", C.c_int), ("seq_count", C.c_int), ("sort_capacity", C.c_int),
                ("queue_capacity", C.c_int), ("max_width", C.c_int), ("max_height", C.c_int)]


class INFO(C.Structure):
    _fields_ = [("channel", C.c_int), ("width", C.c_int), ("height", C.c_int), ("first_seq_nr", C.c_int64),
                ("time_camera", C.c_int64), ("pos2", C.c_int32)]


class STATS(C.Structure):
    _fields_ = [("frames", C.c_int64), ("sequences", C.c_int64), ("missed_sequences", C.c_int64),
                ("dropped_frames", C.c_int64), ("sync_frames", C.c_int64), ("no_sync_windows", C.c_int64)]


def _pixels(seq, bright):
    px = np.full(W * H, 5000 if bright else 100, np.uint16)
    px[0], px[1] = seq & 0xFFFF, (seq >> 16) & 0xFFFF
    return px


def _packet(f, channel=0):
    return ir.write_header(f.seq, W, H, 2, pos1=channel, t_cam=f.t_cam) + f.pixels_np.tobytes()


def _frames(n, mode, raws, seq_count, delay, period, seed, shuffle=0, drop=()):
    """A camera stream: a sync event before every seq_count * raws frames
    (mode 0: a time gap; mode 1: a bright frame, or a gap of `delay`)."""
    rng = np.random.default_rng(seed)
    frames, t, seq = [], 1000, 0
    per = raws * seq_count
    while len(frames) < n:
        if mode == 0:
            t += delay + 500
        else:
            if rng.random() < 0.5:
                bright = ir.Frame(seq, t, None)
                bright.pixels_np = _pixels(seq, True)
                frames.append(bright)
                seq += 1
                t += period
                skip = ir.Frame(seq, t, None)
                skip.pixels_np = _pixels(seq, False)
                frames.append(skip)
                seq += 1
                t += period
            else:
                t += delay - period + int(rng.integers(-40, 41))   # camera-time gap within the jitter of delay
                sync = ir.Frame(seq, t, None)
                sync.pixels_np = _pixels(seq, False)
                frames.append(sync)
                seq += 1
                t += period
        for _ in range(per):
            f = ir.Frame(seq, t, None)
            f.pixels_np = _pixels(seq, False)
            frames.append(f)
            seq += 1
            t += period
    frames = [f for f in frames if f.seq not in drop]
    for f in frames:
        f.pixels = [int(v) for v in f.pixels_np]
    if shuffle:
        # local reordering: swap neighbours within windows of `shuffle`
        for i in range(0, len(frames) - shuffle, shuffle):
            win = frames[i:i + shuffle]
            rng.shuffle(win)
            frames[i:i + shuffle] = win
    return frames


def _native(frames, mode, raws, seq_count, delay, avr, sort_cap=20, queue_cap=810):
    L = _lib().lib()
    cfg = CFG(ndir=raws // 3, nphase=3, sync_mode=mode, sync_delay=delay, sync_avr=avr, seq_count=seq_count,
              sort_capacity=sort_cap, queue_capacity=queue_cap, max_width=64, max_height=64)
    h = C.c_void_p()
    chans = (C.c_int * 1)(0)
    assert L.fsim_ingest_create(C.byref(cfg), chans, 1, C.byref(h)) == 0
    try:
        for f in frames:
            p = _packet(f)
            buf = (C.c_uint8 * len(p)).from_buffer_copy(p)
            assert L.fsim_ingest_push(h, buf, len(p)) == 0, L.fsim_last_error()
        out = np.zeros(raws * W * H, np.uint16)
        seqs = []
        info = INFO()
        while L.fsim_ingest_pop(h, 0, out.ctypes.data_as(C.POINTER(C.c_uint16)), out.size, C.byref(info)) == 1:
            fr = out.reshape(raws, W * H)
            ids = [int(fr[i, 0]) | (int(fr[i, 1]) << 16) for i in range(raws)]
            assert info.first_seq_nr == ids[0] and info.width == W
            seqs.append(ids)
        st = STATS()
        assert L.fsim_ingest_get_stats(h, 0, C.byref(st)) == 0
        return seqs, st
    finally:
        L.fsim_ingest_destroy(h)


def _oracle(frames, mode, raws, seq_count, delay, avr, sort_cap=20, queue_cap=810):
    pcb = ir.PerChannelBuffer(raws, mode, delay, avr, seq_count, sort_cap, queue_cap)
    return pcb.run(frames), pcb


def test_clean_stream_known_answer_mode0():
    raws, sc = 9, 2
    frames = _frames(200, 0, raws, sc, 20000, 1000, 1)
    seqs, st = _native(frames, 0, raws, sc, 20000, 3000)
    # the first sync gap is the one before frame 0 (last time stamp starts
    # at 0); every sequence is 9 consecutive frames, 2 per sync
    assert len(seqs) >= 20
    for j, s in enumerate(seqs):
        assert s == list(range(9 * j, 9 * j + 9))
    assert st.sync_frames == (len(seqs) + 1) // 2 or st.sync_frames == len(seqs) // 2 + 1


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("mode", [0, 1])
def test_random_streams_match_oracle(mode, seed):
    raws, sc = 9, 1 + seed % 3
    rng = np.random.default_rng(100 + seed)
    drop = set(int(v) for v in rng.choice(400, size=seed, replace=False)) if seed % 2 else set()
    frames = _frames(360, mode, raws, sc, 20000, 1000, seed, shuffle=3 + seed % 4, drop=drop)
    got, st = _native(frames, mode, raws, sc, 20000, 3000)
    want, pcb = _oracle(frames, mode, raws, sc, 20000, 3000)
    print(f"mode {mode} seed {seed}: {len(got)} sequences, sync {st.sync_frames}, dropped {st.dropped_frames}")
    assert got == want
    assert st.sync_frames == pcb.syncFrames and st.dropped_frames == pcb.dropped
    assert st.no_sync_windows == pcb.noSync


def test_sort_overflow_resyncs_like_reference():
    """A gap larger than the sort buffer: the reference drops the frame that
    overflows, clears the buffer and restarts at the lowest of the next
    ndir*nphase frames."""
    raws, sc = 9, 1
    frames = _frames(300, 0, raws, sc, 20000, 1000, 3, drop=set(range(40, 70)))
    got, st = _native(frames, 0, raws, sc, 20000, 3000, sort_cap=12)
    want, pcb = _oracle(frames, 0, raws, sc, 20000, 3000, sort_cap=12)
    assert got == want and st.dropped_frames == pcb.dropped


def test_queue_capacity_drops_sequences():
    raws, sc = 9, 3
    frames = _frames(200, 0, raws, sc, 20000, 1000, 2)
    got, st = _native(frames, 0, raws, sc, 20000, 3000, queue_cap=4)
    want, pcb = _oracle(frames, 0, raws, sc, 20000, 3000, queue_cap=4)
    assert got == want and len(got) == 4
    assert st.missed_sequences == pcb.missedSim > 0


def test_eight_bit_frames_sign_extend_and_errors():
    L = _lib().lib()
    cfg = CFG(ndir=1, nphase=3, sync_mode=0, sync_delay=100, sync_avr=3000, seq_count=1, sort_capacity=20,
              queue_capacity=10, max_width=64, max_height=64)
    h = C.c_void_p()
    chans = (C.c_int * 2)(488, 561)
    assert L.fsim_ingest_create(C.byref(cfg), chans, 2, C.byref(h)) == 0
    try:
        pay = bytes([0, 1, 127, 128, 200, 255] + [7] * 10)
        for s in range(6):
            p = ir.write_header(s, W, H, 1, pos1=561, t_cam=1000 + 10 * s) + pay
            buf = (C.c_uint8 * len(p)).from_buffer_copy(p)
            assert L.fsim_ingest_push(h, buf, len(p)) == 0
        out = np.zeros(3 * W * H, np.uint16)
        info = INFO()
        assert L.fsim_ingest_pop(h, 1, out.ctypes.data_as(C.POINTER(C.c_uint16)), out.size, C.byref(info)) == 1
        assert list(out[:16]) == ir.get_pixels(pay, W, H, 1)
        assert list(out[:6]) == [0, 1, 127, 0xFF80, 0xFFC8, 0xFFFF]
        assert info.channel == 1
        # unknown channel, short payload
        p = ir.write_header(9, W, H, 2, pos1=3)
        buf = (C.c_uint8 * (len(p) + 32)).from_buffer_copy(p + bytes(32))
        assert L.fsim_ingest_push(h, buf, len(p) + 32) == -1
        assert "w/o channel" in L.fsim_last_error().decode()
        p = ir.write_header(9, W, H, 2, pos1=488)
        buf = (C.c_uint8 * (len(p) + 8)).from_buffer_copy(p + bytes(8))
        assert L.fsim_ingest_push(h, buf, len(p) + 8) == -1
    finally:
        L.fsim_ingest_destroy(h)


def test_bright_sync_known_answer_mode1():
    """Mode 1, bright sync frames only: the bright frame and the frame after
    it are dropped, the next 9 form the sequence (SimSequenceExtractor.java:
    402-413)."""
    frames, t = [], 5000
    for blk in range(6):
        for j in range(11):
            s = blk * 11 + j
            f = ir.Frame(s, t, None)
            f.pixels_np = _pixels(s, j == 0)
            f.pixels = [int(v) for v in f.pixels_np]
            frames.append(f)
            t += 1000
    got, st = _native(frames, 1, 9, 1, 20000, 3000)
    assert got == [list(range(11 * b + 2, 11 * b + 11)) for b in range(6)]
    assert st.sync_frames == 6
    assert got == _oracle(frames, 1, 9, 1, 20000, 3000)[0]
