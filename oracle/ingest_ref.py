"""CPU restatement of the live-stream ingest -- test infrastructure only.

The wire header of transport/ImageWrapper.java (writeHeader :420-441,
parseHeader :443-476) and one channel of livemode/SimSequenceExtractor.java
(PerChannelBuffer :211-470 with its SortBuffer :472-514) in the reference's
own blocking form: the thread loop pulls frames with getSorted(), which
takes from the raw queue and blocks when it is empty.  Here the raw queue is
the list of frames a test pushes; "blocking" on an empty queue ends the run,
and the sequences completed so far are the result.  Pure Python, for small
streams.
"""
import struct

HEADER = 128
SYNCJITTER = 50


class BrokenHeader(Exception):
    pass


def write_header(seq, w, h, bpp=2, pos_a=0, pos_b=0, pos0=0, pos1=0, pos2=0, t_cam=0, t_cap=0, t_rec=0):
    """ImageWrapper.writeHeader: zero-filled 128 bytes, little endian."""
    b = bytearray(HEADER)
    struct.pack_into("<q", b, 0, seq)
    struct.pack_into("<i", b, 12, w * h)
    b[16] = 1
    b[17] = bpp
    struct.pack_into("<hhhhhh", b, 32, w, h, pos_a, pos_b, pos0, pos1)
    struct.pack_into("<i", b, 44, pos2)
    struct.pack_into("<qqq", b, 48, t_cam, t_cap, t_rec)
    return bytes(b)


def parse_header(b, max_w, max_h):
    """ImageWrapper.parseHeader: the fields, or BrokenHeader."""
    seq, = struct.unpack_from("<q", b, 0)
    ln, = struct.unpack_from("<i", b, 12)
    vers, bpp = struct.unpack_from("<bb", b, 16)
    w, h, pa, pb, p0, p1 = struct.unpack_from("<hhhhhh", b, 32)
    p2, = struct.unpack_from("<i", b, 44)
    tc, tp, tr = struct.unpack_from("<qqq", b, 48)
    if w > max_w or h > max_h:
        raise BrokenHeader("Image too large")
    if ln != w * h:
        raise BrokenHeader("Width x Height != Len")
    if bpp != 1 and bpp != 2:
        raise BrokenHeader("bpp neither 1 nor 2")
    if vers != 1:
        raise BrokenHeader("unsupported version")
    return dict(seq_nr=seq, npix=ln, version=vers, bpp=bpp, width=w, height=h, pos_a=pa, pos_b=pb,
                pos0=p0, pos1=p1, pos2=p2, time_camera=tc, time_capture=tp, time_record=tr)


def get_pixels(payload, w, h, bpp):
    """ImageWrapper.getPixels as unsigned 16-bit patterns: 16-bit LE words,
    or signed bytes widened to short."""
    if bpp == 2:
        return list(struct.unpack_from("<%dH" % (w * h), payload, 0))
    return [(v if v < 128 else v - 256) & 0xFFFF for v in payload[:w * h]]


class _Blocked(Exception):
    """The raw queue is empty: the reference thread would block here."""


class Frame:
    def __init__(self, seq, t_cam, pixels, w=1, h=1):
        self.seq, self.t_cam, self.pixels, self.w, self.h = seq, t_cam, pixels, w, h


class PerChannelBuffer:
    """One channel's sorting and sequence detection."""

    def __init__(self, raws, sync_mode, sync_delay, sync_avr, seq_count, sort_capacity=20, queue_capacity=810):
        self.raws, self.mode = raws, sync_mode
        self.delay, self.avr, self.seq_count = sync_delay, sync_avr, seq_count
        self.sort_cap, self.queue_cap = sort_capacity, queue_capacity
        self.sort = {}
        self.seqNr = 0
        self.simSeq = []
        self.missedSim = 0
        self.dropped = 0
        self.syncFrames = 0
        self.noSync = 0

    # --- SortBuffer ---
    def _sort_add(self, f):
        if len(self.sort) >= self.sort_cap:
            raise OverflowError
        if f.seq in self.sort:
            raise RuntimeError("Same seqNr")
        self.sort[f.seq] = f

    # --- raw queue ---
    def _take(self):
        if self.pos >= len(self.frames):
            raise _Blocked
        f = self.frames[self.pos]
        self.pos += 1
        return f

    def _set_seq_nr(self):
        self.sort.clear()
        self.seqNr = 1 << 63
        while len(self.sort) < self.raws:
            f = self._take()
            self._sort_add(f)
            if f.seq < self.seqNr:
                self.seqNr = f.seq

    def _get_sorted(self):
        if not self.sort:
            return self._sort_from_queue()
        f = self.sort.pop(self.seqNr, None)
        if f is not None:
            self.seqNr += 1
            return f
        return self._sort_from_queue()

    def _sort_from_queue(self):
        f = self._take()
        if f.seq == self.seqNr:
            self.seqNr += 1
            return f
        try:
            self._sort_add(f)
        except OverflowError:
            self.seqNr += 1
            self.dropped += 1
            self._set_seq_nr()
        return self._get_sorted()

    @staticmethod
    def _avr(f):
        return sum(f.pixels) / len(f.pixels)

    def run(self, frames):
        """Feed the frames (arrival order); return the completed sequences
        as lists of frame sequence numbers."""
        self.frames, self.pos = frames, 0
        n_raw = self.raws
        try:
            self._set_seq_nr()
            last = 0
            while True:
                counter = 0
                while True:
                    s = self._get_sorted()
                    cur = s.t_cam
                    if self.mode == 0:
                        if cur - last > self.delay:
                            self._sort_add(s)
                            self.seqNr -= 1
                            self.syncFrames += 1
                            break
                    else:
                        if abs(cur - last - self.delay) < SYNCJITTER:
                            self.syncFrames += 1
                            break
                        if self._avr(s) > self.avr:
                            self.syncFrames += 1
                            self._get_sorted()
                            break
                    last = cur
                    counter += 1
                    if counter >= n_raw * self.seq_count:
                        self.noSync += 1
                        counter = 0
                for _k in range(self.seq_count):
                    seq = []
                    for _i in range(n_raw):
                        f = self._get_sorted()
                        seq.append(f)
                        last = f.t_cam
                    if len(self.simSeq) < self.queue_cap:
                        self.simSeq.append([f.seq for f in seq])
                    else:
                        self.missedSim += 1
        except _Blocked:
            pass
        return self.simSeq
