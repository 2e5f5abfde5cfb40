"""Autotuner of the fused pipeline's launch shapes (fsim_autotune_*).

Per size class -- raw side n, directions D, phases P, stored components Cs
(bands B with the Hermitian pairing, 2B-1 without) and a Hermitian result --
the pipeline picks the separation kernel (all units of a row in one
wavefront, ``sep=1``, or one unit per block, ``sep=0``), the column pass's
elements per thread (the Stockham radix) and columns per block (``col``), and
how a call's stacks are chunked over library streams (``chunk``,
``streams``).
``tune`` times every candidate on a channel's class and caches the fastest
for the process; ``save`` / ``load`` persist the cache as text.
"""
import ctypes as C

from . import _lib
from ._lib import TUNE_RESULT, call


def class_of(reconstructor):
    """(n, D, P, Cs, half) of a channel (fsim_channel_class)."""
    v = [C.c_int() for _ in range(5)]
    call("fsim_channel_class", reconstructor.handle, *[C.byref(x) for x in v])
    return tuple(x.value for x in v)


def tune(reconstructor, count=16, reps=5):
    """Time the candidates for ``reconstructor``'s class on ``count`` stacks
    per launch; returns the choice and the measured times (ms per launch)."""
    _lib.init()
    r = TUNE_RESULT()
    call("fsim_autotune", reconstructor.handle, int(count), int(reps), C.byref(r))
    return {"sep": r.sep, "col": r.col, "col_elems": r.col_elems, "col_width": r.col_width,
            "chunk": r.chunk, "streams": r.streams,
            "candidates": r.candidates, "ms_best": r.ms_best, "ms_default": r.ms_default,
            "ms_whole": r.ms_whole, "ms_chunk_default": r.ms_chunk_default,
            "class": dict(zip(("n", "D", "P", "Cs", "half"), class_of(reconstructor)))}


def set_choice(n, D, P, Cs, half, sep, col):
    call("fsim_autotune_set", int(n), int(D), int(P), int(Cs), int(half), int(sep), int(col))


def get_choice(n, D, P, Cs, half):
    """(sep, col) of the class, or None when it has no entry."""
    s, c = C.c_int(), C.c_int()
    call("fsim_autotune_get", int(n), int(D), int(P), int(Cs), int(half), C.byref(s), C.byref(c))
    return None if s.value < 0 else (s.value, c.value)


def set_chunks(n, D, P, Cs, half, chunk, streams):
    """Chunked calls of the class: ``chunk`` stacks per launch (-1 the size
    default, 0 one launch) on ``streams`` library streams."""
    call("fsim_autotune_set_chunks", int(n), int(D), int(P), int(Cs), int(half), int(chunk), int(streams))


def chunk_plan(n, D, P, Cs, half, count):
    """(stacks per chunk, streams) a call of ``count`` stacks uses."""
    k, s = C.c_int(), C.c_int()
    call("fsim_autotune_get_chunks", int(n), int(D), int(P), int(Cs), int(half), int(count), C.byref(k),
         C.byref(s))
    return k.value, s.value


def clear():
    call("fsim_autotune_clear")


def save(path):
    call("fsim_autotune_save", str(path).encode())


def load(path):
    call("fsim_autotune_load", str(path).encode())
