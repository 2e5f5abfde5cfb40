"""fsim_reconstruct_noalloc: the fused reconstruction with its work buffers in
the library's fixed device array, for callers that cannot allocate.

Through the C ABI: two calls in a row, the second with fewer stacks, each
bitwise equal to fsim_reconstruct of the same stacks (same kernels, same
inputs); the library's device bytes do not change across the calls; a call
that needs more than the fixed array is refused before anything runs.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import fp64, ref

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 128 << 20


def _noalloc(L, rec, raw_dev, out_dev, count):
    f = L.fsim_reconstruct_noalloc
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    f.restype = C.c_int
    return f(rec.handle, C.c_void_p(int(raw_dev)), C.c_void_p(int(out_dev)), int(count), None)


def test_noalloc_matches_reconstruct(fsim_lib):
    from fsim import synthetic
    from fsim._lib import FSIM_ERR_ARG, call, check, lib
    from fsim.linalg import DeviceBuffer
    from fsim.sim_algorithm import Reconstructor
    L = lib()
    n, count = 256, 5
    raws, sps = [], []
    for s in range(count):
        raw, sp = synthetic.sim_stack(n=n, seed=700 + s, nbeads=150, nfil=20)
        raws.append(raw)
        sps.append(sp)
    stk = np.stack(raws)
    rec = Reconstructor(synthetic.make_param(sps[0], n))
    rb = DeviceBuffer(stk.nbytes)
    rb.upload(stk)
    ob_ref = DeviceBuffer(count * rec.out_bytes())
    rec.reconstruct(rb.value, ob_ref.value, count)
    call("fsim_sync")
    want = ob_ref.download((count, 2 * n, 2 * n), np.float32)
    ob = DeviceBuffer(count * rec.out_bytes())
    held = L.fsim_device_bytes()
    print(f"[noalloc] first call: {count} stacks of {n}^2", flush=True)
    check(_noalloc(L, rec, rb.value, ob.value, count))
    call("fsim_sync")
    got = ob.download((count, 2 * n, 2 * n), np.float32)
    assert L.fsim_device_bytes() == held
    for s in range(count):
        assert np.array_equal(got[s], want[s]), f"stack {s}"
    e = fp64.rel_l2(got[0], ref.recon32(ref.recon_dict(sps[0], n), raws[0], threads=8))
    print(f"[noalloc] relL2 vs ref32 {e:.3e}", flush=True)
    assert e < 1e-4
    # second call, fewer stacks (3 and 4), into a fresh output buffer
    ob2 = DeviceBuffer(2 * rec.out_bytes())
    held = L.fsim_device_bytes()
    print("[noalloc] second call: 2 stacks", flush=True)
    check(_noalloc(L, rec, rb.value + 3 * raws[0].nbytes, ob2.value, 2))
    call("fsim_sync")
    got2 = ob2.download((2, 2 * n, 2 * n), np.float32)
    assert L.fsim_device_bytes() == held
    assert np.array_equal(got2[0], want[3]) and np.array_equal(got2[1], want[4])
    # more stacks than the fixed array holds: refused on the host
    per = (3 * 2 * n * n + 4 * n * n) * 8
    too_many = SCRATCH_BYTES // per + 1
    assert _noalloc(L, rec, rb.value, ob.value, too_many) == FSIM_ERR_ARG
    assert b"work buffers" in L.fsim_last_error()
    rb.free()
    ob.free()
    ob2.free()
    ob_ref.free()
