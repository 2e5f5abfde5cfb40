"""GPU parity of the fused reconstruction (fsim_reconstruct) and of the plain
per-op path against the FP32 Java-order oracle and the FP64 re-run.

Tolerance: relative L2 <= 1e-4 against the FP32 oracle (the JTransforms FFT
rounding is not reproducible, so the comparison is to tolerance, not
bitwise).  Each test prints the errors of GPU and oracle against FP64.
"""
import numpy as np
import pytest

from oracle import fp64, ref

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _setup(n, seed, nphase=3, nband=2, harmonic2=0.0, lam=525.0, fade_px=10, apo_bend=0.9):
    from fsim import synthetic
    raw, sp = synthetic.sim_stack(n=n, nphase=nphase, seed=seed, lam=lam,
                                  nbeads=max(30, 300 * n // 512), nfil=max(5, 40 * n // 512),
                                  harmonic2=harmonic2)
    param = synthetic.make_param(sp, n, nphase=nphase, nband=nband, lam=lam)
    p = ref.recon_dict(sp, n, nband=nband, lam=lam, fade_px=fade_px, apo_bend=apo_bend)
    if nband == 3:
        # setPxPy(highest-band shift) stores shift/(bands-1): band 1 sits at k
        p["px"] = [param.dir(d).pX for d in range(len(sp["px"]))]
        p["py"] = [param.dir(d).pY for d in range(len(sp["py"]))]
    return raw, sp, param, p


@pytest.mark.parametrize("n,seed,mode", [(64, 1, 0), (128, 2, 0), (256, 3, 0), (512, 4, 0),
                                         (1024, 5, 0), (64, 1, 1), (512, 4, 1)])
def test_fused_matches_oracle(fsim_lib, n, seed, mode):
    """mode 0: neg components derived from pos (Hermitian pairing), mode 1:
    every component computed."""
    from fsim.sim_algorithm import Reconstructor
    raw, sp, param, p = _setup(n, seed)
    rec = Reconstructor(param, mode=mode)
    got = rec.reconstruct_host(raw)
    r32 = ref.recon32(p, raw, threads=8)
    e = fp64.rel_l2(got, r32)
    if n <= 256:
        r64 = ref.recon64(p, raw)
        print(f"n={n} mode={mode} relL2 gpu-vs-ref32 {e:.3e} gpu-vs-fp64 {fp64.rel_l2(got, r64):.3e} "
              f"ref32-vs-fp64 {fp64.rel_l2(r32, r64):.3e}")
    else:
        print(f"n={n} mode={mode} relL2 gpu-vs-ref32 {e:.3e}")
    assert e < TOL


@pytest.mark.parametrize("nphase,nband,h2,mode", [(5, 2, 0.0, 0), (5, 3, 0.3, 0), (5, 3, 0.3, 1)])
def test_fused_five_phase(fsim_lib, nphase, nband, h2, mode):
    from fsim.sim_algorithm import Reconstructor
    raw, sp, param, p = _setup(128, 11, nphase=nphase, nband=nband, harmonic2=h2)
    got = Reconstructor(param, mode=mode).reconstruct_host(raw)
    r32 = ref.recon32(p, raw, threads=8)
    e = fp64.rel_l2(got, r32)
    print(f"P={nphase} B={nband} mode={mode} relL2 gpu-vs-ref32 {e:.3e}")
    assert e < TOL


def test_fused_batch_equals_single(fsim_lib):
    """Each stack of a batch gives the same image as alone (bitwise: the
    kernels are deterministic and independent per stack)."""
    from fsim import synthetic
    from fsim.sim_algorithm import Reconstructor
    n = 128
    raws = []
    for s in range(4):
        raw, sp = synthetic.sim_stack(n=n, seed=20 + s, nbeads=60, nfil=8)
        raws.append(raw)
    param = synthetic.make_param(sp, n)
    rec = Reconstructor(param)
    batch = rec.reconstruct_host(np.stack(raws))
    for s in range(4):
        single = rec.reconstruct_host(raws[s])
        assert np.array_equal(batch[s], single)


def test_plain_path_matches_oracle(fsim_lib):
    """The per-op vector-API path (ReconstructionRunner's op sequence)."""
    from fsim.linalg import DeviceBuffer
    from fsim.sim_algorithm import PlainPipeline
    n = 128
    raw, sp, param, p = _setup(n, 5, apo_bend=0.9)
    pipe = PlainPipeline(param, wiener=0.05, apo_bend=0.9, apo_cut=2.0, fade_px=10)
    buf = DeviceBuffer(raw.nbytes)
    buf.upload(raw)
    res = pipe.run(buf.value).vectorData()
    r32 = ref.recon32(p, raw, threads=8)
    e = fp64.rel_l2(res, r32)
    print(f"plain relL2 vs ref32 {e:.3e}")
    assert e < TOL


def test_no_fade_live_apodization(fsim_lib):
    """Live-mode filter settings: apodization bend 1 (ReconstructionRunner.java:797), no fade."""
    from fsim.sim_algorithm import Reconstructor
    raw, sp, param, p = _setup(64, 7, fade_px=0, apo_bend=1.0)
    got = Reconstructor(param, apo_bend=1.0, fade_px=0).reconstruct_host(raw)
    assert fp64.rel_l2(got, ref.recon32(p, raw)) < TOL


def test_pairing_matches_full_components(fsim_lib):
    """The Hermitian pairing (mode 0) and the full computation (mode 1) agree
    to FP32 round-off on the same stacks, also with integer shift vectors
    (zero fractional part: no one-pixel offset in the mirrored index)."""
    from fsim import synthetic
    from fsim.sim_algorithm import Reconstructor
    for n, integer in [(128, False), (128, True), (256, False)]:
        raw, sp = synthetic.sim_stack(n=n, seed=60 + n, nbeads=60, nfil=8)
        if integer:
            sp = dict(sp)
            sp["px"] = [round(v) for v in sp["px"]]
            sp["py"] = [round(v) for v in sp["py"]]
        param = synthetic.make_param(sp, n)
        a = Reconstructor(param, mode=0).reconstruct_host(raw)
        b = Reconstructor(param, mode=1).reconstruct_host(raw)
        e = fp64.rel_l2(a, b)
        print(f"n={n} integer={integer} pairing vs full {e:.3e}")
        assert e < 5e-6


def test_unaligned_raw_pointer_matches_aligned(fsim_lib):
    """Raw frames that start off an 8-byte boundary take the 2-byte-load form
    of the separation pass; the image is bitwise the one from an aligned
    buffer (same arithmetic per pixel)."""
    import ctypes as C
    from fsim import synthetic
    from fsim._lib import call
    from fsim.linalg import DeviceBuffer
    from fsim.sim_algorithm import Reconstructor
    n = 128
    raw, sp = synthetic.sim_stack(n=n, seed=71, nbeads=60, nfil=8)
    rec = Reconstructor(synthetic.make_param(sp, n))
    aligned = rec.reconstruct_host(raw)
    rb = DeviceBuffer(raw.nbytes + 8)
    ob = DeviceBuffer(rec.out_bytes())
    call("fsim_memcpy", C.c_void_p(rb.value + 2), raw.ctypes.data_as(C.c_void_p), raw.nbytes, 1)
    rec.reconstruct(rb.value + 2, ob.value, 1)
    call("fsim_sync")
    got = ob.download((2 * n, 2 * n), np.float32)
    rb.free()
    ob.free()
    assert np.array_equal(got, aligned)
