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
                                         (1024, 5, 0), (2048, 6, 0), (64, 1, 1), (512, 4, 1)])
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


@pytest.mark.parametrize("ndir", [2, 4])
def test_fused_other_direction_counts(fsim_lib, ndir):
    """Two directions (2 x 3 frames: the separation with all units of a row in
    one wavefront) and four (4 x 3 frames: one unit per block)."""
    from fsim import synthetic
    from fsim.sim_algorithm import Reconstructor
    n = 128
    raw, sp = synthetic.sim_stack(n=n, ndir=ndir, seed=30 + ndir, nbeads=60, nfil=8)
    param = synthetic.make_param(sp, n, ndir=ndir)
    got = Reconstructor(param).reconstruct_host(raw)
    e = fp64.rel_l2(got, ref.recon32(ref.recon_dict(sp, n), raw, threads=8))
    print(f"ndir={ndir} relL2 gpu-vs-ref32 {e:.3e}")
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


def test_half_spectrum_nyquist_column(fsim_lib):
    """With a real OTF the result spectrum is Hermitian and only its columns
    0..m/2 are formed, followed by a half-length complex-to-real row pass.
    Shift vectors longer than n/2 along x paste contributions into the
    Nyquist column m/2 (and, mirrored, across it); the image still matches
    the FP32 oracle and the full-width computation (mode 1)."""
    from fsim import synthetic
    from fsim.sim_algorithm import Reconstructor
    n = 64
    raw, sp = synthetic.sim_stack(n=n, seed=83, nbeads=40, nfil=6)
    sp = dict(sp)
    sp["px"] = [0.56 * n, -0.61 * n, 0.3 * n]
    sp["py"] = [0.05 * n, 0.1 * n, -0.52 * n]
    param = synthetic.make_param(sp, n)
    a = Reconstructor(param, mode=0).reconstruct_host(raw)
    b = Reconstructor(param, mode=1).reconstruct_host(raw)
    e_full = fp64.rel_l2(a, b)
    e_ref = fp64.rel_l2(a, ref.recon32(ref.recon_dict(sp, n), raw))
    print(f"half vs full {e_full:.3e}, half vs ref32 {e_ref:.3e}")
    assert e_full < 5e-6
    assert e_ref < TOL


def test_unaligned_output_takes_full_width_pass(fsim_lib):
    """The half-spectrum pass stores output pixel pairs; an output pointer
    that is only 4-byte aligned is served by the full-width pass, which
    agrees to FP32 round-off."""
    from fsim import synthetic
    from fsim._lib import call
    from fsim.linalg import DeviceBuffer
    from fsim.sim_algorithm import Reconstructor
    n = 128
    raw, sp = synthetic.sim_stack(n=n, seed=84, nbeads=60, nfil=8)
    rec = Reconstructor(synthetic.make_param(sp, n))
    aligned = rec.reconstruct_host(raw)
    rb = DeviceBuffer(raw.nbytes)
    rb.upload(raw)
    ob = DeviceBuffer(rec.out_bytes() + 8)
    rec.reconstruct(rb.value, ob.value + 4, 1)
    call("fsim_sync")
    whole = ob.download((rec.out_bytes() // 4 + 2,), np.float32)
    got = whole[1:1 + (2 * n) ** 2].reshape(2 * n, 2 * n)
    rb.free()
    ob.free()
    e = fp64.rel_l2(got, aligned)
    print(f"unaligned output vs aligned {e:.3e}")
    assert e < 5e-6


def _multi_stacks(n, cases, nphase=3, nband=2):
    """Stacks with different shifts / phase offsets / modulations; the
    parameters of each as a SimParam on one OTF."""
    from fsim import synthetic
    raws, params, sps = [], [], []
    for s, (k_frac, ang, pha_off, mod) in enumerate(cases):
        raw, sp = synthetic.sim_stack(n=n, nphase=nphase, seed=60 + s, nbeads=max(30, 300 * n // 512),
                                      nfil=max(5, 40 * n // 512), k_frac=k_frac, angle0_deg=ang)
        prm = synthetic.make_param(sp, n, nphase=nphase, nband=nband, modulation=mod)
        for d in range(3):
            prm.dir(d).setPhaOff(pha_off)
        raws.append(raw)
        params.append(prm)
        sps.append(sp)
    return np.stack(raws), params, sps


def test_reconstruct_params_equals_own_channel(fsim_lib):
    """fsim_reconstruct_params: each stack of a batch with its own shifts,
    phases and modulations gives the image of a channel built for exactly
    those parameters (bitwise: same constants, same kernels)."""
    from fsim.linalg import DeviceBuffer
    from fsim.sim_algorithm import Reconstructor
    from fsim._lib import call
    n = 128
    raws, params, _ = _multi_stacks(n, [(0.85, 3.0, 0.0, 0.7), (0.78, 11.0, 0.4, 0.55), (0.8, -7.0, -1.1, 0.9)])
    rec = Reconstructor(params[0])
    rb = DeviceBuffer(raws.nbytes)
    rb.upload(raws)
    ob = DeviceBuffer(len(params) * rec.out_bytes())
    for _ in range(2):   # second call reuses the per-stack workspace
        rec.reconstruct_params(rb.value, ob.value, params)
        call("fsim_sync")
        got = ob.download((len(params), 2 * n, 2 * n), np.float32)
        for s, prm in enumerate(params):
            own = Reconstructor(prm).reconstruct_host(raws[s])
            assert np.array_equal(got[s], own), f"stack {s}"
    # a smaller batch after a larger one
    rec.reconstruct_params(rb.value, ob.value, params[:1])
    call("fsim_sync")
    assert np.array_equal(ob.download((1, 2 * n, 2 * n), np.float32)[0], rec.reconstruct_host(raws[0]))
    rb.free()
    ob.free()


def test_reconstruct_params_chunked_equals_own_channel(fsim_lib):
    """At 512^2 a call of more stacks than one chunk (8) runs as chunks on
    the library streams, each stream with its own work buffers and the
    job list carrying them; every stack is still bitwise the image of a
    channel built for its parameters, in a first call and a repeated one."""
    from fsim.linalg import DeviceBuffer
    from fsim.sim_algorithm import Reconstructor
    from fsim._lib import call
    n = 512
    raws, params, _ = _multi_stacks(n, [(0.85, 3.0, 0.0, 0.7), (0.78, 11.0, 0.4, 0.55), (0.8, -7.0, -1.1, 0.9)])
    order = [0, 1, 2, 2, 0, 1, 1, 0, 2, 0, 1]   # 11 stacks: chunks of 8 and 3
    raw_b = np.ascontiguousarray(raws[order])
    prm_b = [params[k] for k in order]
    rec = Reconstructor(params[0])
    rb = DeviceBuffer(raw_b.nbytes)
    rb.upload(raw_b)
    ob = DeviceBuffer(len(order) * rec.out_bytes())
    own = [Reconstructor(prm).reconstruct_host(raws[k]) for k, prm in enumerate(params)]
    for _ in range(2):
        rec.reconstruct_params(rb.value, ob.value, prm_b)
        call("fsim_sync")
        got = ob.download((len(order), 2 * n, 2 * n), np.float32)
        for s, k in enumerate(order):
            assert np.array_equal(got[s], own[k]), f"stack {s}"
    rb.free()
    ob.free()


def test_estimate_then_reconstruct_matches_oracle(fsim_lib):
    """Estimate-then-reconstruct flow: parameters estimated per stack (fsim_estimate), then
    every stack reconstructed with its own estimate (fsim_reconstruct_params);
    compared with the FP32 oracle reconstructing with the same parameters."""
    from fsim.linalg import DeviceBuffer
    from fsim.sim_algorithm import Reconstructor, estimate_batch, params_from_estimate
    from fsim import synthetic
    from fsim._lib import call
    n = 256
    raws, params, sps = _multi_stacks(n, [(0.85, 3.0, 0.0, 0.7), (0.8, 9.0, 0.0, 0.7)])
    base = synthetic.make_param(sps[0], n)
    est = estimate_batch(base, 1, 0.6, images=raws.reshape(len(sps), 3, 3, n, n).astype(np.float32))
    sim = [params_from_estimate(base, e) for e in est]
    rec = Reconstructor(base)
    rb = DeviceBuffer(raws.nbytes)
    rb.upload(raws)
    ob = DeviceBuffer(len(sim) * rec.out_bytes())
    rec.reconstruct_params(rb.value, ob.value, sim)
    call("fsim_sync")
    got = ob.download((len(sim), 2 * n, 2 * n), np.float32)
    for s in range(len(sim)):
        p = ref.recon_dict(sps[s], n)
        p["px"] = [sim[s].px[d] for d in range(3)]
        p["py"] = [sim[s].py[d] for d in range(3)]
        p["phases"] = [[sim[s].phases[d][i] for i in range(3)] for d in range(3)]
        p["modul"] = [[sim[s].modul[d][b] for b in range(2)] for d in range(3)]
        e = fp64.rel_l2(got[s], ref.recon32(p, raws[s], threads=8))
        k_err = max(abs(sim[s].px[d] - sps[s]["px"][d]) + abs(sim[s].py[d] - sps[s]["py"][d]) for d in range(3))
        print(f"stack {s}: |k - k_true| {k_err:.3f} px, relL2 gpu-vs-ref32 {e:.3e}")
        assert k_err < 0.1
        assert e < TOL
    rb.free()
    ob.free()


@pytest.mark.parametrize("band0", [False, True])
def test_multiband_otf_matches_fp64(fsim_lib, band0, tmp_path):
    """A measured three-band OTF (per-band radial tables, as an OTF file
    holds them) through the fused path, with each band's own OTF (the
    benchmark's order) or band 0's for every component (the live runner),
    against the FP64 re-run with the same tables; and the OTF read back from
    its XML form reconstructs bitwise identically."""
    from fsim import synthetic
    from fsim.conf import Conf
    from fsim.sim_algorithm import OtfProvider, Reconstructor
    n = 128
    raw, sp = synthetic.sim_stack(n=n, nphase=5, seed=12, nbeads=75, nfil=10, harmonic2=0.3, k_frac=0.45)
    ests = [OtfProvider.fromEstimate(1.4, 525.0, a) for a in (1.0, 0.7, 0.45)]
    vals = np.stack([e.band_values()[0] for e in ests])
    otf = OtfProvider.fromData(1.4, 525.0, ests[0].cpm, vals)
    otf.setAttenuation(0.995, 1.2)
    otf.switchAttenuation(True)
    param = synthetic.make_param(sp, n, nphase=5, nband=3)
    param.otf(otf)
    got = Reconstructor(param, otf_band0=band0).reconstruct_host(raw)
    p = ref.recon_dict(sp, n, nband=3)
    p["px"] = [param.dir(d).pX for d in range(3)]
    p["py"] = [param.dir(d).pY for d in range(3)]
    p["otf_band0"] = band0
    want = fp64.recon_fast(p, raw, vals.astype(np.complex128))
    e = fp64.rel_l2(got, want)
    other = fp64.recon_fast(dict(p, otf_band0=not band0), raw, vals.astype(np.complex128))
    print(f"band0={band0}: rel L2 vs FP64 {e:.2e}; the other OTF choice differs by {fp64.rel_l2(other, want):.2e}")
    assert e < TOL and fp64.rel_l2(other, want) > 1e-3
    # the file keeps the band data exactly; "cycles" is a decimal written with
    # six significant digits, as OtfProvider.saveConfig writes it
    c = Conf("fairsim")
    otf.saveConfig(c.r().mk("ch"))
    c.saveFile(str(tmp_path / "otf.xml"))
    loaded = OtfProvider.loadFromConfig(Conf.loadFile(str(tmp_path / "otf.xml")).r().cd("ch"))
    assert loaded.cpm == float("%8.5e" % otf.cpm)
    same = OtfProvider.fromData(1.4, 525.0, loaded.cpm, vals)
    same.setAttenuation(0.995, 1.2)
    same.switchAttenuation(True)
    param.otf(loaded)
    a = Reconstructor(param, otf_band0=band0).reconstruct_host(raw)
    param.otf(same)
    assert np.array_equal(a, Reconstructor(param, otf_band0=band0).reconstruct_host(raw))
    assert fp64.rel_l2(a, got) < 1e-3


def test_complex_otf_takes_all_components(fsim_lib):
    """A measured three-band OTF with non-zero imaginary parts (as an OTF
    file may hold them, OtfProvider.java:592-628): conj(OTF) is then not
    symmetric, so the channel stores every component (no Hermitian pairing,
    full-width result) even when the pairing is asked for, and the image
    matches the FP64 re-run with the same complex tables."""
    import ctypes as C
    from fsim import synthetic
    from fsim._lib import call
    from fsim.sim_algorithm import OtfProvider, Reconstructor
    n = 128
    raw, sp = synthetic.sim_stack(n=n, nphase=5, seed=14, nbeads=75, nfil=10, harmonic2=0.3, k_frac=0.45)
    ests = [OtfProvider.fromEstimate(1.4, 525.0, a) for a in (1.0, 0.7, 0.45)]
    ns = ests[0].band_values().shape[1]
    r = np.arange(ns) / ns
    # a radial phase per band (a defocus-like aberration), |OTF| unchanged
    phase = np.stack([0.6 * r ** 2, -0.4 * r, 0.9 * r ** 3])
    vals = (np.stack([e.band_values()[0] for e in ests]) * np.exp(1j * phase)).astype(np.complex64)
    assert np.abs(vals.imag).max() > 0.05
    otf = OtfProvider.fromData(1.4, 525.0, ests[0].cpm, vals)
    otf.setAttenuation(0.995, 1.2)
    otf.switchAttenuation(True)
    param = synthetic.make_param(sp, n, nphase=5, nband=3)
    param.otf(otf)
    paired = Reconstructor(param, mode=0)
    full = Reconstructor(param, mode=1)
    cls = [C.c_int() for _ in range(5)]
    call("fsim_channel_class", paired.handle, *[C.byref(c) for c in cls])
    n_, D, P, Cs, half = [c.value for c in cls]
    assert (n_, D, P, Cs, half) == (n, 3, 5, 5, 0)
    got = paired.reconstruct_host(raw)
    assert np.array_equal(got, full.reconstruct_host(raw))
    p = ref.recon_dict(sp, n, nband=3)
    p["px"] = [param.dir(d).pX for d in range(3)]
    p["py"] = [param.dir(d).pY for d in range(3)]
    want = fp64.recon_fast(p, raw, vals.astype(np.complex128))
    real_only = fp64.recon_fast(p, raw, vals.real.astype(np.complex128))
    e = fp64.rel_l2(got, want)
    print(f"complex OTF: rel L2 vs FP64 {e:.2e}; the real part alone differs by {fp64.rel_l2(real_only, want):.2e}")
    assert e < TOL and fp64.rel_l2(real_only, want) > 1e-3


def test_channel_create_frees_on_allocation_failure(fsim_lib):
    """Every device allocation of the channel set-up and of the work buffers
    is made to fail in turn (fsim_debug_fail_alloc): each failure raises
    MemoryError (the reference's OutOfMemoryError) and leaves the library's
    device bytes where they were."""
    from fsim import synthetic
    from fsim._lib import lib
    from fsim.linalg import DeviceBuffer
    from fsim.sim_algorithm import Reconstructor
    L = lib()
    n = 128
    raw, sp = synthetic.sim_stack(n=n, seed=15, nbeads=60, nfil=8)
    param = synthetic.make_param(sp, n)
    ok = Reconstructor(param)          # OTF tables and twiddles exist from here on
    want = ok.reconstruct_host(raw)
    del ok
    base = L.fsim_device_bytes()
    try:
        for k in range(9):             # the set-up's nine device buffers
            L.fsim_debug_fail_alloc(k)
            with pytest.raises(MemoryError):
                Reconstructor(param)
            assert L.fsim_device_bytes() == base, f"allocation {k}"
        L.fsim_debug_fail_alloc(-1)
        rec = Reconstructor(param)
        held = L.fsim_device_bytes()
        assert held > base
        rb = DeviceBuffer(raw.nbytes)
        ob = DeviceBuffer(rec.out_bytes())
        rb.upload(raw)
        held = L.fsim_device_bytes()
        for k in range(2):             # the work buffers A, B
            L.fsim_debug_fail_alloc(k)
            with pytest.raises(MemoryError):
                rec.reconstruct(rb.value, ob.value, 1)
            assert L.fsim_device_bytes() == held
        L.fsim_debug_fail_alloc(-1)
        assert np.array_equal(rec.reconstruct_host(raw), want)
        rb.free()
        ob.free()
        del rec
        assert L.fsim_device_bytes() == base
    finally:
        L.fsim_debug_fail_alloc(-1)


def _hip():
    import ctypes as C
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            h = C.CDLL(name)
            break
        except OSError:
            h = None
    assert h is not None, "HIP runtime library not found"
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    return h


def test_streams_have_own_workspaces(fsim_lib):
    """Two streams reconstructing on one channel at the same time (the
    reference's concurrent ReconstructionThreads) each use their own work
    buffers: every image is bitwise the one of a call alone."""
    import ctypes as C
    from fsim import synthetic
    from fsim.linalg import DeviceBuffer
    from fsim.sim_algorithm import Reconstructor
    n = 256
    raws = []
    for s in range(4):
        raw, sp = synthetic.sim_stack(n=n, seed=110 + s, nbeads=150, nfil=20)
        raws.append(raw)
    rec = Reconstructor(synthetic.make_param(sp, n))
    want = [rec.reconstruct_host(r) for r in raws]
    hip = _hip()
    streams = [C.c_void_p(), C.c_void_p()]
    for st in streams:
        assert hip.hipStreamCreate(C.byref(st)) == 0
    try:
        ins = [DeviceBuffer(2 * raws[0].nbytes) for _ in range(2)]
        outs = [DeviceBuffer(2 * rec.out_bytes()) for _ in range(2)]
        ins[0].upload(np.stack(raws[:2]))
        ins[1].upload(np.stack(raws[2:]))
        for _ in range(3):
            for i in range(2):
                rec.reconstruct(ins[i].value, outs[i].value, 2, stream=streams[i].value)
        for st in streams:
            assert hip.hipStreamSynchronize(st) == 0
        got = [outs[i].download((2, 2 * n, 2 * n), np.float32) for i in range(2)]
        for i in range(2):
            for j in range(2):
                assert np.array_equal(got[i][j], want[2 * i + j]), f"stream {i} stack {j}"
    finally:
        for st in streams:
            hip.hipStreamDestroy(st)
