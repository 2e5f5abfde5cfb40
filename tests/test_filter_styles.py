"""Filter styles beyond Wiener: Richardson-Lucy deconvolution
(RLDeconvolution.java) and SimAlgorithm.runReconstruction's RL on input /
output / both and otfBeforeShift branches.

CPU: the FP64 oracle (oracle/fp64.py: rl_deconvolve64,
run_reconstruction64) is pinned -- the Wiener branch equals the slow-shift
reconstruction already pinned in test_oracle; RL leaves data unchanged for a
unit OTF, conserves the total for a normalised symmetric OTF, and lowers the
data residual on a blurred point pattern; the support mask is the cutoff
disc.  GPU: fsim_rl_deconvolve and the device flow (fsim.sim_algorithm.
run_reconstruction) against the oracle, to the tolerances stated per test.
"""
import numpy as np
import pytest

from oracle import fp64, ref

NA, LAM, PX = 1.4, 525.0, 0.08


def _case(n, seed, nphase=3, nband=2, harmonic2=0.0):
    from fsim import synthetic
    raw, sp = synthetic.sim_stack(n=n, nphase=nphase, seed=seed, nbeads=max(30, 300 * n // 512),
                                  nfil=max(5, 40 * n // 512), harmonic2=harmonic2)
    p = ref.recon_dict(sp, n, nband=nband)   # band-1 shift, also with 3 bands
    vals = fp64.otf_from_estimate(NA, LAM, 1.0)
    if isinstance(vals, tuple):
        vals = vals[0]
    return raw, sp, p, vals


def _in_fft64(raw, p):
    n, D, P = p["n"], p["ndir"], p["nphase"]
    fade = fp64.fade_border(n, n, p["fade_px"])
    r = raw.astype(np.float64).reshape(D, P, n, n)
    return [[fp64.fft2(r[d, q] * fade) for q in range(P)] for d in range(D)]


def test_wiener_branch_is_the_slow_shift_reconstruction():
    """At integer shifts the OTF applied after the shift (runReconstruction)
    equals the OTF applied before it (the live loop's order, recon_fast with
    the slow shift): the shift is then an exact circular one."""
    raw, sp, p, vals = _case(64, 3)
    p["px"] = [float(round(v)) for v in p["px"]]
    p["py"] = [float(round(v)) for v in p["py"]]
    want = fp64.recon_fast(dict(p, fast_shift=0), raw, vals)
    got = fp64.run_reconstruction64(p, _in_fft64(raw, p), vals, "Wiener")
    assert fp64.rel_l2(got, want) < 1e-12
    # a wrong sign of the shifted OTF position breaks it
    p2 = dict(p, px=[-v for v in p["px"]])
    bad = fp64.run_reconstruction64(p2, _in_fft64(raw, p), vals, "Wiener")
    assert fp64.rel_l2(bad, fp64.recon_fast(dict(p2, fast_shift=0), raw, vals)) < 1e-12
    assert fp64.rel_l2(bad, want) > 1e-2


def test_rl_oracle_properties():
    rng = np.random.default_rng(0)
    n = 64
    d = rng.uniform(1, 10, (n, n))
    # unit OTF: d / (u * 1) = 1, so every iteration returns d
    assert np.allclose(fp64.rl_deconvolve64(d, np.ones((n, n)), 5), d, rtol=1e-12)
    # a normalised, real, symmetric OTF conserves the total
    xh, yh = fp64.coords(n, n)
    otf = np.exp(-(xh ** 2 + yh ** 2) / (2 * 6.0 ** 2))
    u = fp64.rl_deconvolve64(d, otf, 7)
    assert abs(u.real.sum() - d.sum()) < 1e-9 * d.sum()
    # blurred points: the data residual drops with the iterations
    truth = np.zeros((n, n))
    truth[rng.integers(0, n, 12), rng.integers(0, n, 12)] = 100
    truth += 1
    blur = fp64.ifft2(fp64.fft2(truth) * otf).real
    res = [np.linalg.norm(fp64.ifft2(fp64.fft2(fp64.rl_deconvolve64(blur, otf, k)) * otf).real - blur)
           for k in (0, 2, 8)]
    assert res[0] > res[1] > res[2]
    # and in frequency space (inputIsInFreqSpace) it is the same operator
    f = fp64.rl_deconvolve64(fp64.fft2(blur), otf, 3, True)
    assert np.allclose(fp64.ifft2(f), fp64.rl_deconvolve64(blur, otf, 3), atol=1e-9)


def test_mask_is_the_cutoff_disc():
    n, vcpm, cut = 64, 1 / (64 * PX), 1000 / (LAM / NA / 2)
    m = fp64.mask_otf(n, n, cut, vcpm, 0, 0)
    xh, yh = fp64.coords(n, n)
    r = np.sqrt(xh ** 2 + yh ** 2) * vcpm
    assert np.array_equal(m == 1, r <= cut)
    assert m[0, 0] == 1 and m[n // 2, n // 2] == 0


@pytest.mark.parametrize("style", ["RLin", "RLout", "RLboth"])
def test_filter_styles_run_and_differ(style):
    raw, sp, p, vals = _case(64, 5)
    f = _in_fft64(raw, p)
    a = fp64.run_reconstruction64(p, f, vals, style, rl_iterations=3)
    b = fp64.run_reconstruction64(p, f, vals, "Wiener")
    assert np.isfinite(a).all() and fp64.rel_l2(a, b) > 1e-3


@pytest.mark.gpu
def test_rl_deconvolve_matches_oracle(fsim_lib):
    from fsim.linalg import Vec2dCplx
    from fsim.sim_algorithm import RLDeconvolution
    rng = np.random.default_rng(1)
    n = 128
    xh, yh = fp64.coords(n, n)
    otf = np.exp(-(xh ** 2 + yh ** 2) / (2 * 9.0 ** 2))
    d = fp64.ifft2(fp64.fft2(rng.uniform(1, 50, (n, n))) * otf).real + 5
    for in_freq in (False, True):
        x = fp64.fft2(d) if in_freq else d.astype(np.complex128)
        img, o = Vec2dCplx(n, n), Vec2dCplx(n, n)
        img.setData(np.stack([x.real, x.imag], -1).astype(np.float32))
        o.setData(np.stack([otf, np.zeros_like(otf)], -1).astype(np.float32))
        RLDeconvolution.deconvolve(img, o, 6, in_freq)
        g = img.vectorData()
        got = g[..., 0] + 1j * g[..., 1]
        want = fp64.rl_deconvolve64(x, otf, 6, in_freq)
        e = fp64.rel_l2(got, want)
        print(f"in_freq={in_freq}: rel L2 {e:.2e}")
        assert e < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("w,h,steps", [(256, 64, 3), (64, 512, 1), (1024, 1024, 2), (4096, 16, 2), (8, 4, 2),
                                         (32, 32, 0)])
def test_rl_deconvolve_shapes(fsim_lib, w, h, steps):
    """The fused iteration (column FFT * otf IFFT, row IFFT with the ratio or
    the update and the next row FFT) on non-square images, the largest row
    length, a few pixels (one thread per column transform), one iteration (no
    next-row transform after it) and none (the input back unchanged)."""
    from fsim.linalg import Vec2dCplx
    from fsim.sim_algorithm import RLDeconvolution
    rng = np.random.default_rng(w * 7 + h + steps)
    xh, yh = fp64.coords(w, h)
    otf = np.exp(-(xh ** 2 + yh ** 2) / (2 * 9.0 ** 2))
    d = fp64.ifft2(fp64.fft2(rng.uniform(1, 50, (h, w))) * otf).real + 5
    x = d.astype(np.complex128)
    img, o = Vec2dCplx(w, h), Vec2dCplx(w, h)
    img.setData(np.stack([x.real, x.imag], -1).astype(np.float32))
    o.setData(np.stack([otf, np.zeros_like(otf)], -1).astype(np.float32))
    RLDeconvolution.deconvolve(img, o, steps, False)
    g = img.vectorData()
    got = g[..., 0] + 1j * g[..., 1]
    if steps == 0:
        assert np.array_equal(got, x.astype(np.complex64))
        return
    want = fp64.rl_deconvolve64(x, otf, steps, False)
    e = fp64.rel_l2(got, want)
    print(f"{w}x{h}, {steps} iterations: rel L2 {e:.2e}")
    assert e < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("style,before,clip", [("Wiener", False, "NONE"), ("Wiener", True, "NONE"),
                                               ("RLin", False, "NONE"), ("RLout", False, "NONE"),
                                               ("RLboth", False, "BOTH"), ("Wiener", False, "CLIP")])
def test_run_reconstruction_matches_oracle(fsim_lib, style, before, clip):
    from fsim import synthetic
    from fsim.sim_algorithm import input_ffts, run_reconstruction
    n = 64
    raw, sp, p, vals = _case(n, 7)
    param = synthetic.make_param(sp, n)
    inFFT = input_ffts(raw.reshape(3, 3, n, n).astype(np.float32))
    got = run_reconstruction(param, inFFT, otfBeforeShift=before, clipScale=clip, filterStyle=style,
                             rlIterations=4, apoCutoff=p["apo_cut"]).vectorData()
    want = fp64.run_reconstruction64(p, _in_fft64(raw, p), vals, style, before, 4, clip)
    e = fp64.rel_l2(got, want)
    print(f"{style} otfBeforeShift={before} {clip}: rel L2 vs FP64 {e:.2e}")
    assert e < 1e-4
