"""Pins of the oracle (CPU only).

The reference ships no golden vectors (its tests are print-only mains), so
each oracle function is pinned here by what mathematics fixes: closed forms,
invariants, textbook special cases, brute force on tiny inputs, and the
agreement of two independent restatements (FP32 Java-order C and FP64
NumPy).  Each check is chosen so that a dropped term, a wrong sign or index
or a transposed operand fails it.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import fp64, ref

L = ref.lib()


def c32(z):
    return ref.cplx_f32(z)


def test_fsin_fcos_error_bound():
    """MTool.fsin/fcos claim |err| <= 5.4e-9 (linalg/MTool.java:33-48)."""
    xs = np.linspace(-40, 40, 200001)
    es = max(abs(L.ref_fsin(x) - math.sin(x)) for x in xs[::7])
    ec = max(abs(L.ref_fcos(x) - math.cos(x)) for x in xs[::7])
    assert es < 6e-9 and ec < 6e-9
    # quadrant logic: exact symmetry points
    assert abs(L.ref_fsin(math.pi / 2) - 1) < 6e-9
    assert abs(L.ref_fsin(-math.pi / 2) + 1) < 6e-9


@pytest.mark.parametrize("w,h", [(64, 64), (128, 32), (16, 256)])
def test_fft_closed_forms(w, h):
    a = np.zeros((h, w, 2), np.float32)
    a[0, 0, 0] = 1
    L.ref_fft2d(ref.ptr(a), w, h, 0)
    assert np.allclose(a[..., 0], 1) and np.allclose(a[..., 1], 0)        # delta -> const
    L.ref_fft2d(ref.ptr(a), w, h, 1)
    assert abs(a[0, 0, 0] - 1) < 1e-6 and np.abs(a).sum() - 1 < 1e-4     # and back
    y, x = np.mgrid[0:h, 0:w]
    kx, ky = 3, 5
    c = c32(np.cos(2 * np.pi * (kx * x / w + ky * y / h)))
    L.ref_fft2d(ref.ptr(c), w, h, 0)
    mag = np.hypot(c[..., 0], c[..., 1])
    big = sorted(zip(*np.nonzero(mag > 1)))
    assert big == sorted([(ky, kx), (h - ky, w - kx)])
    assert abs(mag[ky, kx] - w * h / 2) < 1e-3 * w * h
    # forward sign: e^{+2 pi i k x} lands on +k only
    e = c32(np.exp(2j * np.pi * (kx * x / w + ky * y / h)))
    L.ref_fft2d(ref.ptr(e), w, h, 0)
    mag = np.hypot(e[..., 0], e[..., 1])
    assert mag[ky, kx] > 0.99 * w * h and mag[h - ky, w - kx] < 1e-3 * w * h


def test_fft_parseval_and_numpy():
    rng = np.random.default_rng(0)
    z = rng.standard_normal((128, 64)) + 1j * rng.standard_normal((128, 64))
    a = c32(z)
    L.ref_fft2d(ref.ptr(a), 64, 128, 0)
    f = ref.to_complex(a)
    assert fp64.rel_l2(f, np.fft.fft2(z)) < 2e-6
    assert abs(np.sum(np.abs(f) ** 2) / (64 * 128) - np.sum(np.abs(z) ** 2)) < 1e-5 * np.sum(np.abs(z) ** 2)


def test_fourier_shift_integer_is_roll():
    n = 64
    rng = np.random.default_rng(1)
    spec = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    a = c32(np.fft.ifft2(spec))
    L.ref_fourier_shift(ref.ptr(a), n, 5.0, -3.0)
    got = np.fft.fft2(ref.to_complex(a))
    assert fp64.rel_l2(got, np.roll(np.roll(spec, 5, axis=1), -3, axis=0)) < 1e-6


def test_fourier_shift_fractional_matches_exact():
    n = 32
    rng = np.random.default_rng(2)
    z = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    a = c32(z)
    L.ref_fourier_shift(ref.ptr(a), n, 0.37, 0.81)
    assert fp64.rel_l2(ref.to_complex(a), fp64.fourier_shift(z, 0.37, 0.81)) < 1e-6


@pytest.mark.parametrize("wi,xo,yo", [(16, 0, 0), (16, 5, -3), (32, -11, 40), (8, -16, 15)])
def test_paste_freq_is_zero_padding(wi, xo, yo):
    """pasteFreq(0,0) is the zero-padded spectrum: its 2n IFFT sampled at even
    pixels equals the n IFFT / 4; an offset multiplies by a linear ramp."""
    wo = 2 * wi
    rng = np.random.default_rng(wi)
    F = rng.standard_normal((wi, wi)) + 1j * rng.standard_normal((wi, wi))
    a = c32(F)
    out0 = np.zeros((wo, wo, 2), np.float32)
    L.ref_paste_freq(ref.ptr(out0), wo, wo, ref.ptr(a), wi, wi, 0, 0)
    P0 = ref.to_complex(out0)
    assert np.count_nonzero(np.abs(P0)) == wi * wi
    s = np.fft.ifft2(P0)
    assert fp64.rel_l2(4 * s[::2, ::2], np.fft.ifft2(ref.to_complex(a))) < 1e-6
    out = np.zeros((wo, wo, 2), np.float32)
    L.ref_paste_freq(ref.ptr(out), wo, wo, ref.ptr(a), wi, wi, xo, yo)
    y, x = np.mgrid[0:wo, 0:wo]
    ramp = np.exp(2j * np.pi * (xo * x + yo * y) / wo)
    assert fp64.rel_l2(np.fft.ifft2(ref.to_complex(out)), s * ramp) < 1e-6
    assert np.array_equal(ref.to_complex(out), fp64.paste_freq(ref.to_complex(a), wo, wo, xo, yo))


def _sep(phases, B, fac):
    P = len(phases)
    Cc = 2 * B - 1
    out = np.zeros(2 * Cc * P)
    ph = np.asarray(phases, dtype=np.float64)
    fa = None if fac is None else ref.ptr(np.asarray(fac, dtype=np.float64))
    assert L.ref_separation_matrix(ref.ptr(ph), P, B, fa, ref.ptr(out)) == 0
    o = out.reshape(Cc, P, 2)
    return o[..., 0] + 1j * o[..., 1]


@pytest.mark.parametrize("P,B,off", [(3, 2, 0.0), (3, 2, 1.1), (5, 3, 0.4), (5, 2, -0.7), (7, 3, 2.0)])
def test_separation_matrix_inverse(P, B, off):
    phases = [2 * np.pi * p / P + off for p in range(P)]
    fac = [1.0] + [0.7] * (B - 1)
    S = _sep(phases, B, fac)
    # the forward mixing matrix written from the band model (SIM image
    # formation): image_p = f0*b0 + sum_j f_j/2 (e^{+i j phi_p} b_{2j-1} + e^{-i j phi_p} b_{2j})
    Cc = 2 * B - 1
    M = np.zeros((P, Cc), complex)
    for p in range(P):
        M[p, 0] = 1.0
        for j in range(1, B):
            M[p, 2 * j - 1] = 0.35 * np.exp(1j * j * phases[p])
            M[p, 2 * j] = 0.35 * np.exp(-1j * j * phases[p])
    assert np.allclose(S @ M, np.eye(Cc), atol=1e-12)
    assert np.allclose(S, np.linalg.pinv(M), atol=1e-12)
    assert np.allclose(S, fp64.separation_matrix(phases, B, fac), atol=1e-12)


def test_separate_bands_known_answer():
    rng = np.random.default_rng(4)
    n, P, B = 16, 5, 3
    Cc = 2 * B - 1
    phases = [2 * np.pi * p / P + 0.2 for p in range(P)]
    bands = [rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) for _ in range(Cc)]
    ims = []
    for p in range(P):
        im = bands[0].copy()
        for j in range(1, B):
            im = im + 0.5 * np.exp(1j * j * phases[p]) * bands[2 * j - 1] \
                    + 0.5 * np.exp(-1j * j * phases[p]) * bands[2 * j]
        ims.append(c32(im))
    S = _sep(phases, B, None).astype(np.complex64)
    Sf = c32(S)
    outs = [np.zeros((n, n, 2), np.float32) for _ in range(Cc)]
    ip = (C.c_void_p * P)(*[x.ctypes.data for x in ims])
    op = (C.c_void_p * Cc)(*[x.ctypes.data for x in outs])
    L.ref_separate_bands(ip, op, n * n, ref.ptr(Sf), P, Cc)
    for c in range(Cc):
        assert fp64.rel_l2(ref.to_complex(outs[c]), bands[c]) < 1e-5


def test_ideal_otf_is_pupil_autocorrelation():
    """valIdealOTF(d) = overlap area of two unit disks at distance 2d / pi."""
    g = 2000
    y, x = (np.mgrid[0:g, 0:g] - g / 2 + 0.5) / (g / 5.0)   # extent 5: no wrap-around
    disk = (x * x + y * y) <= 1
    for d in (0.1, 0.3, 0.5, 0.8):
        sh = int(round(2 * d * g / 5.0))
        ov = np.logical_and(disk, np.roll(disk, sh, axis=1)).sum() / disk.sum()
        assert abs(L.ref_val_ideal_otf(d) - ov) < 5e-3
    assert L.ref_val_ideal_otf(0) == 1.0 and L.ref_val_ideal_otf(1) == 0.0
    assert L.ref_val_ideal_otf(1.2) == 0.0 and L.ref_val_ideal_otf(-0.1) == 0.0


def test_otf_lookup_quirk_and_cutoffs():
    vals, cut, cpm = ref.otf_estimate(1.4, 525.0, 1.0)
    assert abs(cut - 1000 / (525 / 1.4 / 2)) < 1e-12 and abs(cpm - cut / 512) < 1e-15
    v = vals[:, 0]
    n = 64
    vec_cpm = 1 / (n * 0.08)
    out = np.zeros((n, n, 2), np.float32)
    L.ref_otf_to_vector(ref.ptr(out), n, n, ref.ptr(vals), 512, cpm, cut, vec_cpm, 0.0, 0.0, 1)
    # the lookup uses the lower sample for both weights: v[l](1-f) + v[l] f
    # equals v[l] within one float rounding, never the linear interpolation
    xh = np.where(np.arange(n) < n // 2, np.arange(n), np.arange(n) - n)
    yh = np.where(np.arange(n) < n // 2, -np.arange(n), n - np.arange(n))
    cyc = np.sqrt(xh[None, :] ** 2 + yh[:, None] ** 2) * vec_cpm
    pos = cyc / cpm
    inside = (cyc < cut) & (np.ceil(pos) < 512)
    low = v[np.floor(np.where(inside, pos, 0)).astype(int)]
    got = out[..., 0]
    assert np.all(np.abs(got[inside] - low[inside]) <= 2 * np.spacing(np.abs(low[inside]) + 1e-30))
    assert np.all(got[~inside] == 0)
    assert np.all(out[..., 1] == 0)
    # the double restatement agrees to float rounding
    d64 = fp64.otf_to_vector(n, n, v.astype(np.float64), cpm, cut, vec_cpm)
    assert np.max(np.abs(got - d64)) < 1e-6


def test_fade_border_closed_form():
    n, px = 48, 10
    f = np.ones((n, n), np.float32)
    L.ref_fade_border(ref.ptr(f), n, n, px)
    assert np.all(f[px:n - px, px:n - px] == 1)
    assert np.all(f[0, :] == 0) and np.all(f[:, -1] == 0)
    assert f[5, 20] == np.float32(0.5)                          # sin^2(pi/4)
    assert abs(f[5, 5] - 0.25) < 1e-7                           # corner: product of both factors
    assert abs(f[n - 1 - 5, 20] - 0.5) < 1e-7                   # mirrored bottom
    assert fp64.rel_l2(f, fp64.fade_border(n, n, px)) < 1e-7


def test_wiener_matches_fp64_and_double_counts_band0():
    n = 32
    vals, cut, cpm = ref.otf_estimate(1.4, 525.0, 1.0)
    vatt = np.zeros_like(vals)
    oatt = np.zeros(512, np.float32)
    L.ref_otf_set_attenuation(ref.ptr(vals), 512, cpm, 0.995, 1.2, ref.ptr(vatt), ref.ptr(oatt))
    vec_cpm = 1 / (n * 0.08)
    m = 2 * n
    px = np.array([10.3, -4.2, -6.1])
    py = np.array([-2.2, 9.7, -7.5])
    for use_att in (0, 1):
        out = np.zeros((m, m), np.float32)
        L.ref_wiener_denominator(ref.ptr(out), m, m, ref.ptr(vals), ref.ptr(oatt), 512, cpm, cut, vec_cpm,
                                 3, 2, ref.ptr(px), ref.ptr(py), use_att, 0.05)
        exp = fp64.wiener_denominator(m, m, vals[:, 0].astype(float), oatt.astype(float), cpm, cut, vec_cpm,
                                      px, py, 2, use_att, 0.05)
        assert np.max(np.abs(out - exp) / exp) < 1e-5
    # at DC without attenuation: 3 directions x (band 0 twice: 2*|OTF(0)|^2
    # + band 1 at +-k) -> closed form from the table values
    out = np.zeros((m, m), np.float32)
    L.ref_wiener_denominator(ref.ptr(out), m, m, ref.ptr(vals), ref.ptr(oatt), 512, cpm, cut, vec_cpm,
                             3, 2, ref.ptr(px), ref.ptr(py), 0, 0.05)
    otf_k = [float(fp64.get_otf_val(vals[:, 0].astype(float), cpm, cut,
                                    math.hypot(px[d], py[d]) * vec_cpm)) for d in range(3)]
    den = sum(2 * 1.0 + 2 * otf_k[d] ** 2 for d in range(3)) + 0.05 ** 2
    assert abs(1 / out[0, 0] - den) < 1e-5 * den


def test_locate_peak_known_answers():
    n = 64
    z = np.zeros((n, n), np.complex64)
    z[1, 1] = 100                 # excluded by kmin
    z[10, 50] = 7 + 7j            # (x=50, y=10) -> (-14, -10)
    res = np.zeros(4)
    L.ref_locate_peak(ref.ptr(c32(z)), n, n, 5.0, ref.ptr(res))
    assert list(res[:2]) == [-14.0, -10.0]
    assert abs(res[2] - abs(7 + 7j)) < 1e-5 and abs(res[3] - math.pi / 4) < 1e-6
    z[40, 3] = 20                 # (x=3, y=40) -> (3, 24)
    L.ref_locate_peak(ref.ptr(c32(z)), n, n, 5.0, ref.ptr(res))
    assert list(res[:2]) == [3.0, 24.0]
    z[40, 7] = 20                 # tie later in x: first maximum kept
    z[45, 2] = 20                 # tie in a later row
    L.ref_locate_peak(ref.ptr(c32(z)), n, n, 5.0, ref.ptr(res))
    assert list(res[:2]) == [3.0, 24.0]


def test_norm2_and_sum_double_accumulation():
    rng = np.random.default_rng(9)
    a = c32(rng.standard_normal((64, 64)) + 1j * rng.standard_normal((64, 64)))
    f = a.ravel()
    assert L.ref_norm2(ref.ptr(a), 64 * 64) == pytest.approx(float(np.sum((f * f).astype(np.float64))), rel=1e-13)
    r = np.zeros(2)
    L.ref_sum_elements(ref.ptr(a), 64 * 64, ref.ptr(r))
    assert r[0] == pytest.approx(float(a[..., 0].astype(np.float64).sum()), rel=1e-12)


@pytest.mark.parametrize("n,nphase,nband", [(64, 3, 2), (128, 3, 2), (64, 5, 2)])
def test_recon_fp32_vs_fp64(n, nphase, nband):
    """The two independent restatements of the full dataflow agree."""
    from fsim import synthetic
    raw, sp = synthetic.sim_stack(n=n, nphase=nphase, seed=n + nphase, nbeads=40, nfil=6)
    p = ref.recon_dict(sp, n, nband=nband)
    a = ref.recon32(p, raw, threads=2)
    b = ref.recon64(p, raw)
    assert fp64.rel_l2(a, b) < 5e-6


def test_recon_slow_and_fast_shift_agree_when_shift_is_integer():
    """Slow shift (2N FFTs) and fast shift (N FFTs + paste) coincide for an
    integer shift vector, where the sub-pixel ramp is exactly one."""
    from fsim import synthetic
    n = 64
    raw, sp = synthetic.sim_stack(n=n, seed=5, nbeads=40, nfil=6)
    sp = dict(sp)
    sp["px"] = [round(v) for v in sp["px"]]
    sp["py"] = [round(v) for v in sp["py"]]
    pf = ref.recon_dict(sp, n, fast_shift=1)
    ps = ref.recon_dict(sp, n, fast_shift=0)
    assert fp64.rel_l2(ref.recon32(pf, raw), ref.recon32(ps, raw)) < 1e-5


def test_recon_parameters_known_answer():
    """Noise-free data: the reconstruction with the true shift correlates with
    the object far better than one with the y-sign flipped or the phases
    rotated, which pins the y-flip, shift signs and component pairing."""
    from fsim import synthetic
    n = 128
    raw, sp = synthetic.sim_stack(n=n, seed=3, nbeads=60, nfil=8, peak=1e5, bg=0, offset=0)
    obj = synthetic.make_object(2 * n, np.random.default_rng(3), 60, 8)

    def score(s, dphi=0.0):
        p = ref.recon_dict(s, n, fade_px=0)
        p["phases"] = [[x + dphi for x in ph] for ph in p["phases"]]
        r = ref.recon32(p, raw, threads=2)
        a = (r - r.mean()).ravel()
        b = (obj - obj.mean()).ravel()
        return float(a @ b / np.linalg.norm(a) / np.linalg.norm(b))

    good = score(sp)
    flip = dict(sp)
    flip["py"] = [-v for v in sp["py"]]
    assert good > 0.6
    assert score(flip) < good - 0.2
    assert score(sp, np.pi / 2) < good - 0.2


# --- parameter estimation (Correlation.java, SimAlgorithm.estimateParameters) ---

def _shifted_pair(n, kx, ky, cut_px, seed):
    """band0: a random spectrum inside the OTF support; band1: the same image
    multiplied in real space by exp(-2 pi i (kx x - ky y) / n), i.e. what
    fitPeak's Fourier shift by (kx, -ky) moves back onto band0."""
    rng = np.random.default_rng(seed)
    f = np.fft.fft2(rng.standard_normal((n, n)))
    fr = np.fft.fftfreq(n) * n
    fx, fy = np.meshgrid(fr, fr)
    f[np.hypot(fx, fy) > 0.9 * cut_px] = 0
    s0 = np.fft.ifft2(f)
    y, x = np.mgrid[0:n, 0:n]
    s1 = s0 * np.exp(-2j * np.pi * (kx * x - ky * y) / n)
    return ref.cplx_f32(np.fft.fft2(s0)), ref.cplx_f32(np.fft.fft2(s1))


@pytest.mark.parametrize("kx,ky", [(20.37, -11.62), (-25.8, 9.13)])
def test_fit_peak_recovers_exact_shift(kx, ky):
    """A band and its exactly shifted copy: fitPeak, started within a pixel,
    returns the shift to the last round's grid step (2 * 0.1 / 9 px), and
    getPeak there has phase ~0."""
    n, vec_cpm = 128, 1 / (128 * 0.08)
    otf = ref.Otf32(1.4, 525.0, 1.0)
    b0, b1 = _shifted_pair(n, kx, ky, otf.cutoff / vec_cpm, 5)
    r = ref.fit_peak32(b0, b1, otf, vec_cpm, round(kx) + 0.8, round(ky) - 0.6, 0.05, 2.5)
    assert abs(r[0] - kx) < 0.03 and abs(r[1] - ky) < 0.03
    g = ref.get_peak32(b0, b1, otf, vec_cpm, r[0], r[1], 0.05)
    assert abs(np.angle(g)) < 0.1


def test_common_region_zero_sets():
    """Without weights (limit 0, no division) commonRegion only cuts the DC
    ring 0.15 < |r| / |k| < 0.85 from band0 and the same ring moved by
    (-(int)kx, +(int)ky) (truncated toward zero) from band1."""
    n, vec_cpm = 64, 1 / (64 * 0.08)
    otf = ref.Otf32(1.4, 525.0, 1.0)
    kx, ky = -20.7, 9.4
    b0 = np.ones((n, n, 2), np.float32)
    b1 = np.ones((n, n, 2), np.float32)
    ref.common_region32(b0, b1, otf, vec_cpm, kx, ky, 0.15, 0.0, divide=False)
    fr = np.fft.fftfreq(n) * n
    fx, fy = np.meshgrid(fr, fr)
    ratio = np.hypot(fx, fy) / math.hypot(kx, ky)
    ring = (ratio < 0.15) | (ratio > 0.85)
    assert np.array_equal(b0[..., 0] == 0, ring)
    ikx, iky = int(kx), int(ky)
    moved = np.zeros_like(ring)
    yy, xx = np.nonzero(ring)
    moved[(yy + iky) % n, (xx - ikx) % n] = True
    assert np.array_equal(b1[..., 0] == 0, moved)


def test_estimate_known_answer_two_band():
    """Seeded 3-phase data (|k0| = 0.85 cutoff, phases 0, 2pi/3, 4pi/3,
    modulation 0.7): the estimation returns the true shift vectors, phase
    offset ~0 and a modulation estimate of the right size."""
    from fsim import synthetic
    n = 256
    raw, sp = synthetic.sim_stack(n=n, seed=3, nbeads=150, nfil=20)
    otf = ref.Otf32(1.4, 525.0, 1.0)
    est = ref.estimate32(ref.input_ffts32(raw.reshape(3, 3, n, n)), 2, 1, 0.6, otf, 1 / (n * 0.08))
    for d, r in enumerate(est):
        assert abs(r["px"] - sp["px"][d]) < 0.1 and abs(r["py"] - sp["py"][d]) < 0.1
        assert abs(r["pha_off"]) < 0.2
        assert 0.4 < r["mod1"] < 0.9


@pytest.mark.parametrize("fit_band,exclude", [(1, 0.3), (2, 0.6)])
def test_estimate_known_answer_three_band(fit_band, exclude):
    """5-phase data with a 2 k0 harmonic (|k0| = 0.45 cutoff): both fit bands
    return the band-2 shift 2 k0 (setPxPy argument)."""
    from fsim import synthetic
    n = 256
    raw, sp = synthetic.sim_stack(n=n, nphase=5, seed=4, nbeads=150, nfil=20, harmonic2=0.3, k_frac=0.45)
    otf = ref.Otf32(1.4, 525.0, 1.0)
    est = ref.estimate32(ref.input_ffts32(raw.reshape(3, 5, n, n)), 3, fit_band, exclude, otf, 1 / (n * 0.08))
    for d, r in enumerate(est):
        assert abs(r["px"] - 2 * sp["px"][d]) < 0.1 and abs(r["py"] - 2 * sp["py"][d]) < 0.1
        assert abs(r["pha_off"]) < 0.2
