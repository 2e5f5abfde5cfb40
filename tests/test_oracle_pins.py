"""Pins of the oracle's table and filter-style functions (CPU only).

Each check follows from the mathematics or physics of the quantity, not from
the oracle's own formula, and is chosen so that a plausible mistake fails it:

* attenuation (OtfProvider.valAttenuation / setAttenuation / getAttVal /
  writeAttenuationVector, OtfProvider.java:238-287, 503-529): a Gaussian dip
  of depth `str` whose full width at half depth is `fwhm` (the reference
  uses sigma = fwhm / 2.355), sampled with the lower-sample lookup quirk;
* apodization (writeApoVector, OtfProvider.java:469-492): the ideal OTF
  (pinned as a pupil autocorrelation in test_oracle) raised to `bend`,
  stretched to apo_cut x cutoff;
* the estimated OTF's curvature (fromEstimate, OtfProvider.java:96-127):
  vals / ideal = a^v, cutoff = the Abbe limit 2 NA / lambda;
* maskOtf (OtfProvider.java:538-559): the disc of the cutoff radius around
  (kx, ky) in the y-flipped frequency coordinates, by brute force;
* clipAndScale (SimUtils.java:244-264): max -> 255, negatives -> 0, scale
  before clip;
* the RL-on-output composition of runReconstruction (SimAlgorithm.java:
  587-686): the effective SIM OTF at DC (band 0 weighted 1/D per direction,
  every higher band twice: at +k and -k), its symmetry and support, and the
  filter with 0 iterations returning the unfiltered sum;
* the live refit (ReconstructionRunner.doRefit, ReconstructionRunner.java:
  642-746): known-answer recovery of the highest band's shift, identity with
  estimateParameters where the two flows coincide, and OTF band numbers that
  change the fit correlation with a three-band OTF.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import fp64, ref

L = ref.lib()
L.ref_val_attenuation.argtypes = [C.c_double, C.c_double, C.c_double]
L.ref_val_attenuation.restype = C.c_float
L.ref_get_att_val.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
L.ref_get_att_val.restype = C.c_float


def _coords(w, h):
    """Wrapped, y-flipped frequency coordinates (xh, yh) of every pixel."""
    x = np.arange(w)
    y = np.arange(h)
    xh = np.where(x < w // 2, x, x - w)
    yh = np.where(y < h // 2, -y, h - y)
    return np.meshgrid(xh, yh)


@pytest.mark.parametrize("s,fwhm", [(0.995, 1.2), (0.9, 0.8), (0.5, 2.5)])
def test_attenuation_is_gaussian_of_stated_fwhm(s, fwhm):
    assert L.ref_val_attenuation(0.0, s, fwhm) == np.float32(1 - s)
    # half depth at half the full width (2.355 ~ 2 sqrt(2 ln 2) = 2.35482)
    assert abs(L.ref_val_attenuation(fwhm / 2, s, fwhm) - (1 - s / 2)) < 2e-4
    d = np.linspace(0, 3 * fwhm, 301)
    got = np.array([L.ref_val_attenuation(x, s, fwhm) for x in d])
    gauss = 1 - s * np.exp(-4 * math.log(2) * (d / fwhm) ** 2)
    assert np.max(np.abs(got - gauss)) < 1e-3 * s
    assert np.all(np.diff(got) >= 0) and abs(got[-1] - 1) < 1e-6
    assert np.max(np.abs(fp64.val_attenuation(d, s, fwhm) - got)) < 1e-6


def test_attenuation_tables_and_lookup_quirk():
    vals, cut, cpm = ref.otf_estimate(1.4, 525.0, 1.0)
    vatt = np.zeros_like(vals)
    oatt = np.zeros(512, np.float32)
    L.ref_otf_set_attenuation(ref.ptr(vals), 512, cpm, 0.995, 1.2, ref.ptr(vatt), ref.ptr(oatt))
    # sample v of the attenuation table is the Gaussian at v samples
    for v in (0, 10, 37, 200, 511):
        assert oatt[v] == L.ref_val_attenuation(v * cpm, 0.995, 1.2)
    # valsAtt = vals x valsOnlyAtt, the float product
    assert np.array_equal(vatt[:, 0], (vals[:, 0] * oatt).astype(np.float32))
    assert np.all(vatt[:, 1] == 0)
    # getAttVal returns the lower sample (both weights multiply it), not the
    # linear interpolation: pinned on a random table, where the two differ
    rng = np.random.default_rng(3)
    tab = rng.uniform(0, 1, 512).astype(np.float32)
    for cyc in rng.uniform(0, cut, 200):
        g = L.ref_get_att_val(tab.ctypes.data, cpm, cut, cyc)
        low = tab[int(math.floor(cyc / cpm))]
        assert abs(g - low) <= 2 * np.spacing(np.float32(low))
        assert abs(fp64.get_att_val(tab.astype(float), cpm, cut, cyc) - low) < 1e-6
    # at and beyond the cutoff: 1 (no attenuation)
    for cyc in (cut, cut * 1.0001, 3 * cut):
        assert L.ref_get_att_val(tab.ctypes.data, cpm, cut, cyc) == 1.0


def test_attenuation_vector_centre_and_ring():
    """writeAttenuationVector centred at (kx, ky) in the reference's y-flipped
    coordinates: the dip's minimum 1 - str sits at x = kx, y = -ky (a y-flip
    error puts it at h + ky), and pixels at distance fwhm/2 read 1 - str/2."""
    w = h = 64
    vec_cpm = 0.05
    s, fwhm = 0.99, 0.4
    kx, ky = 7.0, -5.0
    out = np.zeros((h, w), np.float32)
    L.ref_write_attenuation(ref.ptr(out), w, h, vec_cpm, s, fwhm, kx, ky)
    iy, ix = np.unravel_index(np.argmin(out), out.shape)
    assert (ix, iy) == (7, 5) and out[iy, ix] == np.float32(1 - s)
    xh, yh = _coords(w, h)
    r = np.hypot(xh - kx, yh - ky) * vec_cpm
    ring = np.abs(r - fwhm / 2) < 1e-12
    assert ring.any() and np.all(np.abs(out[ring] - (1 - s / 2)) < 3e-4)
    assert fp64.rel_l2(out, fp64.attenuation_vector(w, h, vec_cpm, s, fwhm, kx, ky)) < 1e-7


@pytest.mark.parametrize("bend,apo_cut", [(0.9, 2.0), (1.0, 2.0), (0.9, 1.0)])
def test_apodization_closed_forms(bend, apo_cut):
    w = h = 96
    vec_cpm, cut = 0.1, 3.3
    out = np.zeros((h, w), np.float32)
    L.ref_write_apo(ref.ptr(out), w, h, vec_cpm, cut, bend, apo_cut)
    assert out[0, 0] == 1.0
    xh, yh = _coords(w, h)
    frac = np.hypot(xh, yh) * vec_cpm / (cut * apo_cut)
    assert np.all(out[frac >= 1] == 0) and np.all(out[frac < 0.999] > 0)
    # apo^(1/bend) is the ideal OTF (pinned by the pupil autocorrelation)
    inner = (frac > 0) & (frac < 0.95)
    ideal = np.array([L.ref_val_ideal_otf(f) for f in frac[inner]])
    assert np.max(np.abs(out[inner].astype(float) ** (1 / bend) - ideal)) < 2e-6
    assert fp64.rel_l2(out, fp64.apo_vector(w, h, vec_cpm, cut, bend, apo_cut)) < 1e-7


def test_apodization_bend_and_cutoff_scaling():
    """bend is an exponent (apo_0.9 = apo_1^0.9) and apo_cut stretches the
    radius: the value at (2x, 2y) with apo_cut = 2 is the value at (x, y)
    with apo_cut = 1, exactly (the scalings by 2 are exact in float)."""
    w = h = 128
    vec_cpm, cut = 0.07, 2.9
    a1, a9, c1, c2 = (np.zeros((h, w), np.float32) for _ in range(4))
    L.ref_write_apo(ref.ptr(a1), w, h, vec_cpm, cut, 1.0, 2.0)
    L.ref_write_apo(ref.ptr(a9), w, h, vec_cpm, cut, 0.9, 2.0)
    m = a1 > 1e-6
    assert np.max(np.abs(a9[m] - a1[m].astype(float) ** 0.9)) < 2e-6
    L.ref_write_apo(ref.ptr(c1), w, h, vec_cpm, cut, 0.9, 1.0)
    L.ref_write_apo(ref.ptr(c2), w, h, vec_cpm, cut, 0.9, 2.0)
    for x, y in [(3, 4), (10, 1), (20, 17), (31, 0), (25, 30)]:
        assert c2[2 * y, 2 * x] == c1[y, x]


@pytest.mark.parametrize("a", [1.0, 0.7, 0.3])
def test_estimated_otf_curvature_and_abbe_cutoff(a):
    na, lam = 1.4, 525.0
    vals, cut, cpm = ref.otf_estimate(na, lam, a)
    assert abs(cut - 2 * na / (lam / 1000)) < 1e-12      # Abbe limit 2 NA / lambda, cycles per micron
    assert abs(cpm * 512 - cut) < 1e-12
    v = np.arange(512) / 512
    ideal = np.array([L.ref_val_ideal_otf(x) for x in v])
    ok = ideal > 1e-3
    ratio = vals[ok, 0].astype(float) / ideal[ok]
    assert np.max(np.abs(ratio - a ** v[ok]) / a ** v[ok]) < 2e-6
    assert vals[0, 0] == 1.0 and np.all(vals[:, 1] == 0)
    v64, cut64, cpm64 = fp64.otf_from_estimate(na, lam, a)
    assert cut64 == cut and np.max(np.abs(v64 - vals[:, 0])) < 1e-6


@pytest.mark.parametrize("kx,ky", [(0.0, 0.0), (9.0, -6.0), (-13.0, 11.0)])
def test_mask_otf_is_cutoff_disc(kx, ky):
    w = h = 64
    vec_cpm = 0.1
    cut = math.sqrt(150.5) * vec_cpm      # R^2 = 150.5: no pixel on the circle
    v = np.ones((h, w, 2), np.float32)
    L.ref_mask_otf(ref.ptr(v), w, h, vec_cpm, cut, kx, ky)
    keep = np.zeros((h, w), bool)
    for y in range(h):
        for x in range(w):
            xh = x if x < w // 2 else x - w
            yh = -y if y < h // 2 else h - y
            keep[y, x] = (xh - kx) ** 2 + (yh - ky) ** 2 <= 150.5
    assert np.array_equal(v[..., 0] == 1, keep) and np.array_equal(v[..., 1] == 1, keep)
    assert np.array_equal(fp64.mask_otf(w, h, cut, vec_cpm, kx, ky) == 1, keep)


def test_clip_and_scale_closed_forms():
    rng = np.random.default_rng(5)
    img = rng.normal(0.3, 1.0, (40, 40))
    mx = img.max()
    both = fp64.clip_and_scale(img, True, True)
    assert abs(both.max() - 255) < 1e-9 and both.min() == 0
    pos = img > 0
    assert np.allclose(both[pos], img[pos] * 255 / mx) and np.all(both[~pos] == 0)
    clip = fp64.clip_and_scale(img, True, False)
    assert np.array_equal(clip[pos], img[pos]) and np.all(clip[~pos] == 0)
    scale = fp64.clip_and_scale(img, False, True)
    assert np.allclose(scale, img * 255 / mx) and scale.min() < 0
    # scaling uses the maximum of the values, not of the magnitudes: a
    # strongly negative outlier does not change the scale
    img2 = img.copy()
    img2[0, 0] = -100 * mx
    assert abs(fp64.clip_and_scale(img2, True, True).max() - 255) < 1e-9


def _params(n=32):
    sp = {"px": [5.25, -3.5, -1.75], "py": [-1.5, 4.25, -5.0], "phases": [[0, 2 * math.pi / 3, 4 * math.pi / 3]] * 3,
          "modul": 0.7}
    return ref.recon_dict(sp, n)


def test_sim_otf_dc_symmetry_and_support():
    """RL on the output deconvolves with the effective SIM OTF: per
    direction band 0 at weight 1/D and every higher band twice (at +-k).
    At DC that is OTF(0) + sum over directions and bands >= 1 of 2 OTF(|b k|),
    with the lower-sample lookup (pinned in test_oracle)."""
    p = _params()
    vals, cut, cpm = fp64.otf_from_estimate(p["na"], p["lam"], p["a"], p["otf_n"])
    sim = fp64.sim_otf64(p, vals)
    m = 2 * p["n"]
    dc = fp64.get_otf_val(vals, cpm, cut, 0.0) + sum(
        2 * fp64.get_otf_val(vals, cpm, cut, math.hypot(p["px"][d], p["py"][d]) * p["vec_cpm"]) for d in range(3))
    assert abs(sim[0, 0].real - dc) < 1e-12 and abs(sim[0, 0].real - 3.0) > 0.5   # not D x band 0
    # symmetric under k -> -k: the pixel (x, y) maps to (-x, -y) mod m
    flip = sim[(-np.arange(m)) % m][:, (-np.arange(m)) % m]
    assert np.max(np.abs(sim - flip)) < 1e-12
    # zero outside the union of the band discs
    xh, yh = _coords(m, m)
    R = cut / p["vec_cpm"]
    inside = np.hypot(xh, yh) <= R
    for d in range(3):
        for sgn in (1, -1):
            inside |= np.hypot(xh - sgn * p["px"][d], yh - sgn * p["py"][d]) <= R
    assert np.all(sim[~inside] == 0)


def test_rl_output_with_zero_iterations_is_the_unfiltered_sum():
    """runReconstruction with RL on the output and 0 iterations returns the
    plain sum of the pasted, shifted bands (band 0 divided by the number of
    directions), with no OTF or Wiener factor; with iterations it differs."""
    rng = np.random.default_rng(8)
    p = _params()
    n = p["n"]
    vals, cut, cpm = fp64.otf_from_estimate(p["na"], p["lam"], p["a"], p["otf_n"])
    in_fft = [[fp64.fft2(rng.normal(100, 10, (n, n))) for _ in range(3)] for _ in range(3)]
    got = fp64.run_reconstruction64(p, in_fft, vals, filter_style="RLout", rl_iterations=0)
    full = np.zeros((2 * n, 2 * n), complex)
    for d in range(3):
        S = fp64.separation_matrix(p["phases"][d], 2, [fp64.clamp_mod(x) for x in p["modul"][d]])
        sep = [sum(S[c, ph] * in_fft[d][ph] for ph in range(3)) for c in range(3)]
        full += fp64.paste_freq(sep[0], 2 * n, 2 * n) / 3
        for comp, sgn in ((1, -1.0), (2, 1.0)):
            s = fp64.paste_freq(sep[comp], 2 * n, 2 * n)
            full += fp64.fft2(fp64.fourier_shift(fp64.ifft2(s), sgn * p["px"][d], -sgn * p["py"][d]))
    want = fp64.ifft2(full).real
    assert fp64.rel_l2(got, want) < 1e-12
    two = fp64.run_reconstruction64(p, in_fft, vals, filter_style="RLout", rl_iterations=2)
    assert fp64.rel_l2(two, want) > 1e-3


def _three_band_otf(na=1.4, lam=525.0):
    """A measured-style three-band OTF: bands of different curvature."""
    tabs = [fp64.otf_from_estimate(na, lam, a)[0] for a in (1.0, 0.6, 0.3)]
    cut, cpm = fp64.otf_from_estimate(na, lam, 1.0)[1:]
    return ref.OtfBands32(np.stack(tabs), cut, cpm)


@pytest.mark.parametrize("nphase,nband,k_frac,h2", [(3, 2, 0.85, 0.0), (5, 3, 0.45, 0.3)])
def test_refit_known_answer(nphase, nband, k_frac, h2):
    """The live refit recovers the pattern: the highest band's shift (twice
    the band-1 shift with three bands) within 0.1 px, phase offset near 0
    (the data's phases are 2 pi p / P; the binning of the fine grid moves the
    pattern's phase at the raw pixel centres a little) within 0.2 rad, as for
    estimateParameters."""
    from fsim import synthetic
    n = 128
    raw, sp = synthetic.sim_stack(n=n, nphase=nphase, seed=31, nbeads=60, nfil=8, harmonic2=h2, k_frac=k_frac)
    fft_in = ref.input_ffts32_live(raw.reshape(3, nphase, n, n))
    otf = ref.Otf32(1.4, 525.0, 1.0)
    got = ref.refit32(fft_in, nband, otf, 1 / (n * 0.08))
    scale = nband - 1
    for d in range(3):
        g = got[d]
        assert abs(g["px"] - scale * sp["px"][d]) < 0.1 and abs(g["py"] - scale * sp["py"][d]) < 0.1
        assert abs(g["pha_off"]) < 0.2
        assert 0.3 < g["mod1"] < 1.1
        assert "mod2" not in g


def test_refit_equals_estimate_where_the_flows_coincide():
    """Two bands: the refit is estimateParameters with fit band 1 and
    fitExclude 0.8 (same arithmetic, so identical results).  Three bands with
    a single-band OTF: fit band 2 gives the same shift, phase and modulation
    1 (the OTF band numbers do not matter for one table)."""
    from fsim import synthetic
    n = 128
    otf = ref.Otf32(1.4, 525.0, 1.0)
    raw, _ = synthetic.sim_stack(n=n, seed=32, nbeads=60, nfil=8)
    fft_in = ref.input_ffts32_live(raw.reshape(3, 3, n, n))
    a = ref.refit32(fft_in, 2, otf, 1 / (n * 0.08))
    b = ref.estimate32(fft_in, 2, 1, 0.8, otf, 1 / (n * 0.08))
    for d in range(3):
        for k in ("px", "py", "pha_off", "mod1"):
            assert a[d][k] == b[d][k]
    raw, _ = synthetic.sim_stack(n=n, nphase=5, seed=33, nbeads=60, nfil=8, harmonic2=0.3, k_frac=0.45)
    fft_in = ref.input_ffts32_live(raw.reshape(3, 5, n, n))
    a = ref.refit32(fft_in, 3, otf, 1 / (n * 0.08))
    b = ref.estimate32(fft_in, 3, 2, 0.8, otf, 1 / (n * 0.08))
    for d in range(3):
        for k in ("px", "py", "pha_off", "mod1"):
            assert a[d][k] == b[d][k]
    # the live input differs from the GUI's fadeBorderCos on the image only
    # in float rounding
    gui = ref.input_ffts32(raw.reshape(3, 5, n, n).astype(np.float32))
    assert 0 < fp64.rel_l2(ref.to_complex(fft_in), ref.to_complex(gui)) < 1e-6


def test_three_band_otf_band_numbers_matter():
    """With a three-band OTF, fitPeak's band numbers select the tables of the
    commonRegion weights: the refit's (0, 1) and estimateParameters' (0, 2)
    give different correlations of the same bands."""
    from fsim import synthetic
    n = 128
    otf = _three_band_otf()
    raw, sp = synthetic.sim_stack(n=n, nphase=5, seed=34, nbeads=60, nfil=8, harmonic2=0.3, k_frac=0.45)
    fft_in = ref.input_ffts32_live(raw.reshape(3, 5, n, n))
    S = np.zeros(2 * 5 * 5)
    ph = np.array([2 * np.pi * p / 5 for p in range(5)])
    assert ref.lib().ref_separation_matrix(ref.ptr(ph), 5, 3, None, ref.ptr(S)) == 0
    Sf = S.astype(np.float32)
    sep = np.zeros((5, n, n, 2), np.float32)
    ins = (C.c_void_p * 5)(*[fft_in[0, p].ctypes.data for p in range(5)])
    outs = (C.c_void_p * 5)(*[sep[c].ctypes.data for c in range(5)])
    ref.lib().ref_separate_bands(ins, outs, n * n, ref.ptr(Sf), 5, 5)
    k = (-2 * sp["px"][0], -2 * sp["py"][0])
    f1 = ref.fit_peak32(sep[0], sep[3], otf, 1 / (n * 0.08), k[0], k[1], 0.05, 2.5, bn1=1)
    f2 = ref.fit_peak32(sep[0], sep[3], otf, 1 / (n * 0.08), k[0], k[1], 0.05, 2.5, bn1=2)
    assert abs(f1[3] - f2[3]) > 1e-3 * abs(f2[3])
    got = ref.refit32(fft_in, 3, otf, 1 / (n * 0.08))
    assert abs(got[0]["px"] - 2 * sp["px"][0]) < 0.1
