"""ORACLE -- TEST INFRASTRUCTURE ONLY.

FP64 NumPy restatement of fairSIM's fast-shift 2D SIM reconstruction
(the "FP64 re-run" every parity test states its error against).  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import it.

It follows the same Java text as ``java_order.c`` but with independent
building blocks so that the two restatements check each other:
  * FFTs are numpy's pocketfft in complex128 (forward unscaled, inverse
    scaled 1/(w*h) like JTransforms ``complexInverse(v, true)``,
    linalg/JTransformsConnector.java:76-90);
  * the separation matrix is inverted with ``numpy.linalg.inv`` / ``pinv``
    instead of MatrixComplex's LU (sim_algorithm/BandSeparation.java:40-114);
  * the shift ramp uses the exact exp(i*phi) without the float rounding of
    the phase (linalg/BasicVectors.java:309-339);
  * all tables stay in double (no float casts), keeping the getOtfVal /
    getAttVal lower-sample quirk (sim_algorithm/OtfProvider.java:189-269).
"""
import numpy as np


def val_ideal_otf(dist):
    """OtfProvider.valIdealOTF (OtfProvider.java:74-78), vectorised."""
    d = np.asarray(dist, dtype=np.float64)
    out = np.zeros_like(d)
    ok = (d >= 0) & (d <= 1)
    dd = d[ok]
    out[ok] = (2 / np.pi) * (np.arccos(dd) - dd * np.sqrt(1 - dd * dd))
    return out


def otf_from_estimate(na, lam, a, n=512):
    """OtfProvider.fromEstimate (OtfProvider.java:96-127) in double."""
    cutoff = 1000 / (lam / na / 2)
    cpm = cutoff / n
    v = np.arange(n) / float(n)
    vals = val_ideal_otf(v) * np.power(a, v)
    return vals.astype(np.complex128), cutoff, cpm


def val_attenuation(dist, s, fwhm):
    """OtfProvider.valAttenuation (OtfProvider.java:238-240)."""
    return 1 - s * np.exp(-(np.asarray(dist) ** 2) / (2 * (fwhm / 2.355) ** 2))


def get_otf_val(tab, cpm, cutoff, cycl):
    """OtfProvider.getOtfVal (OtfProvider.java:189-222): v[l]*(1-f)+v[l]*f."""
    cycl = np.asarray(cycl, dtype=np.float64)
    n = len(tab)
    pos = cycl / cpm
    ok = (cycl < cutoff) & (np.ceil(pos) < n)
    l = np.where(ok, np.floor(pos), 0).astype(np.int64)
    f = pos - l
    v = tab[l]
    return np.where(ok, v * (1 - f) + v * f, 0)


def get_att_val(only_att, cpm, cutoff, cycl):
    """OtfProvider.getAttVal (OtfProvider.java:246-269)."""
    cycl = np.asarray(cycl, dtype=np.float64)
    pos = cycl / cpm
    ok = cycl < cutoff
    l = np.where(ok, np.floor(pos), 0).astype(np.int64)
    f = pos - l
    v = only_att[l]
    return np.where(ok, v * f + v * (1 - f), 1.0)


def band_table(vals, b):
    """The radial table of band b: a single-band OTF serves every band
    (OtfProvider.getOtfVal, OtfProvider.java:192-193)."""
    vals = np.asarray(vals)
    return vals[b] if vals.ndim == 2 else vals


def coords(w, h):
    """Wrapped, y-flipped frequency coordinates (OtfProvider.java:345-346)."""
    x = np.arange(w)
    y = np.arange(h)
    xh = np.where(x < w // 2, x, x - w).astype(np.float64)
    yh = np.where(y < h // 2, -y, h - y).astype(np.float64)
    return xh[None, :], yh[:, None]


def otf_to_vector(w, h, tab, cpm, cutoff, vec_cpm, kx=0.0, ky=0.0):
    """OtfProvider.otfToVector in write mode (OtfProvider.java:331-373)."""
    xh, yh = coords(w, h)
    cycl = np.sqrt((xh - kx) ** 2 + (yh - ky) ** 2) * vec_cpm
    val = get_otf_val(tab, cpm, cutoff, cycl)
    return np.where(cycl > cutoff, 0, val)


def apo_vector(w, h, vec_cpm, cutoff, bend, apo_cut):
    """OtfProvider.writeApoVector (OtfProvider.java:437-492)."""
    xh, yh = coords(w, h)
    cycl = np.sqrt(xh ** 2 + yh ** 2) * vec_cpm
    return np.power(val_ideal_otf(cycl / (cutoff * apo_cut)), bend)


def attenuation_vector(w, h, vec_cpm, s, fwhm, kx=0.0, ky=0.0):
    """OtfProvider.writeAttenuationVector (OtfProvider.java:503-529)."""
    xh, yh = coords(w, h)
    cycl = np.sqrt((xh - kx) ** 2 + (yh - ky) ** 2) * vec_cpm
    return val_attenuation(cycl, s, fwhm)


def wiener_denominator(w, h, vals, only_att, cpm, cutoff, vec_cpm, px, py, nband,
                       use_att, wparam):
    """WienerFilter.updateCache/addWienerDenominator/getDenominator
    (WienerFilter.java:46-122); band 0 is counted twice per direction."""
    xh, yh = coords(w, h)
    acc = np.zeros((h, w))
    for d in range(len(px)):
        for b in range(nband):
            bx, by = px[d] * b, py[d] * b
            r1 = np.sqrt((xh - bx) ** 2 + (yh - by) ** 2) * vec_cpm
            r2 = np.sqrt((xh + bx) ** 2 + (yh + by) ** 2) * vec_cpm
            tb, ta = band_table(vals, b), band_table(only_att, b)
            v1 = np.abs(get_otf_val(tb, cpm, cutoff, r1)) ** 2
            v2 = np.abs(get_otf_val(tb, cpm, cutoff, r2)) ** 2
            if use_att:
                v1 = v1 * get_att_val(ta, cpm, cutoff, r1)
                v2 = v2 * get_att_val(ta, cpm, cutoff, r2)
            acc = acc + v1 + v2
    return 1 / (acc + wparam * wparam)


def fade_border(w, h, px):
    """SimUtils.fadeBorderCos (SimUtils.java:162-185) on a vector of ones."""
    d = np.ones((h, w))
    fac = 1.0 / px * np.pi / 2.0
    for y in range(px):
        d[y, :] *= np.sin(y * fac) ** 2
    for y in range(h - px, h):
        d[y, :] *= np.sin((h - y - 1) * fac) ** 2
    for x in range(px):
        d[:, x] *= np.sin(x * fac) ** 2
    for x in range(w - px, w):
        d[:, x] *= np.sin((w - x - 1) * fac) ** 2
    return d


def separation_matrix(phases, nband, fac=None):
    """BandSeparation.createSeparationMatrix (BandSeparation.java:40-114),
    with numpy's inverse in place of MatrixComplex's LU.  Returns C x P."""
    phases = np.asarray(phases, dtype=np.float64)
    P = len(phases)
    C = 2 * nband - 1
    if fac is None:
        f = [1.0] + [0.5] * (nband - 1)
    else:
        f = [fac[0]] + [fac[i] * 0.5 for i in range(1, nband)]
    M = np.zeros((P, C), dtype=np.complex128)
    M[:, 0] = f[0]
    for j in range(nband - 1):
        ph = phases * (j + 1)
        M[:, 2 * j + 1] = f[j + 1] * np.exp(1j * ph)
        M[:, 2 * j + 2] = f[j + 1] * np.exp(-1j * ph)
    if P == C:
        return np.linalg.inv(M)
    return np.linalg.inv(M.conj().T @ M) @ M.conj().T


def paste_freq(inv, wo, ho, xoff=0, yoff=0):
    """Vec2d.Cplx.pasteFreq (BasicVectors.java:259-285), vectorised."""
    hi, wi = inv.shape
    out = np.zeros((ho, wo), dtype=inv.dtype)
    x = np.arange(wi)
    y = np.arange(hi)
    xo = np.where(x < wi // 2, x, x + wo // 2)
    yo = np.where(y < hi // 2, y, y + ho // 2)
    xo = (xo + xoff + wo) % wo
    yo = (yo + yoff + ho) % ho
    out[np.ix_(yo, xo)] = inv
    return out


def fourier_shift(v, kx, ky):
    """Vec2d.Cplx.fourierShift (BasicVectors.java:309-339), exact exp."""
    n = v.shape[0]
    y, x = np.mgrid[0:n, 0:n]
    return v * np.exp(1j * (2 * np.pi * (kx * x + ky * y) / n))


def fft2(v):
    return np.fft.fft2(v)


def ifft2(v):
    return np.fft.ifft2(v)


def clamp_mod(m):
    """SimParam.Dir.getModulations (SimParam.java:598-607)."""
    return min(max(m, 0.3), 1.1)


def build_tables(p, vals):
    """OTF (N^2), reciprocal Wiener denominator (2N^2) and apodization (2N^2)
    as built by ReconstructionRunner.setupFilters (:773-797) and
    TestAccelSpeed (:506-529)."""
    n = p["n"]
    vals = np.asarray(vals)
    only_att = val_attenuation(np.arange(vals.shape[-1]) * p["otf_cpm"], p["att_str"], p["att_fwhm"])
    vals_att = vals * only_att
    tabs = vals_att if p["use_att"] else vals
    if vals.ndim == 2:
        only_att = np.broadcast_to(only_att, vals.shape)
        # per band, or band 0's for every component (the live runner)
        nb = 1 if p.get("otf_band0") else vals.shape[0]
        otf = [otf_to_vector(n, n, tabs[b], p["otf_cpm"], p["otf_cutoff"], p["vec_cpm"]) for b in range(nb)]
    else:
        otf = otf_to_vector(n, n, tabs, p["otf_cpm"], p["otf_cutoff"], p["vec_cpm"])
    wien = wiener_denominator(2 * n, 2 * n, vals, only_att, p["otf_cpm"], p["otf_cutoff"],
                              p["vec_cpm"], p["px"], p["py"], p["nband"], p["use_att"],
                              p["wiener"])
    apo = apo_vector(2 * n, 2 * n, p["vec_cpm"], p["otf_cutoff"], p["apo_bend"], p["apo_cut"])
    return otf, wien, apo


def recon_fast(p, raw, vals, tables=None, return_full=False):
    """Fast-shift reconstruction (tests/TestAccelSpeed.java:237-252, 563-703;
    livemode/ReconstructionRunner.java:456-542) in FP64.
    raw: (D*P, n, n) array, angle-major.  Returns the (2n, 2n) real image."""
    n, D, P, B = p["n"], p["ndir"], p["nphase"], p["nband"]
    C = 2 * B - 1
    otf, wien, apo = tables if tables is not None else build_tables(p, vals)
    fade = fade_border(n, n, p["fade_px"]) if p["fade_px"] > 0 else np.ones((n, n))
    full = np.zeros((2 * n, 2 * n), dtype=np.complex128)
    raw = np.asarray(raw, dtype=np.float64).reshape(D, P, n, n)
    for d in range(D):
        spec = [fft2(raw[d, ph] * fade) for ph in range(P)]
        mods = [clamp_mod(m) for m in p["modul"][d][:B]]
        S = separation_matrix(p["phases"][d][:P], B, mods)
        sep = [sum(S[c, ph] * spec[ph] for ph in range(P)) *
               np.conj(otf[min((c + 1) // 2, len(otf) - 1)] if isinstance(otf, list) else otf) for c in range(C)]
        full += paste_freq(sep[0], 2 * n, 2 * n)
        for b in range(1, B):
            for comp, sgn in ((2 * b, 1.0), (2 * b - 1, -1.0)):
                kx, ky = sgn * p["px"][d] * b, sgn * p["py"][d] * b
                if p.get("fast_shift", 1):
                    xf, yf = kx - np.floor(kx), -ky - np.floor(-ky)
                    s = fft2(fourier_shift(ifft2(sep[comp]), xf, yf))
                    full += paste_freq(s, 2 * n, 2 * n, int(np.floor(kx)), int(np.floor(-ky)))
                else:
                    s = paste_freq(sep[comp], 2 * n, 2 * n)
                    full += fft2(fourier_shift(ifft2(s), kx, -ky))
    full = full * wien * apo
    res = ifft2(full).real
    if return_full:
        return res, full
    return res


def rel_l2(a, b):
    """||a - b|| / ||b|| in double (complex inputs stay complex)."""
    a = np.asarray(a)
    b = np.asarray(b)
    dt = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a = a.astype(dt)
    b = b.astype(dt)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def rl_deconvolve64(img, otf, steps, in_freq=False):
    """RLDeconvolution.deconvolve (RLDeconvolution.java:43-103) in FP64:
    u_{j+1} = u_j x IFFT(FFT(d / IFFT(FFT(u_j) otf)) otf), u_0 = d."""
    d = ifft2(img) if in_freq else np.asarray(img, dtype=np.complex128)
    u = d.copy()
    for _ in range(steps):
        nxt = ifft2(fft2(u) * otf)
        nxt = d / nxt
        nxt = ifft2(fft2(nxt) * otf)
        u = nxt * u
    return fft2(u) if in_freq else u


def mask_otf(w, h, cutoff, vec_cpm, kx, ky):
    """OtfProvider.maskOtf (OtfProvider.java:538-559) as a 0/1 array."""
    xh, yh = coords(w, h)
    return np.where(np.sqrt((xh - kx) ** 2 + (yh - ky) ** 2) * vec_cpm > cutoff, 0.0, 1.0)


def clip_and_scale(img, clip, scale):
    """SimUtils.clipAndScale (SimUtils.java:244-264)."""
    img = img.copy()
    if scale:
        mx = max(0.0, float(img.max()))
        img = img * (255.0 / mx)
    if clip:
        img[img < 0] = 0
    return img


def sim_otf64(p, vals):
    """The effective SIM OTF of runReconstruction's RL-on-output branch
    (SimAlgorithm.java:592-615) on the (2n)^2 grid: per direction band 0 at
    weight 1/D, every higher band b at +-(px, py) * b (writeOtfVector, the
    unattenuated table)."""
    n, D, B = p["n"], p["ndir"], p["nband"]
    cpm, cut, vcpm = p["otf_cpm"], p["otf_cutoff"], p["vec_cpm"]
    sim = np.zeros((2 * n, 2 * n), dtype=np.complex128)
    for d in range(D):
        for b in range(B):
            bx, by = p["px"][d] * b, p["py"][d] * b
            if b == 0:
                sim = sim + otf_to_vector(2 * n, 2 * n, vals, cpm, cut, vcpm, bx, by) / D
            else:
                sim = sim + otf_to_vector(2 * n, 2 * n, vals, cpm, cut, vcpm, bx, by)
                sim = sim + otf_to_vector(2 * n, 2 * n, vals, cpm, cut, vcpm, -bx, -by)
    return sim


def run_reconstruction64(p, in_fft, vals, filter_style="Wiener", otf_before_shift=False, rl_iterations=5,
                         clip_scale="NONE"):
    """SimAlgorithm.runReconstruction (SimAlgorithm.java:268-709, no
    intermediate displays) in FP64: slow-shift reconstruction of FFT'd
    inputs in_fft[d][p] with the filter styles Wiener, RLin, RLout, RLboth
    and the otfBeforeShift switch, including the reference's mask around
    (angIdx, i) in the otfBeforeShift branch (:408-410)."""
    n, D, P, B = p["n"], p["ndir"], p["nphase"], p["nband"]
    C = 2 * B - 1
    use_w = filter_style in ("Wiener", "RLin")
    use_rin = filter_style in ("RLin", "RLboth")
    use_rout = filter_style in ("RLout", "RLboth")
    cpm, cut, vcpm = p["otf_cpm"], p["otf_cutoff"], p["vec_cpm"]
    only_att = val_attenuation(np.arange(len(vals)) * cpm, p["att_str"], p["att_fwhm"])
    tab = vals * only_att if p["use_att"] else vals     # applyOtf uses the attenuated table when on
    full = np.zeros((2 * n, 2 * n), dtype=np.complex128)
    input_otf = otf_to_vector(n, n, vals, cpm, cut, vcpm) if use_rin else None
    for d in range(D):
        mods = [clamp_mod(m) for m in p["modul"][d][:B]]
        S = separation_matrix(p["phases"][d][:P], B, mods)
        inputs = [np.asarray(v, dtype=np.complex128) for v in in_fft[d]]
        if use_rin:
            if P != C:
                raise IndexError("#input images != phases per band")
            inputs = [rl_deconvolve64(inputs[i], input_otf, rl_iterations, True) for i in range(C)]
        sep = [sum(S[c, ph] * inputs[ph] for ph in range(P)) for c in range(C)]
        if otf_before_shift and use_w:
            o = np.conj(otf_to_vector(n, n, tab, cpm, cut, vcpm))
            sep = [s * o for s in sep]
        shifted = [paste_freq(sep[0], 2 * n, 2 * n)]
        kxy = [(0.0, 0.0)]
        for b in range(1, B):
            for comp, sgn in ((2 * b - 1, -1.0), (2 * b, 1.0)):
                kx, ky = sgn * p["px"][d] * b, sgn * p["py"][d] * b
                s = paste_freq(sep[comp], 2 * n, 2 * n)
                shifted.append(fft2(fourier_shift(ifft2(s), kx, -ky)))
                kxy.append((kx, ky))
        if use_w:
            for i in range(C):
                if not otf_before_shift:
                    shifted[i] = shifted[i] * np.conj(otf_to_vector(2 * n, 2 * n, tab, cpm, cut, vcpm, *kxy[i]))
                else:
                    shifted[i] = shifted[i] * mask_otf(2 * n, 2 * n, cut, vcpm, d, i)
        if use_rout:
            shifted[0] = shifted[0] / D
        for s in shifted:
            full = full + s
    img = None
    if use_w:
        wien = wiener_denominator(2 * n, 2 * n, vals, only_att, cpm, cut, vcpm, p["px"], p["py"], B,
                                  p["use_att"], p["wiener"])
        apo = apo_vector(2 * n, 2 * n, vcpm, cut, 0.9, p["apo_cut"])
        img = ifft2(full * wien * apo).real
    if use_rout:
        img = ifft2(rl_deconvolve64(full, sim_otf64(p, vals), rl_iterations, True)).real
    if clip_scale in ("CLIP", "BOTH"):
        img = clip_and_scale(img, True, clip_scale == "BOTH")
    return img
