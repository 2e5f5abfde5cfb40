"""ORACLE -- TEST INFRASTRUCTURE ONLY.

ctypes front end of liboracle.so (java_order.c, the FP32 Java-order CPU
restatement) plus helpers that describe a reconstruction as a plain dict
shared with the FP64 restatement (fp64.py).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


class CF(C.Structure):
    _fields_ = [("re", C.c_float), ("im", C.c_float)]


class RefParams(C.Structure):
    """ref_recon_params in java_order.c"""
    _fields_ = [
        ("n", C.c_int), ("ndir", C.c_int), ("nphase", C.c_int), ("nband", C.c_int),
        ("px", C.c_double * 8), ("py", C.c_double * 8),
        ("phases", (C.c_double * 16) * 8),
        ("modul", (C.c_double * 4) * 8),
        ("otf_n", C.c_int),
        ("otf_cpm", C.c_double), ("otf_cutoff", C.c_double),
        ("vec_cpm", C.c_double),
        ("att_str", C.c_double), ("att_fwhm", C.c_double),
        ("use_att", C.c_int),
        ("wiener", C.c_double), ("apo_bend", C.c_double), ("apo_cut", C.c_double),
        ("fade_px", C.c_int),
        ("fast_shift", C.c_int),
    ]


_lib = None


def build():
    subprocess.run(["make", "-B", "-C", HERE], check=True, capture_output=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.c_void_p
        D = C.c_double
        I = C.c_int
        L.ref_set_threads.argtypes = [I]
        L.ref_fsin.argtypes = [D]
        L.ref_fsin.restype = D
        L.ref_fcos.argtypes = [D]
        L.ref_fcos.restype = D
        L.ref_fade_border.argtypes = [P, I, I, I]
        L.ref_val_ideal_otf.argtypes = [D]
        L.ref_val_ideal_otf.restype = D
        L.ref_otf_from_estimate.argtypes = [D, D, D, I, P, P, P]
        L.ref_otf_set_attenuation.argtypes = [P, I, D, D, D, P, P]
        L.ref_otf_to_vector.argtypes = [P, I, I, P, I, D, D, D, D, D, I]
        L.ref_write_apo.argtypes = [P, I, I, D, D, D, D]
        L.ref_write_attenuation.argtypes = [P, I, I, D, D, D, D, D]
        L.ref_mask_otf.argtypes = [P, I, I, D, D, D, D]
        L.ref_wiener_denominator.argtypes = [P, I, I, P, P, I, D, D, D, I, I, P, P, I, D]
        L.ref_separation_matrix.argtypes = [P, I, I, P, P]
        L.ref_separation_matrix.restype = I
        L.ref_separate_bands.argtypes = [P, P, C.c_long, P, I, I]
        L.ref_paste_freq.argtypes = [P, I, I, P, I, I, I, I]
        L.ref_fourier_shift.argtypes = [P, I, D, D]
        L.ref_fft2d.argtypes = [P, I, I, I]
        L.ref_build_tables.argtypes = [C.POINTER(RefParams), P, P, P, P]
        L.ref_recon_fast.argtypes = [C.POINTER(RefParams), P, P, P, P, P]
        L.ref_recon_fast.restype = I
        L.ref_locate_peak.argtypes = [P, I, I, D, P]
        L.ref_norm2.argtypes = [P, C.c_long]
        L.ref_norm2.restype = D
        L.ref_sum_elements.argtypes = [P, C.c_long, P]
        L.ref_common_region.argtypes = [P, P, I, I, P, P, I, D, D, D, D, D, D, D, I]
        L.ref_fit_peak.argtypes = [P, P, I, P, P, I, D, D, D, D, D, D, D, P, P]
        L.ref_get_peak.argtypes = [P, P, I, P, P, I, D, D, D, D, D, D, P]
        L.ref_shifted_corr.argtypes = [P, P, I, D, D, P]
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def cplx_f32(z):
    """complex array -> interleaved float32 [..., 2]"""
    z = np.asarray(z)
    a = np.empty(z.shape + (2,), dtype=np.float32)
    a[..., 0] = z.real
    a[..., 1] = z.imag
    return a


def to_complex(a):
    return a[..., 0].astype(np.float64) + 1j * a[..., 1].astype(np.float64)


def otf_estimate(na, lam, a, n=512):
    vals = np.zeros((n, 2), dtype=np.float32)
    cut, cpm = C.c_double(), C.c_double()
    lib().ref_otf_from_estimate(na, lam, a, n, ptr(vals), C.byref(cut), C.byref(cpm))
    return vals, cut.value, cpm.value


def recon_dict(sp, n, nband=2, na=1.4, lam=525.0, px_um=0.080, a=1.0, att_str=0.995,
               att_fwhm=1.2, use_att=True, wiener=0.05, apo_bend=0.9, apo_cut=2.0,
               fade_px=10, fast_shift=1, modulation=None, otf_n=512):
    """Plain-dict description of a reconstruction from synthetic parameters
    (fsim.synthetic.sim_stack).  px/py are the band-1 shift pX/pY
    (SimParam.java:566-569: band b sits at b * pX): sim_stack reports the
    band-1 shift, and fsim.synthetic.make_param hands the highest band's
    shift sp['px'] * (nband - 1) to setPxPy, which divides it by nband - 1
    (SimParam.java:536-541), so both describe the same pattern."""
    ndir = len(sp["px"])
    nphase = len(sp["phases"][0])
    cutoff = 1000 / (lam / na / 2)
    m = sp["modul"] if modulation is None else modulation
    return dict(
        n=n, ndir=ndir, nphase=nphase, nband=nband,
        px=[sp["px"][d] for d in range(ndir)],
        py=[sp["py"][d] for d in range(ndir)],
        phases=[list(sp["phases"][d]) for d in range(ndir)],
        modul=[[1.0] + [m] * (nband - 1) for _ in range(ndir)],
        otf_n=otf_n, otf_cpm=cutoff / otf_n, otf_cutoff=cutoff,
        vec_cpm=1.0 / (n * px_um), att_str=att_str, att_fwhm=att_fwhm,
        use_att=1 if use_att else 0, wiener=wiener, apo_bend=apo_bend, apo_cut=apo_cut,
        fade_px=fade_px, fast_shift=fast_shift, na=na, lam=lam, a=a)


def ref_params(p):
    r = RefParams()
    r.n, r.ndir, r.nphase, r.nband = p["n"], p["ndir"], p["nphase"], p["nband"]
    for d in range(p["ndir"]):
        r.px[d], r.py[d] = p["px"][d], p["py"][d]
        for i, ph in enumerate(p["phases"][d]):
            r.phases[d][i] = ph
        for b in range(p["nband"]):
            r.modul[d][b] = p["modul"][d][b]
    r.otf_n = p["otf_n"]
    r.otf_cpm, r.otf_cutoff, r.vec_cpm = p["otf_cpm"], p["otf_cutoff"], p["vec_cpm"]
    r.att_str, r.att_fwhm, r.use_att = p["att_str"], p["att_fwhm"], p["use_att"]
    r.wiener, r.apo_bend, r.apo_cut = p["wiener"], p["apo_bend"], p["apo_cut"]
    r.fade_px, r.fast_shift = p["fade_px"], p["fast_shift"]
    return r


def tables32(p):
    """OTF (n^2 complex), 1/Wiener and apodization ((2n)^2) in FP32."""
    vals, cut, cpm = otf_estimate(p["na"], p["lam"], p["a"], p["otf_n"])
    n = p["n"]
    otf = np.zeros((n, n, 2), dtype=np.float32)
    wien = np.zeros((2 * n, 2 * n), dtype=np.float32)
    apo = np.zeros((2 * n, 2 * n), dtype=np.float32)
    rp = ref_params(p)
    lib().ref_build_tables(C.byref(rp), ptr(vals), ptr(otf), ptr(wien), ptr(apo))
    return vals, otf, wien, apo


def recon32(p, raw, threads=1, tables=None):
    """Fast-shift reconstruction in FP32 Java order -> (2n, 2n) float32."""
    lib().ref_set_threads(threads)
    n = p["n"]
    if tables is None:
        tables = tables32(p)
    vals, otf, wien, apo = tables
    raw = np.ascontiguousarray(raw, dtype=np.uint16)
    out = np.zeros((2 * n, 2 * n), dtype=np.float32)
    rp = ref_params(p)
    rc = lib().ref_recon_fast(C.byref(rp), ptr(raw), ptr(otf), ptr(wien), ptr(apo), ptr(out))
    if rc != 0:
        raise RuntimeError(f"ref_recon_fast failed: {rc}")
    return out


def fp64_params(p):
    """The same description for oracle/fp64.py."""
    return dict(p)


def recon64(p, raw):
    from . import fp64
    vals, cut, cpm = fp64.otf_from_estimate(p["na"], p["lam"], p["a"], p["otf_n"])
    return fp64.recon_fast(p, raw, vals)


# ---------------------------------------------------------------------------
# parameter estimation (sim_algorithm/Correlation.java, SimAlgorithm.java:37-256)
# ---------------------------------------------------------------------------
class Otf32:
    """A single-band OtfProvider.fromEstimate table (no attenuation: the
    estimation's writeOtfVector calls do not use it)."""

    def __init__(self, na=1.4, lam=525.0, a=1.0, n=512):
        self.vals, self.cutoff, self.cpm = otf_estimate(na, lam, a, n)
        self.ns = n

    def tab(self, band):
        """The table of OTF band `band` (a single-band OTF serves every band,
        OtfProvider.getOtfVal: `if (!isMultiBand) band = 0`)."""
        return self.vals


class OtfBands32:
    """A measured multi-band OTF (OtfProvider.fromData / loadFromConfig,
    OtfProvider.java:592-628): one complex table of `ns` samples per band,
    float32 [nb, ns, 2]."""

    def __init__(self, vals, cutoff, cpm):
        vals = np.asarray(vals)
        self.tabs = [np.ascontiguousarray(cplx_f32(v)) for v in vals]
        self.vals = self.tabs[0]
        self.cutoff, self.cpm, self.ns = cutoff, cpm, vals.shape[1]

    def tab(self, band):
        if band >= len(self.tabs):
            raise IndexError("band idx too high or <0")
        return self.tabs[band]


def common_region32(b0, b1, otf, vec_cpm, kx, ky, dist, wlim, divide=True):
    """Correlation.commonRegion on float32 [N, N, 2] arrays, in place."""
    h, w = b0.shape[:2]
    lib().ref_common_region(ptr(b0), ptr(b1), w, h, ptr(otf.vals), ptr(otf.vals), otf.ns, otf.cpm,
                            otf.cutoff, vec_cpm, kx, ky, dist, wlim, 1 if divide else 0)


def fit_peak32(b0, b1, otf, vec_cpm, kx, ky, wlim, search, cntrl=False, bn0=0, bn1=1):
    """Correlation.fitPeak(b0, b1, bn0, bn1, otf, ...): [kx, ky, phase, mag]
    (and the 10 x 30 control image); bn0 / bn1 select the OTF bands."""
    n = b0.shape[0]
    res = np.zeros(4)
    c = np.zeros((10, 30), dtype=np.float32) if cntrl else None
    lib().ref_fit_peak(ptr(b0), ptr(b1), n, ptr(otf.tab(bn0)), ptr(otf.tab(bn1)), otf.ns, otf.cpm, otf.cutoff,
                       vec_cpm, kx, ky, wlim, search, ptr(c) if cntrl else None, ptr(res))
    return (list(res), c) if cntrl else list(res)


def get_peak32(b0, b1, otf, vec_cpm, kx, ky, wlim, bn0=0, bn1=1):
    """Correlation.getPeak(b0, b1, bn0, bn1, otf, ...): complex."""
    n = b0.shape[0]
    res = np.zeros(2)
    lib().ref_get_peak(ptr(b0), ptr(b1), n, ptr(otf.tab(bn0)), ptr(otf.tab(bn1)), otf.ns, otf.cpm, otf.cutoff,
                       vec_cpm, kx, ky, wlim, ptr(res))
    return complex(res[0], res[1])


def shifted_corr32(b0, b1, kx, ky):
    """sum(b1.fourierShift(kx, ky) * conj(b0)) in the reference's arithmetic
    (float phase, MTool polynomial, float products, double sum)."""
    r = np.zeros(2)
    lib().ref_shifted_corr(ptr(b0), ptr(b1), b0.shape[0], float(kx), float(ky), ptr(r))
    return complex(r[0], r[1])


def times_conj32(y, x):
    """y = y * conj(x) in AbstractVectorCplx.timesConj's float order."""
    xr, xi = x[..., 0], -x[..., 1]
    yr, yi = y[..., 0].copy(), y[..., 1].copy()
    y[..., 0] = (xr * yr) - (xi * yi)
    y[..., 1] = (xi * yr) + (xr * yi)


def fft32(v, inverse):
    lib().ref_fft2d(ptr(v), v.shape[1], v.shape[0], 1 if inverse else 0)


def input_ffts32(images, fade_px=10):
    """The estimation input as the GUI builds it (ImageControl.java:816-840):
    float image, fadeBorderCos(10), to complex, forward FFT."""
    n = images.shape[-1]
    out = np.zeros(images.shape + (2,), dtype=np.float32)
    for idx in np.ndindex(images.shape[:-2]):
        img = np.ascontiguousarray(images[idx], dtype=np.float32)
        lib().ref_fade_border(ptr(img), n, n, fade_px)
        out[idx][..., 0] = img
        fft32(out[idx], False)
    return out


def estimate32(in_fft, nband, fit_band, fit_exclude, otf, vec_cpm):
    """SimAlgorithm.estimateParameters (SimAlgorithm.java:37-256, no visual
    feedback) on FFT'd inputs [D, P, N, N, 2]: per direction the arguments it
    stores, {px, py (setPxPy arguments), pha_off, mod1, mod2, peak}."""
    D, P, n = in_fft.shape[0], in_fft.shape[1], in_fft.shape[2]
    att = np.zeros((n, n), dtype=np.float32)
    lib().ref_write_attenuation(ptr(att), n, n, vec_cpm, .99, 0.15 * otf.cutoff, 0.0, 0.0)
    Cc = 2 * nband - 1
    out = []
    for d in range(D):
        lb, hb = 1, (3 if nband == 3 else 1)
        fb = lb if fit_band == 1 else hb
        phases = np.array([(2 * np.pi * p) / P + 0.0 for p in range(P)])
        Sm = np.zeros(2 * Cc * P)
        assert lib().ref_separation_matrix(ptr(phases), P, nband, None, ptr(Sm)) == 0
        Sf = Sm.astype(np.float32)
        sep = np.zeros((Cc, n, n, 2), dtype=np.float32)
        ins = (C.c_void_p * P)(*[in_fft[d, p].ctypes.data for p in range(P)])
        outs = (C.c_void_p * Cc)(*[sep[c].ctypes.data for c in range(Cc)])
        lib().ref_separate_bands(ins, outs, n * n, ptr(Sf), P, Cc)
        c0, c1, c2 = sep[0].copy(), sep[lb].copy(), sep[hb].copy()
        for c in (c0, c1, c2):
            c[..., 0] *= att
            c[..., 1] *= att
            fft32(c, True)
        times_conj32(c1, c0)
        times_conj32(c2, c0)
        fft32(c1, False)
        fft32(c2, False)
        min_dist = fit_exclude * otf.cutoff / vec_cpm
        pk = np.zeros(4)
        lib().ref_locate_peak(ptr(c1 if fit_band == 1 else c2), n, n, min_dist, ptr(pk))
        peak = fit_peak32(sep[0], sep[fb], otf, vec_cpm, -pk[0], -pk[1], 0.05, 2.5, bn1=fit_band)
        r = {"peak": list(pk)}
        if lb != hb:
            if fit_band == 1:
                peak[0] *= 2
                peak[1] *= 2
            p1 = get_peak32(sep[0], sep[lb], otf, vec_cpm, peak[0] / 2, peak[1] / 2, 0.05)
            p2 = get_peak32(sep[0], sep[hb], otf, vec_cpm, peak[0], peak[1], 0.05, bn1=2)
            r.update(px=-peak[0], py=-peak[1], pha_off=float(np.arctan2(p1.imag, p1.real)),
                     mod1=abs(p1), mod2=abs(p2))
        else:
            p1 = get_peak32(sep[0], sep[1], otf, vec_cpm, peak[0], peak[1], 0.05)
            r.update(px=-peak[0], py=-peak[1], pha_off=float(np.arctan2(p1.imag, p1.real)), mod1=abs(p1))
        out.append(r)
    return out


def input_ffts32_live(raw, fade_px=10):
    """The live refit's input (ReconstructionRunner.java:614-617, 661-664):
    each 16-bit frame to a complex vector (setFrom16bitPixels), times
    borderDampen -- a float vector of ones faded by fadeBorderCos -- and the
    forward FFT.  raw: uint16 [D, P, N, N]."""
    n = raw.shape[-1]
    damp = np.ones((n, n), dtype=np.float32)
    lib().ref_fade_border(ptr(damp), n, n, fade_px)
    out = np.zeros(raw.shape + (2,), dtype=np.float32)
    for idx in np.ndindex(raw.shape[:-2]):
        out[idx][..., 0] = (raw[idx] & 0xFFFF).astype(np.float32) * damp
        fft32(out[idx], False)
    return out


def refit32(in_fft, nband, otf, vec_cpm, fit_exclude=0.8):
    """ReconstructionRunner.ParameterRefitThread.doRefit
    (livemode/ReconstructionRunner.java:642-746) on FFT'd inputs
    [D, P, N, N, 2]: per direction the band separation with equidistant
    phases, the coarse correlation of band 0 with the highest band (hb),
    locatePeak outside fit_exclude * cutoff, fitPeak(sep0, sep_hb, 0, 1, ...)
    -- OTF band 1 whatever hb is -- and getPeak(sep0, sep1, 0, 1, peak /
    (bands - 1)); returns {px, py (setPxPy arguments), pha_off, mod1, peak}
    (modulation 2 is not refitted)."""
    D, P, n = in_fft.shape[0], in_fft.shape[1], in_fft.shape[2]
    att = np.zeros((n, n), dtype=np.float32)
    lib().ref_write_attenuation(ptr(att), n, n, vec_cpm, .99, 0.15 * otf.cutoff, 0.0, 0.0)
    Cc = 2 * nband - 1
    out = []
    for d in range(D):
        lb, hb = 1, (3 if nband == 3 else 1)
        nb = nband - 1
        phases = np.array([(2 * np.pi * p) / P + 0.0 for p in range(P)])
        Sm = np.zeros(2 * Cc * P)
        assert lib().ref_separation_matrix(ptr(phases), P, nband, None, ptr(Sm)) == 0
        Sf = Sm.astype(np.float32)
        sep = np.zeros((Cc, n, n, 2), dtype=np.float32)
        ins = (C.c_void_p * P)(*[in_fft[d, p].ctypes.data for p in range(P)])
        outs = (C.c_void_p * Cc)(*[sep[c].ctypes.data for c in range(Cc)])
        lib().ref_separate_bands(ins, outs, n * n, ptr(Sf), P, Cc)
        c0, c1 = sep[0].copy(), sep[hb].copy()
        for c in (c0, c1):
            c[..., 0] *= att
            c[..., 1] *= att
            fft32(c, True)
        times_conj32(c1, c0)
        fft32(c1, False)
        pk = np.zeros(4)
        lib().ref_locate_peak(ptr(c1), n, n, fit_exclude * otf.cutoff / vec_cpm, ptr(pk))
        peak = fit_peak32(sep[0], sep[hb], otf, vec_cpm, -pk[0], -pk[1], 0.05, 2.5, bn0=0, bn1=1)
        p1 = get_peak32(sep[0], sep[lb], otf, vec_cpm, peak[0] / nb, peak[1] / nb, 0.05, bn0=0, bn1=1)
        out.append({"peak": list(pk), "px": -peak[0], "py": -peak[1],
                    "pha_off": float(np.arctan2(p1.imag, p1.real)), "mod1": abs(p1)})
    return out
