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
    (fsim.synthetic.sim_stack); setPxPy divides by nband-1 (SimParam.java:536-541)."""
    ndir = len(sp["px"])
    nphase = len(sp["phases"][0])
    cutoff = 1000 / (lam / na / 2)
    m = sp["modul"] if modulation is None else modulation
    return dict(
        n=n, ndir=ndir, nphase=nphase, nband=nband,
        px=[sp["px"][d] * (nband - 1) / (nband - 1) for d in range(ndir)],
        py=[sp["py"][d] * (nband - 1) / (nband - 1) for d in range(ndir)],
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
