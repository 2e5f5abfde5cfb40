"""Seeded synthetic 2D-SIM raw stacks (the benchmark workload).

Object: Gaussian beads (sigma 0.5 fine pixels) plus random line filaments on a
fine grid of twice the raw resolution.  Each frame is the object times a
sinusoidal illumination 1 + m cos(2 pi k.r + phi), blurred with the ideal
OTF (sim_algorithm/OtfProvider.java:74-78) of the emission wavelength,
binned 2x2, scaled to a photon peak, plus background, Poisson noise and a
camera offset, stored as uint16.  Frames are angle-major
(livemode/ReconstructionRunner.java:456-461).

The illumination wave vector is returned in the reference's parameter
convention: ``setPxPy(px, py)`` with the y axis flipped
(sim_algorithm/OtfProvider.java:345-346, SimUtils.java:48), so that
``px = kx_index`` and ``py = -ky_index`` for a pattern
cos(2 pi (kx_index x + ky_index y) / n + phi) on the raw pixel grid.
"""
import math

import numpy as np


def ideal_otf(dist):
    d = np.clip(dist, 0, 1)
    v = (2 / np.pi) * (np.arccos(d) - d * np.sqrt(1 - d * d))
    return np.where((dist >= 0) & (dist <= 1), v, 0.0)


def make_object(nfine, rng, nbeads=300, nfil=40, bead_sigma=0.5):
    obj = np.zeros((nfine, nfine))
    # beads: point sources accumulated, then blurred by a Gaussian in Fourier space
    xs = rng.uniform(0, nfine, nbeads)
    ys = rng.uniform(0, nfine, nbeads)
    amp = rng.uniform(0.5, 1.0, nbeads)
    np.add.at(obj, (ys.astype(int) % nfine, xs.astype(int) % nfine), amp)
    # filaments: random line segments drawn with sub-pixel steps
    for _ in range(nfil):
        x0, y0 = rng.uniform(0, nfine, 2)
        ang = rng.uniform(0, np.pi)
        length = rng.uniform(0.1, 0.6) * nfine
        t = np.arange(0, length, 0.5)
        lx = (x0 + t * math.cos(ang)).astype(int) % nfine
        ly = (y0 + t * math.sin(ang)).astype(int) % nfine
        np.add.at(obj, (ly, lx), 0.15 * rng.uniform(0.5, 1.0))
    f = np.fft.fftfreq(nfine)
    fx, fy = np.meshgrid(f, f)
    gauss = np.exp(-2 * (np.pi ** 2) * bead_sigma ** 2 * (fx ** 2 + fy ** 2))
    return np.real(np.fft.ifft2(np.fft.fft2(obj) * gauss))


def _fft2(a):
    """2-D FFT; multithreaded for the large grids (n >= 1024)."""
    if a.shape[0] >= 2048:
        import scipy.fft
        return scipy.fft.fft2(a, workers=-1)
    return np.fft.fft2(a)


def _ifft2(a):
    if a.shape[0] >= 2048:
        import scipy.fft
        return scipy.fft.ifft2(a, workers=-1)
    return np.fft.ifft2(a)


def sim_stack(n=512, ndir=3, nphase=3, na=1.4, lam=525.0, px_um=0.080, k_frac=0.85,
              mod=0.7, angle0_deg=3.0, peak=1000.0, bg=100.0, offset=100.0, seed=0,
              nbeads=300, nfil=40, harmonic2=0.0, timepoints=1):
    """Return (raw uint16 [ndir*nphase, n, n], params dict).

    params: 'px','py' (band-1 shift per direction, reference convention),
    'phases' (per direction, absolute), 'modul', 'k0' (|k| in raw pixels),
    'cutoff_px'.  harmonic2 > 0 adds a 2k component (three-beam data).
    timepoints > 1: raw is [timepoints, ndir*nphase, n, n], the same sample
    imaged repeatedly with independent seeded photon noise (the first
    timepoint equals the single stack of the same seed)."""
    rng = np.random.default_rng(seed)
    nf = 2 * n
    fine_um = px_um / 2
    obj = make_object(nf, rng, nbeads, nfil)
    cutoff = 1000 / (lam / na / 2)            # cycles / micron
    cpm_raw = 1.0 / (n * px_um)               # cycles / micron per raw frequency pixel
    k0 = k_frac * cutoff / cpm_raw            # |k| in raw frequency pixels
    # fine pixel centres in raw pixel units: raw pixel x covers fine 2x, 2x+1
    X = (np.arange(nf) - 0.5) / 2.0
    XX, YY = np.meshgrid(X, X)
    f = np.fft.fftfreq(nf, d=fine_um)
    FX, FY = np.meshgrid(f, f)
    otf = ideal_otf(np.sqrt(FX ** 2 + FY ** 2) / cutoff)
    frames = []
    pxs, pys, phs = [], [], []
    for d in range(ndir):
        ang = math.radians(angle0_deg + d * 180.0 / ndir)
        kx, ky = k0 * math.cos(ang), k0 * math.sin(ang)
        pxs.append(kx)
        pys.append(-ky)
        phases = [2 * math.pi * p / nphase for p in range(nphase)]
        phs.append(phases)
        for p in range(nphase):
            arg = 2 * np.pi * (kx * XX + ky * YY) / n + phases[p]
            illum = 1 + mod * np.cos(arg)
            if harmonic2 > 0:
                illum = illum + harmonic2 * np.cos(2 * arg)
            em = np.real(_ifft2(_fft2(obj * illum) * otf))
            binned = em.reshape(n, 2, n, 2).sum(axis=(1, 3))
            frames.append(binned)
    frames = np.array(frames)
    frames = np.clip(frames, 0, None)
    frames = frames * (peak / frames.max()) + bg
    raws = [_camera(rng, frames, offset) for _ in range(timepoints)]
    raw = raws[0] if timepoints == 1 else np.stack(raws)
    params = dict(px=pxs, py=pys, phases=phs, modul=mod, k0=k0,
                  cutoff_px=cutoff / cpm_raw, na=na, lam=lam, px_um=px_um)
    return raw, params


def _camera(rng, frames, offset):
    """Poisson photon counts plus the camera offset, as uint16."""
    noisy = rng.poisson(frames).astype(np.float64) + offset
    return np.clip(noisy, 0, 65535).astype(np.uint16)


def make_param(sp, n, ndir=3, nphase=3, nband=2, na=1.4, lam=525.0, px_um=0.080, a=1.0,
               att_str=0.995, att_fwhm=1.2, use_att=True, modulation=None):
    """Build an fsim SimParam/OtfProvider pair for stacks from ``sim_stack``
    (the 'parameters given' configuration)."""
    from .sim_algorithm import OtfProvider, SimParam
    otf = OtfProvider.fromEstimate(na, lam, a)
    otf.setAttenuation(att_str, att_fwhm)
    otf.switchAttenuation(use_att)
    param = SimParam.create(nband, ndir, nphase, n, px_um, otf)
    for d in range(ndir):
        dd = param.dir(d)
        dd.setPxPy(sp["px"][d] * (nband - 1), sp["py"][d] * (nband - 1))
        dd.setPhases(sp["phases"][d], True)
        m = sp["modul"] if modulation is None else modulation
        for b in range(1, nband):
            dd.setModulation(b, m)
    return param
