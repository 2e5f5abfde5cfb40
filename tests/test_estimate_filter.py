"""The candidate filter of the batched estimation (fsim_estimate), checked on
the CPU against the oracle's reference-arithmetic correlation.

fsim_estimate evaluates fitPeak's 100 candidates approximately (accurate
phasors, double sums), bounds the distance to the value the reference
computes (Correlation.fitPeak -> Vec2d.Cplx.fourierShift with the phase
rounded to float and MTool's sine polynomial, BasicVectors.java:309-339,
MTool.java:34-59; float products; double sum) by

    D = 1.5 * (2^-24 * 2 pi sum_xy |q| (ax x + ay y) / n + 2e-6 sum|q|),
    ax = |kx| + search, ay = |ky| + search (the phase of pixel (x, y) is at
    most 2 pi (ax x + ay y) / n over the grid; the older, looser form used
    2 pi (|kx| + |ky| + 2 search) for every pixel),

and evaluates exactly only the candidates with |A| >= max|A| - 2 D, where
its own A carries a float part (per-thread sums of q * TY_j(y) in single
precision) that adds at most 64 * 2^-24 * sum|q| to D.  Here
the unscaled bound (without the 1.5) is checked against every candidate of
three fitPeak-like grids on a commonRegion'd, inverse-transformed band pair
with a known shift, and the selected set must contain the reference's
argmax; the bound (a worst case: the per-pixel errors mostly cancel) must
still be tight enough that few candidates need the exact evaluation.  A
wrong sign, scale or axis in the approximation fails the bound by orders of
magnitude.
"""
import math

import numpy as np
import pytest

from oracle import ref

NA, LAM, A_OTF, PX_UM = 1.4, 525.0, 1.0, 0.08
APX_ARITH = 64 * 2.0 ** -24   # kApxArith: the float part of the approximate sums


def _pair(n, kx, ky, seed):
    from test_oracle import _shifted_pair
    otf = ref.Otf32(NA, LAM, A_OTF)
    b0, b1 = _shifted_pair(n, kx, ky, otf.cutoff * n * PX_UM, seed)
    ref.common_region32(b0, b1, otf, 1 / (n * PX_UM), kx, ky, 0.15, 0.05, True)
    ref.fft32(b0, True)
    ref.fft32(b1, True)
    return b0, b1


@pytest.mark.parametrize("n,kx,ky,seed", [(128, 40.37, -21.62, 3), (128, -47.3, 8.9, 5)])
def test_candidate_filter_bound(n, kx, ky, seed):
    b0, b1 = _pair(n, kx, ky, seed)
    z0 = b0[..., 0].astype(np.float64) + 1j * b0[..., 1]
    z1 = b1[..., 0].astype(np.float64) + 1j * b1[..., 1]
    q = z1 * np.conj(z0)
    sum_abs = np.abs(q).sum()
    x = np.arange(n)
    k0x, k0y = round(kx) + 0.7, round(ky) - 0.4
    for search in (2.5, 0.5, 0.1):
        pos = [(k0x + ((xi - 4.5) / 4.5) * search, k0y + ((yi - 4.5) / 4.5) * search)
               for yi in range(10) for xi in range(10)]
        J = np.array([ref.shifted_corr32(b0, b1, xp, -yp) for xp, yp in pos])
        A = np.array([(np.exp(2j * np.pi * xp * x / n)[None, :] * np.exp(-2j * np.pi * yp * x / n)[:, None] * q).sum()
                      for xp, yp in pos])
        D = sum_abs * (2.0 ** -24 * 2 * math.pi * (abs(k0x) + abs(k0y) + 2 * search) + 2e-6)
        # the kernel's form: each pixel's largest phase over the grid,
        # 2 pi (ax x + ay y) / n, instead of the largest over the image
        ax, ay = abs(k0x) + search, abs(k0y) + search
        wq = (np.abs(q) * (ax * x[None, :] + ay * x[:, None])).sum()
        Dp = 2.0 ** -24 * 2 * math.pi * wq / n + 2e-6 * sum_abs
        err = np.abs(J - A).max()
        print(f"search {search}: max|J-A| = {err:.3e}, bound {D:.3e} ({err / D:.2f}), "
              f"per-pixel bound {Dp:.3e} ({err / Dp:.2f}), max|J| = {np.abs(J).max():.3e}")
        assert err <= Dp <= D
        D = Dp
        # the selected set holds the reference's first maximum (y-outer, x-inner)
        thr = np.abs(A).max() - 2 * 1.5 * (D + APX_ARITH * sum_abs)
        sel = np.nonzero(np.abs(A) >= thr)[0]
        best = int(np.argmax(np.abs(J)))
        assert best in sel
        print(f"  selected {len(sel)} of 100")
        assert len(sel) <= 25


def _wave_tree(s, lpw):
    """The kernel's cross-lane float sums within a wavefront: the threads
    t' of a column sit in groups of lpw lanes (lane = t' CW + c), and each
    lane adds its partner's value at lane distance CW, 2 CW, ... (xor), in
    float32; returns the group sums, (T / lpw, n)."""
    w = s.reshape(s.shape[0] // lpw, lpw, -1)
    o = 1
    while o < lpw:
        w = (w + w[:, np.arange(lpw) ^ o]).astype(np.float32)
        o <<= 1
    return w[:, 0]


def _float_part_sums(b0, b1, n, xpos0, ypos, step, T, lpw=4):
    """The approximate sums as the column kernel forms them: per column x,
    thread t adds q(x, y) * TY_j(y) for y = t, t + T, ... in float32
    (q = b1 conj(b0) in float32, TY_j rounded from double; each product
    term added with one rounding, as the kernel's fused multiply-adds), the
    sums of the lpw threads of a column in one wavefront are added in
    float32 (xor tree), the wavefronts' sums in double, then
    sum_x TX_i(x) C_j(x) in double."""
    f = np.float32
    ux, uy = b0[..., 0], b0[..., 1]
    vx, vy = b1[..., 0], b1[..., 1]
    qr = (vx * ux).astype(f) + (vy * uy).astype(f)
    qi = (vy * ux).astype(f) - (vx * uy).astype(f)
    y = np.arange(n)
    A = np.zeros(100, complex)
    C = np.zeros((10, n), complex)
    for j in range(10):
        u = ypos[j] * y / n
        u -= np.floor(u)
        ex = np.cos(2 * np.pi * u).astype(f)[:, None]
        ey = (-np.sin(2 * np.pi * u)).astype(f)[:, None]
        d = np.float64
        a1, a2 = qr.astype(d) * ex, -(qi.astype(d) * ey)   # exact products
        b1, b2 = qr.astype(d) * ey, qi.astype(d) * ex
        # rows y = t + T m of a column: thread t's float sum over m, in order,
        # each product added exactly and the sum rounded (fmaf)
        sr = np.zeros((T, n), f)
        si = np.zeros((T, n), f)
        for m in range(n // T):
            k = slice(m * T, (m + 1) * T)
            sr = (sr.astype(d) + a1[k]).astype(f)
            sr = (sr.astype(d) + a2[k]).astype(f)
            si = (si.astype(d) + b1[k]).astype(f)
            si = (si.astype(d) + b2[k]).astype(f)
        C[j] = _wave_tree(sr, lpw).astype(np.float64).sum(0) + 1j * _wave_tree(si, lpw).astype(np.float64).sum(0)
    x = np.arange(n)
    for j in range(10):
        for i in range(10):
            A[j * 10 + i] = (np.exp(2j * np.pi * (xpos0 + i * step) * x / n) * C[j]).sum()
    return A


@pytest.mark.parametrize("n,kx,ky,seed", [(128, 40.37, -21.62, 3), (128, -47.3, 8.9, 5)])
def test_candidate_filter_float_part(n, kx, ky, seed):
    """The float part of the kernel's approximate sums stays within
    64 * 2^-24 * sum|q| of the exact candidate sums (T = 16 threads per
    column at n = 128, eight terms each, 16-column tiles: four threads of a
    column per wavefront, summed in float); a wrong phasor sign, axis or
    table row misses by orders of magnitude."""
    b0, b1 = _pair(n, kx, ky, seed)
    z0 = b0[..., 0].astype(np.float64) + 1j * b0[..., 1]
    z1 = b1[..., 0].astype(np.float64) + 1j * b1[..., 1]
    q = z1 * np.conj(z0)
    sum_abs = np.abs(q).sum()
    x = np.arange(n)
    k0x, k0y = round(kx) + 0.7, round(ky) - 0.4
    for search in (2.5, 0.1):
        pos = [(k0x + ((xi - 4.5) / 4.5) * search, k0y + ((yi - 4.5) / 4.5) * search)
               for yi in range(10) for xi in range(10)]
        A = np.array([(np.exp(2j * np.pi * xp * x / n)[None, :] * np.exp(-2j * np.pi * yp * x / n)[:, None] * q).sum()
                      for xp, yp in pos])
        ypos = [k0y + ((yi - 4.5) / 4.5) * search for yi in range(10)]
        Af = _float_part_sums(b0, b1, n, k0x + (-4.5 / 4.5) * search, ypos, search / 4.5, T=16)
        err = np.abs(Af - A).max()
        print(f"search {search}: max|A_float - A| = {err:.3e} = {err / sum_abs / 2.0 ** -24:.2f} u sum|q|")
        assert err <= APX_ARITH * sum_abs


def test_region_sample_class_margin():
    """fsim_estimate's commonRegion run tables decide the shifted-weight test
    |w| >= 0.05 per OTF sample v without the interpolation fraction f: pass
    when |v| >= 0.05 (1 + 8u), fail when |v| < 0.05 (1 - 8u), u = 2^-24,
    else the pixel is evaluated.  Float emulation of the reference's
    getOtfVal weighting (v (1 - f) + v f, each operation rounded) and of
    Cplx.Float.abs over 2e6 samples within 12 u of the limit: no decided
    sample is contradicted by any f, while a margin of one u would be
    (the argument gives 3.6 u)."""
    f32 = np.float32
    rng = np.random.default_rng(1)
    N = 2_000_000
    u = 2.0 ** -24
    mag = 0.05 * (1 + rng.uniform(-12, 12, N) * u)
    ang = rng.uniform(0, 2 * np.pi, N)
    ang[: N // 4] = 0.0
    vx = (mag * np.cos(ang)).astype(f32)
    vy = (mag * np.sin(ang)).astype(f32)
    f = rng.uniform(0, 1, N).astype(f32)
    f[:1000], f[1000:2000] = 0, 1
    omf = (f32(1) - f).astype(f32)
    wx = ((vx * omf).astype(f32) + (vx * f).astype(f32)).astype(f32)
    wy = ((vy * omf).astype(f32) + (vy * f).astype(f32)).astype(f32)
    ab = np.sqrt(wx.astype(np.float64) ** 2 + wy.astype(np.float64) ** 2).astype(f32).astype(np.float64)
    m = np.sqrt(vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2)

    def contradicted(margin):
        P, F = m >= 0.05 * (1 + margin * u), m < 0.05 * (1 - margin * u)
        return int((P & (ab < 0.05)).sum() + (F & (ab >= 0.05)).sum())

    assert contradicted(8) == 0
    assert contradicted(1) > 0
