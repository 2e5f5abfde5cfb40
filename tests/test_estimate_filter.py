"""The candidate filter of the batched estimation (fsim_estimate), checked on
the CPU against the oracle's reference-arithmetic correlation.

fsim_estimate evaluates fitPeak's 100 candidates approximately (accurate
phasors, double sums), bounds the distance to the value the reference
computes (Correlation.fitPeak -> Vec2d.Cplx.fourierShift with the phase
rounded to float and MTool's sine polynomial, BasicVectors.java:309-339,
MTool.java:34-59; float products; double sum) by

    D = 1.5 * sum|q| * (2^-24 * 2 pi (|kx| + |ky| + 2 search) + 2e-6),

and evaluates exactly only the candidates with |A| >= max|A| - 2 D.  Here
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
        err = np.abs(J - A).max()
        print(f"search {search}: max|J-A| = {err:.3e}, bound {D:.3e} ({err / D:.2f}), "
              f"max|J| = {np.abs(J).max():.3e}")
        assert err <= D
        # the selected set holds the reference's first maximum (y-outer, x-inner)
        thr = np.abs(A).max() - 2 * 1.5 * D
        sel = np.nonzero(np.abs(A) >= thr)[0]
        best = int(np.argmax(np.abs(J)))
        assert best in sel
        print(f"  selected {len(sel)} of 100")
        assert len(sel) <= 25
