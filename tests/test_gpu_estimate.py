"""GPU parity of the parameter estimation (fsim_common_region, fsim_fit_peak,
fsim_get_peak and the estimateParameters flow) against the FP32 Java-order
oracle.

commonRegion is bit-exact (same per-pixel float arithmetic; the GPU kernel
gathers the ring zeros the reference scatters).  fitPeak / getPeak follow the
FFTs, whose rounding differs from the oracle's (JTransforms rounding is not
reproducible, see test_gpu_recon), so their correlation values are compared to
relative 1e-5; fitPeak's result positions are grid points and compare exactly.
The estimation flow is fed the oracle's FFT'd inputs so that both sides start
from the same bands.
"""
import math

import numpy as np
import pytest

from oracle import ref

pytestmark = pytest.mark.gpu

NA, LAM, A_OTF, PX_UM = 1.4, 525.0, 1.0, 0.08


def _otf(n):
    from fsim.sim_algorithm import OtfProvider
    o = OtfProvider.fromEstimate(NA, LAM, A_OTF)
    o.setPixelSize(1 / (n * PX_UM))
    return o


def _dev(a):
    from fsim.linalg import Vec2dCplx
    v = Vec2dCplx(a.shape[1], a.shape[0])
    v.setData(a)
    return v


def _random_bands(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, n, 2)).astype(np.float32),
            rng.standard_normal((n, n, 2)).astype(np.float32))


@pytest.mark.parametrize("kx,ky", [(20.37, -11.62), (-25.8, 9.13), (0.5, 30.2)])
@pytest.mark.parametrize("divide", [True, False])
def test_common_region_matches_oracle(fsim_lib, kx, ky, divide):
    from fsim.sim_algorithm import Correlation
    n = 128
    otf = _otf(n)
    a0, a1 = _random_bands(n, 7)
    v0, v1 = _dev(a0), _dev(a1)
    Correlation.commonRegion(v0, v1, 0, 1, otf, kx, ky, 0.15, 0.05, divide)
    r0, r1 = a0.copy(), a1.copy()
    ref.common_region32(r0, r1, ref.Otf32(NA, LAM, A_OTF), 1 / (n * PX_UM), kx, ky, 0.15, 0.05, divide)
    assert np.array_equal(v0.vectorData(), r0)
    assert np.array_equal(v1.vectorData(), r1)


def _shifted_pair(n, kx, ky, seed):
    from test_oracle import _shifted_pair as pair
    otf = ref.Otf32(NA, LAM, A_OTF)
    return pair(n, kx, ky, otf.cutoff * n * PX_UM, seed)


@pytest.mark.parametrize("n,kx,ky", [(128, 20.37, -11.62), (256, -41.8, 17.13)])
def test_fit_peak_and_get_peak_match_oracle(fsim_lib, n, kx, ky):
    from fsim.sim_algorithm import Correlation
    otf = _otf(n)
    b0, b1 = _shifted_pair(n, kx, ky, 9)
    v0, v1 = _dev(b0), _dev(b1)
    cntrl = np.zeros((10, 30), np.float32)
    kx0, ky0 = round(kx) + 0.8, round(ky) - 0.6
    got = Correlation.fitPeak(v0, v1, 0, 1, otf, kx0, ky0, 0.05, 2.5, cntrl)
    want, wc = ref.fit_peak32(b0, b1, ref.Otf32(NA, LAM, A_OTF), 1 / (n * PX_UM), kx0, ky0, 0.05, 2.5,
                              cntrl=True)
    print(f"n={n} fitPeak gpu {got} oracle {want}")
    assert got[0] == want[0] and got[1] == want[1]
    assert abs(got[3] - want[3]) <= 1e-5 * want[3]
    assert abs(got[2] - want[2]) <= 1e-5
    assert np.max(np.abs(cntrl - wc)) < 1e-4
    # the bands are unchanged
    assert np.array_equal(v0.vectorData(), b0) and np.array_equal(v1.vectorData(), b1)
    g = Correlation.getPeak(v0, v1, 0, 1, otf, got[0], got[1], 0.05)
    w = ref.get_peak32(b0, b1, ref.Otf32(NA, LAM, A_OTF), 1 / (n * PX_UM), got[0], got[1], 0.05)
    print(f"n={n} getPeak gpu {g} oracle {w}")
    assert abs(g - w) <= 1e-5 * abs(w)


def _estimate_case(nphase, nband, fit_band, exclude, k_frac, harmonic2, seed, n=256):
    from fsim import synthetic
    from fsim.linalg import Vec2dCplx
    from fsim.sim_algorithm import SimParam, estimate_parameters
    raw, sp = synthetic.sim_stack(n=n, nphase=nphase, seed=seed, nbeads=150, nfil=20, harmonic2=harmonic2,
                                  k_frac=k_frac)
    fft_in = ref.input_ffts32(raw.reshape(3, nphase, n, n))
    want = ref.estimate32(fft_in, nband, fit_band, exclude, ref.Otf32(NA, LAM, A_OTF), 1 / (n * PX_UM))
    param = SimParam.create(nband, 3, nphase, n, PX_UM, _otf(n))
    in_dev = [[_dev(np.ascontiguousarray(fft_in[d, p])) for p in range(nphase)] for d in range(3)]
    got = estimate_parameters(param, in_dev, fit_band, exclude)
    return sp, param, got, want


@pytest.mark.parametrize("nphase,nband,fit_band,exclude,k_frac,h2",
                         [(3, 2, 1, 0.6, 0.85, 0.0), (5, 3, 2, 0.6, 0.45, 0.3), (5, 3, 1, 0.3, 0.45, 0.3)])
def test_estimate_parameters_matches_oracle(fsim_lib, nphase, nband, fit_band, exclude, k_frac, h2):
    sp, param, got, want = _estimate_case(nphase, nband, fit_band, exclude, k_frac, h2, seed=3)
    for d in range(3):
        g, w = got[d], want[d]
        print(f"dir {d}: gpu {g} oracle {w}")
        assert g["px"] == w["px"] and g["py"] == w["py"]
        assert abs(g["pha_off"] - w["pha_off"]) < 1e-4
        assert abs(g["mod1"] - w["mod1"]) <= 1e-4 * w["mod1"]
        if nband == 3:
            assert abs(g["mod2"] - w["mod2"]) <= 1e-4 * w["mod2"]
        # SimParam.Dir holds the estimate: setPxPy stores the shift / (bands - 1)
        assert param.dir(d).pX == g["px"] / (nband - 1)
        assert param.dir(d).getPhaOff() == g["pha_off"]
        scale = 1 if nband == 2 else 2
        assert abs(param.dir(d).px(nband - 1) - scale * sp["px"][d]) < 0.1


def _stacks(count, nphase, k_fracs, harmonic2, n=256, seed0=11):
    from fsim import synthetic
    imgs, sps = [], []
    for s in range(count):
        raw, sp = synthetic.sim_stack(n=n, nphase=nphase, seed=seed0 + s, nbeads=150, nfil=20,
                                      harmonic2=harmonic2, k_frac=k_fracs[s % len(k_fracs)])
        imgs.append(raw.reshape(3, nphase, n, n).astype(np.float32))
        sps.append(sp)
    return np.stack(imgs), sps


# The batched estimation runs the same correlation kernels as the per-call
# flow (Correlation.* entry points over device vectors); its input pass fuses
# the fade into the row transform (the same transform code, compiled into
# another kernel, where the compiler may form other fused multiply-adds) and
# |b0|^2 is summed in a different order, so positions agree exactly and
# phases/modulations to FP32 rounding of the input transform (observed
# <= 3.2e-9 rad, 3.2e-9 relative).
@pytest.mark.parametrize("nphase,nband,fit_band,exclude,k_fracs,h2",
                         [(3, 2, 1, 0.6, (0.85, 0.8, 0.75), 0.0), (5, 3, 2, 0.6, (0.45, 0.42), 0.3),
                          (5, 3, 1, 0.3, (0.45,), 0.3), (5, 2, 2, 0.6, (0.85,), 0.0)])
def test_estimate_batch_matches_per_call_flow(fsim_lib, nphase, nband, fit_band, exclude, k_fracs, h2):
    from fsim.sim_algorithm import SimParam, apply_estimate, estimate_batch, estimate_parameters, input_ffts
    n, count = 256, 3
    images, sps = _stacks(count, nphase, k_fracs, h2, n=n)
    param = SimParam.create(nband, 3, nphase, n, PX_UM, _otf(n))
    got = estimate_batch(param, fit_band, exclude, images=images)
    assert len(got) == count
    for s in range(count):
        want = estimate_parameters(param, input_ffts(images[s]), fit_band, exclude)
        for d in range(3):
            g, w = got[s][d], want[d]
            print(f"stack {s} dir {d}: batch {g} per-call {w}")
            assert g["px"] == w["px"] and g["py"] == w["py"]
            assert abs(g["pha_off"] - w["pha_off"]) <= 1e-7
            assert abs(g["mod1"] - w["mod1"]) <= 1e-7 * w["mod1"]
            if nband == 3:
                assert abs(g["mod2"] - w["mod2"]) <= 1e-7 * w["mod2"]
        apply_estimate(param, got[s])
        scale = 1 if nband == 2 else 2
        for d in range(3):
            assert abs(param.dir(d).px(nband - 1) - scale * sps[s]["px"][d]) < 0.1
            assert abs(param.dir(d).py(nband - 1) - scale * sps[s]["py"][d]) < 0.1


@pytest.mark.parametrize("n", [64, 128])
def test_estimate_batch_small_sizes_match_per_call_flow(fsim_lib, n):
    """Sizes below the candidate pass's 256-pixel x tile and 32-row blocks."""
    from fsim.sim_algorithm import SimParam, estimate_batch, estimate_parameters, input_ffts
    images, _ = _stacks(2, 3, (0.85, 0.8), 0.0, n=n, seed0=90)
    param = SimParam.create(2, 3, 3, n, PX_UM, _otf(n))
    got = estimate_batch(param, 1, 0.6, images=images)
    for s in range(2):
        want = estimate_parameters(param, input_ffts(images[s]), 1, 0.6)
        for d in range(3):
            g, w = got[s][d], want[d]
            print(f"n={n} stack {s} dir {d}: batch {g} per-call {w}")
            assert g["px"] == w["px"] and g["py"] == w["py"]
            assert abs(g["pha_off"] - w["pha_off"]) <= 1e-6
            assert abs(g["mod1"] - w["mod1"]) <= 1e-6 * w["mod1"]


def test_estimate_batch_alone_equals_in_batch(fsim_lib):
    """Each stack's estimate is independent of the batch it runs in (bitwise)."""
    from fsim.sim_algorithm import SimParam, estimate_batch
    n = 256
    images, _ = _stacks(4, 3, (0.85, 0.7), 0.0, n=n, seed0=40)
    param = SimParam.create(2, 3, 3, n, PX_UM, _otf(n))
    mixed = estimate_batch(param, 1, 0.6, images=images)
    alone = estimate_batch(param, 1, 0.6, images=images[2:3])
    assert alone[0] == mixed[2]


def test_estimate_batch_matches_oracle(fsim_lib):
    from fsim.sim_algorithm import SimParam, estimate_batch
    n = 256
    images, _ = _stacks(2, 3, (0.85, 0.78), 0.0, n=n, seed0=70)
    param = SimParam.create(2, 3, 3, n, PX_UM, _otf(n))
    got = estimate_batch(param, 1, 0.6, images=images)
    for s in range(2):
        want = ref.estimate32(ref.input_ffts32(images[s]), 2, 1, 0.6, ref.Otf32(NA, LAM, A_OTF), 1 / (n * PX_UM))
        for d in range(3):
            g, w = got[s][d], want[d]
            print(f"stack {s} dir {d}: gpu {g} oracle {w}")
            assert g["peak"] == (w["peak"][0], w["peak"][1])   # coarse peak pixel: bit-exact
            assert g["px"] == w["px"] and g["py"] == w["py"]
            assert abs(g["pha_off"] - w["pha_off"]) < 1e-4
            assert abs(g["mod1"] - w["mod1"]) <= 1e-4 * w["mod1"]


def test_estimate_batch_large_size_known_answer(fsim_lib):
    """1024^2 stacks (row transforms of two wavefronts in the estimation's
    FFTs): the fitted shift is within a tenth of a pixel of the pattern's."""
    from fsim import synthetic
    from fsim.sim_algorithm import SimParam, estimate_batch
    n = 1024
    raw, sp = synthetic.sim_stack(n=n, seed=21, nbeads=600, nfil=80)
    images = raw.reshape(1, 3, 3, n, n).astype(np.float32)
    param = SimParam.create(2, 3, 3, n, PX_UM, _otf(n))
    got = estimate_batch(param, 1, 0.6, images=images)
    for d in range(3):
        g = got[0][d]
        print(f"n={n} dir {d}: {g} true ({sp['px'][d]:.3f}, {sp['py'][d]:.3f})")
        assert abs(g["px"] - sp["px"][d]) < 0.1 and abs(g["py"] - sp["py"][d]) < 0.1


def _live_otf(kind, n):
    """fromEstimate (one band) or a measured-style three-band OTF (bands of
    different curvature, so OTF band 1 and band 2 weights differ), with the
    oracle's copy of the same tables."""
    from fsim.sim_algorithm import OtfProvider
    if kind == "estimate":
        o = OtfProvider.fromEstimate(NA, LAM, A_OTF)
        o.setPixelSize(1 / (n * PX_UM))
        return o, ref.Otf32(NA, LAM, A_OTF)
    ests = [OtfProvider.fromEstimate(NA, LAM, a) for a in (1.0, 0.6, 0.3)]
    vals = np.stack([e.band_values()[0] for e in ests])
    o = OtfProvider.fromData(NA, LAM, ests[0].cpm, vals)
    o.setPixelSize(1 / (n * PX_UM))
    return o, ref.OtfBands32(vals, o.getCutoff(), o.cpm)


@pytest.mark.parametrize("nphase,nband,k_frac,h2,kind", [(3, 2, 0.85, 0.0, "estimate"),
                                                         (5, 3, 0.45, 0.3, "three_band"),
                                                         (5, 3, 0.45, 0.3, "estimate")])
def test_live_refit_matches_oracle(fsim_lib, nphase, nband, k_frac, h2, kind):
    """The live refit (fsim_estimate with live = 1, ReconstructionRunner.
    doRefit) on 16-bit stacks against the oracle's restatement: coarse peaks
    and fitted shifts identical, phase and modulation 1 to 1e-4, modulation 2
    not estimated; the per-call flow over device vectors (refit_parameters)
    gives the same shifts."""
    from fsim import synthetic
    from fsim.sim_algorithm import SimParam, estimate_batch, refit_parameters
    n = 256
    raws, sps = [], []
    for s in range(2):
        raw, sp = synthetic.sim_stack(n=n, nphase=nphase, seed=150 + s, nbeads=150, nfil=20, harmonic2=h2,
                                      k_frac=k_frac)
        raws.append(raw)
        sps.append(sp)
    otf, otf32 = _live_otf(kind, n)
    param = SimParam.create(nband, 3, nphase, n, PX_UM, otf)
    images = np.stack(raws).reshape(2, 3, nphase, n, n).astype(np.float32)
    got = estimate_batch(param, 1, 0.8, images=images, live=True)
    scale = nband - 1
    for s in range(2):
        want = ref.refit32(ref.input_ffts32_live(raws[s].reshape(3, nphase, n, n)), nband, otf32, 1 / (n * PX_UM))
        for d in range(3):
            g, w = got[s][d], want[d]
            print(f"{kind} stack {s} dir {d}: gpu {g} oracle {w}")
            assert g["peak"] == (w["peak"][0], w["peak"][1])
            assert g["px"] == w["px"] and g["py"] == w["py"]
            assert abs(g["pha_off"] - w["pha_off"]) < 1e-4
            assert abs(g["mod1"] - w["mod1"]) <= 1e-4 * w["mod1"]
            assert "mod2" not in g
            assert abs(g["px"] - scale * sps[s]["px"][d]) < 0.1 and abs(g["py"] - scale * sps[s]["py"][d]) < 0.1
    # the per-call flow through the vector API: the same shifts; modulation
    # 2 of the parameters is left as it was
    p2 = SimParam.create(nband, 3, nphase, n, PX_UM, otf)
    if nband == 3:
        for d in range(3):
            p2.dir(d).setModulation(2, 0.55)
    per_call = refit_parameters(p2, raws[0])
    for d in range(3):
        assert per_call[d]["px"] == got[0][d]["px"] and per_call[d]["py"] == got[0][d]["py"]
        assert abs(per_call[d]["pha_off"] - got[0][d]["pha_off"]) < 1e-6
        if nband == 3:
            assert p2.dir(d).getModulations()[2] == 0.55


def test_live_refit_keeps_modulation_two(fsim_lib):
    """fsim_params_from_estimate keeps the base modulation 2 where the live
    refit did not estimate it (NaN), and takes the estimate's otherwise."""
    from fsim import synthetic
    from fsim.sim_algorithm import SimParam, estimate_batch, params_from_estimates
    n = 128
    raw, sp = synthetic.sim_stack(n=n, nphase=5, seed=160, nbeads=60, nfil=8, harmonic2=0.3, k_frac=0.45)
    otf, _ = _live_otf("estimate", n)
    param = SimParam.create(3, 3, 5, n, PX_UM, otf)
    for d in range(3):
        param.dir(d).setModulation(2, 0.42)
    imgs = raw.reshape(1, 3, 5, n, n).astype(np.float32)
    res_live = estimate_batch(param, 1, 0.8, images=imgs, raw=True, live=True)
    res_est = estimate_batch(param, 2, 0.6, images=imgs, raw=True)
    live = params_from_estimates(param, res_live, 1)[0]
    est = params_from_estimates(param, res_est, 1)[0]
    for d in range(3):
        assert np.isnan(res_live[d].mod2) and not np.isnan(res_est[d].mod2)
        assert live.modul[d][2] == 0.42 and est.modul[d][2] == res_est[d].mod2
        assert live.modul[d][1] == res_live[d].mod1 and live.px[d] == res_live[d].px / 2


def _region_otf(kind, n):
    """OTFs for the commonRegion run tables: the estimated one (one pass run,
    then fail); a measured three-band table whose magnitude crosses the 0.05
    limit several times and holds samples at the limit and one float step on
    either side (undecided runs), with imaginary parts; and one that crosses
    it every other sample (more runs than the tables hold: every pixel
    evaluated)."""
    from fsim.sim_algorithm import OtfProvider
    est = OtfProvider.fromEstimate(NA, LAM, A_OTF)
    if kind == "estimate":
        o = est
    else:
        base = est.band_values()[0]
        ns = base.shape[0]
        lim = np.float32(0.05)
        bands = []
        for b in range(3):
            v = base.astype(np.complex64).copy()
            if kind == "crossings":
                k = np.arange(ns // 5, ns // 5 + 9)
                v[k] = [0.04, 0.06, lim, np.nextafter(lim, np.float32(1)), np.nextafter(lim, np.float32(0)),
                        0.2, 0.03, 0.05 * np.exp(0.3j), 0.07]
                v[ns // 2 + b] = 0.049999
                v[ns // 2 + 3 + b] = 0.5
            else:   # alternating
                k = np.arange(ns // 6, ns // 6 + 40)
                v[k] = np.where(k % 2 == 0, 0.03, 0.3)
            v = v * np.exp(1j * 0.2 * b * np.arange(ns) / ns)
            bands.append(v.astype(np.complex64))
        o = OtfProvider.fromData(NA, LAM, est.cpm, np.stack(bands))
    o.setPixelSize(1 / (n * PX_UM))
    return o


@pytest.mark.parametrize("kind", ["estimate", "crossings", "alternating"])
def test_region_runs_equal_per_pixel_weight(fsim_lib, kind):
    """fsim_estimate's commonRegion pass classifies pixels by squared distance
    through per-table run starts (pass / fail / undecided); with the test
    hook it forms the shifted weight at every pixel instead.  The estimates
    are bitwise the same for a monotone OTF, for tables with runs of every
    class (samples at the limit, one float step either side, complex
    values, several crossings) and for a table with too many runs."""
    from fsim._lib import lib
    from fsim.sim_algorithm import SimParam, estimate_batch
    n = 256
    images, _ = _stacks(2, 5, (0.45, 0.42), 0.3, n=n, seed0=23)
    param = SimParam.create(3, 3, 5, n, PX_UM, _region_otf(kind, n))
    L = lib()
    outs = []
    for exact in (0, 1):
        L.fsim_debug_region_exact(exact)
        try:
            res = estimate_batch(param, 2, 0.6, images=images, raw=True)
            outs.append(bytes(res))
        finally:
            L.fsim_debug_region_exact(0)
    assert outs[0] == outs[1]
