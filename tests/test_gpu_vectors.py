"""GPU parity of the vector-API operations against the oracle.

Integer maps and tables are compared bit for bit (tables allow 1 ulp where
the value comes from a libm transcendental: acos/pow/exp on host vs device);
floating-point element-wise ops with the reference's operation order must
match exactly; FFTs are compared to FP64 numpy within FP32 round-off.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import ref, fp64

pytestmark = pytest.mark.gpu


def cvec(vf, z):
    v = vf.createCplx2D(z.shape[1], z.shape[0])
    v.from_numpy(z)
    return v


def rnd_c(rng, h, w):
    return (rng.standard_normal((h, w)) + 1j * rng.standard_normal((h, w))).astype(np.complex64)


@pytest.fixture(scope="module")
def vf(fsim_lib):
    from fsim.linalg import AccelVectorFactory
    return AccelVectorFactory.getFactory()


@pytest.mark.parametrize("w,h", [(8, 8), (64, 64), (128, 32), (256, 256), (512, 512), (1024, 1024),
                                 (2048, 16), (16, 4096), (4096, 64)])
def test_fft2d_matches_fp64(vf, w, h):
    rng = np.random.default_rng(w * 7 + h)
    z = rnd_c(rng, h, w)
    v = cvec(vf, z)
    v.fft2d(False)
    f = v.to_numpy()
    ref_f = np.fft.fft2(z.astype(np.complex128))
    assert fp64.rel_l2(f, ref_f) < 3e-6
    v.fft2d(True)
    back = v.to_numpy()
    assert fp64.rel_l2(back, z) < 3e-6


def test_fft_closed_forms(vf):
    n = 512
    z = np.zeros((n, n), np.complex64)
    z[0, 0] = 1
    v = cvec(vf, z)
    v.fft2d(False)
    assert np.allclose(v.to_numpy(), 1.0, atol=1e-6)       # delta -> constant
    v.fft2d(True)
    assert np.allclose(v.to_numpy(), z, atol=1e-6)
    y, x = np.mgrid[0:n, 0:n]
    c = np.cos(2 * np.pi * (5 * x + 9 * y) / n).astype(np.complex64)
    v = cvec(vf, c)
    v.fft2d(False)
    f = np.abs(v.to_numpy())
    peaks = sorted(zip(*np.nonzero(f > 1000)))
    assert peaks == [(9, 5), (n - 9, n - 5)]                # two bins, forward sign
    assert np.allclose(f[9, 5], n * n / 2, rtol=1e-5)


def test_fourier_shift_matches_oracle(vf):
    rng = np.random.default_rng(5)
    for n, kx, ky in [(64, 0.37, -0.81), (512, 0.25, 0.5), (128, 137.44, -140.91), (1024, -185.69, 9.73)]:
        z = rnd_c(rng, n, n)
        v = cvec(vf, z)
        v.fourierShift(kx, ky)
        a = ref.cplx_f32(z)
        ref.lib().ref_fourier_shift(ref.ptr(a), n, kx, ky)
        got = v.vectorData()
        # same float phase and same double polynomial; only the final
        # float complex product may differ by FMA-free rounding (exact here)
        assert np.max(np.abs(got - a)) <= 2e-6 * np.max(np.abs(a))


def test_fourier_shift_integer_is_roll(vf):
    n = 256
    rng = np.random.default_rng(6)
    z = rnd_c(rng, n, n)
    spec = np.fft.fft2(z)
    v = cvec(vf, np.fft.ifft2(spec).astype(np.complex64))
    v.fourierShift(3, -7)
    v.fft2d(False)
    rolled = np.roll(np.roll(spec, 3, axis=1), -7, axis=0)
    assert fp64.rel_l2(v.to_numpy(), rolled) < 1e-5


@pytest.mark.parametrize("wi,wo,xo,yo", [(64, 128, 0, 0), (64, 128, 23, -17), (512, 1024, 185, -186),
                                          (512, 1024, -186, 185), (32, 96, 40, 5), (16, 32, -32, 31)])
def test_paste_freq_bit_exact(vf, wi, wo, xo, yo):
    rng = np.random.default_rng(abs(wi + xo))
    z = rnd_c(rng, wi, wi)
    vi = cvec(vf, z)
    vo = vf.createCplx2D(wo, wo)
    vo.from_numpy(rnd_c(rng, wo, wo))   # must be overwritten (pasteFreq zeroes)
    vo.pasteFreq(vi, xo, yo)
    a = ref.cplx_f32(z)
    out = np.zeros((wo, wo, 2), np.float32)
    ref.lib().ref_paste_freq(ref.ptr(out), wo, wo, ref.ptr(a), wi, wi, xo, yo)
    assert np.array_equal(vo.vectorData(), out)


def test_elementwise_ops_bit_exact(vf):
    rng = np.random.default_rng(7)
    h, w = 96, 160
    a, b = rnd_c(rng, h, w), rnd_c(rng, h, w)
    r = rng.standard_normal((h, w)).astype(np.float32)
    A = ref.cplx_f32(a).astype(np.float32)
    B = ref.cplx_f32(b).astype(np.float32)
    va, vb = cvec(vf, a), cvec(vf, b)
    vr = vf.createReal2D(w, h)
    vr.setData(r)
    # axpy with a complex scalar (AbstractVectorCplx.java:185-195)
    s = np.float32(0.3), np.float32(-1.7)
    va.axpy(complex(s[0], s[1]), vb)
    exp = A.copy()
    exp[..., 0] += (s[0] * B[..., 0]) - (s[1] * B[..., 1])
    exp[..., 1] += (s[1] * B[..., 0]) + (s[0] * B[..., 1])
    assert np.array_equal(va.vectorData(), exp)
    # timesConj (AbstractVectorCplx.java:281-295)
    va.timesConj(vb)
    y = exp.copy()
    xr, xi = B[..., 0], -B[..., 1]
    exp[..., 0] = (xr * y[..., 0]) - (xi * y[..., 1])
    exp[..., 1] = (xi * y[..., 0]) + (xr * y[..., 1])
    assert np.array_equal(va.vectorData(), exp)
    # times real, add, scal
    va.times(vr)
    exp[..., 0] *= r
    exp[..., 1] *= r
    assert np.array_equal(va.vectorData(), exp)
    va.add(vb)
    exp += B
    assert np.array_equal(va.vectorData(), exp)
    # norm2 / sumElements: float squares, double sums
    n2 = va.norm2()
    flat = exp.ravel()
    assert abs(n2 - float(np.sum((flat * flat).astype(np.float64)))) <= 1e-12 * n2
    sre = va.sumElements()
    assert abs(sre.real - float(np.sum(exp[..., 0].astype(np.float64)))) < 1e-9 * np.abs(exp).sum()
    # copy real <- complex, complex <- real
    vr2 = vf.createReal2D(w, h)
    vr2.copy(va)
    assert np.array_equal(vr2.vectorData(), exp[..., 0])
    vc = vf.createCplx2D(w, h)
    vc.copy(vr)
    assert np.array_equal(vc.vectorData()[..., 0], r) and not vc.vectorData()[..., 1].any()


def test_set_u16_and_size_checks(vf):
    from fsim._lib import FsimError
    px = np.arange(64 * 48, dtype=np.uint32).astype(np.uint16).reshape(48, 64) * 13
    v = vf.createCplx2D(64, 48)
    v.setFrom16bitPixels(px)
    d = v.vectorData()
    assert np.array_equal(d[..., 0], px.astype(np.float32)) and not d[..., 1].any()
    with pytest.raises(FsimError):
        v.setFrom16bitPixels(px[:10])
    with pytest.raises(FsimError):
        v.add(vf.createCplx2D(32, 48))
    with pytest.raises(FsimError):
        vf.createCplx2D(64, 48).fourierShift(1, 1)   # not square


def test_separate_bands_bit_exact(vf):
    from fsim.sim_algorithm import separate_bands, separation_matrix
    rng = np.random.default_rng(8)
    n = 64
    for P, B in [(3, 2), (5, 2), (5, 3)]:
        Cc = 2 * B - 1
        phases = [2 * np.pi * p / P + 0.3 for p in range(P)]
        fac = [1.0] + [0.8] * (B - 1)
        ins = [rnd_c(rng, n, n) for _ in range(P)]
        vin = [cvec(vf, z) for z in ins]
        vout = [vf.createCplx2D(n, n) for _ in range(Cc)]
        separate_bands(vin, vout, phases, B, fac)
        # oracle: same matrix restatement + same axpy order
        S = np.zeros(2 * Cc * P)
        ph = np.array(phases)
        ref.lib().ref_separation_matrix(ref.ptr(ph), P, B, ref.ptr(np.array(fac)), ref.ptr(S))
        Sn = separation_matrix(phases, B, fac)
        assert np.array_equal(S.reshape(Cc, P, 2)[..., 0], Sn.real)
        assert np.array_equal(S.reshape(Cc, P, 2)[..., 1], Sn.imag)
        Sf = S.astype(np.float32)
        ia = [ref.cplx_f32(z) for z in ins]
        oa = [np.zeros((n, n, 2), np.float32) for _ in range(Cc)]
        ip = (C.c_void_p * P)(*[x.ctypes.data for x in ia])
        op = (C.c_void_p * Cc)(*[x.ctypes.data for x in oa])
        ref.lib().ref_separate_bands(ip, op, n * n, ref.ptr(Sf), P, Cc)
        for c in range(Cc):
            assert np.array_equal(vout[c].vectorData(), oa[c])


def _otf(att=True):
    from fsim.sim_algorithm import OtfProvider
    o = OtfProvider.fromEstimate(1.4, 525.0, 1.0)
    o.setAttenuation(0.995, 1.2)
    o.switchAttenuation(att)
    return o


def test_otf_tables_match_oracle(vf):
    o = _otf()
    vals, cut, cpm = ref.otf_estimate(1.4, 525.0, 1.0)
    assert np.array_equal(o.vals.reshape(-1, 2), vals)
    assert cut == o.cutoff and cpm == o.cpm
    n = 128
    vec_cpm = 1.0 / (n * 0.08)
    o.setPixelSize(vec_cpm)
    vatt = np.zeros((512, 2), np.float32)
    oatt = np.zeros(512, np.float32)
    ref.lib().ref_otf_set_attenuation(ref.ptr(vals), 512, cpm, 0.995, 1.2, ref.ptr(vatt), ref.ptr(oatt))
    assert np.array_equal(o.vals_att.reshape(-1, 2), vatt)
    assert np.array_equal(o.only_att, oatt)
    for kx, ky, write in [(0, 0, 1), (12.3, -40.7, 1), (-3.5, 7.25, 0)]:
        v = vf.createCplx2D(n, n)
        rng = np.random.default_rng(9)
        z = rnd_c(rng, n, n)
        v.from_numpy(z)
        o.otfToVector(v, 0, kx, ky, True, write)
        exp = ref.cplx_f32(z)
        ref.lib().ref_otf_to_vector(ref.ptr(exp), n, n, ref.ptr(vatt), 512, cpm, cut, vec_cpm, kx, ky, write)
        assert np.array_equal(v.vectorData(), exp)


def test_filters_match_oracle(vf):
    from fsim.sim_algorithm import SimParam, fade_border_cos, wiener_denominator
    o = _otf()
    n = 128
    param = SimParam.create(2, 3, 3, n, 0.08, o)
    k0 = 0.85 * o.cutoff / param.pxlSizeCyclesMicron()
    for d in range(3):
        a = np.radians(3 + 60 * d)
        param.dir(d).setPxPy(k0 * np.cos(a), -k0 * np.sin(a))
    vec_cpm = param.pxlSizeCyclesMicron()
    m = 2 * n
    wd = vf.createReal2D(m, m)
    wiener_denominator(wd, param, 0.05)
    vals, cut, cpm = ref.otf_estimate(1.4, 525.0, 1.0)
    vatt = np.zeros((512, 2), np.float32)
    oatt = np.zeros(512, np.float32)
    ref.lib().ref_otf_set_attenuation(ref.ptr(vals), 512, cpm, 0.995, 1.2, ref.ptr(vatt), ref.ptr(oatt))
    exp = np.zeros((m, m), np.float32)
    px = np.array([param.dir(d).pX for d in range(3)])
    py = np.array([param.dir(d).pY for d in range(3)])
    ref.lib().ref_wiener_denominator(ref.ptr(exp), m, m, ref.ptr(vals), ref.ptr(oatt), 512, cpm, cut, vec_cpm,
                                     3, 2, ref.ptr(px), ref.ptr(py), 1, 0.05)
    assert np.array_equal(wd.vectorData(), exp)
    apo = vf.createReal2D(m, m)
    o.writeApoVector(apo, 0.9, 2.0)
    ea = np.zeros((m, m), np.float32)
    ref.lib().ref_write_apo(ref.ptr(ea), m, m, vec_cpm, cut, 0.9, 2.0)
    ga = apo.vectorData()
    assert np.max(np.abs(ga.view(np.int32) - ea.view(np.int32))) <= 1   # libm pow/acos: 1 ulp
    att = vf.createReal2D(n, n)
    o.writeAttenuationVector(att, 0.99, 0.15 * cut, 0, 0)
    et = np.zeros((n, n), np.float32)
    ref.lib().ref_write_attenuation(ref.ptr(et), n, n, vec_cpm, 0.99, 0.15 * cut, 0, 0)
    assert np.max(np.abs(att.vectorData().view(np.int32) - et.view(np.int32))) <= 1
    f = vf.createReal2D(n, n)
    f.addConst(1.0)
    fade_border_cos(f, 10)
    ef = np.ones((n, n), np.float32)
    ref.lib().ref_fade_border(ref.ptr(ef), n, n, 10)
    assert np.array_equal(f.vectorData(), ef)


@pytest.mark.parametrize("kx,ky", [(0.0, 0.0), (9.25, -6.5), (-40.0, 31.75)])
def test_mask_otf_bit_exact(vf, kx, ky):
    """maskOtf (OtfProvider.java:538-559) on random data, centred and at
    fractional band centres, against the oracle's mask (pinned as the cutoff
    disc by the oracle tests): bitwise (the kept pixels are the input, the
    others zero)."""
    n = 128
    o = _otf()
    vec_cpm = 1 / (n * 0.08)
    o.setPixelSize(vec_cpm)
    rng = np.random.default_rng(21)
    z = rnd_c(rng, n, n)
    v = cvec(vf, z)
    o.maskOtf(v, kx, ky)
    want = np.ascontiguousarray(z.view(np.float32).reshape(n, n, 2))
    ref.lib().ref_mask_otf(ref.ptr(want), n, n, vec_cpm, o.cutoff, kx, ky)
    got = v.vectorData()
    assert np.array_equal(got.view(np.float32).reshape(n, n, 2), want)
    kept = np.count_nonzero(want[..., 0])
    assert 0 < kept < n * n       # both sides of the cutoff occur


def test_locate_peak_bit_exact(vf):
    from fsim.sim_algorithm import locate_peak
    rng = np.random.default_rng(10)
    for n, kmin in [(64, 5.0), (128, 40.5), (512, 92.0)]:
        z = rnd_c(rng, n, n)
        z[17, 3] = 50 + 50j          # inside the excluded disc for the larger kmin
        z[n - 30, n // 2 + 7] = 9 - 9j
        z[n - 30, n // 2 + 8] = 9 - 9j   # tie: first in y-major order wins
        v = cvec(vf, z)
        got = locate_peak(v, kmin)
        exp = np.zeros(4)
        ref.lib().ref_locate_peak(ref.ptr(ref.cplx_f32(z)), n, n, kmin, ref.ptr(exp))
        assert got == list(exp)
