"""C-ABI boundary checks that need no GPU: the library loads, exports every
symbol include/fsim.h declares, its host-side computations (separation
matrix, OTF tables) match the oracle bit for bit, and the host mirror of the
reference's parameter classes behaves like the Java."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "fsim.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fsim_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from fsim import _lib
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) > 40
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    unbound = sorted(set(syms) - set(_lib.declared_symbols()))
    assert not unbound, f"declared in fsim.h but not bound in _lib: {unbound}"


def test_calls_before_init_fail_cleanly():
    from fsim import _lib
    L = _lib.lib()
    if _lib._initialised:
        pytest.skip("library already initialised in this process")
    h = C.c_void_p()
    rc = L.fsim_vec_create(1, 8, 8, C.byref(h))
    assert rc == _lib.FSIM_ERR_STATE
    assert b"fsim_init" in L.fsim_last_error()
    with pytest.raises(_lib.FsimError):
        _lib.check(rc)
    with pytest.raises(MemoryError):
        _lib.check(_lib.FSIM_ERR_OOM)


@pytest.mark.parametrize("P,B,fac", [(3, 2, None), (3, 2, [1.0, 0.7]), (5, 2, [1.0, 0.9]),
                                     (5, 3, [1.0, 0.8, 0.6]), (7, 3, None)])
def test_separation_matrix_host_bit_exact(P, B, fac):
    from fsim.sim_algorithm import separation_matrix
    phases = [2 * np.pi * p / P + 0.37 for p in range(P)]
    got = separation_matrix(phases, B, fac)
    Cc = 2 * B - 1
    exp = np.zeros(2 * Cc * P)
    ph = np.array(phases)
    fa = None if fac is None else ref.ptr(np.array(fac, dtype=np.float64))
    ref.lib().ref_separation_matrix(ref.ptr(ph), P, B, fa, ref.ptr(exp))
    e = exp.reshape(Cc, P, 2)
    assert np.array_equal(got.real, e[..., 0]) and np.array_equal(got.imag, e[..., 1])


def test_separation_matrix_errors():
    from fsim._lib import FsimError
    from fsim.sim_algorithm import separation_matrix
    with pytest.raises(FsimError):
        separation_matrix([0.0, 1.0], 2)        # not enough phases for 2 bands
    with pytest.raises(FsimError):
        separation_matrix([0.0, 1.0, 2.0], 1)   # bands < 2


def test_otf_tables_host_bit_exact():
    from fsim.sim_algorithm import OtfProvider
    for na, lam, a in [(1.4, 525.0, 1.0), (1.2, 670.0, 0.3), (1.49, 600.0, 0.7)]:
        o = OtfProvider.fromEstimate(na, lam, a)
        o.setAttenuation(0.995, 1.2)
        vals, cut, cpm = ref.otf_estimate(na, lam, a)
        assert np.array_equal(o.vals.reshape(-1, 2), vals)
        assert o.cutoff == cut and o.cpm == cpm
        va = np.zeros_like(vals)
        oa = np.zeros(512, np.float32)
        ref.lib().ref_otf_set_attenuation(ref.ptr(vals), 512, cpm, 0.995, 1.2, ref.ptr(va), ref.ptr(oa))
        assert np.array_equal(o.vals_att.reshape(-1, 2), va)
        assert np.array_equal(o.only_att, oa)


def test_otf_estimate_rejects_unphysical():
    from fsim.sim_algorithm import OtfProvider
    from fsim._lib import FsimError
    with pytest.raises(FsimError):
        OtfProvider.fromEstimate(3.0, 525.0, 1.0)
    with pytest.raises(FsimError):
        OtfProvider.fromEstimate(1.4, 200.0, 1.0)


def test_simparam_mirror():
    """SimParam.Dir semantics (SimParam.java:472-607)."""
    from fsim.sim_algorithm import Dir, OtfProvider, SimParam
    d = Dir(3, 5)
    d.setPxPy(100.0, -40.0)
    assert d.px(0) == 0 and d.px(1) == 50.0 and d.py(2) == -40.0
    assert d.getModulations() == [1.0, 1.0, 1.0]
    d.setModulation(1, 0.1)
    d.setModulation(2, 1.5)
    assert d.getModulations() == [1.0, 0.3, 1.1]
    d.setPhaOff(0.5)
    assert d.getPhases()[1] == pytest.approx(2 * np.pi / 5 + 0.5)
    with pytest.raises(RuntimeError):
        Dir(2, 2)
    o = OtfProvider.fromEstimate(1.4, 525.0, 1.0)
    sp = SimParam.create(2, 3, 3, 512, 0.08, o)
    assert sp.pxlSizeCyclesMicron() == 1 / (512 * 0.08)
    assert o.vec_cpm == sp.pxlSizeCyclesMicron()


def test_recon_params_packing():
    from fsim import synthetic
    from fsim.sim_algorithm import recon_params
    raw, sp = synthetic.sim_stack(n=64, seed=0, nbeads=10, nfil=2)
    param = synthetic.make_param(sp, 64)
    p = recon_params(param)
    assert p.n == 64 and p.ndir == 3 and p.nphase == 3 and p.nband == 2
    assert p.px[1] == pytest.approx(sp["px"][1]) and p.py[2] == pytest.approx(sp["py"][2])
    assert p.modul[0][0] == 1.0 and p.modul[0][1] == pytest.approx(0.7)
    assert p.use_att == 1 and p.fade_px == 10


def test_synthetic_is_seeded_and_shaped():
    from fsim import synthetic
    a, spa = synthetic.sim_stack(n=64, seed=7, nbeads=20, nfil=3)
    b, _ = synthetic.sim_stack(n=64, seed=7, nbeads=20, nfil=3)
    c, _ = synthetic.sim_stack(n=64, seed=8, nbeads=20, nfil=3)
    assert a.shape == (9, 64, 64) and a.dtype == np.uint16
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert 100 <= a.min() and a.max() < 2000                  # offset + background + peak ~1000
    assert spa["k0"] == pytest.approx(0.85 * spa["cutoff_px"])


@pytest.mark.parametrize("nband", [2, 3])
def test_params_from_estimate_matches_simparam(nband):
    """fsim_params_from_estimate leaves the parameters SimParam.Dir holds
    after estimateParameters' setPxPy / setPhaOff / setModulation
    (SimAlgorithm.java:139-142, 155-158), as the host mirror computes them."""
    from fsim import _lib
    from fsim.sim_algorithm import OtfProvider, SimParam, apply_estimate, params_from_estimates, sim_params
    P, D, count = 5, 3, 3
    otf = OtfProvider.fromEstimate(1.4, 525.0, 1.0)
    param = SimParam.create(nband, D, P, 256, 0.08, otf)
    param.dir(1).setModulation(1, 0.8)
    rng = np.random.default_rng(4)
    res = (_lib.ESTIMATE_RESULT * (count * D))()
    dicts = []
    for s in range(count):
        dirs = []
        for d in range(D):
            r = res[s * D + d]
            r.px, r.py = rng.uniform(-120, 120, 2)
            r.pha_off, r.mod1, r.mod2 = rng.uniform(-3, 3), rng.uniform(0.2, 1.2), rng.uniform(0.1, 0.9)
            e = dict(px=r.px, py=r.py, pha_off=r.pha_off, mod1=r.mod1)
            if nband == 3:
                e["mod2"] = r.mod2
            dirs.append(e)
        dicts.append(dirs)
    got = params_from_estimates(param, res, count)
    for s in range(count):
        want_param = SimParam.create(nband, D, P, 256, 0.08, otf)
        want_param.dir(1).setModulation(1, 0.8)
        apply_estimate(want_param, dicts[s])
        want = sim_params(want_param)
        assert bytes(got[s]) == bytes(want), f"stack {s}"
