"""XML parameter files (utils/Conf.java) and what reads them:
SimParam.loadConfig / saveConfig, OtfProvider.loadFromConfig / saveConfig
(estimates and measured multi-band data), the live channel registry.

Fixtures: the reference's example configurations (tests/golden/*.xml, copied
by tests/golden/make_xml_fixtures.py); known answers are read off the XML
text.  Measured OTF data is pinned by the format (base64 of big-endian
float32 (re, im) pairs per band) and by round trips.
"""
import os
import struct

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _conf():
    from fsim.conf import Conf
    return Conf.loadFile(os.path.join(GOLD, "vigor-fastsim-example.xml"))


def test_load_reference_config_known_answers():
    cfg = _conf()
    assert cfg.namespace == "fairsim"
    vs = cfg.r().cd("vigor-settings")
    assert vs.getInt("TCPPort").val() == 32322
    assert vs.getDbl("RawPxlSize").val() == 0.078
    assert vs.getStr("Microscope").val() == "fastSIM"
    assert vs.getInt("SyncFrameDelay").val() == 5000 and vs.getInt("SyncMode").val() == 0
    ch = vs.cd("channel-488")
    assert ch.getDbl("offset").val() == 140 and ch.getInt("ChannelNumber").val() == 488
    sp = ch.cd("sim-param")
    assert sp.getDbl("microns-per-pxl").val() == 0.078
    assert sp.getTDE("timestamp").val() == 1481205767740
    assert sp.cd("dir-2").getDbl("shift").vals == [-23.7889, -172.878]


def test_simparam_and_otf_from_reference_config():
    from fsim.sim_algorithm import OtfProvider, SimParam
    ch = _conf().r().cd("vigor-settings").cd("channel-488")
    p = SimParam.loadConfig(ch)
    assert (p.nrDir(), p.nrBand(), p.dir(0).nrPha(), p.imgSize) == (3, 2, 3, 512)
    assert p.imgSeq == "PZA" and p.getWienerFilter() == 0.05 and p.getApoCutoff() == 2.0
    d0 = p.dir(0)
    assert (d0.pX, d0.pY) == (137.478, -106.278)       # setPxPy(shift), 2 bands
    assert d0.getPhaOff() == -0.307837 and d0.modul == [1.0, 0.63947]
    otf = OtfProvider.loadFromConfig(ch)
    ref = OtfProvider.fromEstimate(1.4, 525, 0.3)
    assert otf.is_estimate and np.array_equal(otf.vals, ref.vals) and otf.cutoff == ref.cutoff
    assert not otf.isAttenuate()


def test_live_channel_registry():
    from fsim.livemode import load_live_channels, sequence_extractor_from_conf
    vs = _conf().r().cd("vigor-settings")
    chans = load_live_channels(vs, ["488", "568"])
    assert [c.chNumber for c in chans] == [488, 568]
    # the global pixel size overrides the channel's own (ReconstructionRunner.PerChannel)
    assert all(c.param.micronsPerPixel == 0.078 and c.param.imgSize == 512 for c in chans)
    assert chans[1].offset == 140.0
    ing = sequence_extractor_from_conf(vs, chans, max_size=64)
    assert ing.nch == 2 and ing.raws == 9
    ing.close()


def test_conf_round_trip(tmp_path):
    from fsim.conf import Conf
    c = Conf("fairsim")
    f = c.r().mk("a").mk("b")
    f.newInt("i", 3, -7)
    f.newDbl("d", 1.4, -2.5e-7, 123456.789)
    f.newDbl("x", np.pi, exact=True)
    f.newStr("s", "PZA")
    f.newData("raw", bytes(range(256)))
    f.newTDE("t", 1481205767740)
    path = tmp_path / "c.xml"
    c.saveFile(str(path))
    txt = path.read_text()
    assert 'type="decimal"> 1.40000e+00 -2.50000e-07 1.23457e+05 <' in txt
    assert " BASE64:" in txt and "<?xml" not in txt
    g = Conf.loadFile(str(path)).r().cd("a").cd("b")
    assert g.getInt("i").vals == [3, -7]
    assert g.getDbl("d").vals == [1.4, -2.5e-7, 123457.0]    # %8.5e keeps six digits
    assert g.getDbl("x").vals == [np.pi]                       # exact: base64 doubles
    assert g.getStr("s").val() == "PZA" and g.getData("raw").val() == bytes(range(256))
    assert g.getTDE("t").val() == 1481205767740


def test_measured_multiband_otf_round_trip(tmp_path):
    from fsim import _lib
    from fsim.conf import Conf
    from fsim.sim_algorithm import OtfProvider, SimParam
    rng = np.random.default_rng(4)
    nb, ns = 3, 256
    vals = (rng.uniform(0, 1, (nb, ns)) + 1j * rng.uniform(-0.1, 0.1, (nb, ns))).astype(np.complex64)
    otf = OtfProvider.fromData(1.4, 600, 0.0215, vals)
    otf.setAttenuation(0.995, 1.2)
    otf.switchAttenuation(True)
    c = Conf("fairsim")
    ch = c.r().mk("channel-600")
    otf.saveConfig(ch)
    param = SimParam.create(3, 3, 5, 256, 0.08, otf)
    param.dir(1).setPhases([0.1 * i for i in range(5)], False)
    param.saveConfig(ch)
    path = tmp_path / "otf.xml"
    c.saveFile(str(path))
    ch2 = Conf.loadFile(str(path)).r().cd("channel-600")
    # the band payload: base64 of big-endian float32 (re, im) pairs
    raw = ch2.cd("otf2d").cd("data").getData("band-1").val()
    assert struct.unpack(">2f", raw[8:16]) == (float(vals[1, 1].real), float(vals[1, 1].imag))
    back = OtfProvider.loadFromConfig(ch2)
    assert (back.nband, back.samples, back.cpm) == (3, 256, 0.0215)
    assert np.array_equal(back.band_values(), vals)
    assert back.isAttenuate() and (back.att_str, back.att_fwhm) == (0.995, 1.2)
    assert np.array_equal(back.vals_att, otf.vals_att) and np.array_equal(back.only_att, otf.only_att)
    # the attenuation tables follow the same formula as the estimate's
    est = OtfProvider.fromEstimate(1.4, 600, 1.0, samples=ns)
    one = OtfProvider.fromData(1.4, 600, est.cpm, est.band_values())
    assert np.array_equal(one.only_att, est.only_att) and np.array_equal(one.vals_att, est.vals_att)
    p2 = SimParam.loadConfig(ch2)
    # decimals are written as %8.5e (six significant digits), as Conf does
    assert p2.dir(1).phases == [float("%8.5e" % (0.1 * i)) for i in range(5)] and p2.dir(1).hasIndividualPhases
    assert not p2.dir(0).hasIndividualPhases
    assert p2.dir(2).pX == float("%8.5e" % param.dir(2).pX)
