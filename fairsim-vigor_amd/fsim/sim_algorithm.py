"""Host side of the SIM path: parameters, OTF provider and reconstructions.

Mirrors org.fairsim.sim_algorithm for the 2D fast-shift reconstruction:
``OtfProvider`` (sim_algorithm/OtfProvider.java), ``SimParam`` /
``SimParam.Dir`` (sim_algorithm/SimParam.java), the ``SimUtils`` helpers
used by the reconstruction (sim_algorithm/SimUtils.java:33-91) and two ways
of running a reconstruction on the GPU:

* ``reconstruct_plain`` -- the op sequence of
  ``ReconstructionRunner.ReconstructionThread.run``
  (livemode/ReconstructionRunner.java:456-542) issued one vector operation at
  a time through the vector API, as the reference's accelerated path does;
* ``Reconstructor`` -- the fused pipeline (fsim_channel / fsim_reconstruct)
  that runs the same dataflow in five HBM passes for a batch of stacks.

Parameter estimation: ``Correlation`` (sim_algorithm/Correlation.java) with
its per-pixel loops on the device, and ``estimate_parameters``
(SimAlgorithm.estimateParameters, SimAlgorithm.java:37-256) over device
vectors.
"""
import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import RECON_PARAMS, call, lib
from .linalg import AccelVectorFactory, DeviceBuffer, Vec2dCplx, Vec2dReal


class OtfProvider:
    """OTF tables from an estimate (OtfProvider.fromEstimate,
    OtfProvider.java:96-127) or from measured per-band data (loadFromConfig,
    :574-631), with attenuation (:276-287)."""

    def __init__(self):
        self.samples = 512
        self.nband = 1
        self.is_estimate = True
        self.na = self.lam = self.a = None
        self.att_str, self.att_fwhm = 0.99, 1.2
        self.use_att = False
        self.vec_cpm = -1.0
        self._dev = None

    @staticmethod
    def fromEstimate(na, lam, a, samples=512):
        o = OtfProvider()
        o.na, o.lam, o.a, o.samples = float(na), float(lam), float(a), int(samples)
        o._compute()
        o.use_att = False
        return o

    @staticmethod
    def fromData(na, lam, cycles_per_micron, vals):
        """A measured OTF: vals complex [bands, samples] (one radial profile
        per band, sample i at i * cycles_per_micron); cutoff 1000 / (lambda /
        NA / 2) as the reference sets it (OtfProvider.java:593-597)."""
        vals = np.asarray(vals, dtype=np.complex64)
        if vals.ndim == 1:
            vals = vals[None, :]
        o = OtfProvider()
        o.na, o.lam = float(na), float(lam)
        o.is_estimate = False
        o.nband, o.samples = vals.shape
        o.cpm = float(cycles_per_micron)
        o.cutoff = 1000 / (o.lam / o.na / 2)
        o.vals = np.ascontiguousarray(np.stack([vals.real, vals.imag], -1).reshape(-1), dtype=np.float32)
        o._compute()
        o.use_att = False
        return o

    def band_values(self):
        """vals as complex [bands, samples]."""
        v = self.vals.reshape(self.nband, self.samples, 2)
        return v[..., 0] + 1j * v[..., 1]

    def _compute(self):
        if not self.is_estimate:
            self.vals_att = np.zeros_like(self.vals)
            self.only_att = np.zeros(self.nband * self.samples, dtype=np.float32)
            call("fsim_otf_attenuation", self.nband, self.samples, self.cpm, self.att_str, self.att_fwhm,
                 self.vals.ctypes.data_as(C.c_void_p), self.vals_att.ctypes.data_as(C.c_void_p),
                 self.only_att.ctypes.data_as(C.c_void_p))
            self._release()
            return
        n = self.samples
        self.vals = np.zeros(2 * n, dtype=np.float32)
        self.vals_att = np.zeros(2 * n, dtype=np.float32)
        self.only_att = np.zeros(n, dtype=np.float32)
        cut, cpm = C.c_double(), C.c_double()
        call("fsim_otf_from_estimate", self.na, self.lam, self.a, n, self.att_str, self.att_fwhm,
             self.vals.ctypes.data_as(C.c_void_p), self.vals_att.ctypes.data_as(C.c_void_p),
             self.only_att.ctypes.data_as(C.c_void_p), C.byref(cut), C.byref(cpm))
        self.cutoff, self.cpm = cut.value, cpm.value
        self._release()

    def _release(self):
        if self._dev is not None and _lib._lib is not None:
            lib().fsim_otf_destroy(self._dev)
        self._dev = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def setAttenuation(self, strength, fwhm):
        self.att_str, self.att_fwhm = float(strength), float(fwhm)
        self._compute()

    def switchAttenuation(self, on):
        self.use_att = bool(on)

    def isAttenuate(self):
        return self.use_att

    def getCutoff(self):
        return self.cutoff

    def setPixelSize(self, cycles_per_micron):
        if cycles_per_micron <= 0:
            raise ValueError("pxl size must be positive")
        self.vec_cpm = float(cycles_per_micron)

    def device(self):
        """Device copy of the 1-D tables (fsim_otf)."""
        if self._dev is None:
            _lib.init()
            h = C.c_void_p()
            call("fsim_otf_create", self.nband, self.samples, self.cutoff, self.cpm,
                 self.vals.ctypes.data_as(C.c_void_p), self.vals_att.ctypes.data_as(C.c_void_p),
                 self.only_att.ctypes.data_as(C.c_void_p), C.byref(h))
            self._dev = h
        return self._dev

    # --- table writers over device vectors (OtfProvider.java:331-529) ---
    def otfToVector(self, vec, band, kx, ky, useAtt, write):
        call("fsim_otf_to_vector", vec.handle, self.device(), int(band), float(kx), float(ky),
             self.vec_cpm, 1 if useAtt else 0, 1 if write else 0)

    def applyOtf(self, vec, band, kx=0.0, ky=0.0):
        self.otfToVector(vec, band, kx, ky, self.use_att, False)

    def writeOtfVector(self, vec, band, kx, ky):
        self.otfToVector(vec, band, kx, ky, False, True)

    def writeOtfWithAttVector(self, vec, band, kx, ky):
        self.otfToVector(vec, band, kx, ky, self.use_att, True)

    def writeApoVector(self, vec, bend, cutoff):
        call("fsim_write_apo", vec.handle, self.device(), self.vec_cpm, float(bend), float(cutoff))

    def writeAttenuationVector(self, vec, strength, fwhm, kx, ky):
        call("fsim_write_attenuation", vec.handle, self.vec_cpm, float(strength), float(fwhm),
             float(kx), float(ky))

    def maskOtf(self, vec, kx, ky):
        call("fsim_mask_otf", vec.handle, self.device(), float(kx), float(ky), self.vec_cpm)

    # --- configuration files (OtfProvider.java:566-664) ---
    @staticmethod
    def loadFromConfig(cfg):
        """OtfProvider.loadFromConfig from a Conf folder holding "otf2d":
        an estimate (NA, emission, a-estimate) or data (bands, samples,
        cycles, band-<b> as big-endian float32 (re, im) pairs); then the
        attenuation folder, when present, sets and switches it on."""
        from . import conf
        fld = cfg.cd("otf2d")
        if not fld.contains("data"):
            ret = OtfProvider.fromEstimate(fld.getDbl("NA").val(), fld.getInt("emission").val(),
                                           fld.getDbl("a-estimate").val())
        else:
            data = fld.cd("data")
            nb, ns = data.getInt("bands").val(), data.getInt("samples").val()
            vals = np.zeros((nb, ns), np.complex64)
            for b in range(nb):
                f = np.asarray(conf.from_byte(data.getData("band-%d" % b).val()), np.float32)
                vals[b] = f[0:2 * ns:2] + 1j * f[1:2 * ns:2]
            ret = OtfProvider.fromData(fld.getDbl("NA").val(), fld.getInt("emission").val(),
                                       data.getDbl("cycles").val(), vals)
        if fld.contains("attenuation"):
            att = fld.cd("attenuation")
            ret.setAttenuation(att.getDbl("strength").val(), att.getDbl("FWHM").val())
            ret.switchAttenuation(True)
        return ret

    def saveConfig(self, cfg):
        """OtfProvider.saveConfig into a Conf folder (creates "otf2d")."""
        from . import conf
        fld = cfg.mk("otf2d")
        fld.newDbl("NA", self.na)
        fld.newInt("emission", int(self.lam))
        if self.is_estimate:
            fld.newDbl("a-estimate", self.a)
        else:
            data = fld.mk("data")
            data.newInt("bands", self.nband)
            data.newInt("samples", self.samples)
            data.newDbl("cycles", self.cpm)
            for b in range(self.nband):
                data.newData("band-%d" % b, conf.to_byte(
                    self.vals[2 * b * self.samples:2 * (b + 1) * self.samples].tolist()))
        if self.use_att:
            att = fld.mk("attenuation")
            att.newDbl("strength", self.att_str)
            att.newDbl("FWHM", self.att_fwhm)


class Dir:
    """SimParam.Dir (SimParam.java:378-615)."""

    def __init__(self, bands, phases):
        if bands < 2:
            raise RuntimeError("bands < 2, not useful")
        if phases < 2 * bands - 1:
            raise RuntimeError(f"not enough phases for bands:{bands} {phases}")
        self.nrBands, self.nrPhases = bands, phases
        self.pX = self.pY = 0.0
        self.phaOff = 0.0
        self.phases = [(2 * math.pi / phases) * i for i in range(phases)]
        self.modul = [1.0] * bands
        self.hasIndividualPhases = False

    def nrBand(self):
        return self.nrBands

    def nrComp(self):
        return 2 * self.nrBands - 1

    def nrPha(self):
        return self.nrPhases

    def setPxPy(self, x, y):
        self.pX = x / (self.nrBands - 1)
        self.pY = y / (self.nrBands - 1)

    def px(self, band):
        return self.pX * band

    def py(self, band):
        return self.pY * band

    def setPhaOff(self, p):
        self.phaOff = float(p)

    def getPhaOff(self):
        return self.phaOff

    def setPhases(self, pha, reset):
        if len(pha) != self.nrPhases:
            raise RuntimeError("Length mismatch")
        if reset:
            self.phaOff = 0.0
        self.phases = [float(p) for p in pha]
        self.hasIndividualPhases = True

    def getPhases(self):
        return [p + self.phaOff for p in self.phases]

    def setModulation(self, b, m):
        self.modul[b] = float(m)
        return 0.3 < m < 1.1

    def getModulations(self):
        return [min(max(m, 0.3), 1.1) for m in self.modul]


class SimParam:
    """SimParam.create(bands, dirs, phases, size, micronsPerPxl, otf)
    (SimParam.java:103-130, 250-256)."""

    IMGSEQ = ("PAZ", "PZA", "ZAP")   # SimParam.IMGSEQ (SimParam.java:753-832)

    def __init__(self, bands, dirs, phases, size, microns, otf):
        self.nrBands, self.nrDirs, self.nrPhases = bands, dirs, phases
        self.directions = [Dir(bands, phases) for _ in range(dirs)]
        self.otf_ = None
        self.wienerFilterParameter = 0.05
        self.apoCutOff = 2.0
        self.imgSeq = "PAZ"
        self.paramDate = 0
        self.setPxlSize(size, microns)
        if otf is not None:
            self.otf(otf)

    @staticmethod
    def create(bands, dirs, phases, size, microns, otf):
        return SimParam(bands, dirs, phases, size, microns, otf)

    def setPxlSize(self, size, microns):
        self.imgSize = int(size)
        self.micronsPerPixel = float(microns)
        self.cyclesPerMicron = 1 / (size * microns)
        if self.otf_ is not None:
            self.otf_.setPixelSize(self.cyclesPerMicron)

    def getWienerFilter(self):
        return self.wienerFilterParameter

    def getApoCutoff(self):
        return self.apoCutOff

    @staticmethod
    def loadConfig(cfg):
        """SimParam.loadConfig (SimParam.java:676-726) from a Conf folder
        holding "sim-param"."""
        fd = cfg.cd("sim-param")
        ret = SimParam(fd.getInt("nr-bands").val(), fd.getInt("nr-angles").val(), fd.getInt("nr-phases").val(),
                       fd.getInt("img-size-pxl").val(), fd.getDbl("microns-per-pxl").val(), None)
        if fd.contains("timestamp"):
            ret.paramDate = fd.getTDE("timestamp").val()
        seq = fd.getStr("img-seq").val().strip()
        ret.imgSeq = seq if seq in SimParam.IMGSEQ else None
        ret.wienerFilterParameter = fd.getDbl("wiener-parameter").val()
        ret.apoCutOff = fd.getDbl("apodization-cutoff").val()
        for d in range(ret.nrDirs):
            df = fd.cd("dir-%d" % d)
            dd = ret.dir(d)
            shift = df.getDbl("shift").vals
            if len(shift) < 2:
                raise RuntimeError("Array too short")
            dd.setPxPy(shift[0], shift[1])
            dd.setPhaOff(df.getDbl("phase-offset").val())
            for i, m in enumerate(df.getDbl("modulations").vals):
                dd.setModulation(i, m)
            if df.contains("phases"):
                dd.setPhases(df.getDbl("phases").vals, False)
        return ret

    def saveConfig(self, cfg):
        """SimParam.saveConfig (SimParam.java:645-674) into a Conf folder."""
        fd = cfg.mk("sim-param")
        fd.newInt("nr-angles", self.nrDirs)
        fd.newInt("nr-bands", self.nrBands)
        fd.newInt("nr-phases", self.nrPhases)
        fd.newStr("img-seq", self.imgSeq)
        fd.newInt("img-size-pxl", self.imgSize)
        fd.newDbl("microns-per-pxl", self.micronsPerPixel)
        fd.newDbl("wiener-parameter", self.wienerFilterParameter)
        fd.newDbl("apodization-cutoff", self.apoCutOff)
        fd.newTDE("timestamp", self.paramDate)
        for d in range(self.nrDirs):
            df = fd.mk("dir-%d" % d)
            dd = self.dir(d)
            df.newDbl("shift", dd.pX * (dd.nrBands - 1), dd.pY * (dd.nrBands - 1))
            df.newDbl("phase-offset", dd.phaOff)
            df.newDbl("modulations", *dd.modul)
            if dd.hasIndividualPhases:
                df.newDbl("phases", *dd.phases)

    def dir(self, i):
        return self.directions[i]

    def nrDir(self):
        return self.nrDirs

    def nrBand(self):
        return self.nrBands

    def otf(self, o=None):
        """getter, or SimParam.otf(OtfProvider): set it, with this pixel size"""
        if o is None:
            return self.otf_
        self.otf_ = o
        o.setPixelSize(self.cyclesPerMicron)
        return o

    def pxlSizeCyclesMicron(self):
        return self.cyclesPerMicron

    def vectorWidth(self):
        return self.imgSize

    def vectorHeight(self):
        return self.imgSize


def separation_matrix(phases, bands, fac=None):
    """BandSeparation.createSeparationMatrix (BandSeparation.java:40-114),
    computed natively; returns a (C, P) complex128 array."""
    P = len(phases)
    Cc = 2 * bands - 1
    ph = (C.c_double * P)(*phases)
    out = (C.c_double * (2 * Cc * P))()
    facp = (C.c_double * bands)(*fac) if fac is not None else None
    call("fsim_separation_matrix", ph, P, bands, facp, out)
    a = np.frombuffer(out, dtype=np.float64).reshape(Cc, P, 2)
    return a[..., 0] + 1j * a[..., 1]


def separate_bands(inv, out, phases, bands, fac):
    """BandSeparation.separateBands (BandSeparation.java:131-165) in one pass."""
    S = separation_matrix(phases, bands, fac)
    Sf = np.empty(S.shape + (2,), dtype=np.float32)
    Sf[..., 0] = S.real
    Sf[..., 1] = S.imag
    ins = (C.c_void_p * len(inv))(*[v.handle.value for v in inv])
    outs = (C.c_void_p * len(out))(*[v.handle.value for v in out])
    call("fsim_separate_bands", ins, len(inv), outs, len(out),
         Sf.ctypes.data_as(C.POINTER(C.c_float)))


def wiener_denominator(vec, param, wparam):
    """WienerFilter(param).getDenominator(wparam) (WienerFilter.java:32-122)."""
    otf = param.otf()
    D = param.nrDir()
    px = (C.c_double * D)(*[param.dir(d).pX for d in range(D)])
    py = (C.c_double * D)(*[param.dir(d).pY for d in range(D)])
    call("fsim_wiener_denominator", vec.handle, otf.device(), param.pxlSizeCyclesMicron(), D,
         param.nrBand(), px, py, 1 if otf.isAttenuate() else 0, float(wparam))


def fade_border_cos(vec, px):
    """SimUtils.fadeBorderCos (SimUtils.java:162-185)."""
    call("fsim_fade_border", vec.handle, int(px))


def locate_peak(vec, kmin):
    """Correlation.locatePeak (Correlation.java:144-167)."""
    r = (C.c_double * 4)()
    call("fsim_locate_peak", vec.handle, float(kmin), r)
    return list(r)


class Correlation:
    """sim_algorithm/Correlation.java with the pixel loops on the device.
    Bands are complex device vectors in Fourier space; the OTF's vector pixel
    size (setPixelSize) must be set."""

    @staticmethod
    def locatePeak(vec, kMin):
        """Correlation.locatePeak (Correlation.java:144-167): [x, y, mag, phase]."""
        return locate_peak(vec, kMin)

    @staticmethod
    def commonRegion(band0, band1, bn0, bn1, otf, kx, ky, dist, weightLimit, divideByOtf):
        """Correlation.commonRegion (Correlation.java:227-288), in place."""
        call("fsim_common_region", band0.handle, band1.handle, int(bn0), int(bn1), otf.device(),
             float(kx), float(ky), float(dist), float(weightLimit), 1 if divideByOtf else 0, otf.vec_cpm)

    @staticmethod
    def fitPeak(band0, band1, bn0, bn1, otf, kx, ky, weightLimit, search, cntrl=None):
        """Correlation.fitPeak (Correlation.java:43-136): [kx, ky, phase, mag].
        cntrl: None or a float32 array of shape (10, 30) to receive the
        normalised correlation magnitudes of the three rounds."""
        res = (C.c_double * 4)()
        cp = None
        if cntrl is not None:
            if cntrl.dtype != np.float32 or cntrl.size != 300 or not cntrl.flags.c_contiguous:
                raise ValueError("cntrl must be a contiguous float32 array of 10 x 30")
            cp = cntrl.ctypes.data_as(C.POINTER(C.c_float))
        call("fsim_fit_peak", band0.handle, band1.handle, int(bn0), int(bn1), otf.device(), float(kx),
             float(ky), float(weightLimit), float(search), otf.vec_cpm, cp, res)
        return list(res)

    @staticmethod
    def getPeak(band0, band1, bn0, bn1, otf, kx, ky, weightLimit):
        """Correlation.getPeak (Correlation.java:183-209): a complex number
        (phase = phase offset, magnitude = modulation estimate)."""
        res = (C.c_double * 2)()
        call("fsim_get_peak", band0.handle, band1.handle, int(bn0), int(bn1), otf.device(), float(kx),
             float(ky), float(weightLimit), otf.vec_cpm, res)
        return complex(res[0], res[1])


def estimate_parameters(param, inFFT, fitBand, fitExclude):
    """SimAlgorithm.estimateParameters (SimAlgorithm.java:37-256) without the
    visual feedback: inFFT[d][p] are the FFT'd input images (complex device
    vectors); per direction the shift vector (setPxPy), phase offset and
    modulations of ``param`` are set.  Returns the per-direction results."""
    w, h = inFFT[0][0].vectorWidth(), inFFT[0][0].vectorHeight()
    otfPr = param.otf()
    if fitBand != 1 and fitBand != 2:
        raise RuntimeError("Fitband neither 1 nor 2")
    # the attenuation vector fades out the DC component for the correlation
    otfAtt = Vec2dReal(w, h)
    otfPr.writeAttenuationVector(otfAtt, .99, 0.15 * otfPr.getCutoff(), 0, 0)
    results = []
    for angIdx in range(param.nrDir()):
        d = param.dir(angIdx)
        lb = 1
        hb = 3 if d.nrBand() == 3 else 1
        fb = lb if fitBand == 1 else hb
        # band separation with equidistant phases (BandSeparation.java:183-197)
        P = len(inFFT[angIdx])
        separate = [Vec2dCplx(w, h) for _ in range(d.nrComp())]
        separate_bands(inFFT[angIdx], separate, [(2 * math.pi * p) / P + 0.0 for p in range(P)],
                       d.nrBand(), None)
        # coarse correlation: ifft, multiply in real space, fft back
        c0, c1, c2 = separate[0].duplicate(), separate[lb].duplicate(), separate[hb].duplicate()
        for c in (c0, c1, c2):
            c.times(otfAtt)
            c.fft2d(True)
        c1.timesConj(c0)
        c2.timesConj(c0)
        c1.fft2d(False)
        c2.fft2d(False)
        minDist = fitExclude * otfPr.getCutoff() / param.pxlSizeCyclesMicron()
        peak = Correlation.locatePeak(c1 if fitBand == 1 else c2, minDist)
        peak = Correlation.fitPeak(separate[0], separate[fb], 0, fitBand, otfPr, -peak[0], -peak[1], 0.05, 2.5)
        r = {}
        if lb != hb:
            if fitBand == 1:
                peak[0] *= 2
                peak[1] *= 2
            p1 = Correlation.getPeak(separate[0], separate[lb], 0, 1, otfPr, peak[0] / 2, peak[1] / 2, 0.05)
            p2 = Correlation.getPeak(separate[0], separate[hb], 0, 2, otfPr, peak[0], peak[1], 0.05)
            d.setPxPy(-peak[0], -peak[1])
            d.setPhaOff(math.atan2(p1.imag, p1.real))
            d.setModulation(1, abs(p1))
            d.setModulation(2, abs(p2))
            r.update(px=-peak[0], py=-peak[1], pha_off=math.atan2(p1.imag, p1.real), mod1=abs(p1), mod2=abs(p2))
        else:
            p1 = Correlation.getPeak(separate[0], separate[1], 0, 1, otfPr, peak[0], peak[1], 0.05)
            d.setPxPy(-peak[0], -peak[1])
            d.setPhaOff(math.atan2(p1.imag, p1.real))
            d.setModulation(1, abs(p1))
            r.update(px=-peak[0], py=-peak[1], pha_off=math.atan2(p1.imag, p1.real), mod1=abs(p1))
        results.append(r)
    return results


def refit_parameters(param, raw_stack, fitExclude=0.8, fade_px=10):
    """ReconstructionRunner.ParameterRefitThread.doRefit
    (livemode/ReconstructionRunner.java:642-746) over device vectors: the
    live-mode refit of one channel from its latest raw stack (uint16 [D*P,
    N, N], direction-major).  Each frame is set from its 16-bit pixels,
    multiplied by borderDampen (a vector of ones faded by fadeBorderCos) and
    transformed; per direction the bands are separated with equidistant
    phases, the coarse correlation of band 0 with the highest band is
    searched outside fitExclude * cutoff (0.8 in the reference), fitPeak
    refines it with OTF band numbers (0, 1) whatever the band, and getPeak of
    band 1 at peak / (bands - 1) gives the phase offset and modulation 1;
    modulation 2 is left as it was.  ``param`` receives setPxPy, setPhaOff
    and setModulation(1) per direction; returns the per-direction results."""
    raw_stack = np.ascontiguousarray(raw_stack, dtype=np.uint16)
    D, P = param.nrDir(), param.dir(0).nrPha()
    n = param.vectorWidth()
    otfPr = param.otf()
    borderDampen = Vec2dReal(n, n)
    borderDampen.addConst(1)
    fade_border_cos(borderDampen, fade_px)
    inFFT = []
    count = 0
    for a in range(D):
        row = []
        for _ in range(P):
            v = Vec2dCplx(n, n)
            v.setFrom16bitPixels(raw_stack[count])
            count += 1
            v.times(borderDampen)
            v.fft2d(False)
            row.append(v)
        inFFT.append(row)
    otfAtt = Vec2dReal(n, n)
    otfPr.writeAttenuationVector(otfAtt, .99, 0.15 * otfPr.getCutoff(), 0, 0)
    results = []
    for angIdx in range(D):
        par = param.dir(angIdx)
        separate = [Vec2dCplx(n, n) for _ in range(par.nrComp())]
        separate_bands(inFFT[angIdx], separate, [(2 * math.pi * p) / P + 0.0 for p in range(P)], par.nrBand(), None)
        lb = 1
        hb = 3 if par.nrBand() == 3 else 1
        nBand = par.nrBand() - 1
        c0, c1 = separate[0].duplicate(), separate[hb].duplicate()
        c0.times(otfAtt)
        c1.times(otfAtt)
        c0.fft2d(True)
        c1.fft2d(True)
        c1.timesConj(c0)
        c1.fft2d(False)
        minDist = fitExclude * otfPr.getCutoff() / param.pxlSizeCyclesMicron()
        peak = Correlation.locatePeak(c1, minDist)
        peak = Correlation.fitPeak(separate[0], separate[hb], 0, 1, otfPr, -peak[0], -peak[1], 0.05, 2.5)
        p1 = Correlation.getPeak(separate[0], separate[lb], 0, 1, otfPr, peak[0] / nBand, peak[1] / nBand, 0.05)
        par.setPxPy(-peak[0], -peak[1])
        par.setPhaOff(math.atan2(p1.imag, p1.real))
        par.setModulation(1, abs(p1))
        results.append(dict(px=-peak[0], py=-peak[1], pha_off=math.atan2(p1.imag, p1.real), mod1=abs(p1)))
    return results


def estimate_batch(param, fitBand, fitExclude, images=None, images_dev=None, count=None, fade_px=10, raw=False,
                   live=False):
    """fsim_estimate: SimAlgorithm.estimateParameters (SimAlgorithm.java:37-256)
    for a batch of stacks in one native call, with the GUI's input preparation
    (input_ffts: fadeBorderCos, FFT) on the device.  ``images``: host float32
    [count, D, P, N, N]; or ``images_dev``: the device address of such a block
    plus ``count``.  Returns, per stack, the per-direction result dicts of
    estimate_parameters (plus the coarse ``peak``); ``param`` is not changed
    (apply_estimate stores one stack's results in it); ``raw``: return the
    fsim_estimate_result array instead (for params_from_estimates).
    ``live``: the live-mode refit instead (refit_parameters' semantics,
    ReconstructionRunner.java:642-746): highest band, OTF band numbers (0, 1),
    modulation 1 only, input faded by the borderDampen product; ``fitBand``
    is ignored (pass fitExclude = 0.8 for the reference's constant)."""
    _lib.init()
    D, P = param.nrDir(), param.dir(0).nrPha()
    n = param.vectorWidth()
    buf = None
    if images is not None:
        images = np.ascontiguousarray(images, dtype=np.float32)
        count = images.size // (D * P * n * n)
        if images.size != count * D * P * n * n or count < 1:
            raise ValueError("images must hold count x D x P x N x N values")
        buf = DeviceBuffer(images.nbytes)
        buf.upload(images)
        images_dev = buf.value
    elif images_dev is None or count is None:
        raise ValueError("give images, or images_dev and count")
    p = _lib.ESTIMATE_PARAMS(n=n, ndir=D, nphase=P, nband=param.nrBand(), fit_band=int(fitBand),
                             fit_exclude=float(fitExclude), vec_cpm=param.pxlSizeCyclesMicron(),
                             fade_px=int(fade_px), live=1 if live else 0)
    res = (_lib.ESTIMATE_RESULT * (count * D))()
    try:
        call("fsim_estimate", C.c_void_p(int(images_dev)), int(count), C.byref(p), param.otf().device(), res)
    finally:
        if buf is not None:
            buf.free()
    if raw:
        return res
    out = []
    for s in range(count):
        dirs = []
        for d in range(D):
            r = res[s * D + d]
            e = dict(px=r.px, py=r.py, pha_off=r.pha_off, mod1=r.mod1, peak=(r.peak_x, r.peak_y))
            if param.nrBand() == 3 and not live:
                e["mod2"] = r.mod2
            dirs.append(e)
        out.append(dirs)
    return out


def params_from_estimates(param, res, count):
    """fsim_params_from_estimate: per-stack fsim_sim_params (an array of
    ``count``) from estimate_batch(..., raw=True) results, on top of
    ``param``'s phases (without offset) and modulations."""
    base = sim_params(param)
    for d in range(param.nrDir()):
        for i, ph in enumerate(param.dir(d).phases):
            base.phases[d][i] = ph
    out = (_lib.SIM_PARAMS * count)()
    call("fsim_params_from_estimate", res, int(count), param.nrDir(), param.dir(0).nrPha(), param.nrBand(),
         C.byref(base), out)
    return out


def apply_estimate(param, dirs):
    """Store one stack's estimate in ``param`` as estimateParameters does
    (SimAlgorithm.java:139-142, 155-158): setPxPy, setPhaOff, setModulation."""
    for d, r in enumerate(dirs):
        dd = param.dir(d)
        dd.setPxPy(r["px"], r["py"])
        dd.setPhaOff(r["pha_off"])
        dd.setModulation(1, r["mod1"])
        if "mod2" in r:
            dd.setModulation(2, r["mod2"])


def input_ffts(images, fade_px=10):
    """The estimation input as the reference's GUI builds it
    (sim_gui/ImageControl.java:816-840): each float image faded at the border
    (fadeBorderCos), copied to a complex vector and transformed.  images:
    float array [D, P, N, N] on the host; returns inFFT[d][p] device vectors."""
    images = np.asarray(images, dtype=np.float32)
    D, P, h, w = images.shape
    out = []
    for d in range(D):
        row = []
        for p in range(P):
            r = Vec2dReal(w, h)
            r.setData(images[d, p])
            fade_border_cos(r, fade_px)
            c = Vec2dCplx(w, h)
            c.copy(r)
            c.fft2d(False)
            row.append(c)
        out.append(row)
    return out


def paste_and_fourier_shift(inV, outV, kx, ky, fast=True):
    """SimUtils.pasteAndFourierShift (SimUtils.java:64-91) on device vectors."""
    if not fast:
        outV.pasteFreq(inV, 0, 0)
        outV.fft2d(True)
        outV.fourierShift(kx, -ky)
        outV.fft2d(False)
    else:
        xf = kx - math.floor(kx)
        yf = -ky - math.floor(-ky)
        inV.fft2d(True)
        inV.fourierShift(xf, yf)
        inV.fft2d(False)
        outV.pasteFreq(inV, int(math.floor(kx)), int(math.floor(-ky)))


class RLDeconvolution:
    """sim_algorithm/RLDeconvolution.java on the device (fsim_rl_deconvolve)."""

    @staticmethod
    def deconvolve(img, otf, steps, inputIsInFreqSpace=False):
        """Richardson-Lucy iterations on ``img`` in place (RLDeconvolution.java:43-103)."""
        call("fsim_rl_deconvolve", img.handle, otf.handle, int(steps), 1 if inputIsInFreqSpace else 0)


FILTER_STYLES = ("Wiener", "RLin", "RLout", "RLboth")   # SimParam.FilterStyle (SimParam.java:33-43)


def spatial(vec, clipScale="NONE"):
    """SimUtils.spatial (SimUtils.java:124-153): inverse FFT of a copy, real
    part; clipScale CLIP or BOTH runs clipAndScale (clip, or scale and clip)."""
    tmp = vec.duplicate()
    tmp.fft2d(True)
    out = Vec2dReal(vec.vectorWidth(), vec.vectorHeight())
    out.copy(tmp)
    if clipScale in ("CLIP", "BOTH"):
        call("fsim_clip_scale", out.handle, 1, 1 if clipScale == "BOTH" else 0)
    return out


def run_reconstruction(param, inFFT, otfBeforeShift=False, clipScale="NONE", filterStyle="Wiener",
                       rlIterations=5, wiener=0.05, apoCutoff=2.0):
    """SimAlgorithm.runReconstruction (SimAlgorithm.java:268-709) without the
    intermediate displays: slow-shift reconstruction of FFT'd inputs
    inFFT[d][p] (device vectors) with the filter style of SimParam
    (Wiener, RL on input, RL on output, RL on both) and the otfBeforeShift
    switch.  Returns the (2N)^2 real result on the device.

    As in the reference: with otfBeforeShift the shifted bands are masked by
    maskOtf(shifted[i], angIdx, i), i.e. around the pixel (angIdx, i)
    (SimAlgorithm.java:408-410); RL on input deconvolves and separates the
    first 2*bands-1 phase images, so it needs nphases == 2*bands-1 (the
    reference fails with an index exception otherwise)."""
    if filterStyle not in FILTER_STYLES:
        raise ValueError(f"filter style {filterStyle} not in {FILTER_STYLES}")
    useWiener = filterStyle in ("Wiener", "RLin")
    useRLin = filterStyle in ("RLin", "RLboth")
    useRLout = filterStyle in ("RLout", "RLboth")
    w, h = inFFT[0][0].vectorWidth(), inFFT[0][0].vectorHeight()
    apoB, apoF = .9, apoCutoff
    otfPr = param.otf()
    fullResult = Vec2dCplx(2 * w, 2 * h)
    inputOtf = None
    if useRLin:
        inputOtf = Vec2dCplx(w, h)
        otfPr.writeOtfVector(inputOtf, 0, 0, 0)
    for angIdx in range(param.nrDir()):
        par = param.dir(angIdx)
        B = par.nrBand()
        nc = 2 * B - 1
        separate = [Vec2dCplx(w, h) for _ in range(nc)]
        if useRLin:
            if len(par.getPhases()) != nc:
                raise IndexError("#input images != phases per band")
            tmp = [inFFT[angIdx][i].duplicate() for i in range(nc)]
            for t in tmp:
                RLDeconvolution.deconvolve(t, inputOtf, rlIterations, True)
            separate_bands(tmp, separate, par.getPhases(), B, par.getModulations())
        else:
            separate_bands(inFFT[angIdx], separate, par.getPhases(), B, par.getModulations())
        if otfBeforeShift and useWiener:
            for i in range(nc):
                otfPr.applyOtf(separate[i], (i + 1) // 2)
        shifted = [Vec2dCplx(2 * w, 2 * h) for _ in range(nc)]
        shifted[0].pasteFreq(separate[0], 0, 0)
        for b in range(1, B):
            pos, neg = 2 * b, 2 * b - 1
            paste_and_fourier_shift(separate[pos], shifted[pos], par.px(b), par.py(b), fast=False)
            paste_and_fourier_shift(separate[neg], shifted[neg], -par.px(b), -par.py(b), fast=False)
        if useWiener:
            if not otfBeforeShift:
                otfPr.applyOtf(shifted[0], 0)
                for b in range(1, B):
                    otfPr.applyOtf(shifted[2 * b], b, par.px(b), par.py(b))
                    otfPr.applyOtf(shifted[2 * b - 1], b, -par.px(b), -par.py(b))
            else:
                for i in range(nc):
                    otfPr.maskOtf(shifted[i], angIdx, i)
        if useRLout:
            shifted[0].scal(np.float32(1) / np.float32(param.nrDir()))
        for i in range(nc):
            fullResult.add(shifted[i])
    image = None
    if useWiener:
        denom = Vec2dReal(2 * w, 2 * h)
        wiener_denominator(denom, param, wiener)
        fullResult.times(denom)
        apo = Vec2dCplx(2 * w, 2 * h)
        otfPr.writeApoVector(apo, apoB, apoF)
        fullResult.times(apo)
        image = spatial(fullResult, clipScale)
    if useRLout:
        otfSim = Vec2dCplx(2 * w, 2 * h)
        pos_v, neg_v = Vec2dCplx(2 * w, 2 * h), Vec2dCplx(2 * w, 2 * h)
        for ang in range(param.nrDir()):
            par = param.dir(ang)
            for b in range(param.nrBand()):
                otfPr.writeOtfVector(pos_v, b, par.px(b), par.py(b))
                otfPr.writeOtfVector(neg_v, b, -par.px(b), -par.py(b))
                if b == 0:
                    pos_v.scal(np.float32(1) / np.float32(param.nrDir()))
                    otfSim.add(pos_v)
                else:
                    otfSim.add(pos_v)
                    otfSim.add(neg_v)
        RLDeconvolution.deconvolve(fullResult, otfSim, rlIterations, True)
        image = spatial(fullResult, clipScale)
    return image


class PlainPipeline:
    """Pre-allocated vectors and filters of one ReconstructionThread
    (ReconstructionRunner.java:380-413) for one channel; ``run`` issues the
    reconstruction loop body one vector op at a time."""

    def __init__(self, param, wiener=0.05, apo_bend=0.9, apo_cut=2.0, fade_px=10):
        vf = AccelVectorFactory.getFactory()
        self.param = param
        n = param.vectorWidth()
        D, P, B = param.nrDir(), param.dir(0).nrPha(), param.nrBand()
        Cc = 2 * B - 1
        self.n = n
        otf = param.otf()
        # FilterUpdateThread.setupFilters (ReconstructionRunner.java:773-813)
        self.damp = vf.createReal2D(n, n)
        self.damp.addConst(1.0)
        if fade_px > 0:
            fade_border_cos(self.damp, fade_px)
        self.otfV = vf.createCplx2D(n, n)
        otf.writeOtfWithAttVector(self.otfV, 0, 0, 0)
        self.wien = vf.createReal2D(2 * n, 2 * n)
        wiener_denominator(self.wien, param, wiener)
        self.apo = vf.createReal2D(2 * n, 2 * n)
        otf.writeApoVector(self.apo, apo_bend, apo_cut)
        self.inFFT = vf.createArrayCplx2D(D, P, n, n)
        self.sep = vf.createArrayCplx2D(D, Cc, n, n)
        self.shifted = vf.createArrayCplx2D(D, Cc, 2 * n, 2 * n)
        self.full = vf.createCplx2D(2 * n, 2 * n)
        self.result = vf.createReal2D(2 * n, 2 * n)

    def run(self, raw_dev_ptr):
        """One reconstruction of a device-resident raw stack (uint16,
        angle-major); the result stays in ``self.result``."""
        par = self.param
        n, D = self.n, par.nrDir()
        P, B = par.dir(0).nrPha(), par.nrBand()
        self.full.zero()
        frame = n * n * 2
        for a in range(D):
            for p in range(P):
                v = self.inFFT[a][p]
                v.setFrom16bitPixelsDevice(raw_dev_ptr + (a * P + p) * frame)
                v.times(self.damp)
                v.fft2d(False)
        for a in range(D):
            d = par.dir(a)
            separate_bands(self.inFFT[a], self.sep[a], d.getPhases(), B, d.getModulations())
            for c in range(2 * B - 1):
                self.sep[a][c].timesConj(self.otfV)
            self.shifted[a][0].pasteFreq(self.sep[a][0])
            for b in range(1, B):
                pos, neg = 2 * b, 2 * b - 1
                paste_and_fourier_shift(self.sep[a][pos], self.shifted[a][pos], d.px(b), d.py(b), True)
                paste_and_fourier_shift(self.sep[a][neg], self.shifted[a][neg], -d.px(b), -d.py(b), True)
            for c in range(2 * B - 1):
                self.full.add(self.shifted[a][c])
        self.full.times(self.wien)
        self.full.times(self.apo)
        self.full.fft2d(True)
        self.result.copy(self.full)
        return self.result


def sim_params(param):
    """Pack the per-direction SIM parameters of a SimParam (pX, pY,
    getPhases, modulations) into fsim_sim_params."""
    p = _lib.SIM_PARAMS()
    for d in range(param.nrDir()):
        dd = param.dir(d)
        p.px[d], p.py[d] = dd.pX, dd.pY
        for i, ph in enumerate(dd.getPhases()):
            p.phases[d][i] = ph
        for b in range(param.nrBand()):
            p.modul[d][b] = dd.modul[b]
    return p


def params_from_estimate(param, dirs):
    """fsim_sim_params of one stack from its estimate (estimate_batch's
    per-direction dicts) on top of ``param``'s phases and modulations, as
    estimateParameters leaves a SimParam (setPxPy, setPhaOff, setModulation)."""
    p = sim_params(param)
    B = param.nrBand()
    for d, r in enumerate(dirs):
        dd = param.dir(d)
        p.px[d], p.py[d] = r["px"] / (B - 1), r["py"] / (B - 1)
        for i, ph in enumerate(dd.phases):
            p.phases[d][i] = ph + r["pha_off"]
        p.modul[d][1] = r["mod1"]
        if B == 3 and "mod2" in r:
            p.modul[d][2] = r["mod2"]
    return p


def recon_params(param, wiener=0.05, apo_bend=0.9, apo_cut=2.0, fade_px=10, mode=0, otf_band0=False):
    """Pack a SimParam into fsim_recon_params."""
    p = RECON_PARAMS()
    p.n = param.vectorWidth()
    p.ndir = param.nrDir()
    p.nphase = param.dir(0).nrPha()
    p.nband = param.nrBand()
    if p.ndir > 5 or p.nphase > 7 or p.nband > 3:
        raise _lib.FsimError(_lib.FSIM_ERR_ARG, "at most 5 directions, 7 phases, 3 bands")
    for d in range(p.ndir):
        dd = param.dir(d)
        p.px[d], p.py[d] = dd.pX, dd.pY
        for i, ph in enumerate(dd.getPhases()):
            p.phases[d][i] = ph
        for b in range(p.nband):
            p.modul[d][b] = dd.modul[b]
    p.vec_cpm = param.pxlSizeCyclesMicron()
    p.use_att = 1 if param.otf().isAttenuate() else 0
    p.wiener = wiener
    p.apo_bend = apo_bend
    p.apo_cut = apo_cut
    p.fade_px = fade_px
    p.mode = mode
    p.otf_band0 = 1 if otf_band0 else 0
    return p


class Reconstructor:
    """Fused reconstruction of one channel (fsim_channel): tables resident on
    the device, ``reconstruct`` runs a batch of stacks in five passes."""

    def __init__(self, param, wiener=0.05, apo_bend=0.9, apo_cut=2.0, fade_px=10, mode=0, otf_band0=False):
        """mode 0 derives each band's neg component from its pos component
        when the OTF is real (exact identity); mode 1 computes all.
        otf_band0: band 0's OTF for every component (the live runner)."""
        _lib.init()
        self.param = param
        self.n = param.vectorWidth()
        self.D, self.P = param.nrDir(), param.dir(0).nrPha()
        self.p = recon_params(param, wiener, apo_bend, apo_cut, fade_px, mode, otf_band0)
        self._h = C.c_void_p()
        call("fsim_channel_create", C.byref(self.p), param.otf().device(), C.byref(self._h))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            lib().fsim_channel_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def raw_bytes(self):
        return self.D * self.P * self.n * self.n * 2

    def out_bytes(self):
        return 4 * self.n * self.n * 4

    def reconstruct(self, raw_dev, out_dev, count=1, stream=None):
        """raw_dev / out_dev: device addresses (int) of count stacks/images."""
        call("fsim_reconstruct", self._h, C.c_void_p(int(raw_dev)), C.c_void_p(int(out_dev)),
             int(count), C.c_void_p(stream) if stream else None)

    def reconstruct_params(self, raw_dev, out_dev, params, stream=None):
        """fsim_reconstruct_params: len(params) stacks, stack i with the SIM
        parameters params[i] (a SimParam, or an fsim_sim_params; or a ctypes
        array of fsim_sim_params)."""
        count = len(params)
        if isinstance(params, C.Array):
            arr = params
        else:
            arr = (_lib.SIM_PARAMS * count)()
            for i, prm in enumerate(params):
                arr[i] = prm if isinstance(prm, _lib.SIM_PARAMS) else sim_params(prm)
        call("fsim_reconstruct_params", self._h, C.c_void_p(int(raw_dev)), C.c_void_p(int(out_dev)), count, arr,
             C.c_void_p(stream) if stream else None)

    def widefield(self, raw_dev, out_dev, count=1, raw_output=0, stream=None):
        """fsim_widefield: the widefield images (n x n float) of count stacks."""
        call("fsim_widefield", self._h, C.c_void_p(int(raw_dev)), C.c_void_p(int(out_dev)), int(count),
             int(raw_output), C.c_void_p(stream) if stream else None)

    def reconstruct_host(self, raw):
        """Convenience: host uint16 stack(s) in, host float32 image(s) out."""
        raw = np.ascontiguousarray(raw, dtype=np.uint16)
        count = raw.size // (self.D * self.P * self.n * self.n)
        rb = DeviceBuffer(raw.nbytes)
        ob = DeviceBuffer(count * self.out_bytes())
        rb.upload(raw)
        self.reconstruct(rb.value, ob.value, count)
        call("fsim_sync")
        out = ob.download((count, 2 * self.n, 2 * self.n), np.float32)
        rb.free()
        ob.free()
        return out[0] if count == 1 else out
