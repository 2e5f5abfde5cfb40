"""CPU restatement of the live-mode side outputs -- test infrastructure only.

widefield32: ReconstructionRunner.java:452-481 (setFrom16bitPixels, times
dampBorder, add in (direction, phase) order, scal(1.f / (D*P)), real part).
register32: Registration.registerImage / transPixel (Registration.java:
196-268).  NumPy float32 operations, each rounded to float32 as the Java
float arithmetic is; the dampBorder vector comes from the Java-order C
restatement of fadeBorderCos (java_order.c).
"""
import numpy as np

from . import ref


def damp_border(n, fade_px):
    """fadeBorderCos applied to a vector of ones (ReconstructionRunner.java:775-780)."""
    d = np.ones((n, n), np.float32)
    if fade_px > 0:
        ref.lib().ref_fade_border(ref.ptr(d), n, n, fade_px)
    return d


def widefield32(raw, ndir, nphase, fade_px, raw_output=0):
    """raw: [ndir*nphase, n, n] uint16 of one stack."""
    nf = ndir * nphase
    if raw_output > 0:
        return raw[raw_output - 1].astype(np.float32)
    d = damp_border(raw.shape[-1], fade_px)
    acc = np.zeros(raw.shape[1:], np.float32)
    for f in range(nf):
        acc = acc + raw[f].astype(np.float32) * d
    return acc * np.float32(np.float32(1.0) / np.float32(nf))


def register32(src, xt, yt):
    """transPixel for every target pixel; xt, yt: tables (th, tw) in which the
    (h, w) image is centred."""
    h, w = src.shape
    th, tw = xt.shape
    ox, oy = (tw - w) // 2, (th - h) // 2
    out = np.zeros((h, w), np.float32)
    for y in range(h):
        for x in range(w):
            sx = np.float32(xt[y + oy, x + ox]) - np.float32(ox)
            sy = np.float32(yt[y + oy, x + ox]) - np.float32(oy)
            if sx >= 0 and sy >= 0 and sx < w - 1 and sy < h - 1:
                left, top = int(sx), int(sy)
                right, bottom = left + 1, top + 1
                xf = sx - np.float32(left)
                yf = sy - np.float32(top)
                one = np.float32(1)
                tlf, trf = (one - xf) * (one - yf), xf * (one - yf)
                blf, brf = (one - xf) * yf, xf * yf
                v = src[top, left] * tlf + src[top, right] * trf
                v = v + src[bottom, left] * blf
                v = v + src[bottom, right] * brf
                out[y, x] = v
    return out
