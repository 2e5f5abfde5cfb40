"""Live-mode side outputs: the widefield image (ReconstructionRunner.java:
452-481) and the registration remap (registration/Registration.java).

CPU: the oracle restatements (oracle/sideout_ref.py) are pinned by known
answers -- the identity table copies the image except its last row and
column (the reference's strict source bounds), an integer table translates
it, a half-pixel table averages neighbours; equal frames average to the
frame times the fade; raw_output picks the frame -- and the registration
file reader follows the reference's format.  GPU: fsim_widefield and
fsim_register_images equal the oracle bit for bit.
"""
import numpy as np
import pytest

from oracle import sideout_ref as so


def _grid(h, w):
    y, x = np.mgrid[0:h, 0:w]
    return x.astype(np.float32), y.astype(np.float32)


def test_register_oracle_known_answers():
    rng = np.random.default_rng(1)
    src = rng.uniform(0, 100, (12, 10)).astype(np.float32)
    xt, yt = _grid(12, 10)
    out = so.register32(src, xt, yt)
    assert np.array_equal(out[:-1, :-1], src[:-1, :-1])
    assert not out[-1].any() and not out[:, -1].any()
    # integer translation by (2, 1): target (x, y) reads (x + 2, y + 1)
    out = so.register32(src, xt + 2, yt + 1)
    assert np.array_equal(out[:-2, :-3], src[1:-1, 2:-1])
    # half a pixel in x: the mean of horizontal neighbours
    out = so.register32(src, xt + 0.5, yt)
    np.testing.assert_allclose(out[:-1, :-2], 0.5 * (src[:-1, :-2] + src[:-1, 1:-1]), rtol=1e-6)
    # an image smaller than the tables sits in their centre
    big_x, big_y = _grid(16, 14)
    out = so.register32(src, big_x, big_y)
    assert np.array_equal(out[:-1, :-1], src[:-1, :-1])


def test_widefield_oracle_known_answers():
    n, fade = 32, 4
    raw = np.full((9, n, n), 1000, np.uint16)
    wf = so.widefield32(raw, 3, 3, fade)
    d = so.damp_border(n, fade)
    np.testing.assert_allclose(wf, 1000 * d, rtol=1e-6, atol=1e-3)
    assert wf[0, 5] == 0 and wf[n // 2, n // 2] == pytest.approx(1000, rel=1e-6)
    # the border follows sin^2 of the fade (SimUtils.java:162-185)
    np.testing.assert_allclose(d[fade // 2, n // 2], np.sin(fade // 2 * np.pi / 2 / fade) ** 2, rtol=1e-6)
    raw = np.random.default_rng(2).integers(0, 65535, (9, n, n)).astype(np.uint16)
    assert np.array_equal(so.widefield32(raw, 3, 3, fade, raw_output=4), raw[3].astype(np.float32))
    np.testing.assert_allclose(so.widefield32(raw, 3, 3, 0), raw.astype(np.float64).mean(0), rtol=2e-6)


def test_registration_file_reader(tmp_path):
    from fsim.livemode import read_registration_file
    W, H = 6, 4
    rng = np.random.default_rng(3)
    xt = rng.uniform(-1, 7, (H, W)).astype(np.float32)
    yt = rng.uniform(-1, 5, (H, W)).astype(np.float32)
    lines = [f"Width={W}", f"Height={H}", "", "X Trans -----------------------------------"]
    lines += [" ".join(repr(float(v)) for v in row) for row in xt] + [""]
    lines += ["Y Trans -----------------------------------"]
    lines += [" " + " ".join(repr(float(v)) for v in row) + " " for row in yt] + [""]
    p = tmp_path / "ch0.txt"
    p.write_text("\n".join(lines) + "\n")
    gx, gy = read_registration_file(str(p))
    assert np.array_equal(gx, xt) and np.array_equal(gy, yt)


@pytest.mark.gpu
@pytest.mark.parametrize("n,fade,raw_output,nphase", [(64, 10, 0, 3), (128, 0, 0, 3), (64, 10, 5, 3),
                                                     (64, 10, 0, 5)])
def test_widefield_matches_oracle(fsim_lib, n, fade, raw_output, nphase):
    """3 x 3 frames, and 3 x 5 (more frames than one group of loads)."""
    from fsim import synthetic
    from fsim.linalg import DeviceBuffer
    from fsim.sim_algorithm import Reconstructor
    from fsim._lib import call
    kw = dict(nbeads=40, nfil=6, nphase=nphase)
    stacks = [synthetic.sim_stack(n=n, seed=s, **kw)[0] for s in range(3)]
    _, sp = synthetic.sim_stack(n=n, seed=0, **kw)
    rec = Reconstructor(synthetic.make_param(sp, n, nphase=nphase), fade_px=fade)
    raw = np.stack(stacks)
    rb, ob = DeviceBuffer(raw.nbytes), DeviceBuffer(3 * n * n * 4)
    rb.upload(raw)
    rec.widefield(rb.value, ob.value, 3, raw_output)
    call("fsim_sync")
    got = ob.download((3, n, n), np.float32)
    for s in range(3):
        assert np.array_equal(got[s], so.widefield32(stacks[s], 3, nphase, fade, raw_output)), s
    rb.free()
    ob.free()


@pytest.mark.gpu
@pytest.mark.parametrize("widefield,count", [(False, 5), (True, 5), (False, 67)])
def test_registration_matches_oracle(fsim_lib, widefield, count):
    from fsim.linalg import DeviceBuffer
    from fsim.livemode import Registration
    from fsim._lib import call
    rng = np.random.default_rng(5)
    TW, TH = 96, 80
    gx, gy = _grid(TH, TW)
    # a smooth warp: rotation by 1 degree, a small shear and sub-pixel shifts
    a = np.deg2rad(1.0)
    xt = (np.cos(a) * (gx - TW / 2) - np.sin(a) * (gy - TH / 2) + TW / 2 + 0.37 + 0.01 * gy).astype(np.float32)
    yt = (np.sin(a) * (gx - TW / 2) + np.cos(a) * (gy - TH / 2) + TH / 2 - 0.61).astype(np.float32)
    reg = Registration(xt, yt)
    # widefield tables: the even positions, halved
    txt = xt[::2, ::2] / np.float32(2) if widefield else xt
    tyt = yt[::2, ::2] / np.float32(2) if widefield else yt
    h, w = (36, 44) if widefield else (70, 90)   # smaller than the tables: centred
    # five images: a partial group (the kernel takes up to 64 images per
    # group and reads the tables once for it); 67: a full group and a
    # partial one
    src = rng.uniform(0, 1000, (count, h, w)).astype(np.float32)
    sb, db = DeviceBuffer(src.nbytes), DeviceBuffer(src.nbytes)
    sb.upload(src)
    reg.register(sb.value, db.value, w, h, count, widefield)
    call("fsim_sync")
    got = db.download((count, h, w), np.float32)
    for k in range(count):
        want = so.register32(src[k], txt, tyt)
        assert np.array_equal(got[k], want), k
        assert (want != 0).mean() > 0.8
    reg.close()
    sb.free()
    db.free()
