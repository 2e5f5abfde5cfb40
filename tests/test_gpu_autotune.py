"""GPU test of the autotuner: every candidate launch shape gives the same
image (to FP32 round-off: the candidates differ in FFT radix), the tuned
choice is one of them and the tuned pipeline matches the FP32 oracle."""
import numpy as np
import pytest

from oracle import fp64, ref

pytestmark = pytest.mark.gpu


@pytest.fixture
def tuned_channel(fsim_lib):
    from fsim import autotune, synthetic
    from fsim.sim_algorithm import Reconstructor
    autotune.clear()
    n = 128
    raw, sp = synthetic.sim_stack(n=n, seed=95, nbeads=60, nfil=8)
    rec = Reconstructor(synthetic.make_param(sp, n))
    yield n, raw, sp, rec
    autotune.clear()


@pytest.mark.parametrize("n", [1024, 2048])
def test_candidates_agree_large(fsim_lib, n):
    """At the sizes whose column transforms take 16 elements per thread by
    default, the 8-element candidate runs four Stockham passes."""
    from fsim import autotune, synthetic
    from fsim.sim_algorithm import Reconstructor
    autotune.clear()
    raw, sp = synthetic.sim_stack(n=n, seed=96, nbeads=300 * n // 512, nfil=40 * n // 512)
    rec = Reconstructor(synthetic.make_param(sp, n))
    cls = autotune.class_of(rec)
    default = rec.reconstruct_host(raw)
    try:
        for col in range(8):
            print(f"[autotune] n={n} col={col}", flush=True)
            autotune.set_choice(*cls, 0, col)
            e = fp64.rel_l2(rec.reconstruct_host(raw), default)
            print(f"  vs default {e:.3e}", flush=True)
            assert e < 5e-6
    finally:
        autotune.clear()


def test_candidates_agree_and_tuned_matches_oracle(tuned_channel):
    from fsim import autotune
    n, raw, sp, rec = tuned_channel
    cls = autotune.class_of(rec)
    assert cls == (n, 3, 3, 2, 1)
    default = rec.reconstruct_host(raw)
    for sep in (0, 1):
        for col in range(8):
            print(f"[autotune] n={n} sep={sep} col={col}", flush=True)
            autotune.set_choice(*cls, sep, col)
            got = rec.reconstruct_host(raw)
            e = fp64.rel_l2(got, default)
            print(f"  vs default {e:.3e}")
            assert e < 5e-6
    autotune.clear()
    res = autotune.tune(rec, count=4, reps=2)
    print(res)
    assert res["candidates"] >= 2 and 0 <= res["col"] < 8 and res["sep"] in (0, 1)
    assert res["ms_best"] > 0 and res["ms_best"] <= res["ms_default"]
    assert autotune.get_choice(*cls) == (res["sep"], res["col"])
    got = rec.reconstruct_host(raw)
    assert fp64.rel_l2(got, ref.recon32(ref.recon_dict(sp, n), raw)) < 1e-4


def test_init_loads_cache_named_by_environment(fsim_lib, tmp_path):
    """fsim_init merges the cache file named by FSIM_AUTOTUNE_CACHE (a fresh
    process: the library initialises once per process)."""
    import os
    import subprocess
    import sys
    path = tmp_path / "tune.txt"
    path.write_text("fsim-autotune 1\n256 3 3 2 1 0 5 16 0.25\n")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import sys; sys.path[:0] = [%r, %r]\n"
            "from fsim import _lib, autotune\n"
            "_lib.init(0)\n"
            "print('choice', autotune.get_choice(256, 3, 3, 2, 1))\n") % (root, os.path.join(root, "fairsim-vigor_amd"))
    env = dict(os.environ, FSIM_AUTOTUNE_CACHE=str(path))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr[-500:])
    assert "choice (0, 5)" in out.stdout


def test_chunked_calls_bitwise_equal_one_launch(fsim_lib):
    """A call's stacks run as chunks on library streams (each stream with its
    own work buffers): for every chunking, ragged last chunk included, the
    images are bitwise those of one launch; two channels interleave on the
    shared chunk streams."""
    from fsim import autotune, synthetic
    from fsim._lib import call
    from fsim.linalg import DeviceBuffer
    from fsim.sim_algorithm import Reconstructor
    autotune.clear()
    n, count = 256, 21
    raws = [synthetic.sim_stack(n=n, seed=120 + i, nbeads=150, nfil=20)[0] for i in range(3)]
    sp = synthetic.sim_stack(n=n, seed=120, nbeads=150, nfil=20)[1]
    recs = [Reconstructor(synthetic.make_param(sp, n)), Reconstructor(synthetic.make_param(sp, n, lam=600.0))]
    cls = autotune.class_of(recs[0])
    stk = np.stack([raws[i % 3] for i in range(count)])
    rb = DeviceBuffer(stk.nbytes)
    rb.upload(stk)
    obs = [DeviceBuffer(count * recs[0].out_bytes()) for _ in recs]
    try:
        autotune.set_chunks(*cls, 0, 1)
        assert autotune.chunk_plan(*cls, count) == (count, 1)
        want = []
        for r, ob in zip(recs, obs):
            r.reconstruct(rb.value, ob.value, count)
            call("fsim_sync")
            want.append(ob.download((count, 2 * n, 2 * n), np.float32))
        for chunk, streams in [(8, 2), (4, 3), (5, 4), (16, 1), (1, 2), (-1, 2)]:
            autotune.set_chunks(*cls, chunk, streams)
            print(f"[chunks] {chunk} x {streams}: plan {autotune.chunk_plan(*cls, count)}", flush=True)
            for ob in obs:
                ob.upload(np.full((count, 2 * n, 2 * n), np.nan, np.float32))
            for rep in range(2):
                for r, ob in zip(recs, obs):
                    r.reconstruct(rb.value, ob.value, count)
            call("fsim_sync")
            for i, ob in enumerate(obs):
                got = ob.download((count, 2 * n, 2 * n), np.float32)
                assert np.array_equal(got, want[i]), f"chunk {chunk} streams {streams} channel {i}"
        # the chunked image is the single-stack image
        assert np.array_equal(want[0][4], recs[0].reconstruct_host(raws[1]))
    finally:
        autotune.clear()


def test_tune_reports_chunking(fsim_lib):
    from fsim import autotune, synthetic
    from fsim.sim_algorithm import Reconstructor
    autotune.clear()
    n = 256
    raw, sp = synthetic.sim_stack(n=n, seed=121, nbeads=150, nfil=20)
    rec = Reconstructor(synthetic.make_param(sp, n))
    cls = autotune.class_of(rec)
    try:
        res = autotune.tune(rec, count=32, reps=2)
        print(res)
        assert res["chunk"] >= 0 and 1 <= res["streams"] <= 4
        assert res["ms_best"] <= res["ms_chunk_default"] and res["ms_whole"] > 0
        k, s = autotune.chunk_plan(*cls, 32)
        assert (k, s) == ((32, 1) if res["chunk"] == 0 else (res["chunk"], min(res["streams"], -(-32 // res["chunk"]))))
        assert np.array_equal(rec.reconstruct_host(np.stack([raw] * 3))[1], rec.reconstruct_host(raw))
    finally:
        autotune.clear()
