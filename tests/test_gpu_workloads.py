"""GPU parity at the benchmark workloads' own shapes, through the C ABI,
against the FP32 Java-order oracle (relative L2 <= 1e-4; coarse peaks and
fitted grid positions exactly):

* 512^2 parameter estimation (SimAlgorithm.estimateParameters,
  SimAlgorithm.java:37-256), 3 phases / 2 bands fitting band 1 and 5 phases /
  3 bands fitting band 2, batched (fsim_estimate);
* a mixed batch of 256^2, 512^2 and 1024^2 stacks over three emission
  channels (525 / 600 / 670 nm, each with its own OTF and k0), one fixed-shape
  batch replayed as a hipGraph (fsim_batch_*), with per-class launches and
  through the persistent work-list kernels;
* a served Poisson stream of those sizes and channels (fsim_server_*), the
  concurrency ReconstructionRunner's threads model
  (ReconstructionRunner.java:243-251);
* 2048^2 raw stacks at 65 nm raw pixels (k0 = 603.5 px at 525 nm), plain and
  with the k0 fit on the first timepoint applied to the next one.

Every image is also compared bitwise with fsim_reconstruct of the same stack
alone: batch composition and concurrency do not change results.
"""
import time

import numpy as np
import pytest

from oracle import fp64, ref

pytestmark = pytest.mark.gpu

TOL = 1e-4
NA, A_OTF, PX_UM = 1.4, 1.0, 0.08
LAMS = (525.0, 600.0, 670.0)


def _otf(n, lam=525.0, px_um=PX_UM):
    from fsim.sim_algorithm import OtfProvider
    o = OtfProvider.fromEstimate(NA, lam, A_OTF)
    o.setPixelSize(1 / (n * px_um))
    return o


@pytest.mark.parametrize("nphase,nband,fit_band,k_fracs,h2",
                         [(3, 2, 1, (0.85, 0.78), 0.0), (5, 3, 2, (0.45, 0.42), 0.3)])
def test_estimate_512_matches_oracle(fsim_lib, nphase, nband, fit_band, k_fracs, h2):
    from fsim import synthetic
    from fsim.sim_algorithm import SimParam, estimate_batch
    n = 512
    imgs, sps = [], []
    for s, kf in enumerate(k_fracs):
        raw, sp = synthetic.sim_stack(n=n, nphase=nphase, seed=200 + s, nbeads=300, nfil=40, harmonic2=h2,
                                      k_frac=kf)
        imgs.append(raw.reshape(3, nphase, n, n).astype(np.float32))
        sps.append(sp)
    images = np.stack(imgs)
    param = SimParam.create(nband, 3, nphase, n, PX_UM, _otf(n))
    print(f"[estimate] n={n} P={nphase} B={nband} fit band {fit_band}, {len(imgs)} stacks", flush=True)
    got = estimate_batch(param, fit_band, 0.6, images=images)
    scale = 1 if nband == 2 else 2
    for s in range(len(imgs)):
        want = ref.estimate32(ref.input_ffts32(images[s]), nband, fit_band, 0.6, ref.Otf32(NA, 525.0, A_OTF),
                              1 / (n * PX_UM))
        for d in range(3):
            g, w = got[s][d], want[d]
            print(f"stack {s} dir {d}: gpu {g} oracle {w}")
            assert g["peak"] == (w["peak"][0], w["peak"][1])     # coarse peak pixel: exact
            assert g["px"] == w["px"] and g["py"] == w["py"]      # fitted grid point: exact
            assert abs(g["pha_off"] - w["pha_off"]) < 1e-4
            assert abs(g["mod1"] - w["mod1"]) <= 1e-4 * w["mod1"]
            if nband == 3:
                assert abs(g["mod2"] - w["mod2"]) <= 1e-4 * w["mod2"]
            # the fit recovers the pattern (known answer)
            assert abs(g["px"] - scale * sps[s]["px"][d]) < 0.1 and abs(g["py"] - scale * sps[s]["py"][d]) < 0.1


@pytest.fixture(scope="module")
def mixed(fsim_lib):
    """Stacks of 256^2, 512^2 (two per channel) and 1024^2 (one per channel)
    for three emission channels; a channel per (size, wavelength); the
    single-call image and the FP32 oracle's image of every stack."""
    from fsim import synthetic
    from fsim.sim_algorithm import Reconstructor
    stacks, recs, single, oracle = {}, {}, {}, {}
    for n in (256, 512, 1024):
        for i, lam in enumerate(LAMS):
            for t in range(1 if n == 1024 else 2):
                raw, sp = synthetic.sim_stack(n=n, seed=500 + 37 * i + 7 * t + n // 256, lam=lam,
                                              nbeads=300 * n // 512, nfil=40 * n // 512)
                stacks[(n, lam, t)] = raw
                if (n, lam) not in recs:
                    recs[(n, lam)] = Reconstructor(synthetic.make_param(sp, n, lam=lam))
                oracle[(n, lam, t)] = ref.recon32(ref.recon_dict(sp, n, lam=lam), raw, threads=8)
    for key, raw in stacks.items():
        single[key] = recs[key[:2]].reconstruct_host(raw)
        e = fp64.rel_l2(single[key], oracle[key])
        print(f"[mixed] n={key[0]} lam={key[1]} t={key[2]}: single call vs ref32 {e:.3e}", flush=True)
        assert e < TOL
    return stacks, recs, single, oracle


@pytest.mark.parametrize("persistent", [False, True])
def test_mixed_size_batch_three_channels(mixed, persistent):
    """One fixed-shape batch of 15 distinct stacks (256^2 / 512^2 / 1024^2 x
    525 / 600 / 670 nm) in mixed slot order, replayed as a graph with the
    slots' stacks permuted between replays, with per-class launches or
    through the persistent work-list kernels; every slot bitwise equal to its
    stack's single call and within 1e-4 of the oracle."""
    from fsim._lib import call
    from fsim.linalg import DeviceBuffer
    from fsim.livemode import ReconstructionBatch
    stacks, recs, single, oracle = mixed
    keys = sorted(stacks, key=lambda k: (k[2], k[1], -k[0]))
    rng = np.random.default_rng(3)
    order = [keys[i] for i in rng.permutation(len(keys))]
    bufs = {}
    for k in order:
        bufs[k] = DeviceBuffer(stacks[k].nbytes)
        bufs[k].upload(stacks[k])
    outs = [DeviceBuffer(16 * k[0] * k[0]) for k in order]
    b = ReconstructionBatch([recs[k[:2]] for k in order], use_graph=True, persistent=persistent)
    for rep in range(3):
        # slots of one (size, channel) exchange their stacks between replays
        perm = list(order)
        if rep % 2 == 1:
            for i, k in enumerate(order):
                if k[2] == 0 and (k[0], k[1], 1) in stacks:
                    perm[i] = (k[0], k[1], 1)
                elif k[2] == 1:
                    perm[i] = (k[0], k[1], 0)
        print(f"[batch] persistent={persistent} replay {rep}: {len(order)} slots", flush=True)
        b.run([bufs[k].value for k in perm], [o.value for o in outs])
        call("fsim_sync")
        for i, k in enumerate(perm):
            got = outs[i].download((2 * k[0], 2 * k[0]), np.float32)
            assert np.array_equal(got, single[k]), f"replay {rep} slot {i} {k}"
            assert fp64.rel_l2(got, oracle[k]) < TOL
    info = b.info()
    assert info["captured"] and info["classes"] == 3
    b.close()


def test_served_poisson_mix_three_channels(mixed):
    """A Poisson stream (2000 req/s mean rate, so requests overlap on the
    device) of 256^2 / 512^2 / 1024^2 requests (50 / 35 / 15 %) from three
    channels through the continuous-batching server: every result bitwise its
    stack's single call, none refused or failed."""
    from fsim.livemode import ReconstructionServer
    stacks, recs, single, oracle = mixed
    srv = ReconstructionServer(mem_budget=8 << 30, max_batch=16, max_queue=1024, nslots=4)
    ids = {k: srv.add_channel(r) for k, r in recs.items()}
    rng = np.random.default_rng(11)
    sizes = rng.choice([256, 512, 1024], size=90, p=[0.5, 0.35, 0.15])
    gaps = rng.exponential(1 / 2000.0, size=90)
    tickets = []
    t0 = time.perf_counter()
    due = t0
    for i, n in enumerate(sizes):
        lam = LAMS[rng.integers(3)]
        t = rng.integers(1 if n == 1024 else 2)
        key = (int(n), lam, int(t))
        due += gaps[i]
        while time.perf_counter() < due:
            pass
        out = np.zeros((2 * n, 2 * n), np.float32)
        tickets.append((key, srv.submit(ids[key[:2]], stacks[key], out), out))
    lats = [srv.wait(tk, timeout_ms=120000) for _, tk, _ in tickets]
    print(f"[served] {len(tickets)} requests in {1e3 * (time.perf_counter() - t0):.1f} ms, "
          f"latency p50 {np.percentile(lats, 50):.2f} ms p99 {np.percentile(lats, 99):.2f} ms", flush=True)
    for key, _, out in tickets:
        assert np.array_equal(out, single[key]), f"{key}"
    st = srv.stats()
    assert st["completed"] == len(tickets) and st["failed"] == 0 and st["rejected"] == 0
    srv.close()


def test_large_fov_2048_at_65nm(fsim_lib):
    """2048^2 raw stacks at 65 nm raw pixels (4096^2 output): the fused
    reconstruction with the pattern's parameters against the oracle; then the
    k0 / phase fit on the first timepoint (fsim_estimate) and both timepoints
    reconstructed with that fit (fsim_reconstruct_params, as the served
    large-FOV requests do), against the oracle with the same parameters."""
    from fsim import synthetic
    from fsim._lib import call
    from fsim.linalg import DeviceBuffer
    from fsim.sim_algorithm import Reconstructor, estimate_batch, params_from_estimate
    n, px = 2048, 0.065
    raws, sps = [], []
    for t in range(2):
        raw, sp = synthetic.sim_stack(n=n, px_um=px, seed=640 + t, nbeads=1200, nfil=160)
        raws.append(raw)
        sps.append(sp)
    assert abs(sps[0]["k0"] - 603.5) < 0.1
    param = synthetic.make_param(sps[0], n, px_um=px)
    rec = Reconstructor(param)
    print(f"[large] n={n} px={px} um: fused reconstruction", flush=True)
    got = rec.reconstruct_host(raws[0])
    p = ref.recon_dict(sps[0], n, px_um=px)
    e = fp64.rel_l2(got, ref.recon32(p, raws[0], threads=16))
    print(f"[large] relL2 gpu-vs-ref32 {e:.3e}", flush=True)
    assert e < TOL
    # fit on timepoint 0, both timepoints reconstructed with it
    img0 = raws[0].reshape(1, 3, 3, n, n).astype(np.float32)
    est = estimate_batch(param, 1, 0.6, images=img0)
    for d in range(3):
        g = est[0][d]
        print(f"[large] fit dir {d}: {g} true ({sps[0]['px'][d]:.3f}, {sps[0]['py'][d]:.3f})", flush=True)
        assert abs(g["px"] - sps[0]["px"][d]) < 0.1 and abs(g["py"] - sps[0]["py"][d]) < 0.1
    sim = params_from_estimate(param, est[0])
    stk = np.stack(raws)
    rb = DeviceBuffer(stk.nbytes)
    rb.upload(stk)
    ob = DeviceBuffer(2 * rec.out_bytes())
    rec.reconstruct_params(rb.value, ob.value, [sim, sim])
    call("fsim_sync")
    out = ob.download((2, 2 * n, 2 * n), np.float32)
    rb.free()
    ob.free()
    q = dict(p)
    q["px"] = [sim.px[d] for d in range(3)]
    q["py"] = [sim.py[d] for d in range(3)]
    q["phases"] = [[sim.phases[d][i] for i in range(3)] for d in range(3)]
    q["modul"] = [[sim.modul[d][b] for b in range(2)] for d in range(3)]
    for t in range(2):
        e = fp64.rel_l2(out[t], ref.recon32(q, raws[t], threads=16))
        print(f"[large] timepoint {t} with the t=0 fit: relL2 gpu-vs-ref32 {e:.3e}", flush=True)
        assert e < TOL
