"""GPU tests of the continuous-batching server (fsim_server_*)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _channel(n, lam, seed):
    from fsim import synthetic
    from fsim.sim_algorithm import Reconstructor
    raw, sp = synthetic.sim_stack(n=n, seed=seed, lam=lam, nbeads=max(30, 300 * n // 512),
                                  nfil=max(5, 40 * n // 512))
    return raw, Reconstructor(synthetic.make_param(sp, n, lam=lam))


@pytest.fixture(scope="module")
def setup(fsim_lib):
    chans = {}
    for i, (n, lam) in enumerate([(64, 525.0), (128, 600.0), (128, 670.0), (256, 525.0)]):
        chans[(n, lam)] = _channel(n, lam, 40 + i)
    return chans


def test_result_same_alone_and_in_mix(setup):
    from fsim.livemode import ReconstructionServer
    srv = ReconstructionServer(mem_budget=2 << 30, max_batch=8, max_queue=256, nslots=3)
    ids = {k: srv.add_channel(v[1]) for k, v in setup.items()}
    target = (128, 600.0)
    raw_t, rec_t = setup[target]
    alone = np.zeros((256, 256), np.float32)
    srv.wait(srv.submit(ids[target], raw_t, alone))
    # the same request inside a stream of other sizes and channels
    rng = np.random.default_rng(0)
    keys = list(setup)
    outs, tickets = [], []
    for i in range(40):
        k = keys[rng.integers(len(keys))] if i != 17 else target
        raw = setup[k][0]
        o = np.zeros((2 * k[0], 2 * k[0]), np.float32)
        tickets.append((k, srv.submit(ids[k], raw, o), o))
    for k, t, o in tickets:
        srv.wait(t)
    st = srv.stats()
    assert st["completed"] >= 41 and st["failed"] == 0
    assert st["mean_batch"] > 1.0          # requests were batched
    in_mix = [o for (k, t, o) in tickets if k == target]
    for o in in_mix:
        assert np.array_equal(o, alone)
    # and equal to the single-channel fused call
    assert np.array_equal(alone, rec_t.reconstruct_host(raw_t))
    for k, t, o in tickets:
        assert np.array_equal(o, setup[k][1].reconstruct_host(setup[k][0]))
    srv.close()


def test_queue_full_is_refused(setup):
    from fsim._lib import FsimBusy
    from fsim.livemode import ReconstructionServer
    srv = ReconstructionServer(mem_budget=1 << 30, max_batch=4, max_queue=3, nslots=1)
    raw, rec = setup[(256, 525.0)]
    cid = srv.add_channel(rec)
    outs = [np.zeros((512, 512), np.float32) for _ in range(200)]
    accepted, refused = [], 0
    for o in outs:
        try:
            accepted.append(srv.submit(cid, raw, o))
        except FsimBusy:
            refused += 1
    for t in accepted:
        srv.wait(t)
    st = srv.stats()
    assert refused > 0 and st["rejected"] == refused
    assert st["completed"] == len(accepted)
    t = srv.queueImage(cid, raw, outs[0])
    assert t is not False
    srv.wait(t)
    srv.close()


def test_small_memory_budget_evicts_and_completes(setup):
    """A budget that holds one slot at a time still serves every size class."""
    from fsim.livemode import ReconstructionServer
    srv = ReconstructionServer(mem_budget=64 << 20, max_batch=16, max_queue=64, nslots=2)
    ids = {k: srv.add_channel(v[1]) for k, v in setup.items()}
    res = []
    for i in range(12):
        k = list(setup)[i % len(setup)]
        o = np.zeros((2 * k[0], 2 * k[0]), np.float32)
        res.append((k, srv.submit(ids[k], setup[k][0], o), o))
    for k, t, o in res:
        assert srv.wait(t, timeout_ms=60000) >= 0
        assert np.array_equal(o, setup[k][1].reconstruct_host(setup[k][0]))
    st = srv.stats()
    assert st["device_bytes"] <= 64 << 20
    srv.close()


def test_device_resident_submit(setup):
    from fsim.linalg import DeviceBuffer
    from fsim.livemode import ReconstructionServer
    raw, rec = setup[(64, 525.0)]
    rb = DeviceBuffer(raw.nbytes)
    rb.upload(raw)
    ob = DeviceBuffer(128 * 128 * 4)
    srv = ReconstructionServer(mem_budget=256 << 20, max_batch=4, max_queue=16, nslots=2)
    cid = srv.add_channel(rec)
    srv.wait(srv.submit(cid, rb.value, ob.value, device=True))
    got = ob.download((128, 128), np.float32)
    assert np.array_equal(got, rec.reconstruct_host(raw))
    srv.close()
