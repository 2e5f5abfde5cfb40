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


def test_add_channel_warms_a_slot_within_budget(setup):
    """add_channel allocates a slot for its class when the budget has room
    (one per class, not per channel) and none when it has not."""
    from fsim.livemode import ReconstructionServer
    srv = ReconstructionServer(mem_budget=1 << 30, max_batch=4, max_queue=16, nslots=2)
    assert srv.stats()["device_bytes"] == 0
    srv.add_channel(setup[(128, 600.0)][1])
    one = srv.stats()["device_bytes"]
    assert one > 0
    srv.add_channel(setup[(128, 670.0)][1])     # same size class
    assert srv.stats()["device_bytes"] == one
    srv.add_channel(setup[(256, 525.0)][1])
    assert srv.stats()["device_bytes"] > one
    srv.close()
    tiny = ReconstructionServer(mem_budget=1 << 20, max_batch=4, max_queue=16, nslots=1)
    tiny.add_channel(setup[(256, 525.0)][1])
    assert tiny.stats()["device_bytes"] == 0
    tiny.close()


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


def _shifted_params(rec_param, scale):
    """The channel's parameters with the shifts scaled and the phases offset:
    a stand-in for a per-acquisition fit."""
    from fsim.sim_algorithm import sim_params
    p = sim_params(rec_param)
    for d in range(3):
        p.px[d] *= scale
        p.py[d] *= scale
        for i in range(3):
            p.phases[d][i] += 0.1 * (d + 1)
    return p


def test_params_request_equals_reconstruct_params(setup):
    """A request with its own parameters (3 timepoints) gives bitwise the
    images of fsim_reconstruct_params with those parameters repeated, from
    host and from device buffers, alone and among ordinary requests."""
    from fsim import synthetic
    from fsim.linalg import DeviceBuffer
    from fsim.livemode import ReconstructionServer
    n, lam = 128, 600.0
    raw0, rec = setup[(n, lam)]
    stacks = np.stack([raw0] + [synthetic.sim_stack(n=n, seed=300 + i, lam=lam, nbeads=60, nfil=8)[0]
                                for i in range(2)])
    sp = synthetic.sim_stack(n=n, seed=41, lam=lam, nbeads=60, nfil=8)[1]
    prm = _shifted_params(synthetic.make_param(sp, n, lam=lam), 1.01)
    rb = DeviceBuffer(stacks.nbytes)
    rb.upload(stacks)
    ob = DeviceBuffer(3 * rec.out_bytes())
    rec.reconstruct_params(rb.value, ob.value, [prm] * 3)
    from fsim._lib import call
    call("fsim_sync")
    exp = ob.download((3, 2 * n, 2 * n), np.float32)
    # each timepoint alone with the same parameters
    for t in range(3):
        o1 = DeviceBuffer(rec.out_bytes())
        rec.reconstruct_params(rb.value + t * stacks[0].nbytes, o1.value, [prm])
        call("fsim_sync")
        assert np.array_equal(o1.download((2 * n, 2 * n), np.float32), exp[t])
    srv = ReconstructionServer(mem_budget=2 << 30, max_batch=8, max_queue=256, nslots=3)
    ids = {k: srv.add_channel(v[1]) for k, v in setup.items()}
    out_h = np.zeros((3, 2 * n, 2 * n), np.float32)
    ob2 = DeviceBuffer(3 * rec.out_bytes())
    tickets = []
    for i, k in enumerate(list(setup) * 3):
        if i == 4:
            tickets.append(("h", srv.submit_params(ids[(n, lam)], stacks, out_h, 3, prm)))
        if i == 7:
            tickets.append(("d", srv.submit_params(ids[(n, lam)], rb.value, ob2.value, 3, prm, device=True)))
        o = np.zeros((2 * k[0], 2 * k[0]), np.float32)
        tickets.append((k, srv.submit(ids[k], setup[k][0], o), o))
    for tk in tickets:
        srv.wait(tk[1])
    for tk in tickets:
        if tk[0] not in ("h", "d"):
            assert np.array_equal(tk[2], setup[tk[0]][1].reconstruct_host(setup[tk[0]][0]))
    assert np.array_equal(out_h, exp)
    assert np.array_equal(ob2.download((3, 2 * n, 2 * n), np.float32), exp)
    st = srv.stats()
    assert st["failed"] == 0
    srv.close()


def test_params_request_larger_than_budget_is_refused(setup):
    from fsim._lib import FsimBusy
    from fsim.livemode import ReconstructionServer
    from fsim.sim_algorithm import sim_params
    from fsim import synthetic
    n, lam = 256, 525.0
    raw, rec = setup[(n, lam)]
    # 8 stacks of 256^2 need about 60 MB of work buffers
    srv = ReconstructionServer(mem_budget=32 << 20, max_batch=4, max_queue=16, nslots=1)
    cid = srv.add_channel(rec)
    sp = synthetic.sim_stack(n=n, seed=43, lam=lam, nbeads=30, nfil=5)[1]
    prm = sim_params(synthetic.make_param(sp, n, lam=lam))
    big = np.zeros((8,) + raw.shape, np.uint16)
    out = np.zeros((8, 2 * n, 2 * n), np.float32)
    with pytest.raises(FsimBusy):
        srv.submit_params(cid, big, out, 8, prm)
    srv.close()


def test_reconstruct_params_shares_tables_of_equal_sets(setup):
    """Runs of equal parameter sets share one set of tables; every stack's
    image is bitwise the one computed with its set alone."""
    from fsim import synthetic
    from fsim._lib import call
    from fsim.linalg import DeviceBuffer
    n, lam = 64, 525.0
    raw, rec = setup[(n, lam)]
    sp = synthetic.sim_stack(n=n, seed=40, lam=lam, nbeads=30, nfil=5)[1]
    base = synthetic.make_param(sp, n, lam=lam)
    p1, p2 = _shifted_params(base, 1.0), _shifted_params(base, 1.02)
    order = [p1, p1, p2, p2, p1]
    stacks = np.stack([raw] * 5)
    rb = DeviceBuffer(stacks.nbytes)
    rb.upload(stacks)
    ob = DeviceBuffer(5 * rec.out_bytes())
    rec.reconstruct_params(rb.value, ob.value, order)
    call("fsim_sync")
    got = ob.download((5, 2 * n, 2 * n), np.float32)
    for t, p in enumerate(order):
        o1 = DeviceBuffer(rec.out_bytes())
        rec.reconstruct_params(rb.value, o1.value, [p])
        call("fsim_sync")
        assert np.array_equal(o1.download((2 * n, 2 * n), np.float32), got[t])
    assert not np.array_equal(got[0], got[2])


def test_request_larger_than_budget_is_refused(setup):
    """A plain request whose slot alone exceeds the memory budget is refused
    at submit (FsimBusy) instead of waiting at the head of the queue for
    memory that no completion frees; smaller requests behind it still run."""
    from fsim._lib import FsimBusy
    from fsim.livemode import ReconstructionServer
    # one 256^2 stack needs about 7.3 MB of slot memory, a 64^2 one 0.5 MB
    srv = ReconstructionServer(mem_budget=4 << 20, max_batch=4, max_queue=16, nslots=1)
    raw, rec = setup[(256, 525.0)]
    big = srv.add_channel(rec)
    with pytest.raises(FsimBusy):
        srv.submit(big, raw, np.zeros((512, 512), np.float32))
    raw64, rec64 = setup[(64, 525.0)]
    small = srv.add_channel(rec64)
    o = np.zeros((128, 128), np.float32)
    srv.wait(srv.submit(small, raw64, o), timeout_ms=60000)
    assert np.array_equal(o, rec64.reconstruct_host(raw64))
    st = srv.stats()
    assert st["rejected"] == 1 and st["completed"] == 1 and st["failed"] == 0
    srv.close()
