"""GPU tests of fixed-shape batches (fsim_batch_*): mixed sizes and channels in
one launch, eager and as a replayed hipGraph.

Every batch result is compared bitwise with fsim_reconstruct of the same stack
on the same channel (the same kernels run on the same data).  The graph test
replays the captured launch several times with new raw and output addresses,
with eager library work and synchronisations between the replays, and checks
every replay.
"""
import ctypes as C

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
    from fsim import synthetic
    chans = {}
    for i, (n, lam) in enumerate([(64, 525.0), (64, 600.0), (128, 670.0)]):
        chans[(n, lam)] = _channel(n, lam, 70 + i)
    # a second stack per channel (same optics, other object)
    extra = {}
    for i, k in enumerate(chans):
        raw2, _ = synthetic.sim_stack(n=k[0], seed=90 + i, lam=k[1], nbeads=40, nfil=6)
        extra[k] = raw2
    return chans, extra


def _slots(setup):
    chans, extra = setup
    order = [(64, 525.0), (128, 670.0), (64, 600.0), (64, 525.0), (128, 670.0), (64, 600.0), (64, 525.0)]
    stacks = [chans[k][0] if i % 2 == 0 else extra[k] for i, k in enumerate(order)]
    return order, stacks


def _upload(stacks):
    from fsim.linalg import DeviceBuffer
    bufs = []
    for s in stacks:
        b = DeviceBuffer(s.nbytes)
        b.upload(s)
        bufs.append(b)
    return bufs


def _expected(setup, order, stacks):
    chans, _ = setup
    return [chans[k][1].reconstruct_host(s) for k, s in zip(order, stacks)]


def _download(outs, order):
    return [o.download((2 * k[0], 2 * k[0]), np.float32) for o, k in zip(outs, order)]


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_matches_single_bitwise(setup, use_graph):
    from fsim._lib import call
    from fsim.linalg import DeviceBuffer
    from fsim.livemode import ReconstructionBatch
    chans, _ = setup
    order, stacks = _slots(setup)
    exp = _expected(setup, order, stacks)
    raws = _upload(stacks)
    outs = [DeviceBuffer(16 * k[0] * k[0]) for k in order]
    print(f"[batch] {len(order)} slots, graph={use_graph}", flush=True)
    b = ReconstructionBatch([chans[k][1] for k in order], use_graph=use_graph)
    b.run([r.value for r in raws], [o.value for o in outs])
    call("fsim_sync")
    info = b.info()
    assert info["classes"] == 2
    assert info["captured"] == use_graph
    for got, e in zip(_download(outs, order), exp):
        assert np.array_equal(got, e)
    b.close()


def test_graph_replays_with_new_addresses(setup):
    """Replays 1-4 alternate two input sets and three output sets; eager
    reconstructions and synchronisations run between the replays."""
    from fsim._lib import call
    from fsim.linalg import DeviceBuffer
    from fsim.livemode import ReconstructionBatch
    chans, extra = setup
    order, stacks_a = _slots(setup)
    stacks_b = [extra[k] if i % 2 == 0 else chans[k][0] for i, k in enumerate(order)]
    exp_a, exp_b = _expected(setup, order, stacks_a), _expected(setup, order, stacks_b)
    raw_a, raw_b = _upload(stacks_a), _upload(stacks_b)
    out_sets = [[DeviceBuffer(16 * k[0] * k[0]) for k in order] for _ in range(3)]
    b = ReconstructionBatch([chans[k][1] for k in order], use_graph=True)
    plan = [(raw_a, 0, exp_a), (raw_b, 1, exp_b), (raw_a, 2, exp_a), (raw_b, 0, exp_b)]
    side_raw, side_rec = chans[(64, 600.0)]
    for i, (raws, oi, exp) in enumerate(plan):
        print(f"[batch] graph replay {i}", flush=True)
        b.run([r.value for r in raws], [o.value for o in out_sets[oi]])
        call("fsim_sync")
        assert b.info()["captured"]
        for got, e in zip(_download(out_sets[oi], order), exp):
            assert np.array_equal(got, e)
        # eager work of the library between replays
        assert np.array_equal(side_rec.reconstruct_host(side_raw), side_rec.reconstruct_host(side_raw))
    b.close()


def test_batch_rejects_unaligned_addresses(setup):
    from fsim._lib import FsimError
    from fsim.linalg import DeviceBuffer
    from fsim.livemode import ReconstructionBatch
    chans, _ = setup
    raw, rec = chans[(64, 525.0)]
    rb = DeviceBuffer(raw.nbytes + 8)
    ob = DeviceBuffer(16 * 64 * 64 + 8)
    b = ReconstructionBatch([rec], use_graph=True)
    with pytest.raises(FsimError):
        b.run([rb.value + 2], [ob.value])
    with pytest.raises(FsimError):
        b.run([rb.value], [ob.value + 4])
    assert not b.info()["captured"]
    b.close()


def test_batch_and_params_argument_errors(setup):
    """Refusals before any device work: a null channel or an empty slot list,
    and server requests with their own parameters that name no channel or no
    stacks."""
    from fsim._lib import FsimError, call, SIM_PARAMS
    from fsim.livemode import ReconstructionServer
    chans, _ = setup
    raw, rec = chans[(64, 525.0)]
    h = C.c_void_p()
    with pytest.raises(FsimError):
        call("fsim_batch_create", (C.c_void_p * 1)(None), 1, 1, C.byref(h))
    with pytest.raises(FsimError):
        call("fsim_batch_create", (C.c_void_p * 1)(rec.handle.value), 0, 1, C.byref(h))
    srv = ReconstructionServer(mem_budget=1 << 30, max_batch=4, max_queue=16, nslots=1)
    cid = srv.add_channel(rec)
    out = np.zeros((2 * 64, 2 * 64), np.float32)
    with pytest.raises(FsimError):
        srv.submit_params(cid + 5, raw, out, 1, SIM_PARAMS())
    with pytest.raises(FsimError):
        srv.submit_params(cid, raw, out, 0, SIM_PARAMS())
    srv.close()


def test_persistent_batch_every_group_and_variant(fsim_lib):
    """Persistent work-list kernels (FSIM_BATCH_PERSISTENT): one batch with
    both launch groups (256/512 and 1024/2048 sides), every list segment
    (all-units and one-unit separation; Hermitian and full-width gather and
    last pass), a class of nine slots (the gather's cross-stack block order)
    and a 128^2 slot that keeps its own class launches.  Eager and as a
    replayed graph with the stacks swapped between replays.  Every slot is
    bitwise equal to fsim_reconstruct of its stack: the items run the block
    functions of the per-class kernels, and multiply-adds fuse within one
    source expression only (-ffp-contract=on), so they round alike wherever
    they are inlined."""
    from fsim import synthetic
    from fsim._lib import call
    from fsim.linalg import DeviceBuffer
    from fsim.livemode import ReconstructionBatch
    from fsim.sim_algorithm import Reconstructor
    # (n, wavelength, phases, bands, mode, slots); mode 1: no Hermitian pairing
    specs = [(2048, 525.0, 3, 2, 0, 1), (1024, 600.0, 3, 2, 0, 2), (1024, 525.0, 3, 2, 1, 1),
             (512, 670.0, 3, 2, 0, 2), (512, 525.0, 5, 3, 0, 2), (256, 600.0, 3, 2, 0, 9),
             (256, 525.0, 3, 2, 1, 2), (128, 525.0, 3, 2, 0, 2)]
    recs, pairs, order = [], [], []
    for i, (n, lam, nph, nb, mode, nslots) in enumerate(specs):
        k = 0.45 if nb == 3 else 0.85
        stacks = []
        for t in range(2):
            raw, sp = synthetic.sim_stack(n=n, nphase=nph, seed=700 + 10 * i + t, lam=lam, nbeads=max(30, 300 * n // 512),
                                          nfil=max(5, 40 * n // 512), harmonic2=0.3 if nb == 3 else 0.0, k_frac=k)
            stacks.append(raw)
        rec = Reconstructor(synthetic.make_param(sp, n, lam=lam, nphase=nph, nband=nb), mode=mode)
        recs.append(rec)
        exp = [rec.reconstruct_host(s) for s in stacks]
        bufs = []
        for s in stacks:
            b = DeviceBuffer(s.nbytes)
            b.upload(s)
            bufs.append(b)
        pairs.append((bufs, exp))
        order += [(i, j % 2) for j in range(nslots)]
    rng = np.random.default_rng(5)
    order = [order[j] for j in rng.permutation(len(order))]
    outs = [DeviceBuffer(16 * specs[i][0] ** 2) for i, _ in order]
    worst = {}
    for use_graph in (False, True):
        b = ReconstructionBatch([recs[i] for i, _ in order], use_graph=use_graph, persistent=True)
        for rep in range(3 if use_graph else 1):
            sel = [(i, t ^ (rep % 2)) for i, t in order]
            print(f"[persistent] graph={use_graph} run {rep}: {len(sel)} slots", flush=True)
            b.run([pairs[i][0][t].value for i, t in sel], [o.value for o in outs])
            call("fsim_sync")
            bad = []
            for o, (i, t) in zip(outs, sel):
                n = specs[i][0]
                got = o.download((2 * n, 2 * n), np.float32)
                want = pairs[i][1][t]
                e = float(np.abs(got - want).max() / np.abs(want).max())
                worst[specs[i][0]] = max(worst.get(specs[i][0], 0.0), e)
                if not np.array_equal(got, want):
                    bad.append((specs[i], t, e))
            assert not bad, f"graph={use_graph} run {rep}: {bad}"
        assert b.info()["captured"] == use_graph
        b.close()
    print(f"[persistent] max difference from the single calls per side: {worst}")
