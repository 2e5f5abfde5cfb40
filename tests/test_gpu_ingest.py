"""The live stream feeding the request server directly (fsim_ingest_attach_
server / fsim_ingest_result): camera packets of two channels, interleaved
and locally out of order, are assembled into SIM sequences and
reconstructed; every image equals the channel's reconstruction of the same
stack (bitwise: the server's results do not depend on batch composition)."""
import numpy as np
import pytest

from oracle import ingest_ref as ir

pytestmark = pytest.mark.gpu


def _stream(stacks, channel, t0, seq0, delay, period):
    """Packets of the stacks, a camera-time gap > delay before each stack
    (sync mode 0: the frame after the gap starts the sequence)."""
    out, t, s = [], t0, seq0
    for st in stacks:
        t += delay + 1000
        for fr in st:
            n = fr.shape[0]
            out.append((s, ir.write_header(s, n, n, 2, pos1=channel, t_cam=t) + fr.tobytes()))
            s += 1
            t += period
    return out


def test_stream_to_server_matches_direct(fsim_lib):
    from fsim import synthetic
    from fsim.livemode import ReconstructionServer, SimSequenceExtractor
    from fsim.sim_algorithm import Reconstructor
    n = 128
    recs, stacks = [], []
    for c, lam in enumerate((525.0, 600.0)):
        sts = []
        for k in range(3):
            raw, sp = synthetic.sim_stack(n=n, seed=80 + 10 * c + k, lam=lam, nbeads=60, nfil=8)
            sts.append(raw)
        stacks.append(sts)
        recs.append(Reconstructor(synthetic.make_param(sp, n, lam=lam)))
    srv = ReconstructionServer(mem_budget=1 << 30, max_batch=4, max_queue=64, nslots=2)
    ids = [srv.add_channel(r) for r in recs]
    ing = SimSequenceExtractor([488, 561], sync_mode=0, sync_delay=20000, seq_count=1)
    ing.attach(srv, ids, n, max_inflight=3)
    a = _stream(stacks[0], 488, 0, 100, 20000, 1000)
    b = _stream(stacks[1], 561, 500, 7000, 20000, 1000)
    pk = [p for pair in zip(a, b) for p in pair]
    rng = np.random.default_rng(1)
    for i in range(0, len(pk) - 4, 4):   # local reordering
        w = pk[i:i + 4]
        rng.shuffle(w)
        pk[i:i + 4] = w
    got = {}
    for _, p in pk:
        ing.push(p)
    for _ in range(6):
        img, info = ing.take(timeout_ms=60000)
        got[(info.channel, info.first_seq_nr)] = img
    for c, base in enumerate((100, 7000)):
        for k in range(3):
            want = recs[c].reconstruct_host(stacks[c][k])
            assert np.array_equal(got[(c, base + 9 * k)], want), (c, k)
    for c in range(2):
        st = ing.stats(c)
        assert st["sequences"] == 3 and st["missed_sequences"] == 0
    ing.close()
    srv.close()
