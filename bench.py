#!/usr/bin/env python3
"""Benchmark: 2D SIM reconstructions per second on one MI355X.

Workload: 512 x 512 raw, 3 angles x 3 phases uint16 (seeded synthetic beads
and filaments, NA 1.4, 525 nm, 80 nm pixels, |k0| = 0.85 cutoff), parameters
given, fast-shift reconstruction to 1024 x 1024 float32 with Wiener 0.05,
apodization 2.0 / 0.9, attenuation 0.995 / 1.2, border fade 10 px.

A step is one pass of the fused pipeline over a batch of stacks already
resident in HBM.  Also measured: the plain per-operation vector-API path
(the reference's accelerated op sequence, one reconstruction per call),
single-stack latency, live per-kernel HIP-event times (roofline of the
dominant kernel) and the FP32 Java-order CPU restatement on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "fairsim-vigor_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
STAGES = ["k_sep_rows", "k_col_shift", "k_row_shift", "k_gather", "k_row_real"]
TUNE = True   # run the autotuner for each workload's size class (--no-tune: defaults)


def profiled_traffic(kernel, n, per_launch):
    """HBM bytes per launch of `kernel` (read + write) from the committed
    rocprofv3 counter summary of this workload (profiles/), or None when no
    summary covers this n and number of stacks per launch.  Read bytes are
    2 x FETCH_SIZE and write bytes WRITE_SIZE, the factors the summary's
    copy-kernel calibration gives."""
    path = os.path.join(ROOT, "profiles", f"recon_pipeline_{n}x3x3_{per_launch}_stacks_per_launch.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        name = cells[0].strip("`") if cells else ""
        if len(cells) == 7 and name.split("<")[0] in (f"fsim::{kernel}", f"fsim::{kernel}1", f"fsim::{kernel}_all"):
            try:
                return int(round((float(cells[4]) + float(cells[6])) * 2 ** 20))
            except ValueError:
                return None
    return None


def stage_bytes(n, D=3, P=3, C=3, Cs=None, half=False, dc_sum=False):
    """HBM bytes per stack each pipeline kernel must move.  Cs = stored
    components per direction: C for the reference dataflow (the sum is
    B_alg = 434 n^2 at D = P = C = 3), B when each band's neg component is
    derived from its pos component.  dc_sum: the separation pass writes one
    DC component (the sum over the directions) instead of one per direction.
    half: the result spectrum is Hermitian and only its columns 0..n are
    formed -- half of every contribution lands there (the DC component's n/2
    columns; of a pos/neg pair at offset kx, n/2 + kx and n/2 - kx columns)
    and B is half as wide."""
    Cs = C if Cs is None else Cs
    nn = n * n
    sh = D * (Cs - 1)
    ndc = 1 if dc_sum else D                    # DC components written and read
    dc_cols = 0.5 if half else 1.0              # share of a DC row spectrum kept
    sep = 2 * D * P * nn + int(8 * ndc * dc_cols * nn) + 8 * sh * nn
    if half:
        gather = 4 * D * (C - 1) * nn + 4 * ndc * nn + 8 * 2 * nn   # half of each contribution, 2n x n out
        last = 8 * 2 * nn + 4 * 4 * nn           # 2n x n complex in, (2n)^2 real out
    else:
        gather = 8 * D * (C - 1) * nn + 8 * ndc * nn + 8 * 4 * nn   # every contribution once, (2n)^2 out
        last = 8 * 4 * nn + 4 * 4 * nn           # (2n)^2 complex in, real out
    return [sep,                                 # raw u16 in, stored components out
            16 * sh * nn,                        # shifted components in + out
            16 * sh * nn,
            gather, last]


def stage_names(half):
    """Kernel of each pipeline stage: the last pass is k_row_c2r when the
    result spectrum is Hermitian, k_row_real otherwise."""
    return STAGES[:4] + (["k_row_c2r"] if half else ["k_row_real"])


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pct(a, q):
    import numpy as np
    return float(np.percentile(np.asarray(a), q))


def probe_copy(reps):
    """y += x over two 64 MiB float vectors: 128 MiB read + 64 MiB written
    per launch (k_add_f), the known-size kernel that calibrates FETCH_SIZE /
    WRITE_SIZE."""
    from fsim.linalg import Vec2dReal
    from fsim import _lib
    a, b = Vec2dReal(4096, 4096), Vec2dReal(4096, 4096)
    for _ in range(reps):
        a.add(b)
    _lib.call("fsim_sync")
    print(json.dumps({"probe_copy": True, "launches": reps, "read_bytes": 2 * 4096 * 4096 * 4,
                      "write_bytes": 4096 * 4096 * 4}), flush=True)


def serve_saturation(rec, raw_buf, out_buf, B, raw_bytes, out_bytes, nreq):
    """Closed loop: keep up to B requests in flight on the server, stacks and
    outputs resident in HBM; throughput and submit-to-done latency.  Batches
    of up to 8 stacks on 4 slots: the size the chunked call uses (a batch's
    work buffers stay in the Infinity Cache between the passes, and the
    slots' batches overlap); measured against 16 or 32 stacks per batch and
    2-8 slots, this was the fastest."""
    from fsim.livemode import ReconstructionServer
    srv = ReconstructionServer(mem_budget=16 << 30, max_batch=8, max_queue=4 * B, nslots=4)
    cid = srv.add_channel(rec)
    inflight = []
    for i in range(min(64, nreq)):                      # warm the slots
        inflight.append(srv.submit(cid, raw_buf.value + (i % B) * raw_bytes,
                                   out_buf.value + (i % B) * out_bytes, device=True))
    for t in inflight:
        srv.wait(t)
    srv.stats(reset=True)
    inflight = []
    t0 = time.perf_counter()
    for i in range(nreq):
        if len(inflight) >= B:
            srv.wait(inflight.pop(0))
        inflight.append(srv.submit(cid, raw_buf.value + (i % B) * raw_bytes,
                                   out_buf.value + (i % B) * out_bytes, device=True))
    for t in inflight:
        srv.wait(t)
    el = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    return {"form": "server, closed loop, 512^2 stacks resident in HBM, up to %d in flight" % B,
            "recons_per_s": round(nreq / el, 2), "p50_ms": round(st["p50_ms"], 3),
            "p99_ms": round(st["p99_ms"], 3), "mean_batch": round(st["mean_batch"], 2)}


def serve_mixed_stream(nreq, rate):
    """Open loop: Poisson arrivals at `rate` req/s, sizes {256: .5, 512: .35,
    1024: .15}, 3 channels (525/600/670 nm), pinned host buffers (so the
    latency includes both PCIe copies).  One request per channel runs first
    (the first launches of each kernel shape load code; their worst latency
    is reported as first_request_per_channel_max_ms)."""
    import numpy as np
    from fsim import synthetic
    from fsim.livemode import ReconstructionServer, pinned
    from fsim.sim_algorithm import Reconstructor
    rng = np.random.default_rng(1234)
    sizes, probs = [256, 512, 1024], [0.5, 0.35, 0.15]
    lams = [525.0, 600.0, 670.0]
    srv = ReconstructionServer(mem_budget=16 << 30, max_batch=16, max_queue=4096, nslots=4)
    chans, raws, outs = {}, {}, {}
    for n in sizes:
        for ci, lam in enumerate(lams):
            raw, sp = synthetic.sim_stack(n=n, seed=100 + ci, lam=lam)
            chans[(n, ci)] = Reconstructor(synthetic.make_param(sp, n, lam=lam))
            buf, _ = pinned(raw.shape, np.uint16)
            buf[...] = raw
            raws[(n, ci)] = buf
    ids = {k: srv.add_channel(v) for k, v in chans.items()}
    pool = {n: [pinned((2 * n, 2 * n), np.float32)[0] for _ in range(8)] for n in sizes}
    # one request per channel before the timed stream: the first launches of
    # each kernel shape (code loading) are reported apart, not in p99
    first = [srv.wait(srv.submit(ids[k], raws[k], pool[k[0]][0])) for k in chans]
    srv.stats(reset=True)
    gaps = rng.exponential(1.0 / rate, nreq)
    kinds = [(int(rng.choice(sizes, p=probs)), int(rng.integers(3))) for _ in range(nreq)]
    tickets = []
    t0 = time.perf_counter()
    t_next = t0
    for i, k in enumerate(kinds):
        t_next += gaps[i]
        while time.perf_counter() < t_next:
            pass
        o = pool[k[0]][i % 8]
        tickets.append(srv.submit(ids[k], raws[k], o))
    for t in tickets:
        srv.wait(t)
    el = time.perf_counter() - t0
    st = srv.stats(reset=True)
    # saturation: the same mix submitted as fast as the server admits it
    t0 = time.perf_counter()
    tickets = []
    for i, k in enumerate(kinds):
        tickets.append(srv.submit(ids[k], raws[k], pool[k[0]][i % 8]))
    for t in tickets:
        srv.wait(t)
    sat_el = time.perf_counter() - t0
    st2 = srv.stats()
    srv.close()
    mean_n2 = sum(p * n * n for n, p in zip(sizes, probs))
    return {"arrivals": f"Poisson {rate:g} req/s, {nreq} requests, sizes 256/512/1024 at .5/.35/.15, "
                        "3 channels, pinned host buffers",
            "first_request_per_channel_max_ms": round(max(first), 3),
            "offered_req_per_s": round(nreq / el, 1), "p50_ms": round(st["p50_ms"], 3),
            "p99_ms": round(st["p99_ms"], 3), "mean_batch": round(st["mean_batch"], 2),
            "saturation_req_per_s": round(nreq / sat_el, 1),
            "saturation_p50_ms": round(st2["p50_ms"], 3), "saturation_p99_ms": round(st2["p99_ms"], 3),
            "saturation_roofline_frac_pcie_inclusive": round(nreq / sat_el * 434 * mean_n2 / 8e12, 4)}


_STACKS = {}


def parity_errors(got, sp, n, raw, lam=525.0, threads=None, px_um=0.080):
    """rel-L2 of the GPU image against the FP32 Java-order oracle and the FP64
    re-run of the same dataflow, and of the FP32 oracle against FP64 (the
    error budget of the reference arithmetic itself)."""
    from oracle import fp64, ref
    p = ref.recon_dict(sp, n, lam=lam, px_um=px_um)
    threads = threads or max(1, min(16, os.cpu_count() or 1))
    r32 = ref.recon32(p, raw, threads=threads)
    r64 = ref.recon64(p, raw)
    return {"gpu_vs_ref32": fp64.rel_l2(got, r32), "gpu_vs_fp64": fp64.rel_l2(got, r64),
            "ref32_vs_fp64": fp64.rel_l2(r32, r64)}


def _stack(n, seed, **kw):
    """Seeded synthetic stack, generated once per process (2048^2 takes
    seconds on the host)."""
    from fsim import synthetic
    key = (n, seed, tuple(sorted(kw.items())))
    if key not in _STACKS:
        _STACKS[key] = synthetic.sim_stack(n=n, seed=seed, **kw)
    return _STACKS[key]


def serve_large_fov_mix(nreq=400, rate=200.0, frac=0.05, tps=8):
    """Served large-FOV mix: the mixed Poisson request stream (`rate` req/s, sizes
    256/512/1024 at .5/.35/.15, 3 channels, pinned host buffers) with a share
    `frac` of large-FOV requests: 2048^2 x 3 x 3 uint16 stacks, `tps`
    timepoints each, resident on the device; the k0/phase fit on the first
    timepoint (fsim_estimate on its float32 frames), then the request's
    timepoints through fsim_server_submit_params with that fit, admitted
    against the server's memory budget.  A worker thread runs the large
    requests (fit, submit, wait) so the arrival loop keeps its schedule;
    latency of a large request = arrival -> its last image ready."""
    import queue
    import threading
    import numpy as np
    from fsim import synthetic
    from fsim.linalg import DeviceBuffer
    from fsim.livemode import ReconstructionServer, pinned
    from fsim.sim_algorithm import Reconstructor, estimate_batch, params_from_estimates
    rng = np.random.default_rng(4321)
    sizes, probs = [256, 512, 1024], [0.5, 0.35, 0.15]
    lams = [525.0, 600.0, 670.0]
    srv = ReconstructionServer(mem_budget=24 << 30, max_batch=16, max_queue=4096, nslots=5)
    chans, raws = {}, {}
    for n in sizes:
        for ci, lam in enumerate(lams):
            raw, sp = synthetic.sim_stack(n=n, seed=100 + ci, lam=lam)
            chans[(n, ci)] = Reconstructor(synthetic.make_param(sp, n, lam=lam))
            buf, _ = pinned(raw.shape, np.uint16)
            buf[...] = raw
            raws[(n, ci)] = buf
    ids = {k: srv.add_channel(v) for k, v in chans.items()}
    pool = {n: [pinned((2 * n, 2 * n), np.float32)[0] for _ in range(8)] for n in sizes}
    # the large-FOV channel: 65 nm pixels, 2048^2; its parameters come from
    # the fit of each request's first timepoint
    nl = 2048
    series_l, sp_l = _stack(nl, 7, nbeads=1200, nfil=160, px_um=0.065, timepoints=8)
    raw_l = series_l[0]
    param_l = synthetic.make_param(sp_l, nl, px_um=0.065)
    rec_l = Reconstructor(param_l)
    lid = srv.add_channel(rec_l)
    rb_l = DeviceBuffer(tps * raw_l.nbytes)
    rb_l.upload(np.ascontiguousarray(series_l[np.arange(tps) % len(series_l)]))
    ib_l = DeviceBuffer(raw_l.size * 4)
    ib_l.upload(raw_l.astype(np.float32))
    ob_l = DeviceBuffer(tps * rec_l.out_bytes())
    jobs = queue.Queue()
    large_lat, k_err = [], []

    def worker():
        while True:
            item = jobs.get()
            if item is None:
                return
            t_arr = item
            res = estimate_batch(param_l, 1, 0.6, images_dev=ib_l.value, count=1, raw=True)
            k_err.append(max(abs(res[d].px - sp_l["px"][d]) + abs(res[d].py - sp_l["py"][d]) for d in range(3)))
            prm = params_from_estimates(param_l, res, 1)[0]
            t = srv.submit_params(lid, rb_l.value, ob_l.value, tps, prm, device=True)
            srv.wait(t)
            large_lat.append((time.perf_counter() - t_arr) * 1e3)

    th = threading.Thread(target=worker)
    th.start()
    # one large request before the stream: the first fit allocates the
    # estimation buffers, the first submit the server slot
    jobs.put(time.perf_counter())
    while not large_lat:
        time.sleep(0.01)
    large_lat.clear()
    k_err.clear()
    srv.stats(reset=True)
    gaps = rng.exponential(1.0 / rate, nreq)
    kinds = []
    for _ in range(nreq):
        if rng.random() < frac:
            kinds.append(None)
        else:
            kinds.append((int(rng.choice(sizes, p=probs)), int(rng.integers(3))))
    tickets = []
    t0 = time.perf_counter()
    t_next = t0
    for i, k in enumerate(kinds):
        t_next += gaps[i]
        while time.perf_counter() < t_next:
            pass
        if k is None:
            jobs.put(time.perf_counter())
        else:
            tickets.append(srv.submit(ids[k], raws[k], pool[k[0]][i % 8]))
    small_lat = [srv.wait(t) for t in tickets]
    jobs.put(None)
    th.join()
    el = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    nlarge = sum(1 for k in kinds if k is None)
    return {"arrivals": f"Poisson {rate:g} req/s, {nreq} requests: {nlarge} large-FOV ({tps} x 2048^2 timepoints "
                        f"at 65 nm, k0 fit on the first) and {nreq - nlarge} of sizes 256/512/1024 at "
                        f".5/.35/.15, 3 channels",
            "offered_req_per_s": round(nreq / el, 1),
            "small_p50_ms": round(float(np.percentile(small_lat, 50)), 3),
            "small_p99_ms": round(float(np.percentile(small_lat, 99)), 3),
            "large_p50_ms": round(float(np.percentile(large_lat, 50)), 3) if large_lat else None,
            "large_p99_ms": round(float(np.percentile(large_lat, 99)), 3) if large_lat else None,
            "large_fit_max_k_error_px": round(float(max(k_err)), 4) if k_err else None,
            "server_device_bytes": st.get("device_bytes"), "failed": st["failed"],
            "rejected": st["rejected"]}


def batched_mixed(nper=32, reps=10):
    """Mixed-size batch: one batch of 96 stacks -- 3 channels (525/600/670 nm) x sizes
    256/512/1024, 32 of each size, stacks and outputs resident in HBM.  The
    headline form is one fixed-shape batch (fsim_batch_*: the three size
    classes run concurrently, captured once as a hipGraph and replayed; the
    slot addresses travel in the device job lists); time per replay from
    host timestamps around run + sync, median of `reps`.  Beside it the same
    96 stacks submitted one by one to the request server (time to the last
    result)."""
    import numpy as np
    from fsim import synthetic
    from fsim._lib import call
    from fsim.linalg import DeviceBuffer
    from fsim.livemode import ReconstructionBatch, ReconstructionServer
    from fsim.sim_algorithm import Reconstructor
    sizes, lams = [256, 512, 1024], [525.0, 600.0, 670.0]
    recs, bufs, keep, host = {}, {}, [], {}
    ntp = (nper + 2) // 3                       # timepoints per (size, channel)
    for n in sizes:
        for ci, lam in enumerate(lams):
            # one sample per (size, channel) imaged ntp times with independent
            # seeded photon noise: every slot of the batch has its own stack
            raw, sp = synthetic.sim_stack(n=n, seed=200 + ci, lam=lam, nbeads=300 * n // 512,
                                          nfil=40 * n // 512, timepoints=ntp)
            host[(n, ci)] = (raw[0], sp, lam)
            rec = Reconstructor(synthetic.make_param(sp, n, lam=lam))
            keep.append(rec)
            recs[(n, ci)] = rec
            for t in range(ntp):
                rb = DeviceBuffer(raw[t].nbytes)
                rb.upload(raw[t])
                bufs[(n, ci, t)] = rb
    tuned = {}
    if TUNE:
        from fsim import autotune
        for n in sizes:
            tuned[n] = autotune.tune(recs[(n, 0)], count=nper, reps=3)
    slots = [(n, i % 3, i // 3) for n in sizes for i in range(nper)]
    jobs = [s[:2] for s in slots]
    outs = {n: [DeviceBuffer(16 * n * n) for _ in range(nper)] for n in sizes}
    out_ptrs = [outs[k[0]][j % nper].value for j, k in enumerate(jobs)]
    raw_ptrs = [bufs[s].value for s in slots]
    assert len(set(raw_ptrs)) == len(slots)     # distinct raw stacks
    raw_bytes = sum(bufs[s].nbytes for s in slots)
    b_alg = sum(434 * n * n for n, _ in jobs)
    b_moved = sum(sum(stage_bytes(n, Cs=2, half=True, dc_sum=True)) for n, _ in jobs)

    def time_batch(persistent):
        batch = ReconstructionBatch([recs[k] for k in jobs], use_graph=True, persistent=persistent)
        batch.run(raw_ptrs, out_ptrs)            # captures the graph
        call("fsim_sync")
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            batch.run(raw_ptrs, out_ptrs)
            call("fsim_sync")
            times.append(time.perf_counter() - t0)
        info = batch.info()
        batch.close()
        return float(np.median(times)), info

    # persistent work-list kernels: one launch per stage and group of sizes;
    # its images must be the per-class form's bits
    el_p, _ = time_batch(True)
    firsts = {n: outs[n][0].download((2 * n, 2 * n), np.float32) for n in sizes}
    el, info = time_batch(False)
    same_bits = all(np.array_equal(firsts[n], outs[n][0].download((2 * n, 2 * n), np.float32)) for n in sizes)
    # parity of the first slot of each size (channel 0) from the batch output
    parity = {}
    for n in sizes:
        raw, sp, lam = host[(n, 0)]
        got = outs[n][0].download((2 * n, 2 * n), np.float32)
        parity[n] = {k: float(f"{v:.3e}") for k, v in parity_errors(got, sp, n, raw, lam=lam).items()}

    srv = ReconstructionServer(mem_budget=32 << 30, max_batch=32, max_queue=4096, nslots=4)
    ids = {k: srv.add_channel(r) for k, r in recs.items()}

    def run_served():
        t0 = time.perf_counter()
        tickets = [srv.submit(ids[k], raw_ptrs[j], outs[k[0]][j % nper].value, device=True)
                   for j, k in enumerate(jobs)]
        for t in tickets:
            srv.wait(t)
        return time.perf_counter() - t0

    run_served()
    srv.stats(reset=True)
    el_s = min(run_served() for _ in range(3))
    srv.close()
    return {"batch": f"{len(jobs)} distinct stacks: sizes 256/512/1024 x {nper}, 3 channels, resident in HBM "
                     f"({raw_bytes / 1e6:.0f} MB raw in; per size and channel one sample imaged {ntp} times with "
                     f"independent seeded photon noise)",
            "form": f"one fixed-shape batch, {info['classes']} size classes on concurrent streams, "
                    f"replayed hipGraph (captured={info['captured']})",
            "ms_per_batch": round(el * 1e3, 3), "recons_per_s": round(len(jobs) / el, 1),
            "executed_bytes_per_batch": b_moved, "executed_bytes_frac": round(b_moved / el / 8e12, 4),
            "reference_dataflow_bytes_per_batch": b_alg,
            "reference_dataflow_equivalent_frac": round(b_alg / el / 8e12, 4),
            "autotune": {n: (t["sep"], t["col_elems"], t["col_width"]) for n, t in tuned.items()},
            "rel_l2": parity,
            "persistent_work_list": {"form": "the same batch through persistent work-list kernels (per stage one "
                                             "launch for 256/512 and one for 1024), replayed hipGraph",
                                     "ms_per_batch": round(el_p * 1e3, 3),
                                     "recons_per_s": round(len(jobs) / el_p, 1),
                                     "executed_bytes_frac": round(b_moved / el_p / 8e12, 4),
                                     "bitwise_equal_to_class_launches":
                                         same_bits if all(t["col"] == 0 and (t["sep"] == 1 or n > 512)
                                                          for n, t in tuned.items())
                                         else "not compared: tuned column choices differ from the defaults"},
            "served_one_by_one": {"ms_per_batch": round(el_s * 1e3, 3),
                                  "recons_per_s": round(len(jobs) / el_s, 1),
                                  "executed_bytes_frac": round(b_moved / el_s / 8e12, 4),
                                  "reference_dataflow_equivalent_frac": round(b_alg / el_s / 8e12, 4)}}


def large_fov(batch=8, steps=20):
    """Large field of view: 2048^2 x 3 x 3 uint16 stacks at 65 nm raw pixels -> 4096^2, fused pipeline,
    `batch` stacks per launch resident in HBM; parity of one stack against the
    FP32 oracle."""
    import numpy as np
    from fsim import synthetic
    from fsim.linalg import DeviceBuffer
    from fsim.sim_algorithm import Reconstructor
    from oracle import fp64, ref
    n, px = 2048, 0.065
    series, sp = _stack(n, 7, nbeads=1200, nfil=160, px_um=px, timepoints=8)
    raw = series[0]
    rec = Reconstructor(synthetic.make_param(sp, n, px_um=px))
    tuned = None
    if TUNE:
        from fsim import autotune
        tuned = autotune.tune(rec, count=batch, reps=2)
    rb = DeviceBuffer(raw.nbytes * batch)
    host = np.ascontiguousarray(series[np.arange(batch) % len(series)])
    rb.upload(host)
    ob = DeviceBuffer(rec.out_bytes() * batch)
    for _ in range(3):   # untimed
        rec.reconstruct(rb.value, ob.value, batch)
    _lib_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        rec.reconstruct(rb.value, ob.value, batch)
    _lib_sync()
    el = time.perf_counter() - t0
    value = batch * steps / el
    got = rec.reconstruct_host(raw)
    threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), os.cpu_count() or 1))
    par = parity_errors(got, sp, n, raw, threads=threads, px_um=px)
    return {"workload": f"2048^2 x 3 x 3 uint16 at 65 nm raw pixels (|k0| = {sp['k0']:.1f} px) -> 4096^2 "
                        f"float32, batch {batch} distinct timepoints",
            "recons_per_s": round(value, 2), "ms_per_recon": round(1e3 / value, 3),
            "executed_bytes_frac": round(value * sum(stage_bytes(n, Cs=2, half=True, dc_sum=True)) / 8e12, 4),
            "reference_dataflow_equivalent_frac": round(value * 434 * n * n / 8e12, 4),
            "rel_l2_vs_oracle": par["gpu_vs_ref32"],
            "rel_l2": par, "autotune": tuned}


def estimation(runs=3, batch=64):
    """Parameter estimation (SimAlgorithm.estimateParameters) of one
    512^2 stack given as float images: 3 phases / 2 bands (fit band 1) and
    5 phases / 3 bands with a 2 k0 harmonic (|k0| = 0.45 cutoff, fit band 2),
    GPU (device vector flow) vs the FP32 oracle on the host cores; the
    estimate + fused reconstruction of the same stack; and the batched form
    (estimation_batched)."""
    import numpy as np
    from fsim import synthetic
    from fsim.sim_algorithm import OtfProvider, Reconstructor, SimParam, estimate_parameters, input_ffts
    from oracle import ref
    n = 512
    threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), os.cpu_count() or 1))
    ref.lib().ref_set_threads(threads)
    out = {}
    for tag, nphase, nband, fb, k_frac, h2 in (("p3b2", 3, 2, 1, 0.85, 0.0), ("p5b3", 5, 3, 2, 0.45, 0.3)):
        raw, sp = synthetic.sim_stack(n=n, nphase=nphase, seed=11, harmonic2=h2, k_frac=k_frac)
        imgs = raw.reshape(3, nphase, n, n).astype(np.float32)

        def est():
            otf = OtfProvider.fromEstimate(1.4, 525.0, 1.0)
            param = SimParam.create(nband, 3, nphase, n, 0.08, otf)
            r = estimate_parameters(param, input_ffts(imgs), fb, 0.6)
            _lib_sync()
            return param, r

        est()
        t = []
        for _ in range(runs):
            a = time.perf_counter()
            param, r = est()
            t.append((time.perf_counter() - a) * 1e3)
        err_k = max(abs(r[d]["px"] - (nband - 1) * sp["px"][d]) + abs(r[d]["py"] - (nband - 1) * sp["py"][d])
                    for d in range(3))
        # estimate + reconstruct (uint16 stack, fused path, estimated parameters)
        a = time.perf_counter()
        param, r = est()
        rec = Reconstructor(param)
        rec.reconstruct_host(raw)
        er_ms = (time.perf_counter() - a) * 1e3
        a = time.perf_counter()
        otf32 = ref.Otf32(1.4, 525.0, 1.0)
        ref.estimate32(ref.input_ffts32(imgs), nband, fb, 0.6, otf32, 1 / (n * 0.08))
        cpu_ms = (time.perf_counter() - a) * 1e3
        out[tag] = {"gpu_ms_per_estimate": round(float(np.median(t)), 2),
                    "estimate_plus_recon_ms": round(er_ms, 2),
                    "cpu_oracle_ms_per_estimate": round(cpu_ms, 1), "cpu_threads": threads,
                    "max_k_error_px": round(err_k, 4)}
        out[tag]["batched"] = estimation_batched(nphase, nband, fb, k_frac, h2, batch)
    return out


def estimation_batched(nphase, nband, fb, k_frac, h2, batch=64, runs=10):
    """Estimation plus reconstruction as a throughput: `batch` distinct 512^2 stacks per call
    (4 seeded samples, each imaged batch / 4 times with independent seeded
    photon noise; stack s is sample s % 4), float32 images and uint16 raw
    frames resident on the device; fsim_estimate for all of them, then
    fsim_reconstruct_params with every stack's own estimate.  Timed: the
    whole sequence including the host step that turns estimates into
    per-stack parameters."""
    import numpy as np
    from fsim import synthetic
    from fsim.linalg import DeviceBuffer
    from fsim.sim_algorithm import Reconstructor, estimate_batch, params_from_estimates
    n = 512
    tp = batch // 4
    raws, sps = [], []
    for s in range(4):
        raw, sp = synthetic.sim_stack(n=n, nphase=nphase, seed=11 + s, harmonic2=h2, k_frac=k_frac, timepoints=tp)
        raws.append(raw)
        sps.append(sp)
    # stack k = timepoint k // 4 of sample k % 4
    raw_b = np.ascontiguousarray(np.stack(raws, 1).reshape(batch, *raws[0].shape[1:]))
    raws = [r[0] for r in raws]
    img_b = raw_b.reshape(batch, 3, nphase, n, n).astype(np.float32)
    param = synthetic.make_param(sps[0], n, nphase=nphase, nband=nband)
    rec = Reconstructor(param)
    rb, ib = DeviceBuffer(raw_b.nbytes), DeviceBuffer(img_b.nbytes)
    ob = DeviceBuffer(batch * rec.out_bytes())
    rb.upload(raw_b)
    ib.upload(img_b)
    del img_b

    def run(recon):
        res = estimate_batch(param, fb, 0.6, images_dev=ib.value, count=batch, raw=True)
        if recon:
            rec.reconstruct_params(rb.value, ob.value, params_from_estimates(param, res, batch))
        _lib_sync()
        return res

    res0 = run(True)
    # parity of stack 0's reconstruction with its own estimate
    from oracle import fp64, ref
    prm = params_from_estimates(param, res0, 1)[0]
    got0 = ob.download((2 * n, 2 * n), np.float32)
    p = ref.recon_dict(sps[0], n, nband=nband)
    p["px"] = [prm.px[d] for d in range(3)]
    p["py"] = [prm.py[d] for d in range(3)]
    p["phases"] = [[prm.phases[d][i] for i in range(nphase)] for d in range(3)]
    p["modul"] = [[prm.modul[d][b] for b in range(nband)] for d in range(3)]
    r32 = ref.recon32(p, raws[0], threads=max(1, min(16, os.cpu_count() or 1)))
    r64 = ref.recon64(p, raws[0])
    parity = {"gpu_vs_ref32": float(f"{fp64.rel_l2(got0, r32):.3e}"),
              "gpu_vs_fp64": float(f"{fp64.rel_l2(got0, r64):.3e}"),
              "ref32_vs_fp64": float(f"{fp64.rel_l2(r32, r64):.3e}")}
    est = [[{"px": res0[s * 3 + d].px, "py": res0[s * 3 + d].py} for d in range(3)] for s in range(4)]
    scale = 1 if nband == 2 else 2
    err_k = max(abs(est[s][d]["px"] - scale * sps[s % 4]["px"][d]) + abs(est[s][d]["py"] - scale * sps[s % 4]["py"][d])
                for s in range(4) for d in range(3))
    res = {"stacks_per_call": batch}
    for recon, key in ((False, "estimates_per_s"), (True, "estimate_plus_recon_per_s")):
        run(recon)   # untimed
        t = []
        for _ in range(runs):
            a = time.perf_counter()
            run(recon)
            t.append(time.perf_counter() - a)
        res[key] = round(batch / float(np.median(t)), 1)
    # estimate + reconstruction algorithmic bytes with float input: 1268 n^2
    res["roofline_frac_estimate_plus_recon"] = round(res["estimate_plus_recon_per_s"] * 1268 * n * n / 8e12, 4)
    res["max_k_error_px"] = round(err_k, 4)
    res["rel_l2_stack0"] = parity
    for b in (rb, ib, ob):
        b.free()
    return res


def stream_ingest(n=512, per_channel=48, runs=2):
    """Live stream front end (fsim_ingest): camera packets of two channels
    (128-byte header + 512^2 uint16 pixels, a sync gap before every stack,
    interleaved) pushed from host memory.  Reported: the host-side assembly
    rate alone (frames/s, sequences popped), and the end-to-end rate with the
    request server attached (pinned staging, PCIe included), with a bitwise
    check of one image against the channel's own reconstruction."""
    import numpy as np
    from fsim import synthetic
    from fsim.livemode import ReconstructionServer, SimSequenceExtractor
    from fsim.sim_algorithm import Reconstructor
    from oracle import ingest_ref as ir
    recs, packets, first = [], [], []
    for c, lam in enumerate((525.0, 600.0)):
        raws = []
        for k in range(4):
            raw, sp = synthetic.sim_stack(n=n, seed=200 + 10 * c + k, lam=lam)
            raws.append(raw)
        recs.append(Reconstructor(synthetic.make_param(sp, n, lam=lam)))
        first.append(raws[0])
        seq, t, pk = 1000 * (c + 1), 0, []
        for k in range(per_channel):
            t += 30000
            for fr in raws[k % 4]:
                pk.append(ir.write_header(seq, n, n, 2, pos1=c, t_cam=t) + fr.tobytes())
                seq += 1
                t += 1000
        packets.append(pk)
    stream = [np.frombuffer(p, np.uint8) for pair in zip(*packets) for p in pair]
    nseq = 2 * per_channel

    def assemble_only():
        ing = SimSequenceExtractor([0, 1], sync_mode=0, sync_delay=20000, max_size=n)
        a = time.perf_counter()
        popped = 0
        for i, p in enumerate(stream):
            ing.push(p)
            if i % 18 == 17:   # a consumer taking sequences as they complete
                popped += sum(1 for c in (0, 1) for _ in iter(lambda: ing.pop(c), None))
        popped += sum(1 for c in (0, 1) for _ in iter(lambda: ing.pop(c), None))
        el = time.perf_counter() - a
        ing.close()
        return el, popped

    def end_to_end():
        srv = ReconstructionServer(mem_budget=8 << 30, max_batch=16, max_queue=256, nslots=4)
        ids = [srv.add_channel(r) for r in recs]
        ing = SimSequenceExtractor([0, 1], sync_mode=0, sync_delay=20000, max_size=n)
        ing.attach(srv, ids, n, max_inflight=32)
        a = time.perf_counter()
        got = None
        taken = 0
        for i, p in enumerate(stream):
            ing.push(p)
            if i % 64 == 63:
                while True:   # drain what is ready without blocking the stream
                    try:
                        img, info = ing.take(timeout_ms=0)
                    except Exception:
                        break
                    taken += 1
                    if info.channel == 0 and info.first_seq_nr == 1000:
                        got = img
        while taken < nseq:
            img, info = ing.take(timeout_ms=60000)
            taken += 1
            if info.channel == 0 and info.first_seq_nr == 1000:
                got = img
        el = time.perf_counter() - a
        ing.close()
        srv.close()
        return el, got

    assemble_only()
    ta = min(assemble_only()[0] for _ in range(runs))
    end_to_end()
    te, img = min((end_to_end() for _ in range(runs)), key=lambda r: r[0])
    exact = bool(img is not None and np.array_equal(img, recs[0].reconstruct_host(first[0])))
    return {"workload": f"2 channels x {per_channel} stacks of 9 x {n}^2 uint16 packets, interleaved",
            "assembly_frames_per_s": round(len(stream) / ta, 1),
            "assembly_sequences_per_s": round(nseq / ta, 1),
            "end_to_end_recons_per_s": round(nseq / te, 1),
            "pcie_inclusive": True, "bitwise_equal_to_direct": exact}


def side_outputs(n=512, batch=64, reps=20):
    """Live-mode side outputs on the device: widefield images of `batch`
    512^2 stacks (9 uint16 frames read, one float image written: 22 B per
    pixel) and the registration of `batch` reconstructed 1024^2 images
    (source and result 8 B per pixel, the two tables 8 B per pixel once per
    call), as HBM rates."""
    import numpy as np
    from fsim import synthetic
    from fsim.linalg import DeviceBuffer
    from fsim.livemode import Registration
    from fsim.sim_algorithm import Reconstructor
    raw, sp = synthetic.sim_stack(n=n, seed=5)
    rec = Reconstructor(synthetic.make_param(sp, n))
    rb = DeviceBuffer(batch * raw.nbytes)
    rb.upload(np.ascontiguousarray(np.broadcast_to(raw, (batch,) + raw.shape)))
    wb = DeviceBuffer(batch * n * n * 4)
    rec.widefield(rb.value, wb.value, batch)
    _lib_sync()
    a = time.perf_counter()
    for _ in range(reps):
        rec.widefield(rb.value, wb.value, batch)
    _lib_sync()
    t_wf = (time.perf_counter() - a) / reps
    M = 2 * n
    gy, gx = np.mgrid[0:M, 0:M].astype(np.float32)
    reg = Registration(gx + np.float32(0.3), gy - np.float32(0.45))
    sb, db = DeviceBuffer(batch * M * M * 4), DeviceBuffer(batch * M * M * 4)
    reg.register(sb.value, db.value, M, M, batch)
    _lib_sync()
    a = time.perf_counter()
    for _ in range(reps):
        reg.register(sb.value, db.value, M, M, batch)
    _lib_sync()
    t_reg = (time.perf_counter() - a) / reps
    for b in (rb, wb, sb, db):
        b.free()
    reg.close()
    wf_gbs = batch * n * n * 22 / t_wf / 1e9
    reg_gbs = (batch * M * M * 8 + M * M * 8) / t_reg / 1e9   # images in and out; the tables once
    return {"widefield": {"stacks_per_s": round(batch / t_wf, 1), "GB_per_s": round(wf_gbs, 1),
                          "hbm_frac": round(wf_gbs / 8000, 4)},
            "registration_1024": {"images_per_s": round(batch / t_reg, 1), "GB_per_s": round(reg_gbs, 1),
                                  "hbm_frac": round(reg_gbs / 8000, 4)}}


def filter_styles(n=512, iters=10, reps=5):
    """Filter styles beyond Wiener: fsim_rl_deconvolve on a (2n)^2 spectrum
    (the RL-on-output size) -- iterations/s and the HBM rate on the 104 B
    per pixel and iteration the fused steps move (two column passes: read,
    write, OTF; two row passes: read, write and the image, or the deconvolved
    estimate read and written), beside the 232 B the transforms and
    element-wise passes one at a time would move -- and
    SimAlgorithm.runReconstruction's slow-shift flow per filter style through
    the vector API (one 512^2 3x3 stack)."""
    import numpy as np
    from fsim import synthetic
    from fsim.linalg import Vec2dCplx
    from fsim.sim_algorithm import RLDeconvolution, input_ffts, run_reconstruction
    M = 2 * n
    rng = np.random.default_rng(0)
    img, otf = Vec2dCplx(M, M), Vec2dCplx(M, M)
    img.setData(rng.uniform(1, 2, (M, M, 2)).astype(np.float32))
    otf.setData(rng.uniform(0, 1, (M, M, 2)).astype(np.float32))
    RLDeconvolution.deconvolve(img, otf, 1)
    a = time.perf_counter()
    for _ in range(reps):
        RLDeconvolution.deconvolve(img, otf, iters)
    t = (time.perf_counter() - a) / (reps * iters)
    gbs = 104 * M * M / t / 1e9
    raw, sp = synthetic.sim_stack(n=n, seed=9)
    param = synthetic.make_param(sp, n)
    inFFT = input_ffts(raw.reshape(3, 3, n, n).astype(np.float32))
    styles = {}
    for style in ("Wiener", "RLin", "RLout", "RLboth"):
        run_reconstruction(param, inFFT, filterStyle=style, rlIterations=5)
        _lib_sync()
        a = time.perf_counter()
        run_reconstruction(param, inFFT, filterStyle=style, rlIterations=5)
        _lib_sync()
        styles[style] = round((time.perf_counter() - a) * 1e3, 2)
    return {"rl_1024": {"iterations_per_s": round(1 / t, 1), "GB_per_s": round(gbs, 1),
                        "hbm_frac": round(gbs / 8000, 4), "bytes_per_pixel_iteration": 104,
                        "unfused_equivalent_GB_per_s": round(232 * M * M / t / 1e9, 1)},
            "run_reconstruction_512_ms": styles}


def config_io(runs=5):
    """XML parameter files: load the example live configuration
    (tests/golden/vigor-fastsim-example.xml), build its three channels
    (SimParam, OTF) and their device tables (fused reconstruction channels),
    and write it back -- milliseconds per step."""
    from fsim.conf import Conf
    from fsim.livemode import load_live_channels
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "vigor-fastsim-example.xml")
    t_load, t_chan = [], []
    for _ in range(runs):
        a = time.perf_counter()
        cfg = Conf.loadFile(path)
        vs = cfg.r().cd("vigor-settings")
        chans = load_live_channels(vs, ["488", "568", "647"])
        t_load.append((time.perf_counter() - a) * 1e3)
        a = time.perf_counter()
        recs = [c.reconstructor() for c in chans]
        _lib_sync()
        t_chan.append((time.perf_counter() - a) * 1e3)
        del recs
    return {"load_config_and_params_ms": round(min(t_load), 2), "channel_tables_3x512_ms": round(min(t_chan), 2)}


def _lib_sync():
    from fsim import _lib
    _lib.call("fsim_sync")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the warm-up also brings the device out of the idle state the host-side
    # parity check leaves it in (clocks ramp over tens of milliseconds)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--px-um", type=float, default=0.080,
                    help="raw pixel size of the synthetic stacks (micron)")
    ap.add_argument("--plain-runs", type=int, default=20)
    ap.add_argument("--latency-runs", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--served-requests", type=int, default=2048)
    ap.add_argument("--mix-requests", type=int, default=400)
    ap.add_argument("--mix-rate", type=float, default=200.0)
    ap.add_argument("--no-mix", action="store_true")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the estimation, mixed-batch and 2048^2 legs")
    ap.add_argument("--profile-only", action="store_true",
                    help="only the timed batched loop (for rocprofv3 runs)")
    ap.add_argument("--one-launch", action="store_true",
                    help="with --profile-only: the batch as one launch on one stream (no chunks)")
    ap.add_argument("--stream-only", action="store_true",
                    help="only the live stream ingest, side-output and filter-style legs")
    ap.add_argument("--estimation-only", action="store_true",
                    help="only the batched 512^2 estimation legs (for rocprofv3 runs)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (default launch shapes)")
    ap.add_argument("--serve-fov-only", action="store_true",
                    help="only the served stream with large-FOV requests")
    ap.add_argument("--mixed-only", action="store_true",
                    help="only the 96-stack mixed-size batch")
    ap.add_argument("--probe-copy", action="store_true",
                    help="run a plain streaming kernel of known size (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    from fsim import _lib, synthetic
    from fsim.linalg import DeviceBuffer
    from fsim.sim_algorithm import PlainPipeline, Reconstructor

    n, B = args.n, args.batch
    global TUNE
    TUNE = not args.no_tune
    _lib.init(0)
    if args.probe_copy:
        probe_copy(args.steps)
        return
    if args.stream_only:
        print(json.dumps({"stream_only": True, "stream_ingest_512": stream_ingest(),
                          "side_outputs_512": side_outputs(), "filter_styles_512": filter_styles(),
                          "config_io": config_io()}),
              flush=True)
        return
    if args.serve_fov_only:
        print(json.dumps({"serve_fov_only": True, "served_large_fov_mix": serve_large_fov_mix()}), flush=True)
        return
    if args.mixed_only:
        print(json.dumps({"mixed_only": True, "batched_mixed_96": batched_mixed()}), flush=True)
        return
    if args.estimation_only:
        out = {}
        for tag, nphase, nband, fb, k_frac, h2 in (("p3b2", 3, 2, 1, 0.85, 0.0), ("p5b3", 5, 3, 2, 0.45, 0.3)):
            log(f"batched estimation {tag}")
            out[tag] = estimation_batched(nphase, nband, fb, k_frac, h2, B)
        print(json.dumps({"estimation_only": True, "n": 512, "batched": out}), flush=True)
        return
    log(f"generating synthetic stacks n={n}")
    # B distinct stacks: 4 seeded samples, each imaged B/4 times with
    # independent photon noise (stacks[i] is stack i of the batch)
    per = max(2, (B + 3) // 4)
    series = [synthetic.sim_stack(n=n, seed=s, timepoints=per, px_um=args.px_um) for s in range(4)]
    nst = max(B, 4)
    stacks = [series[i % 4][0][i // 4] for i in range(nst)]
    sps = [series[i % 4][1] for i in range(nst)]
    param = synthetic.make_param(sps[0], n, px_um=args.px_um)
    rec = Reconstructor(param)
    tuned = None
    if TUNE:
        from fsim import autotune
        tuned = autotune.tune(rec, count=B, reps=3)
        log(f"autotune: {tuned}")
    frame_stack = np.stack([stacks[i % len(stacks)] for i in range(B)])
    raw_buf = DeviceBuffer(frame_stack.nbytes)
    raw_buf.upload(frame_stack)
    out_buf = DeviceBuffer(B * rec.out_bytes())

    if args.profile_only:
        if args.one_launch:
            from fsim import autotune
            autotune.set_chunks(*autotune.class_of(rec), 0, 1)
        for _ in range(args.warmup):
            rec.reconstruct(raw_buf.value, out_buf.value, B)
        _lib.call("fsim_sync")
        _lib.call("fsim_stage_timing", 1)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            rec.reconstruct(raw_buf.value, out_buf.value, B)
        _lib.call("fsim_sync")
        el = time.perf_counter() - t0
        _lib.call("fsim_stage_timing", 0)
        import ctypes as C
        ms = (C.c_double * 5)()
        cnt = (C.c_int * 5)()
        _lib.call("fsim_stage_times", ms, cnt, 5)
        print(json.dumps({"profile_only": True, "batch": B, "steps": args.steps, "n": n,
                          "recons_per_s": round(B * args.steps / el, 2),
                          "stage_ms": {STAGES[i]: round(ms[i] / max(cnt[i], 1), 4) for i in range(5)},
                          "autotune": tuned}),
              flush=True)
        return

    # parity of the measured path on stack 0
    from oracle import fp64, ref
    log("parity check vs FP32 oracle")
    got = rec.reconstruct_host(stacks[0])
    p32 = ref.recon_dict(sps[0], n, px_um=args.px_um)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    threads = max(1, min(threads, os.cpu_count() or 1))
    t0 = time.perf_counter()
    r32 = ref.recon32(p32, stacks[0], threads=threads)
    first_cpu = time.perf_counter() - t0
    parity = fp64.rel_l2(got, r32)
    r64 = ref.recon64(p32, stacks[0])
    parity64 = {"gpu_vs_ref32": parity, "gpu_vs_fp64": fp64.rel_l2(got, r64), "ref32_vs_fp64": fp64.rel_l2(r32, r64)}
    log(f"rel-L2 gpu vs oracle {parity:.3e}, vs FP64 {parity64['gpu_vs_fp64']:.3e}")

    # batched fused pipeline: the timed region runs without per-kernel
    # events; an identical second region with HIP events around every kernel
    # (on the stream each kernel is launched on) gives the per-kernel times
    from fsim import autotune
    cls = autotune.class_of(rec)
    per_launch, nstreams = autotune.chunk_plan(*cls, B)
    log(f"warmup {args.warmup} steps of batch {B} ({per_launch} stacks per launch on {nstreams} streams)")
    for _ in range(args.warmup):
        rec.reconstruct(raw_buf.value, out_buf.value, B)
    _lib.call("fsim_sync")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rec.reconstruct(raw_buf.value, out_buf.value, B)
    _lib.call("fsim_sync")
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    value = B * args.steps / elapsed
    log(f"batched: {value:.1f} recons/s, {ms_per_step:.3f} ms/step")
    import ctypes as C

    def stage_timed(steps):
        """`steps` calls with a HIP event around every kernel on its launch
        stream: (recons/s, ms per launch and launches per stage)."""
        _lib.call("fsim_stage_timing", 1)
        a = time.perf_counter()
        for _ in range(steps):
            rec.reconstruct(raw_buf.value, out_buf.value, B)
        _lib.call("fsim_sync")
        b = time.perf_counter()
        _lib.call("fsim_stage_timing", 0)
        ms = (C.c_double * 5)()
        cnt = (C.c_int * 5)()
        _lib.call("fsim_stage_times", ms, cnt, 5)
        return B * steps / (b - a), [ms[i] / max(cnt[i], 1) for i in range(5)], [int(c) for c in cnt]

    half = rec.p.mode == 0                    # real estimated OTF: the pairing is used in mode 0
    cs = 2 if half else 3
    sb = stage_bytes(n, Cs=cs, half=half, dc_sum=True)
    names = stage_names(half)

    def stage_table(stage_ms, cnt, k):
        return {names[i]: {"ms_per_launch": round(stage_ms[i], 4), "launches": cnt[i], "stacks_per_launch": k,
                           "bytes_per_stack": sb[i],
                           "GBps": round(sb[i] * k / (stage_ms[i] * 1e-3) / 1e9, 1) if stage_ms[i] else None}
                for i in range(5)}

    # the timed configuration with kernel events (concurrent chunks: a
    # kernel's event time includes the other stream's share of the chip)
    value_timed, ms_chunked, cnt_chunked = stage_timed(args.steps)
    stages_chunked = stage_table(ms_chunked, cnt_chunked, per_launch)
    # the kernels alone: the same B stacks as one launch on one stream, each
    # kernel's event time is its own execution; the roofline of the dominant
    # kernel comes from this region (profiles/ holds rocprofv3 of the same)
    autotune.set_chunks(*cls, 0, 1)
    for _ in range(10):
        rec.reconstruct(raw_buf.value, out_buf.value, B)
    value_whole, stage_ms, cnt_whole = stage_timed(args.steps)
    if tuned:
        autotune.set_chunks(*cls, tuned["chunk"], tuned["streams"])
    else:
        autotune.set_chunks(*cls, -1, 2)
    dom = int(np.argmax(stage_ms))
    dom_bytes = sb[dom] * B
    achieved = dom_bytes / (stage_ms[dom] * 1e-3) / 1e9
    stages = stage_table(stage_ms, cnt_whole, B)
    b_alg = sum(stage_bytes(n))               # reference dataflow: 434 n^2
    pipeline_frac = value * b_alg / (HBM_PEAK_GBS * 1e9)
    hbm_frac_moved = value * sum(sb) / (HBM_PEAK_GBS * 1e9)

    # single-stack latency of the fused path
    lat = []
    for _ in range(args.latency_runs):
        a = time.perf_counter()
        rec.reconstruct(raw_buf.value, out_buf.value, 1)
        _lib.call("fsim_sync")
        lat.append((time.perf_counter() - a) * 1000)
    # batch latency (every stack of a batch completes with the batch)
    blat = []
    for _ in range(10):
        a = time.perf_counter()
        rec.reconstruct(raw_buf.value, out_buf.value, B)
        _lib.call("fsim_sync")
        blat.append((time.perf_counter() - a) * 1000)

    # plain per-op path through the vector API
    log("plain per-op path")
    pipe = PlainPipeline(param)
    for _ in range(3):
        pipe.run(raw_buf.value)
    _lib.call("fsim_sync")
    plat = []
    for i in range(args.plain_runs):
        a = time.perf_counter()
        pipe.run(raw_buf.value + (i % 4) * stacks[0].nbytes)
        _lib.call("fsim_sync")
        plat.append((time.perf_counter() - a) * 1000)
    plain_value = 1000.0 / float(np.mean(plat))
    plain_err = fp64.rel_l2(pipe.run(raw_buf.value).vectorData(), r32)

    # served form: continuous batching through the request server
    log("served: 512^2 closed loop, device-resident stacks")
    served = serve_saturation(rec, raw_buf, out_buf, B, stacks[0].nbytes, rec.out_bytes(),
                              args.served_requests)
    log(f"served saturation {served['recons_per_s']:.1f} recons/s")
    if not args.no_mix:
        log("served: open-loop mixed stream at 200 req/s")
        served["mixed_stream"] = serve_mixed_stream(args.mix_requests, args.mix_rate)

    workloads = None
    if not args.no_workloads:
        log("512^2 parameter estimation")
        workloads = {"estimation_512": estimation()}
        log("mixed batch of 96 stacks (256/512/1024, 3 channels)")
        workloads["batched_mixed_96"] = batched_mixed()
        log("2048^2 large field of view at 65 nm")
        workloads["large_fov_2048"] = large_fov()
        log("served: large-FOV requests in the mixed Poisson stream")
        workloads["served_large_fov_mix"] = serve_large_fov_mix()
        log("live stream ingest: packets -> sequences -> server")
        workloads["stream_ingest_512"] = stream_ingest()
        log("live-mode side outputs: widefield, registration")
        workloads["side_outputs_512"] = side_outputs()
        log("filter styles: RL deconvolution, runReconstruction per style")
        workloads["filter_styles_512"] = filter_styles()
        log("configuration files: load, channel set-up")
        workloads["config_io"] = config_io()

    # CPU baseline: the oracle restatement on the host cores, bounded sample
    cpu = None
    if not args.no_cpu:
        def cpu_rate(nthreads, seconds, first):
            runs = min(200, max(3, int(seconds / max(first, 1e-3))))
            a = time.perf_counter()
            for i in range(runs):
                ref.recon32(p32, stacks[i % len(stacks)], threads=nthreads)
            return runs, time.perf_counter() - a
        log(f"cpu baseline, {threads} threads")
        runs, cpu_el = cpu_rate(threads, args.cpu_seconds, first_cpu)
        cpu = {"value": round(runs / cpu_el, 3), "unit": "recons/s", "cores": threads, "kind": "port",
               "sample": f"{runs} reconstructions of the 512^2 3x3 stacks of the GPU workload, FP32 Java-order "
                         f"C restatement of the reference CPU path (oracle/java_order.c, not the JVM), "
                         f"OpenMP {threads} threads, {cpu_el:.1f} s"}
        if threads > 1:
            log("cpu baseline, 1 thread")
            a = time.perf_counter()
            ref.recon32(p32, stacks[0], threads=1)
            first1 = time.perf_counter() - a
            runs1, el1 = cpu_rate(1, args.cpu_seconds, first1)
            cpu["one_core"] = {"value": round(runs1 / el1, 3), "unit": "recons/s", "cores": 1,
                               "sample": f"{runs1} reconstructions, 1 thread, {el1:.1f} s"}

    res = {
        "metric": "SIM recons/s + p50/p99 latency, 512² 3×3 raw; served vs plain GPU",
        "value": round(value, 2),
        "unit": "recons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded beads + filaments, Poisson noise), parameters given",
        "config": {"workload": f"fast-shift 2D SIM reconstruction, {n}^2 x 3 angles x 3 phases uint16 -> "
                               f"{2 * n}^2 float32, batched {B} stacks per step",
                   "n": n, "angles": 3, "phases": 3, "bands": 2, "batch": B},
        "roofline": {"bound": "hbm", "kernel": names[dom], "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(names[dom], n, B),
                     "bytes_per_launch": dom_bytes, "stacks_per_launch": B,
                     "ms_per_launch": round(stage_ms[dom], 4),
                     "measured": f"HIP events around each kernel, {B} stacks as one launch on one stream "
                                 f"(the kernel alone on the GPU)"},
        "pipeline_executed_bytes_frac": round(hbm_frac_moved, 4),
        "bytes_moved_per_recon": sum(sb),
        "reference_dataflow_equivalent_frac": round(pipeline_frac, 4),
        "reference_dataflow_bytes_per_recon": b_alg,
        "chunking": {"stacks_per_launch": per_launch, "streams": nstreams},
        "stages": stages,
        "one_launch": {"recons_per_s_with_kernel_events": round(value_whole, 2)},
        "stages_in_timed_configuration": stages_chunked,
        "value_with_kernel_events": round(value_timed, 2),
        "autotune": tuned,
        "latency_single_ms": {"p50": round(pct(lat, 50), 3), "p99": round(pct(lat, 99), 3)},
        "latency_batch_ms": {"p50": round(pct(blat, 50), 3), "p99": round(pct(blat, 99), 3)},
        "plain": {"recons_per_s": round(plain_value, 2), "p50_ms": round(pct(plat, 50), 3),
                  "p99_ms": round(pct(plat, 99), 3), "rel_l2_vs_oracle": plain_err},
        "served": served,
        "workloads": workloads,
        "speedup_vs_plain": round(value / plain_value, 2),
        "served_speedup_vs_plain": round(served["recons_per_s"] / plain_value, 2),
        "rel_l2_vs_oracle": parity,
        "rel_l2": parity64,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
