// Request server: continuous batching of reconstructions on one GPU.
//
// Modelled on the VIGOR live mode (livemode/ReconstructionRunner.java):
// queueImage() offers a stack to a bounded queue and refuses when it is full
// (:140-150, counted as missed frames), reconstruction threads take stacks and
// run the pipeline with per-channel filters kept on the device (:380-413),
// results go to an output queue (:57-58, 564).  Here one scheduler thread
// forms batches per size class (raw side and D/P/B), admits them when a
// device slot of that class is free or the memory budget allows a new one,
// and launches each batch on its own stream, so copies and kernels of
// different batches overlap; a completion thread polls the batch events and
// releases slots, which immediately admits the next requests.
#include "fsim_internal.h"
#include "recon.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>
#include <vector>

namespace fsim {

static double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
    uint64_t ticket = 0;
    int chan = 0;
    const uint16_t *raw = nullptr;
    float *out = nullptr;
    bool device = false;
    // a request with its own SIM parameters: `count` stacks (timepoints),
    // all reconstructed with sp
    bool params = false;
    int count = 1;
    fsim_sim_params sp{};
    double t_submit = 0, t_done = 0;
    int state = 0;  // 0 pending, 1 running, 2 done, 3 failed
    int rc = 0;
};

struct ClassKey {
    int n, D, P, B, Cs;   // Cs: stored components per direction
    int half;             // Hermitian result (ReconConst::half)
    int pcount;           // requests with their own parameters: stacks per request (one request per slot)
    bool operator==(const ClassKey &o) const {
        return n == o.n && D == o.D && P == o.P && B == o.B && Cs == o.Cs && half == o.half && pcount == o.pcount;
    }
};

struct Slot {
    ClassKey key{};
    int cap = 0;
    size_t bytes = 0;
    uint16_t *raw = nullptr;
    float *out = nullptr;
    float2 *A = nullptr, *B = nullptr;
    Job *jobs_d = nullptr, *jobs_h = nullptr;
    MultiWork *mw = nullptr;   // slots of requests with their own parameters
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::vector<Request *> reqs;
    bool busy = false;
    double t_launch = 0;
    uint64_t last_use = 0;     // order of the slot's last batch (eviction picks the oldest)
};

static size_t per_stack_bytes(const ClassKey &k) {
    const size_t nn = (size_t)k.n * k.n;
    const int C = k.Cs;
    return 2 * (size_t)k.D * k.P * nn + 8 * (size_t)k.D * C * nn + 32 * nn + 16 * nn;
}

// Device bytes of a slot of class k holding cap requests.
static size_t slot_bytes(const ClassKey &k, int cap) {
    if (k.pcount > 0) return per_stack_bytes(k) * k.pcount + multiwork_bytes(k.n, k.pcount, 1);
    return (per_stack_bytes(k) + sizeof(Job)) * cap;
}

struct Server {
    fsim_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> pending;
    std::unordered_map<uint64_t, Request *> reqs;
    std::vector<const fsim_channel_s *> channels;
    std::vector<Slot *> slots;
    std::vector<Slot *> inflight;
    size_t bytes = 0;
    uint64_t next_ticket = 1;
    uint64_t use_clock = 0;
    bool stop = false;
    std::thread sched, compl_;
    // statistics
    int64_t submitted = 0, completed = 0, rejected = 0, failed = 0, batches = 0, batched_reqs = 0;
    std::vector<double> lat;

    ~Server() {
        for (Slot *s : slots) free_slot(s);
        for (auto &kv : reqs) delete kv.second;
    }

    static void free_slot(Slot *s) {
        if (s->stream) (void)hipStreamSynchronize(s->stream);
        (void)hipFree(s->raw);
        (void)hipFree(s->out);
        (void)hipFree(s->A);
        (void)hipFree(s->B);
        (void)hipFree(s->jobs_d);
        if (s->jobs_h) (void)hipHostFree(s->jobs_h);
        if (s->mw) {
            multiwork_release(*s->mw);
            delete s->mw;
        }
        if (s->done) (void)hipEventDestroy(s->done);
        if (s->stream) (void)hipStreamDestroy(s->stream);
        delete s;
    }

    int cap_for(const ClassKey &k) const {
        if (k.pcount > 0) return 1;
        const size_t per = per_stack_bytes(k);
        const size_t share = (size_t)cfg.mem_budget / (size_t)std::max(cfg.nslots, 1);
        int cap = (int)std::min<size_t>((size_t)cfg.max_batch, std::max<size_t>(share / per, 1));
        return cap;
    }

    Slot *alloc_slot(const ClassKey &k) {
        Slot *s = new Slot();
        s->key = k;
        s->cap = cap_for(k);
        const size_t nn = (size_t)k.n * k.n;
        const int C = k.Cs;
        const size_t st = (size_t)s->cap * (k.pcount > 0 ? k.pcount : 1);   // stacks
        s->bytes = slot_bytes(k, s->cap);
        if (k.pcount > 0) s->mw = new MultiWork();
        bool ok = hipMalloc(&s->raw, 2 * (size_t)k.D * k.P * nn * st) == hipSuccess &&
                  hipMalloc(&s->out, 16 * nn * st) == hipSuccess &&
                  hipMalloc(&s->A, 8 * (size_t)k.D * C * nn * st) == hipSuccess &&
                  hipMalloc(&s->B, 32 * nn * st) == hipSuccess &&
                  hipMalloc(&s->jobs_d, sizeof(Job) * s->cap) == hipSuccess &&
                  hipHostMalloc(&s->jobs_h, sizeof(Job) * s->cap, hipHostMallocDefault) == hipSuccess &&
                  hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) == hipSuccess &&
                  hipEventCreateWithFlags(&s->done, hipEventDisableTiming) == hipSuccess;
        if (!ok) {
            (void)hipGetLastError();
            free_slot(s);
            return nullptr;
        }
        return s;
    }

    int busy_slots() const {
        int b = 0;
        for (const Slot *s : slots) b += s->busy ? 1 : 0;
        return b;
    }

    // Pick a free slot for class k: at most nslots batches are in flight; an
    // idle slot of class k is reused; else a new one is allocated within
    // mem_budget, evicting idle slots of other classes, least recently used
    // first, only as far as the budget needs (freeing device memory waits
    // for the device, so idle slots keep their memory while it fits).
    // Called with mu held.
    Slot *acquire_slot(const ClassKey &k) {
        if (busy_slots() >= cfg.nslots) return nullptr;
        for (Slot *s : slots)
            if (!s->busy && s->key == k) return s;
        const size_t need = slot_bytes(k, cap_for(k));
        while (bytes + need > (size_t)cfg.mem_budget) {
            Slot *victim = nullptr;
            for (Slot *s : slots)
                if (!s->busy && !(s->key == k) && (!victim || s->last_use < victim->last_use)) victim = s;
            if (!victim) return nullptr;
            bytes -= victim->bytes;
            slots.erase(std::find(slots.begin(), slots.end(), victim));
            free_slot(victim);
        }
        Slot *s = alloc_slot(k);
        if (!s) return nullptr;
        slots.push_back(s);
        bytes += s->bytes;
        return s;
    }

    // A slot for the class of a newly added channel, so that its first
    // requests do not wait for the allocation (skipped when the budget has
    // no room for it beside the existing slots).  Called with mu held.
    void warm_slot(const ClassKey &k) {
        for (Slot *s : slots)
            if (s->key == k) return;
        if (bytes + slot_bytes(k, cap_for(k)) > (size_t)cfg.mem_budget) return;
        Slot *s = alloc_slot(k);
        if (!s) return;
        slots.push_back(s);
        bytes += s->bytes;
    }

    ClassKey key_of(const Request *r) const {
        const ReconConst &K = channel_host_const(channels[r->chan]);
        return ClassKey{K.n, K.D, K.P, K.B, K.Cs, K.half, r->params ? r->count : 0};
    }

    // One request with its own parameters: its stacks through the
    // per-parameter pipeline on the slot's buffers.
    void launch_params(Slot *s) {
        const ClassKey k = s->key;
        const size_t nn = (size_t)k.n * k.n;
        const size_t raw_elems = (size_t)k.D * k.P * nn * k.pcount, out_elems = 4 * nn * k.pcount;
        Request *r = s->reqs[0];
        int rc = FSIM_OK;
        const uint16_t *raw = r->raw;
        float *out = r->out;
        if (!r->device) {
            raw = s->raw;
            out = s->out;
            if (hipMemcpyAsync((void *)raw, r->raw, raw_elems * 2, hipMemcpyHostToDevice, s->stream) != hipSuccess)
                rc = FSIM_ERR_HIP;
        }
        if (!rc)
            rc = reconstruct_params_with(channels[r->chan], *s->mw, s->A, s->B, raw, out, k.pcount, &r->sp, 0,
                                         s->stream);
        if (!rc && !r->device &&
            hipMemcpyAsync(r->out, out, out_elems * 4, hipMemcpyDeviceToHost, s->stream) != hipSuccess)
            rc = FSIM_ERR_HIP;
        if (hipEventRecord(s->done, s->stream) != hipSuccess) rc = FSIM_ERR_HIP;
        if (rc) r->rc = rc;
    }

    void launch(Slot *s) {
        if (s->key.pcount > 0) {
            launch_params(s);
            return;
        }
        const ClassKey k = s->key;
        const size_t nn = (size_t)k.n * k.n;
        const int C = k.Cs;
        const size_t raw_elems = (size_t)k.D * k.P * nn;
        int rc = FSIM_OK;
        for (size_t i = 0; i < s->reqs.size(); i++) {
            Request *r = s->reqs[i];
            Job &j = s->jobs_h[i];
            j.c = channel_device_const(channels[r->chan]);
            j.A = s->A + i * (size_t)k.D * C * nn;
            j.B = s->B + i * 4 * nn;
            if (r->device) {
                j.raw = r->raw;
                j.out = r->out;
            } else {
                j.raw = s->raw + i * raw_elems;
                j.out = s->out + i * 4 * nn;
                if (hipMemcpyAsync((void *)j.raw, r->raw, raw_elems * 2, hipMemcpyHostToDevice, s->stream) !=
                    hipSuccess)
                    rc = FSIM_ERR_HIP;
            }
        }
        const int count = (int)s->reqs.size();
        if (hipMemcpyAsync(s->jobs_d, s->jobs_h, sizeof(Job) * count, hipMemcpyHostToDevice, s->stream) != hipSuccess)
            rc = FSIM_ERR_HIP;
        JobList jl{};
        jl.jobs = s->jobs_d;
        jl.raw_aligned = true;
        // the half-spectrum pass writes output pixel pairs (8-byte stores)
        bool half = k.half != 0;
        for (int i = 0; i < count; i++) {
            jl.raw_aligned &= ((uintptr_t)s->jobs_h[i].raw & 7) == 0;
            half &= ((uintptr_t)s->jobs_h[i].out & 7) == 0;
        }
        if (!rc) rc = launch_pipeline(k.n, jl, count, k.D, k.P, C, half, s->stream);
        for (int i = 0; i < count && !rc; i++) {
            Request *r = s->reqs[i];
            if (!r->device &&
                hipMemcpyAsync(r->out, s->jobs_h[i].out, 16 * nn, hipMemcpyDeviceToHost, s->stream) != hipSuccess)
                rc = FSIM_ERR_HIP;
        }
        if (hipEventRecord(s->done, s->stream) != hipSuccess) rc = FSIM_ERR_HIP;
        if (rc)
            for (Request *r : s->reqs) r->rc = rc;
    }

    void scheduler() {
        std::unique_lock<std::mutex> lk(mu);
        while (!stop) {
            if (pending.empty()) {
                cv_work.wait(lk);
                continue;
            }
            const ClassKey k = key_of(pending.front());
            Slot *s = acquire_slot(k);
            if (!s) {
                if (inflight.empty()) {
                    // nothing on the device will free a slot or memory: the
                    // slot could not be allocated, so the request fails
                    // (the reference turns a failed allocation into
                    // OutOfMemoryError, AccelVectorCplx.java:45-54)
                    Request *r = pending.front();
                    pending.pop_front();
                    r->rc = FSIM_ERR_OOM;
                    r->state = 3;
                    r->t_done = now_ms();
                    failed++;
                    cv_done.notify_all();
                    continue;
                }
                cv_work.wait(lk);  // a slot completes and notifies
                continue;
            }
            s->busy = true;
            s->last_use = ++use_clock;
            s->reqs.clear();
            for (auto it = pending.begin(); it != pending.end() && (int)s->reqs.size() < s->cap;) {
                if (key_of(*it) == k) {
                    (*it)->state = 1;
                    s->reqs.push_back(*it);
                    it = pending.erase(it);
                } else {
                    ++it;
                }
            }
            batches++;
            batched_reqs += (int64_t)s->reqs.size();
            s->t_launch = now_ms();
            lk.unlock();
            launch(s);
            lk.lock();
            inflight.push_back(s);
            cv_done.notify_all();
        }
    }

    void completer() {
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            if (inflight.empty()) {
                if (stop) break;
                cv_done.wait_for(lk, std::chrono::milliseconds(20));
                continue;
            }
            std::vector<Slot *> fl = inflight;
            lk.unlock();
            std::vector<Slot *> fin;
            for (Slot *s : fl) {
                hipError_t e = hipEventQuery(s->done);
                if (e == hipSuccess) fin.push_back(s);
                else if (e != hipErrorNotReady) {
                    for (Request *r : s->reqs) r->rc = FSIM_ERR_HIP;
                    fin.push_back(s);
                }
            }
            const double t = now_ms();
            lk.lock();
            if (fin.empty()) {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(50));
                lk.lock();
                continue;
            }
            for (Slot *s : fin) {
                for (Request *r : s->reqs) {
                    r->t_done = t;
                    r->state = r->rc ? 3 : 2;
                    if (r->rc) failed++;
                    else {
                        completed++;
                        lat.push_back(t - r->t_submit);
                    }
                }
                s->reqs.clear();
                s->busy = false;
                inflight.erase(std::find(inflight.begin(), inflight.end(), s));
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

}  // namespace fsim

using namespace fsim;

struct fsim_server_s {
    Server s;
};

extern "C" {

int fsim_server_create(const fsim_server_config *cfg, fsim_server *out) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!cfg || !out || cfg->mem_budget <= 0 || cfg->max_batch < 1 || cfg->max_queue < 1 || cfg->nslots < 1)
        return fail(FSIM_ERR_ARG, "server_create: bad configuration");
    fsim_server_s *h = new fsim_server_s();
    h->s.cfg = *cfg;
    h->s.sched = std::thread([h] { h->s.scheduler(); });
    h->s.compl_ = std::thread([h] { h->s.completer(); });
    *out = h;
    return FSIM_OK;
}

int fsim_server_destroy(fsim_server h) {
    if (!h) return FSIM_OK;
    {
        std::lock_guard<std::mutex> lk(h->s.mu);
        h->s.stop = true;
    }
    h->s.cv_work.notify_all();
    h->s.cv_done.notify_all();
    h->s.sched.join();
    h->s.compl_.join();
    delete h;
    return FSIM_OK;
}

int fsim_server_add_channel(fsim_server h, fsim_channel ch, int *id) {
    if (!h || !ch || !id) return fail(FSIM_ERR_ARG, "server_add_channel: null argument");
    std::lock_guard<std::mutex> lk(h->s.mu);
    h->s.channels.push_back(ch);
    *id = (int)h->s.channels.size() - 1;
    Request probe;
    probe.chan = *id;
    h->s.warm_slot(h->s.key_of(&probe));
    return FSIM_OK;
}

int fsim_server_submit(fsim_server h, int chan, const uint16_t *raw, float *out, int flags, uint64_t *ticket) {
    if (!h || !raw || !out || !ticket) return fail(FSIM_ERR_ARG, "server_submit: null argument");
    Server &s = h->s;
    std::lock_guard<std::mutex> lk(s.mu);
    if (chan < 0 || chan >= (int)s.channels.size()) return fail(FSIM_ERR_ARG, "server_submit: unknown channel");
    if ((int)s.pending.size() >= s.cfg.max_queue) {
        s.rejected++;
        return fail(FSIM_ERR_BUSY, "server_submit: queue full, request refused");
    }
    {
        // a request whose slot can never fit the budget is refused now (it
        // would otherwise wait at the head of the queue for memory that no
        // completion frees)
        Request probe;
        probe.chan = chan;
        if (slot_bytes(s.key_of(&probe), 1) > (size_t)s.cfg.mem_budget) {
            s.rejected++;
            return fail(FSIM_ERR_BUSY, "server_submit: request larger than the memory budget");
        }
    }
    Request *r = new Request();
    r->ticket = s.next_ticket++;
    r->chan = chan;
    r->raw = raw;
    r->out = out;
    r->device = (flags & FSIM_SUBMIT_DEVICE) != 0;
    r->t_submit = now_ms();
    s.reqs[r->ticket] = r;
    s.pending.push_back(r);
    s.submitted++;
    *ticket = r->ticket;
    s.cv_work.notify_one();
    return FSIM_OK;
}

int fsim_server_submit_params(fsim_server h, int chan, const uint16_t *raw, float *out, int count,
                              const fsim_sim_params *sp, int flags, uint64_t *ticket) {
    if (!h || !raw || !out || !sp || !ticket) return fail(FSIM_ERR_ARG, "server_submit_params: null argument");
    if (count < 1 || count > 4096) return fail(FSIM_ERR_ARG, "server_submit_params: count must be in [1, 4096]");
    Server &s = h->s;
    std::lock_guard<std::mutex> lk(s.mu);
    if (chan < 0 || chan >= (int)s.channels.size())
        return fail(FSIM_ERR_ARG, "server_submit_params: unknown channel");
    if ((int)s.pending.size() >= s.cfg.max_queue) {
        s.rejected++;
        return fail(FSIM_ERR_BUSY, "server_submit_params: queue full, request refused");
    }
    Request probe;
    probe.chan = chan;
    probe.params = true;
    probe.count = count;
    if (slot_bytes(s.key_of(&probe), 1) > (size_t)s.cfg.mem_budget) {
        s.rejected++;
        return fail(FSIM_ERR_BUSY, "server_submit_params: request larger than the memory budget");
    }
    Request *r = new Request();
    r->ticket = s.next_ticket++;
    r->chan = chan;
    r->raw = raw;
    r->out = out;
    r->device = (flags & FSIM_SUBMIT_DEVICE) != 0;
    r->params = true;
    r->count = count;
    r->sp = *sp;
    r->t_submit = now_ms();
    s.reqs[r->ticket] = r;
    s.pending.push_back(r);
    s.submitted++;
    *ticket = r->ticket;
    s.cv_work.notify_one();
    return FSIM_OK;
}

int fsim_server_wait(fsim_server h, uint64_t ticket, int timeout_ms, double *latency_ms) {
    if (!h) return fail(FSIM_ERR_ARG, "server_wait: null server");
    Server &s = h->s;
    std::unique_lock<std::mutex> lk(s.mu);
    auto it = s.reqs.find(ticket);
    if (it == s.reqs.end()) return fail(FSIM_ERR_ARG, "server_wait: unknown ticket");
    Request *r = it->second;
    auto pred = [r] { return r->state >= 2; };
    if (timeout_ms < 0)
        s.cv_done.wait(lk, pred);
    else if (!s.cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred))
        return fail(FSIM_ERR_BUSY, "server_wait: timed out");
    if (latency_ms) *latency_ms = r->t_done - r->t_submit;
    const int rc = r->rc;
    s.reqs.erase(it);
    delete r;
    if (rc == FSIM_ERR_OOM) return fail(rc, "server_wait: no device memory for the request");
    if (rc) return fail(rc, "server_wait: reconstruction failed on the device");
    return FSIM_OK;
}

int fsim_server_get_stats(fsim_server h, fsim_server_stats *st, int reset) {
    if (!h || !st) return fail(FSIM_ERR_ARG, "server_stats: null argument");
    Server &s = h->s;
    std::lock_guard<std::mutex> lk(s.mu);
    st->submitted = s.submitted;
    st->completed = s.completed;
    st->rejected = s.rejected;
    st->failed = s.failed;
    st->batches = s.batches;
    st->mean_batch = s.batches ? (double)s.batched_reqs / s.batches : 0;
    st->device_bytes = (int64_t)s.bytes;
    st->pending = (int64_t)s.pending.size();
    std::vector<double> l = s.lat;
    std::sort(l.begin(), l.end());
    auto q = [&](double p) { return l.empty() ? 0.0 : l[std::min(l.size() - 1, (size_t)(p * (l.size() - 1) + 0.5))]; };
    st->p50_ms = q(0.50);
    st->p99_ms = q(0.99);
    st->max_ms = l.empty() ? 0.0 : l.back();
    if (reset) {
        s.submitted = s.completed = s.rejected = s.failed = s.batches = s.batched_reqs = 0;
        s.lat.clear();
    }
    return FSIM_OK;
}

}  // extern "C"
