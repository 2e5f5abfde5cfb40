// Fixed-shape batches of reconstructions of mixed sizes and channels, run as
// one graph launch.
//
// A batch is a list of slots, each bound to a channel (its raw side n, its
// directions/phases/bands and tables).  At creation the slots are grouped into
// size classes (n, D, P, stored components, Hermitian result); each class gets
// its work buffers, a device job list and a stream.  A run forks the caller's
// stream into the class streams, runs each class's pipeline over its job list
// (every kernel walks (stack, tile) or (stack, row) work items of its class),
// and joins them back, so the classes' kernels overlap on the device.  The
// whole launch sequence is captured once as a hipGraph and replayed by each
// run; the raw and output addresses of the slots may change from run to run:
// they live in the device job lists, which a run rewrites (one copy from
// pinned host memory, stream-ordered before the graph) -- the graph holds only
// kernels, no memory set nodes.
//
// With FSIM_BATCH_PERSISTENT the classes of sizes 256 to 2048 run instead
// through persistent work-list kernels: per launch group (n = 256 and 512;
// n = 1024 and 2048) and stage, one kernel walks a device list of (class,
// block) items that holds the blocks of every class's own launch of that
// stage, larger size first, claiming them through a cursor that a small
// kernel zeroes at the start of each run (see launch_persistent).  The two groups run on
// streams of their own, concurrently; other sizes keep their class launches.
//
// The reference runs the same reconstruction per image from several threads
// (livemode/ReconstructionRunner.java:243-251, 380-413); batching images of
// different sizes and channels into one device launch has no counterpart
// there.
#include "fsim_internal.h"
#include "recon.h"

#include <algorithm>
#include <vector>

namespace fsim {

struct BatchClass {
    int n = 0, D = 0, P = 0, Cs = 0;
    bool half = false;
    std::vector<int> items;   // slot indices, in order
    int off = 0;              // first job of the class in the job arrays
    float2 *A = nullptr, *B = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
};

// A launch group of the persistent form: its classes (device copies, which
// the item lists index), one item list per stage, a stream.
struct PGroupRun {
    std::vector<PClass> hcls;
    PClass *cls_d = nullptr;
    int4 *items_d = nullptr;
    int *cur_d = nullptr;        // the stage lists' cursors
    int off[kPStages] = {};
    int ends[kPStages][4] = {};
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
};

bool stage_timing_enabled();

}  // namespace fsim

using namespace fsim;

struct fsim_batch_s {
    int count = 0;
    std::vector<fsim_channel> chans;
    std::vector<BatchClass> cls;
    std::vector<int> job_of;       // slot -> job index
    Job *jobs_h = nullptr;         // pinned staging of the job arrays
    Job *jobs_d = nullptr;
    hipStream_t origin = nullptr;
    hipEvent_t fork = nullptr, staged = nullptr, joined = nullptr, start = nullptr;
    bool use_graph = false, captured = false, persistent = false;
    PGroupRun pg[kPGroups];
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int64_t device_bytes = 0;
};

static void batch_free(fsim_batch_s *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (BatchClass &c : b->cls) {
        (void)hipFree(c.A);
        (void)hipFree(c.B);
        if (c.stream) (void)hipStreamDestroy(c.stream);
        if (c.done) (void)hipEventDestroy(c.done);
    }
    for (PGroupRun &g : b->pg) {
        (void)hipFree(g.cls_d);
        (void)hipFree(g.items_d);
        (void)hipFree(g.cur_d);
        if (g.stream) (void)hipStreamDestroy(g.stream);
        if (g.done) (void)hipEventDestroy(g.done);
    }
    (void)hipFree(b->jobs_d);
    if (b->jobs_h) (void)hipHostFree(b->jobs_h);
    if (b->origin) (void)hipStreamDestroy(b->origin);
    for (hipEvent_t e : {b->fork, b->staged, b->joined, b->start})
        if (e) (void)hipEventDestroy(e);
    delete b;
}

static bool in_persistent_group(const fsim_batch_s *b, const BatchClass &c) {
    return b->persistent && persistent_group(c.n) >= 0;
}

// The launch sequence of one run on b->origin: fork into the group and class
// streams, the persistent stages per group and one pipeline per other class,
// join.
static int batch_launch(fsim_batch_s *b) {
    FSIM_HIP(hipEventRecord(b->start, b->origin));
    for (int gi = 0; gi < kPGroups; gi++) {
        PGroupRun &g = b->pg[gi];
        if (g.hcls.empty()) continue;
        FSIM_HIP(hipStreamWaitEvent(g.stream, b->start, 0));
        int rc = launch_persistent_reset(g.cur_d, kPStages * kPListCursors, g.stream);
        for (int st = 0; st < kPStages && !rc; st++)
            rc = launch_persistent(gi, st, g.cls_d, g.items_d + g.off[st], g.ends[st], g.cur_d + st * kPListCursors,
                                   g.stream);
        if (rc) return rc;
        FSIM_HIP(hipEventRecord(g.done, g.stream));
        FSIM_HIP(hipStreamWaitEvent(b->origin, g.done, 0));
    }
    for (BatchClass &c : b->cls) {
        if (in_persistent_group(b, c)) continue;
        FSIM_HIP(hipStreamWaitEvent(c.stream, b->start, 0));
        JobList jl{};
        jl.jobs = b->jobs_d + c.off;
        jl.raw_aligned = true;
        const int rc = launch_pipeline(c.n, jl, (int)c.items.size(), c.D, c.P, c.Cs, c.half, c.stream);
        if (rc) return rc;
        FSIM_HIP(hipEventRecord(c.done, c.stream));
        FSIM_HIP(hipStreamWaitEvent(b->origin, c.done, 0));
    }
    return FSIM_OK;
}

extern "C" {

int fsim_batch_create(const fsim_channel *channels, int count, int flags, fsim_batch *out) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!channels || !out || count < 1 || count > 65535) return fail(FSIM_ERR_ARG, "batch_create: bad arguments");
    for (int i = 0; i < count; i++)
        if (!channels[i]) return fail(FSIM_ERR_ARG, "batch_create: null channel");
    fsim_batch_s *b = new fsim_batch_s();
    b->count = count;
    b->chans.assign(channels, channels + count);
    if (flags & ~(FSIM_BATCH_GRAPH | FSIM_BATCH_PERSISTENT)) {
        delete b;
        return fail(FSIM_ERR_ARG, "batch_create: unknown flags");
    }
    b->use_graph = (flags & FSIM_BATCH_GRAPH) != 0;
    b->persistent = (flags & FSIM_BATCH_PERSISTENT) != 0;
    // size classes, in order of first appearance
    for (int i = 0; i < count; i++) {
        const ReconConst &K = channel_host_const(channels[i]);
        BatchClass *hit = nullptr;
        for (BatchClass &c : b->cls)
            if (c.n == K.n && c.D == K.D && c.P == K.P && c.Cs == K.Cs && c.half == (K.half != 0)) hit = &c;
        if (!hit) {
            b->cls.emplace_back();
            hit = &b->cls.back();
            hit->n = K.n;
            hit->D = K.D;
            hit->P = K.P;
            hit->Cs = K.Cs;
            hit->half = K.half != 0;
        }
        hit->items.push_back(i);
    }
    b->job_of.assign(count, 0);
    int off = 0;
    for (BatchClass &c : b->cls) {
        c.off = off;
        for (size_t j = 0; j < c.items.size(); j++) b->job_of[c.items[j]] = off + (int)j;
        off += (int)c.items.size();
    }
    bool ok = hipHostMalloc(&b->jobs_h, sizeof(Job) * count, hipHostMallocDefault) == hipSuccess &&
              hipMalloc(&b->jobs_d, sizeof(Job) * count) == hipSuccess &&
              hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking) == hipSuccess;
    for (hipEvent_t *e : {&b->fork, &b->staged, &b->joined, &b->start})
        ok = ok && hipEventCreateWithFlags(e, hipEventDisableTiming) == hipSuccess;
    for (BatchClass &c : b->cls) {
        if (!ok) break;
        const size_t k = c.items.size();
        const size_t ae = workspace_a_elems(c.n, c.D, c.Cs), be = workspace_b_elems(c.n);
        ok = hipMalloc(&c.A, ae * k * sizeof(float2)) == hipSuccess &&
             hipMalloc(&c.B, be * k * sizeof(float2)) == hipSuccess &&
             hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&c.done, hipEventDisableTiming) == hipSuccess;
        b->device_bytes += (int64_t)((ae + be) * k * sizeof(float2));
        // the twiddle tables are allocated here, outside any capture
        ok = ok && twiddles(c.n) && twiddles(2 * c.n);
        for (size_t j = 0; ok && j < k; j++) {
            Job &J = b->jobs_h[c.off + j];
            J.c = channel_device_const(channels[c.items[j]]);
            J.raw = nullptr;
            J.out = nullptr;
            J.A = c.A + ae * j;
            J.B = c.B + be * j;
        }
    }
    // persistent groups: class table and the stages' item lists
    for (BatchClass &c : b->cls) {
        if (!ok || !in_persistent_group(b, c)) continue;
        PGroupRun &g = b->pg[persistent_group(c.n)];
        PClass pc{};
        pc.jl.jobs = b->jobs_d + c.off;
        pc.jl.raw_aligned = true;
        pc.n = c.n;
        pc.njobs = (int)c.items.size();
        pc.D = c.D;
        pc.C = c.Cs;
        pc.half = c.half ? 1 : 0;
        // the work-list kernels hold up to 9 frames per row (the one-unit
        // separation gives the same bits otherwise)
        pc.sep_all = c.n <= 512 && c.D * c.P <= 9 ? 1 : 0;
        pc.twn = twiddles(c.n);
        pc.twm = twiddles(2 * c.n);
        ok = pc.twn && pc.twm;
        g.hcls.push_back(pc);
    }
    for (PGroupRun &g : b->pg) {
        if (!ok || g.hcls.empty()) continue;
        std::vector<int4> all, st;
        for (int k = 0; k < kPStages; k++) {
            persistent_list(g.hcls.data(), (int)g.hcls.size(), k, st, g.ends[k]);
            g.off[k] = (int)all.size();
            all.insert(all.end(), st.begin(), st.end());
        }
        const size_t cb = sizeof(PClass) * g.hcls.size(), ib = sizeof(int4) * std::max<size_t>(all.size(), 1);
        const size_t kb = sizeof(int) * kPStages * kPListCursors;
        ok = hipMalloc(&g.cls_d, cb) == hipSuccess && hipMalloc(&g.items_d, ib) == hipSuccess &&
             hipMalloc(&g.cur_d, kb) == hipSuccess &&
             hipMemcpy(g.cls_d, g.hcls.data(), cb, hipMemcpyHostToDevice) == hipSuccess &&
             (all.empty() || hipMemcpy(g.items_d, all.data(), sizeof(int4) * all.size(), hipMemcpyHostToDevice) ==
                                 hipSuccess) &&
             hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&g.done, hipEventDisableTiming) == hipSuccess;
        b->device_bytes += (int64_t)(cb + ib + kb);
    }
    if (!ok) {
        (void)hipGetLastError();
        batch_free(b);
        return fail(FSIM_ERR_OOM, "batch_create: device allocation failed");
    }
    b->device_bytes += (int64_t)(sizeof(Job) * count);
    *out = b;
    return FSIM_OK;
}

int fsim_batch_run(fsim_batch b, const uint16_t *const *raw, float *const *out, void *strm) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!b || !raw || !out) return fail(FSIM_ERR_ARG, "batch_run: bad arguments");
    for (int i = 0; i < b->count; i++) {
        if (!raw[i] || !out[i]) return fail(FSIM_ERR_ARG, "batch_run: null raw or output address");
        // the captured launch uses the 4-pixel raw loads and the pixel-pair
        // stores of the half-spectrum pass
        if (((uintptr_t)raw[i] & 7) || ((uintptr_t)out[i] & 7))
            return fail(FSIM_ERR_ARG, "batch_run: raw and output addresses must be 8-byte aligned");
    }
    hipStream_t s = strm ? (hipStream_t)strm : stream();
    // the previous run's copy of the staging array has completed
    FSIM_HIP(hipEventSynchronize(b->staged));
    for (int i = 0; i < b->count; i++) {
        Job &J = b->jobs_h[b->job_of[i]];
        J.raw = raw[i];
        J.out = out[i];
    }
    FSIM_HIP(hipEventRecord(b->fork, s));
    FSIM_HIP(hipStreamWaitEvent(b->origin, b->fork, 0));
    FSIM_HIP(hipMemcpyAsync(b->jobs_d, b->jobs_h, sizeof(Job) * b->count, hipMemcpyHostToDevice, b->origin));
    FSIM_HIP(hipEventRecord(b->staged, b->origin));
    if (b->use_graph && !stage_timing_enabled()) {
        if (!b->captured) {
            FSIM_HIP(hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal));
            rc = batch_launch(b);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(b->origin, &g);
            if (rc || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                (void)hipGetLastError();
                return rc ? rc : fail(FSIM_ERR_HIP, std::string("batch_run: graph capture failed: ") +
                                                        hipGetErrorString(e));
            }
            if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
                (void)hipGraphDestroy(g);
                (void)hipGetLastError();
                return fail(FSIM_ERR_HIP, "batch_run: graph instantiation failed");
            }
            b->graph = g;
            b->captured = true;
        }
        FSIM_HIP(hipGraphLaunch(b->exec, b->origin));
    } else {
        rc = batch_launch(b);
        if (rc) return rc;
    }
    FSIM_HIP(hipEventRecord(b->joined, b->origin));
    FSIM_HIP(hipStreamWaitEvent(s, b->joined, 0));
    return FSIM_OK;
}

int fsim_batch_info(fsim_batch b, int *nclasses, int *captured, int64_t *device_bytes) {
    if (!b) return fail(FSIM_ERR_ARG, "batch_info: null batch");
    if (nclasses) *nclasses = (int)b->cls.size();
    if (captured) *captured = b->captured ? 1 : 0;
    if (device_bytes) *device_bytes = b->device_bytes;
    return FSIM_OK;
}

int fsim_batch_destroy(fsim_batch b) {
    if (!b) return FSIM_OK;
    (void)hipStreamSynchronize(b->origin);
    batch_free(b);
    return FSIM_OK;
}

}  // extern "C"
