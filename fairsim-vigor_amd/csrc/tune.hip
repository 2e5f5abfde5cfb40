// Autotuner of the fused pipeline's launch shapes, with a persisted cache.
//
// Per size class (raw side n, directions D, phases P, stored components Cs,
// Hermitian result) the pipeline has launch choices that change speed, not
// the algorithm: the separation pass with all units of a row in one wavefront
// or one unit per block, the column pass's elements per thread (the
// Stockham radix: 8 or 16) and columns per block (8 or 16), and how a call's
// stacks are split into chunks run round-robin on library streams (each
// stream reusing its own work buffers).  fsim_autotune times every candidate
// on the channel's class with HIP events (zero-filled stacks: the kernels'
// work does not depend on the pixel values), after a warm-up and in
// interleaved rounds so that clock ramps and cache states do not favour a
// position in the list, and keeps the fastest in a process-wide cache that
// the launches consult; the cache is saved to and loaded from a text file.
// Without an entry the defaults apply.
//
// No counterpart in the reference (its cuFFT plans are fixed per size,
// accel/FFTProvider.java:63-102).
#include "fsim_internal.h"
#include "recon.h"

#include <algorithm>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

namespace fsim {

using TuneKey = std::tuple<int, int, int, int, int>;   // n, D, P, Cs, half

struct TuneEntry {
    PipeCfg cfg;
    double ms = 0;   // measured time per launch of the tuned batch, 0 if set by hand
    int count = 0;   // stacks per launch when tuned
};

static std::mutex g_tune_mu;
static std::map<TuneKey, TuneEntry> g_tune;

bool pipe_cfg_lookup(int n, int D, int P, int Cs, bool half, PipeCfg &cfg) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(TuneKey{n, D, P, Cs, half ? 1 : 0});
    if (it == g_tune.end()) return false;
    cfg = it->second.cfg;
    return true;
}

static int check_cfg(int n, const PipeCfg &c) {
    if (!recon_size_ok(n)) return fail(FSIM_ERR_ARG, "autotune: n must be a power of two in [64, 2048]");
    if (c.sep < 0 || c.sep > 1 || c.col < 0 || c.col >= kColCands)
        return fail(FSIM_ERR_ARG, "autotune: sep must be 0 or 1 and col in [0, " + std::to_string(kColCands) + ")");
    if (c.chunk < -1 || c.chunk > 4096 || c.streams < 1 || c.streams > kMaxChunkStreams)
        return fail(FSIM_ERR_ARG, "autotune: chunk must be in [-1, 4096] and streams in [1, " +
                                      std::to_string(kMaxChunkStreams) + "]");
    return FSIM_OK;
}

}  // namespace fsim

using namespace fsim;

extern "C" {

int fsim_autotune_set(int n, int D, int P, int Cs, int half, int sep, int col) {
    PipeCfg c;
    c.sep = sep;
    c.col = col;
    const int rc = check_cfg(n, c);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(TuneKey{n, D, P, Cs, half ? 1 : 0});
    if (it != g_tune.end()) {
        // keep the class's chunking choice
        c.chunk = it->second.cfg.chunk;
        c.streams = it->second.cfg.streams;
    }
    TuneEntry &e = g_tune[TuneKey{n, D, P, Cs, half ? 1 : 0}];
    e.cfg = c;
    e.ms = 0;
    e.count = 0;
    return FSIM_OK;
}

int fsim_autotune_set_chunks(int n, int D, int P, int Cs, int half, int chunk, int streams) {
    PipeCfg c;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(TuneKey{n, D, P, Cs, half ? 1 : 0});
    if (it != g_tune.end()) c = it->second.cfg;
    c.chunk = chunk;
    c.streams = streams;
    const int rc = check_cfg(n, c);
    if (rc) return rc;
    TuneEntry &e = g_tune[TuneKey{n, D, P, Cs, half ? 1 : 0}];
    e.cfg = c;
    e.ms = 0;
    e.count = 0;
    return FSIM_OK;
}

int fsim_autotune_get_chunks(int n, int D, int P, int Cs, int half, int count, int *chunk, int *streams) {
    if (!chunk || !streams) return fail(FSIM_ERR_ARG, "autotune_get_chunks: null output");
    if (!recon_size_ok(n) || D < 1 || D > kMaxDir || Cs < 1 || Cs > kMaxComp || count < 1)
        return fail(FSIM_ERR_ARG, "autotune_get_chunks: bad class");
    PipeCfg c;
    if (!pipe_cfg_lookup(n, D, P, Cs, half != 0, c)) c = PipeCfg{};
    int K = count, S = 1;
    chunk_plan(n, D, Cs, count, c, K, S);
    *chunk = K;
    *streams = S;
    return FSIM_OK;
}

int fsim_autotune_get(int n, int D, int P, int Cs, int half, int *sep, int *col) {
    if (!sep || !col) return fail(FSIM_ERR_ARG, "autotune_get: null output");
    PipeCfg c;
    if (pipe_cfg_lookup(n, D, P, Cs, half != 0, c)) {
        *sep = c.sep;
        *col = c.col;
    } else {
        *sep = -1;
        *col = -1;
    }
    return FSIM_OK;
}

int fsim_autotune_clear(void) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.clear();
    return FSIM_OK;
}

// Text file: a header line "fsim-autotune 2", then one line per class:
// "n D P Cs half sep col chunk streams count ms".  Version 1 files (without
// chunk and streams: the size defaults apply) are read too.
int fsim_autotune_save(const char *path) {
    if (!path) return fail(FSIM_ERR_ARG, "autotune_save: null path");
    FILE *f = std::fopen(path, "w");
    if (!f) return fail(FSIM_ERR_ARG, std::string("autotune_save: cannot write ") + path);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::fprintf(f, "fsim-autotune 2\n");
    for (const auto &kv : g_tune) {
        const TuneKey &k = kv.first;
        const TuneEntry &e = kv.second;
        std::fprintf(f, "%d %d %d %d %d %d %d %d %d %d %.6f\n", std::get<0>(k), std::get<1>(k), std::get<2>(k),
                     std::get<3>(k), std::get<4>(k), e.cfg.sep, e.cfg.col, e.cfg.chunk, e.cfg.streams, e.count,
                     e.ms);
    }
    const bool ok = std::fclose(f) == 0;
    return ok ? FSIM_OK : fail(FSIM_ERR_ARG, std::string("autotune_save: write failed: ") + path);
}

// Merges the file's entries into the cache (an entry replaces the one of its
// class).  A malformed file changes nothing.
int fsim_autotune_load(const char *path) {
    if (!path) return fail(FSIM_ERR_ARG, "autotune_load: null path");
    FILE *f = std::fopen(path, "r");
    if (!f) return fail(FSIM_ERR_ARG, std::string("autotune_load: cannot read ") + path);
    char magic[32] = {0};
    int ver = 0;
    if (std::fscanf(f, "%31s %d", magic, &ver) != 2 || std::string(magic) != "fsim-autotune" ||
        (ver != 1 && ver != 2)) {
        std::fclose(f);
        return fail(FSIM_ERR_ARG, "autotune_load: not an fsim-autotune version 1 or 2 file");
    }
    std::vector<std::pair<TuneKey, TuneEntry>> got;
    while (true) {
        int n, D, P, Cs, half, sep, col, count, chunk = -1, streams = 2;
        double ms;
        int r, want;
        if (ver == 1) {
            r = std::fscanf(f, "%d %d %d %d %d %d %d %d %lf", &n, &D, &P, &Cs, &half, &sep, &col, &count, &ms);
            want = 9;
        } else {
            r = std::fscanf(f, "%d %d %d %d %d %d %d %d %d %d %lf", &n, &D, &P, &Cs, &half, &sep, &col, &chunk,
                            &streams, &count, &ms);
            want = 11;
        }
        if (r == EOF) break;
        TuneEntry e;
        e.cfg.sep = sep;
        e.cfg.col = col;
        e.cfg.chunk = chunk;
        e.cfg.streams = streams;
        e.count = count;
        e.ms = ms;
        if (r != want || check_cfg(n, e.cfg) || D < 1 || D > kMaxDir || P < 1 || P > kMaxPhase || Cs < 1 ||
            Cs > kMaxComp || (half != 0 && half != 1)) {
            std::fclose(f);
            return fail(FSIM_ERR_ARG, std::string("autotune_load: malformed entry in ") + path);
        }
        got.push_back({TuneKey{n, D, P, Cs, half}, e});
    }
    std::fclose(f);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const auto &kv : got) g_tune[kv.first] = kv.second;
    return FSIM_OK;
}

int fsim_autotune(fsim_channel ch, int count, int reps, fsim_tune_result *res) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!ch || count < 1 || count > 4096 || reps < 1 || reps > 1000 || !res)
        return fail(FSIM_ERR_ARG, "autotune: bad arguments");
    const ReconConst &K = channel_host_const(ch);
    const int n = K.n;
    const size_t nn = (size_t)n * n;
    const size_t raw_e = (size_t)K.D * K.P * nn, a_e = workspace_a_elems(n, K.D, K.Cs), b_e = workspace_b_elems(n);
    hipStream_t s = stream();
    uint16_t *raw = nullptr;
    float2 *A = nullptr, *B = nullptr;
    float *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto release = [&]() {
        (void)hipStreamSynchronize(s);
        (void)hipFree(raw);
        (void)hipFree(A);
        (void)hipFree(B);
        (void)hipFree(out);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    if (hipMalloc(&raw, raw_e * count * 2) != hipSuccess || hipMalloc(&A, a_e * count * sizeof(float2)) != hipSuccess ||
        hipMalloc(&B, b_e * count * sizeof(float2)) != hipSuccess ||
        hipMalloc(&out, 4 * nn * count * sizeof(float)) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        (void)hipGetLastError();
        release();
        return fail(FSIM_ERR_OOM, "autotune: scratch allocation failed");
    }
    FSIM_HIP(hipMemsetAsync(raw, 0, raw_e * count * 2, s));
    JobList jl{};
    jl.jobs = nullptr;
    jl.base.c = channel_device_const(ch);
    jl.base.raw = raw;
    jl.base.A = A;
    jl.base.B = B;
    jl.base.out = out;
    jl.raw_stride = raw_e;
    jl.a_stride = a_e;
    jl.b_stride = b_e;
    jl.out_stride = 4 * nn;
    jl.raw_aligned = true;
    const bool half = K.half != 0;
    // candidates: sep in {1 where eligible, 0} x distinct column shapes
    std::vector<PipeCfg> cands;
    std::vector<std::pair<int, int>> shapes;
    for (int col = 0; col < kColCands; col++) {
        int e = 0, cw = 0;
        rc = col_shape(n, col, &e, &cw);
        if (rc) {
            release();
            return rc;
        }
        bool dup = false;
        for (auto &sh : shapes) dup |= (sh.first == e && sh.second == cw);
        if (dup) continue;
        shapes.push_back({e, cw});
        for (int sep = 1; sep >= 0; sep--) {
            if (sep == 1 && !sep_all_eligible(n, K.D, K.P)) continue;
            PipeCfg c;
            c.sep = sep;
            c.col = col;
            cands.push_back(c);
        }
    }
    // Warm-up: the default shape until the device has run for 30 ms (clocks
    // and caches settle), then rounds over all candidates, `reps` timed
    // launches each; a candidate's time is its best round.  The default is
    // kept unless a candidate beats it by more than 2%.
    auto time_cfg = [&](const PipeCfg &c, int nl, double &per) -> int {
        FSIM_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < nl; r++) {
            const int q = launch_pipeline_cfg(n, jl, count, K.D, K.P, K.Cs, half, c, s);
            if (q) return q;
        }
        FSIM_HIP(hipEventRecord(e1, s));
        FSIM_HIP(hipEventSynchronize(e1));
        float ms = 0;
        FSIM_HIP(hipEventElapsedTime(&ms, e0, e1));
        per = ms / nl;
        return FSIM_OK;
    };
    double warm = 0, per = 0;
    for (int i = 0; i < 200 && warm < 30.0 && !rc; i++) {
        rc = time_cfg(cands[0], 1, per);
        warm += per;
    }
    std::vector<double> t(cands.size(), 1e30);
    constexpr int kRounds = 3;
    for (int round = 0; round < kRounds && !rc; round++)
        for (size_t i = 0; i < cands.size() && !rc; i++) {
            rc = time_cfg(cands[i], reps, per);
            if (!rc && per < t[i]) t[i] = per;
        }
    const double dflt = t[0];   // the first candidate is the default choice
    double best = dflt;
    PipeCfg pick = cands[0];
    for (size_t i = 1; i < cands.size(); i++)
        if (t[i] < 0.98 * dflt && t[i] < best) {
            best = t[i];
            pick = cands[i];
        }
    const double whole = best;   // all stacks in one launch, best shape
    // Chunks: with the chosen shapes, `count` stacks as chunks of k stacks
    // round-robin on s streams (each stream with its own k-stack work
    // buffers) against one launch; k a power of two up to count / 2 whose
    // buffers take at most 1 GiB.  The size default is kept unless another
    // choice is more than 2% faster.
    std::vector<std::pair<int, int>> chunk_cands;   // (k, streams); (0, 1) = one launch
    chunk_cands.push_back({0, 1});
    const size_t stack_ws = (a_e + b_e) * sizeof(float2);
    for (int k = 1; 2 * k <= count && (size_t)k * stack_ws <= ((size_t)1 << 30); k *= 2)
        for (int st = 1; st <= kMaxChunkStreams; st++)
            if (st <= (count + k - 1) / k) chunk_cands.push_back({k, st});
    int dK = count, dS = 1;
    {
        PipeCfg d = pick;
        d.chunk = -1;
        d.streams = 2;
        chunk_plan(n, K.D, K.Cs, count, d, dK, dS);
    }
    std::vector<double> tc(chunk_cands.size(), 1e30);
    std::vector<float2 *> wa(kMaxChunkStreams, nullptr), wb(kMaxChunkStreams, nullptr);
    int kmax = 0;
    for (auto &cc : chunk_cands) kmax = std::max(kmax, cc.first);
    if (!rc && kmax > 0) {
        for (int i = 0; i < kMaxChunkStreams && !rc; i++)
            if (dev_alloc_n(&wa[i], a_e * kmax) || dev_alloc_n(&wb[i], b_e * kmax))
                rc = fail(FSIM_ERR_OOM, "autotune: chunk work buffer allocation failed");
        if (!rc) {
            std::lock_guard<std::mutex> lk(chunk_mutex());
            if (!chunk_streams_locked()) rc = fail(FSIM_ERR_HIP, "autotune: chunk streams could not be created");
            for (int round = 0; round < kRounds && !rc; round++)
                for (size_t i = 0; i < chunk_cands.size() && !rc; i++) {
                    const int k = chunk_cands[i].first, st = chunk_cands[i].second;
                    if (k == 0) {
                        tc[i] = whole;
                        continue;
                    }
                    if (hipEventRecord(e0, s) != hipSuccess) rc = fail(FSIM_ERR_HIP, "autotune: event record failed");
                    for (int r = 0; r < reps && !rc; r++)
                        rc = launch_chunked(n, jl, count, K.D, K.P, K.Cs, half, pick, k, st, s, wa.data(), wb.data());
                    float ms = 0;
                    if (!rc && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                        rc = fail(FSIM_ERR_HIP, "autotune: event timing failed");
                    if (!rc && ms / reps < tc[i]) tc[i] = ms / reps;
                }
        }
    }
    (void)hipStreamSynchronize(s);
    for (int i = 0; i < kMaxChunkStreams; i++) {
        dev_free(wa[i]);
        dev_free(wb[i]);
    }
    size_t di = 0;   // the size default among the candidates
    for (size_t i = 0; i < chunk_cands.size(); i++)
        if ((dK >= count && chunk_cands[i].first == 0) || (chunk_cands[i].first == dK && chunk_cands[i].second == dS))
            di = i;
    size_t bi = di;
    for (size_t i = 0; i < chunk_cands.size(); i++)
        if (tc[i] < 0.98 * tc[di] && tc[i] < tc[bi]) bi = i;
    pick.chunk = chunk_cands[bi].first;
    pick.streams = chunk_cands[bi].second;
    best = tc[bi];
    release();
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        TuneEntry &e = g_tune[TuneKey{n, K.D, K.P, K.Cs, half ? 1 : 0}];
        e.cfg = pick;
        e.ms = best;
        e.count = count;
    }
    res->sep = pick.sep;
    res->col = pick.col;
    col_shape(n, pick.col, &res->col_elems, &res->col_width);
    res->candidates = (int)(cands.size() + chunk_cands.size() - 1);
    res->ms_best = best;
    res->ms_default = dflt;
    res->chunk = pick.chunk;
    res->streams = pick.streams;
    res->ms_whole = whole;
    res->ms_chunk_default = tc[di];
    return FSIM_OK;
}

}  // extern "C"
