// Fused reconstruction: per-channel constants and job descriptors.
#pragma once
#include <functional>
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

#include "filters.h"

namespace fsim {

// Per-channel constants read by every kernel of the pipeline (device copy).
struct ReconConst {
    int n, D, P, B, C;
    int ns, otf_nb;
    double otf_cpm, cutoff, vec_cpm;
    const float2 *otf_tab;            // vals or valsAtt, otf_nb bands x ns
    const float2 *otf2d;              // otfToVector(band, 0, 0) on the n^2 grid, per OTF band
    const float *wa;                  // (2n)^2 of (1/Wiener denominator) * apodization, tile-major:
                                      // pixel (x, y) at ((x / cw) * 2n + y) * cw + x % cw, cw = gather_cw(n)
    const double *fade_lo, *fade_hi;  // fadeBorderCos factors, n each
    int fade_px;                      // border width; factors are 1.0 inside
    const float2 *ramp;               // [kMaxDir][kMaxComp][x|y][n] sub-pixel ramp factors
    float2 S[kMaxDir][kMaxComp][kMaxPhase];   // separation matrices (toFlt)
    int offx[kMaxDir][kMaxComp], offy[kMaxDir][kMaxComp];   // integer paste offsets
    // Stored components: Cs per direction (C, or B with the Hermitian pairing:
    // DC and the pos component of each band).  srow/sband/sramp: separation
    // row, OTF band and ramp of stored component j; src[c]: stored component
    // that contribution c reads; mir/mdx/mdy: contribution c is the mirrored,
    // conjugated pos spectrum Z[(delta - k) mod n] with delta = (mdx, mdy).
    int Cs;
    int srow[kMaxComp], sband[kMaxComp], sramp[kMaxComp], src[kMaxComp];
    int mir[kMaxDir][kMaxComp], mdx[kMaxDir][kMaxComp], mdy[kMaxDir][kMaxComp];
    double fx[kMaxDir][kMaxComp], fy[kMaxDir][kMaxComp];    // sub-pixel ramp shifts
    // half: the m x m result spectrum is Hermitian (real raw data, real OTF,
    // the Hermitian pairing), so the gather forms columns 0..n of it and the
    // last pass is a half-length complex-to-real row transform.
    int half;
    // otf_real: every OTF table value is real, so conj(OTF) multiplies as a
    // real scale (the same values as the complex product without its zero
    // terms)
    int otf_real;
    // columns x of the n x n grid with zlo <= x <= zhi lie beyond the OTF
    // cutoff along x (|xh| * vec_cpm > cutoff, so every OTF value of the
    // column is zero): the column pass writes zeros there (empty: zlo > zhi)
    int zlo, zhi;
};

// One stack: channel constants, raw input, work buffers and output.
struct Job {
    const ReconConst *c;
    const uint16_t *raw;  // D*P frames of n*n
    float2 *A;            // D*C components of n*n
    float2 *B;            // (2n)^2
    float *out;           // (2n)^2
};

// Either an explicit device array of jobs or a base job plus strides.
// raw_aligned: every job's raw frames start on an 8-byte boundary (the
// separation pass then reads 4 pixels per load).
struct JobList {
    const Job *jobs;
    Job base;
    size_t raw_stride, a_stride, b_stride, out_stride;
    bool raw_aligned;
    hipEvent_t gather_after;   // host side: the gather's stream waits for it (nullptr: no wait)
};

// Launch choices of the pipeline that the autotuner picks per size class:
// sep = 1: the separation with all units of a row in one wavefront (where the
// frames fit: n <= 512, D*P <= 9), 0: one unit per block; col: column-pass
// candidate (elements per thread, columns per block; col_shape reports the
// columns negated for the one-column-per-wavefront layout).
constexpr int kColCands = 8;
// chunk: stacks per launch when a call's stacks run as consecutive chunks
// (-1: the size's default, 0: all in one launch), streams: chunks run
// round-robin on this many library streams, each reusing its own work
// buffers (kept in the 256 MB Infinity Cache between the passes when a
// chunk's buffers are small enough, and one chunk's tail overlaps the next
// chunk's head).
constexpr int kMaxChunkStreams = 4;
struct PipeCfg {
    int sep = 1;
    int col = 0;
    int chunk = -1;
    int streams = 2;
};
// chunk size and stream count a call of `count` stacks uses: (count, 1) when
// it runs as one launch
void chunk_plan(int n, int D, int Cs, int count, const PipeCfg &cfg, int &K, int &S);
// The library streams of chunked launches and the lock a caller holds from
// looking up their work buffers until its chunks are enqueued (the fork /
// join events are shared, and a call that grows a stream's buffers waits for
// that stream first).  chunk_streams_locked: with the lock held, the
// kMaxChunkStreams streams (created on first use), nullptr if they could not
// be created.
std::mutex &chunk_mutex();
const hipStream_t *chunk_streams_locked();
// With chunk_mutex() held: runs `count` stacks (jl: base job plus strides,
// its A/B unused; or a device job list whose entries carry their chunk's
// work buffers) as chunks of K on the first S chunk streams, forked from
// and joined into s; ws_a / ws_b: the work buffers (K stacks each) of stream
// i.  gather_input: work enqueued on s after the fork (the Wiener tables the
// gathers read), which the chunks' first passes overlap: each chunk's gather
// waits for it.
int launch_chunked(int n, const JobList &jl, int count, int D, int P, int Cs, bool half, const PipeCfg &cfg, int K,
                   int S, hipStream_t s, float2 *const *ws_a, float2 *const *ws_b,
                   const std::function<int(hipStream_t)> *gather_input = nullptr);
bool pipe_cfg_lookup(int n, int D, int P, int Cs, bool half, PipeCfg &cfg);
int col_shape(int n, int col, int *e, int *cw);
bool sep_all_eligible(int n, int D, int P);
int launch_pipeline_cfg(int n, const JobList &jl, int njobs, int D, int P, int Cs, bool half, const PipeCfg &cfg,
                        hipStream_t s);

bool recon_size_ok(int n);
int gather_cw(int n);   // column tile width of the gather at raw side n
int launch_pipeline(int n, const JobList &jl, int njobs, int D, int P, int Cs, bool half, hipStream_t s);
// Persistent work-list execution of a batch's pipeline.  A launch group is
// a set of sizes whose stage kernels share a block shape and about the same
// registers and LDS: group 0 holds n = 256 and 512, group 1 n = 1024 and
// 2048 (other sizes run their own launches).  Per group and stage one device
// list of work items (class, bx, by, unused) holds the blocks of every
// class's own launch of that stage, and one persistent kernel runs the list.
struct PClass {
    JobList jl;                 // the class's jobs (device job array)
    int n, njobs, D, C, half;   // C: stored components per direction
    int sep_all;                // separation with all units of a row in one wavefront
    const float2 *twn, *twm;    // twiddle tables of n and 2n points
};
// stages: separation, column pass, row shift, gather, last row pass
constexpr int kPStages = 5;
constexpr int kPGroups = 2;
int persistent_group(int n);    // -1: n runs its own launches
// The item list of a stage over classes cls[0, ncls) of one group (host
// copies; item class indices refer to this array): four segments, the
// group's larger size (variants 0, 1) then the smaller, each class's run
// padded with (-1, 0, 0, 0) to a multiple of 8; ends: the segment ends.
void persistent_list(const PClass *cls, int ncls, int stage, std::vector<int4> &items, int ends[4]);
// One persistent launch of a group's stage over a list (device copies);
// cur: the list's 8 cursors (one per XCD, kPCursorStride ints apart), zero
// before the launch (launch_persistent_reset).
int launch_persistent(int group, int stage, const PClass *cls_d, const int4 *items_d, const int ends[4], int *cur,
                      hipStream_t s);
int launch_persistent_reset(int *cur, int count, hipStream_t s);
// ints between two list cursors (256 bytes), and per list
constexpr int kPCursorStride = 64;
constexpr int kPListCursors = 8 * kPCursorStride;
size_t workspace_a_elems(int n, int D, int C);
size_t workspace_b_elems(int n);

}  // namespace fsim

struct fsim_channel_s;
namespace fsim {
const ReconConst *channel_device_const(const fsim_channel_s *ch);
const ReconConst &channel_host_const(const fsim_channel_s *ch);

// Device state of reconstructions with their own SIM parameters, grown on
// demand: per parameter set constants, (1/Wiener) * apodization table, ramp
// tables; per stack jobs.
struct MultiWork {
    ReconConst *dK = nullptr;
    float *wa = nullptr;
    float2 *ramp = nullptr;
    double *f = nullptr;
    ShiftSet *sh = nullptr;
    Job *jobs = nullptr;
    int cap_sets = 0, cap_jobs = 0;
    hipEvent_t uploaded = nullptr;   // recorded after the uploads from the host arrays below
    std::vector<ReconConst> hK;
    std::vector<double> hf;
    std::vector<ShiftSet> hsh;
    std::vector<Job> hjobs;
};
void multiwork_release(MultiWork &w);
size_t multiwork_bytes(int n, int count, int nsets);
// Chunked execution of reconstruct_params_with: the stacks run as chunks of
// k on ns library streams (chunk_streams_locked, the chunk mutex held),
// chunk c on stream c % ns with that stream's work buffers a, b of k stacks.
struct ParamsChunks {
    int k = 0, ns = 0;
    float2 *a[kMaxChunkStreams] = {}, *b[kMaxChunkStreams] = {};
    PipeCfg cfg;
};
// pc: nullptr for one launch with the work buffers A and B of count stacks.
int reconstruct_params_with(const fsim_channel_s *ch, MultiWork &w, float2 *A, float2 *B, const uint16_t *raw, float *out,
                            int count, const fsim_sim_params *sp, int sp_step, hipStream_t s,
                            const ParamsChunks *pc = nullptr);
}  // namespace fsim
