// Fused fast-shift SIM reconstruction.
//
// Dataflow of one stack (livemode/ReconstructionRunner.java:456-542 and
// tests/TestAccelSpeed.java:237-252, 563-703), regrouped into five HBM passes
// over the raw side n and the doubled side m = 2n:
//
//  K1 k_sep_rows    raw u16 -> fade -> band separation (linear, so it moves in
//                   front of the FFT) -> row FFT            writes A[d][c]
//  K2 k_col_shift   shifted components only: column FFT -> * conj(OTF) ->
//                   column IFFT                               in place on A
//  K3 k_row_shift   shifted components only: row IFFT -> fast-shift phase ramp
//                   (SimUtils.java:74-81) -> row FFT          in place on A
//  K5 k_gather      per column tile of the m x m result: for every direction
//                   and component, column FFT of the pasted column (the DC
//                   component also * conj(OTF)), pasteFreq as a gather into a
//                   register accumulator (SimUtils.java:83-86), then
//                   * Wiener * apodization and the column IFFT  writes B
//  K6 k_row_real    row IFFT of B, real part                  writes out
//
// HBM bytes per stack: 2*D*P*n^2 (raw) + 16*n^2 per complex pass per
// component + 32*m^2 for B + 4*m^2 for the output -- the algorithmic count
// of the fast-shift dataflow.
#include "fsim_internal.h"
#include "fft_core.h"
#include "fft_launch.h"
#include "filters.h"
#include "recon.h"
#include "simmath.h"

#include <cmath>
#include <mutex>
#include <set>
#include <utility>
#include <vector>

#ifndef FSIM_RAMP_TABLE
#define FSIM_RAMP_TABLE 1
#endif
#ifndef FSIM_K1_ALLC
#define FSIM_K1_ALLC 0
#endif
#ifndef FSIM_GATHER_CW
#define FSIM_GATHER_CW 8
#endif
#ifndef FSIM_COL_CW
#define FSIM_COL_CW 16
#endif

namespace fsim {

__device__ __forceinline__ Job get_job(const JobList &jl, int j) {
    if (jl.jobs) return jl.jobs[j];
    Job J = jl.base;
    J.raw += (size_t)j * jl.raw_stride;
    J.A += (size_t)j * jl.a_stride;
    J.B += (size_t)j * jl.b_stride;
    J.out += (size_t)j * jl.out_stride;
    return J;
}

// Fade factor of pixel (x, y): fadeBorderCos on a vector of ones
// (SimUtils.java:162-185, ReconstructionRunner.java:781-783).
__device__ __forceinline__ float fade_at(const double *__restrict__ lo, const double *__restrict__ hi, int x,
                                         int y) {
    float f = 1.0f;
    f = (float)((double)f * lo[y]);
    f = (float)((double)f * hi[y]);
    f = (float)((double)f * lo[x]);
    f = (float)((double)f * hi[x]);
    return f;
}

// ---------------------------------------------------------------------------
// K1: band separation + row FFT.  One wavefront (T = n/E threads) per row of
// one direction: the P raw rows are read once, faded, and every component
// c = sum_p S[d][c][p] * raw_p is transformed and stored in turn.
// ---------------------------------------------------------------------------
#if !FSIM_K1_ALLC
template <int N>
__global__ void __launch_bounds__(256) k_sep_rows1(JobList jl, int ngroups, const float2 *__restrict__ tw) {
    constexpr int E = FftE<N>::value, T = N / E;
    constexpr int RW = T >= 256 ? 1 : 256 / T;
    constexpr int LP = FftLds<N, E>::len;
    __shared__ float2 lds[RW * LP];
    const Job J = get_job(jl, blockIdx.y);
    const ReconConst &K = *J.c;
    const int C = K.Cs, P = K.P;
    const int L = blockIdx.x;
    const int g = (L / 8 / C) * 8 + (L % 8);
    const int c = (L / 8) % C;           // stored component
    const int crow = K.srow[c];          // its separation-matrix row
    if (g >= ngroups) return;
    constexpr int rows_per_dir = N / RW;
    const int d = g / rows_per_dir, rb = g % rows_per_dir;
    const int t = threadIdx.x % T, rr = threadIdx.x / T;
    const int y = rb * RW + rr;
    const size_t nn = (size_t)N * N;
    const uint16_t *raw = J.raw + (size_t)d * P * nn + (size_t)y * N;
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; m++) v[m] = make_float2(0.f, 0.f);
    float fade[E];
#pragma unroll
    for (int m = 0; m < E; m++) fade[m] = fade_at(K.fade_lo, K.fade_hi, t + T * m, y);
    for (int p = 0; p < P; p++) {
        const float2 s = K.S[d][crow][p];
        const uint16_t *rp = raw + (size_t)p * nn;
#pragma unroll
        for (int m = 0; m < E; m++) {
            const float px = (float)rp[t + T * m] * fade[m];
            v[m].x += s.x * px;
            v[m].y += s.y * px;
        }
    }
    fft_regs<N, E, false, RowWave<N>::value>(v, lds + rr * LP, 1, t, tw);
    float2 *out = J.A + ((size_t)(d * C + c) * N + y) * N;
#pragma unroll
    for (int m = 0; m < E; m++) out[t + T * m] = v[m];
}
#endif

template <int N>
__global__ void __launch_bounds__(256) k_sep_rows(JobList jl, const float2 *__restrict__ tw) {
    constexpr int E = FftE<N>::value, T = N / E;
    constexpr int RW = T >= 256 ? 1 : 256 / T;
    constexpr int LP = FftLds<N, E>::len;
    constexpr int rows_per_dir = N / RW;
    __shared__ float2 lds[RW * LP];
    const Job J = get_job(jl, blockIdx.y);
    const ReconConst &K = *J.c;
    const int C = K.Cs, P = K.P;
    const int d = blockIdx.x / rows_per_dir, rb = blockIdx.x % rows_per_dir;
    const int t = threadIdx.x % T, rr = threadIdx.x / T;
    const int y = rb * RW + rr;
    const size_t nn = (size_t)N * N;
    const uint16_t *raw = J.raw + (size_t)d * P * nn + (size_t)y * N;
    float px[kMaxPhase][E];
#pragma unroll
    for (int m = 0; m < E; m++) {
        const float f = fade_at(K.fade_lo, K.fade_hi, t + T * m, y);
#pragma unroll
        for (int p = 0; p < kMaxPhase; p++)
            if (p < P) px[p][m] = (float)raw[(size_t)p * nn + t + T * m] * f;
    }
    for (int c = 0; c < C; c++) {
        float2 v[E];
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = make_float2(0.f, 0.f);
#pragma unroll
        for (int p = 0; p < kMaxPhase; p++) {
            if (p < P) {
                const float2 sp = K.S[d][K.srow[c]][p];
#pragma unroll
                for (int m = 0; m < E; m++) {
                    v[m].x += sp.x * px[p][m];
                    v[m].y += sp.y * px[p][m];
                }
            }
        }
        fft_regs<N, E, false, RowWave<N>::value>(v, lds + rr * LP, 1, t, tw);
        if constexpr (!RowWave<N>::value) __syncthreads();
        float2 *out = J.A + ((size_t)(d * C + c) * N + y) * N;
#pragma unroll
        for (int m = 0; m < E; m++) out[t + T * m] = v[m];
    }
}

// ---------------------------------------------------------------------------
// K2: shifted components: column FFT, * conj(OTF_band), column IFFT (1/n).
// ---------------------------------------------------------------------------
template <int N>
struct ColShiftTile {
    static constexpr int e = ColE<N>::value, t = N / e;
    static constexpr int cw0 = ColTile<N, t, e>::value;
    static constexpr int cw = cw0 < FSIM_COL_CW ? cw0 : FSIM_COL_CW;
    static constexpr int value = t * cw;
};

template <int N>
__global__ void __launch_bounds__(ColShiftTile<N>::value) k_col_shift(JobList jl, const float2 *__restrict__ tw) {
    constexpr int E = ColE<N>::value, T = N / E;
    constexpr int CW = ColShiftTile<N>::cw;
    constexpr int tiles = N / CW;
    __shared__ float2 lds[FftLds<N, E>::len * CW];
    const Job J = get_job(jl, blockIdx.y);
    const ReconConst &K = *J.c;
    const int C = K.Cs, nsh = C - 1;
    const int tile = blockIdx.x % tiles, dc = blockIdx.x / tiles;
    const int d = dc / nsh, c = 1 + dc % nsh;   // stored shifted component
    const int cl = threadIdx.x % CW, t = threadIdx.x / CW;
    const int x = tile * CW + cl;
    float2 *col = J.A + (size_t)(d * C + c) * N * N + x;
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; m++) v[m] = col[(size_t)(t + T * m) * N];
    fft_regs<N, E, false>(v, lds + cl, CW, t, tw);
    const int band = K.sband[c];
    const float2 *otf = K.otf2d + (size_t)(K.otf_nb > 1 ? band : 0) * N * N + x;
#pragma unroll
    for (int m = 0; m < E; m++) v[m] = times_conj_java(v[m], otf[(size_t)(t + T * m) * N]);
    fft_regs<N, E, true>(v, lds + cl, CW, t, tw);
    const float sc = 1.0f / N;
#pragma unroll
    for (int m = 0; m < E; m++) col[(size_t)(t + T * m) * N] = cscale(v[m], sc);
}

// ---------------------------------------------------------------------------
// K3: shifted components: row IFFT (1/n), sub-pixel ramp, row FFT.
// ---------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(256) k_row_shift(JobList jl, const float2 *__restrict__ tw) {
    constexpr int E = FftE<N>::value, T = N / E;
    constexpr int RW = T >= 256 ? 1 : 256 / T;
    constexpr int rblocks = N / RW;
    constexpr int LP = FftLds<N, E>::len;
    __shared__ float2 lds[RW * LP];
    const Job J = get_job(jl, blockIdx.y);
    const ReconConst &K = *J.c;
    const int C = K.Cs, nsh = C - 1;
    const int rb = blockIdx.x % rblocks, dc = blockIdx.x / rblocks;
    const int d = dc / nsh, c = 1 + dc % nsh;   // stored shifted component
    const int t = threadIdx.x % T, rr = threadIdx.x / T;
    const int y = rb * RW + rr;
    float2 *row = J.A + ((size_t)(d * C + c) * N + y) * N;
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; m++) v[m] = row[t + T * m];
    fft_regs<N, E, true, RowWave<N>::value>(v, lds + rr * LP, 1, t, tw);
    // sub-pixel ramp exp(2 pi i (fx x + fy y) / n) as a product of the
    // per-channel row and column tables (SimUtils.java:74-81 shift phases)
    const float sc = 1.0f / N;
#if FSIM_RAMP_TABLE
    const float2 *rx = K.ramp + ((size_t)(d * kMaxComp + K.sramp[c]) * 2) * N;
    const float2 ry = rx[N + y];
#pragma unroll
    for (int m = 0; m < E; m++) {
        const float2 r = cmul(rx[t + T * m], ry);
        v[m] = rot_java(cscale(v[m], sc), r.x, r.y);
    }
#else
    const double fx = K.fx[d][K.sramp[c]], fy = K.fy[d][K.sramp[c]];
#pragma unroll
    for (int m = 0; m < E; m++) {
        float co, si;
        shift_cos_sin(fx, fy, t + T * m, y, N, co, si);
        v[m] = rot_java(cscale(v[m], sc), co, si);
    }
#endif
    fft_regs<N, E, false, RowWave<N>::value>(v, lds + rr * LP, 1, t, tw);
#pragma unroll
    for (int m = 0; m < E; m++) row[t + T * m] = v[m];
}

// ---------------------------------------------------------------------------
// K5: gather into the doubled grid, filter, column IFFT.
// Elements per thread for the n-point (en) and 2n-point (em) column
// transforms and the tile width: up to n = 512 a 1024-thread block with 16
// elements per thread fits 128 VGPRs; from n = 1024 on, 32 elements per thread
// in 512-thread blocks.
// ---------------------------------------------------------------------------
template <int N>
struct GatherE {
    static constexpr int en = N >= 1024 ? 32 : FftE<N>::value;
    static constexpr int em = N >= 1024 ? 32 : FftE<2 * N>::value;
    static constexpr int tn = N / en, tm = 2 * N / em;
    static constexpr int tk = tn > tm ? tn : tm;
    static constexpr int cw0 = ColTile<2 * N, tk, (en > em ? en : em)>::value;
    static constexpr int cw = cw0 < FSIM_GATHER_CW ? cw0 : FSIM_GATHER_CW;
};
template <int N>
__global__ void __launch_bounds__(GatherE<N>::tk * GatherE<N>::cw) k_gather(JobList jl, const float2 *__restrict__ twn,
                                                 const float2 *__restrict__ twm) {
    constexpr int M = 2 * N;
    constexpr int EN = GatherE<N>::en, TN = N / EN;
    constexpr int EM = GatherE<N>::em, TM = M / EM;
    constexpr int CW = GatherE<N>::cw;
    constexpr int LPN = FftLds<N, EN>::len, LPM = FftLds<M, EM>::len;
    constexpr int LPK = LPM > LPN ? LPM : LPN;
    __shared__ float2 lds[(LPK > N ? LPK : N) * CW];
    const Job J = get_job(jl, blockIdx.y);
    const ReconConst &K = *J.c;
    const int D = K.D, C = K.C, Cs = K.Cs;
    const int cl = threadIdx.x % CW, t = threadIdx.x / CW;
    // blocks are dealt to the 8 XCDs round robin; put adjacent tile pairs on
    // the same XCD (bx and bx + 8 run back to back there) so that source rows
    // shared by neighbouring tiles (paste offsets are not tile aligned) are
    // fetched once into its L2, while the pairs stay spread over all XCDs
    // (tiles differ in how many components land in them)
    constexpr int tiles = M / CW;
    const int bx = blockIdx.x;
    const int tile = tiles % 16 ? bx : (bx / 16) * 16 + (bx % 8) * 2 + (bx / 8) % 2;
    const int xo = tile * CW + cl;
    float2 acc[EM];
#pragma unroll
    for (int m = 0; m < EM; m++) acc[m] = make_float2(0.f, 0.f);
    const size_t nn = (size_t)N * N;
    for (int d = 0; d < D; d++) {
        for (int c = 0; c < C; c++) {
            // contribution c of direction d, in the reference's summation order
            const int mir = K.mir[d][c];
            int xs = paste_src(xo, M, N, K.offx[d][c]);
            if (mir && xs >= 0) xs = (K.mdx[d][c] - xs + N) & (N - 1);
            const bool act = (t < TN) && (xs >= 0);
            float2 v[EN];
            if (act) {
                const float2 *col = J.A + (size_t)(d * Cs + K.src[c]) * nn + xs;
#pragma unroll
                for (int m = 0; m < EN; m++) v[m] = col[(size_t)(t + TN * m) * N];
            }
            fft_regs<N, EN, false>(v, lds + cl, CW, t, twn, act);
            if (act) {
                if (c == 0) {
                    const float2 *otf = K.otf2d + xs;
#pragma unroll
                    for (int m = 0; m < EN; m++) v[m] = times_conj_java(v[m], otf[(size_t)(t + TN * m) * N]);
                }
#pragma unroll
                for (int m = 0; m < EN; m++) lds[(t + TN * m) * CW + cl] = v[m];
            }
            __syncthreads();
            if (t < TM && xs >= 0) {
                const int oy = K.offy[d][c];
                const int my = K.mdy[d][c];
#pragma unroll
                for (int m = 0; m < EM; m++) {
                    int ys = paste_src(t + TM * m, M, N, oy);
                    if (ys >= 0) {
                        if (mir) ys = (my - ys + N) & (N - 1);
                        const float2 s = lds[ys * CW + cl];
                        acc[m].x += s.x;
                        acc[m].y += mir ? -s.y : s.y;
                    }
                }
            }
            __syncthreads();
        }
    }
    // fullResult.times(wienerDenom); fullResult.times(apo)
    if (t < TM) {
#pragma unroll
        for (int m = 0; m < EM; m++) {
            const float2 wa = K.wa[(size_t)(t + TM * m) * M + xo];
            acc[m] = make_float2(acc[m].x * wa.x, acc[m].y * wa.x);
            acc[m] = make_float2(acc[m].x * wa.y, acc[m].y * wa.y);
        }
    }
    fft_regs<M, EM, true>(acc, lds + cl, CW, t, twm, t < TM);
    if (t < TM) {
        const float sc = 1.0f / M;
#pragma unroll
        for (int m = 0; m < EM; m++) J.B[(size_t)(t + TM * m) * M + xo] = cscale(acc[m], sc);
    }
}

// ---------------------------------------------------------------------------
// K6: row IFFT of the doubled grid, real part out.
// ---------------------------------------------------------------------------
template <int M>
__global__ void __launch_bounds__(256) k_row_real(JobList jl, const float2 *__restrict__ tw) {
    constexpr int E = FftE<M>::value, T = M / E;
    constexpr int RW = T >= 256 ? 1 : 256 / T;
    constexpr int LP = FftLds<M, E>::len;
    __shared__ float2 lds[RW * LP];
    const Job J = get_job(jl, blockIdx.y);
    const int t = threadIdx.x % T, rr = threadIdx.x / T;
    const int y = blockIdx.x * RW + rr;
    const float2 *row = J.B + (size_t)y * M;
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; m++) v[m] = row[t + T * m];
    fft_regs<M, E, true, RowWave<M>::value>(v, lds + rr * LP, 1, t, tw);
    const float sc = 1.0f / M;
    float *o = J.out + (size_t)y * M;
#pragma unroll
    for (int m = 0; m < E; m++) o[t + T * m] = v[m].x * sc;
}

// pack Wiener reciprocal and apodization into one (w, a) table
__global__ void k_pack_wa(float2 *__restrict__ wa, const float *__restrict__ w, const float *__restrict__ a,
                          size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        wa[i] = make_float2(w[i], a[i]);
}

// Optional stage timing: HIP events recorded around each pipeline kernel on
// the launch stream, summed per stage when read back.
struct StageTimer {
    std::mutex mu;
    bool on = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
};
static StageTimer g_timer;

static void stage_mark(int stage, hipStream_t s, hipEvent_t &start) {
    if (!g_timer.on) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, s);
    if (stage >= 0) {
        std::lock_guard<std::mutex> lk(g_timer.mu);
        g_timer.ev.push_back({stage, {start, e}});
    }
    start = e;
}

template <int N>
static int launch_pipeline_t(const JobList &jl, int njobs, int D, int C, hipStream_t s) {
    // C: stored components per direction
    constexpr int M = 2 * N;
    const float2 *twn = twiddles(N), *twm = twiddles(M);
    if (!twn || !twm) return fail(FSIM_ERR_OOM, "twiddle table allocation failed");
    hipEvent_t ev = nullptr;
    stage_mark(-1, s, ev);
    {
        constexpr int E = FftE<N>::value, T = N / E, RW = T >= 256 ? 1 : 256 / T;
#if FSIM_K1_ALLC
        hipLaunchKernelGGL(k_sep_rows<N>, dim3(D * (N / RW), njobs), dim3(T * RW), 0, s, jl, twn);
#else
        const int ngroups = D * (N / RW);
        const int gx = ((ngroups + 7) / 8) * 8 * C;
        hipLaunchKernelGGL(k_sep_rows1<N>, dim3(gx, njobs), dim3(T * RW), 0, s, jl, ngroups, twn);
#endif
        stage_mark(0, s, ev);
    }
    if (C > 1) {
        {
            constexpr int E = ColE<N>::value, T = N / E, CW = ColShiftTile<N>::cw;
            hipLaunchKernelGGL(k_col_shift<N>, dim3(D * (C - 1) * (N / CW), njobs), dim3(T * CW), 0, s, jl, twn);
            stage_mark(1, s, ev);
        }
        constexpr int E = FftE<N>::value, T = N / E;
        constexpr int RW = T >= 256 ? 1 : 256 / T;
        hipLaunchKernelGGL(k_row_shift<N>, dim3(D * (C - 1) * (N / RW), njobs), dim3(T * RW), 0, s, jl, twn);
        stage_mark(2, s, ev);
    }
    {
        constexpr int TK = GatherE<N>::tk, CW = GatherE<N>::cw;
        hipLaunchKernelGGL(k_gather<N>, dim3(M / CW, njobs), dim3(TK * CW), 0, s, jl, twn, twm);
        stage_mark(3, s, ev);
    }
    {
        constexpr int E = FftE<M>::value, T = M / E, RW = T >= 256 ? 1 : 256 / T;
        hipLaunchKernelGGL(k_row_real<M>, dim3(M / RW, njobs), dim3(T * RW), 0, s, jl, twm);
        stage_mark(4, s, ev);
    }
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

bool recon_size_ok(int n) { return n == 64 || n == 128 || n == 256 || n == 512 || n == 1024 || n == 2048; }

int launch_pipeline(int n, const JobList &jl, int njobs, int D, int C, hipStream_t s) {
    // C: stored components per direction (ReconConst::Cs)
    switch (n) {
        case 64: return launch_pipeline_t<64>(jl, njobs, D, C, s);
        case 128: return launch_pipeline_t<128>(jl, njobs, D, C, s);
        case 256: return launch_pipeline_t<256>(jl, njobs, D, C, s);
        case 512: return launch_pipeline_t<512>(jl, njobs, D, C, s);
        case 1024: return launch_pipeline_t<1024>(jl, njobs, D, C, s);
        case 2048: return launch_pipeline_t<2048>(jl, njobs, D, C, s);
        default: return fail(FSIM_ERR_ARG, "reconstruct: n must be a power of two in [64, 2048]");
    }
}

size_t workspace_a_elems(int n, int D, int C) { return (size_t)D * C * n * n; }
size_t workspace_b_elems(int n) { return (size_t)4 * n * n; }

// ---------------------------------------------------------------------------
// channel set-up: the FilterUpdateThread work (ReconstructionRunner.java:
// 773-828) plus the per-direction constants of the reconstruction loop.
// ---------------------------------------------------------------------------
int build_channel_const(const fsim_recon_params &p, const Otf &o, bool sym, ReconConst &K) {
    const int n = p.n, D = p.ndir, P = p.nphase, B = p.nband, C = 2 * B - 1;
    if (!recon_size_ok(n)) return fail(FSIM_ERR_ARG, "channel: n must be a power of two in [64, 2048]");
    if (D < 1 || D > kMaxDir || B < 2 || B > kMaxBand || P < C || P > kMaxPhase)
        return fail(FSIM_ERR_ARG, "channel: unsupported directions/phases/bands (phases >= 2*bands-1)");
    if (p.vec_cpm <= 0) return fail(FSIM_ERR_ARG, "channel: vec_cpm must be positive");
    if (o.nband != 1 && o.nband < B) return fail(FSIM_ERR_ARG, "channel: OTF has fewer bands than requested");
    if (p.fade_px < 0 || 2 * p.fade_px > n) return fail(FSIM_ERR_ARG, "channel: fade border too wide");
    K = ReconConst{};
    K.n = n;
    K.D = D;
    K.P = P;
    K.B = B;
    K.C = C;
    K.ns = o.ns;
    K.otf_nb = o.nband;
    K.otf_cpm = o.cpm;
    K.cutoff = o.cutoff;
    K.vec_cpm = p.vec_cpm;
    K.otf_tab = p.use_att ? o.vals_att : o.vals;
    K.Cs = sym ? B : C;
    for (int j = 0; j < kMaxComp; j++) K.srow[j] = K.sband[j] = K.sramp[j] = K.src[j] = 0;
    for (int j = 0; j < K.Cs; j++) {
        K.srow[j] = sym ? 2 * j : j;          // DC, then pos of band j
        K.sband[j] = sym ? j : (j + 1) / 2;
        K.sramp[j] = K.srow[j];
    }
    for (int c = 0; c < C; c++) K.src[c] = sym ? (c + 1) / 2 : c;
    for (int d = 0; d < D; d++) {
        // SimParam.Dir.getModulations clamp (SimParam.java:598-607)
        double mods[kMaxBand];
        for (int b = 0; b < B; b++) {
            double m = p.modul[d][b];
            if (m < 0.3) m = 0.3;
            if (m > 1.1) m = 1.1;
            mods[b] = m;
        }
        double S[2 * kMaxComp * kMaxPhase];
        int rc = separation_matrix_host(p.phases[d], P, B, mods, S);
        if (rc) return rc;
        for (int c = 0; c < C; c++)
            for (int ph = 0; ph < P; ph++)
                K.S[d][c][ph] = make_float2((float)S[2 * (c * P + ph)], (float)S[2 * (c * P + ph) + 1]);
        K.offx[d][0] = K.offy[d][0] = 0;
        K.fx[d][0] = K.fy[d][0] = 0;
        for (int b = 1; b < B; b++) {
            for (int s = 0; s < 2; s++) {
                const int comp = s == 0 ? 2 * b : 2 * b - 1;  // pos, neg (SimAlgorithm.java:385-391)
                const double kx = (s == 0 ? 1 : -1) * p.px[d] * b;
                const double ky = (s == 0 ? 1 : -1) * p.py[d] * b;
                // SimUtils.pasteAndFourierShift fast mode (SimUtils.java:74-86)
                K.fx[d][comp] = kx - std::floor(kx);
                K.fy[d][comp] = -ky - std::floor(-ky);
                K.offx[d][comp] = (int)std::floor(kx);
                K.offy[d][comp] = (int)std::floor(-ky);
                K.mir[d][comp] = (sym && s == 1) ? 1 : 0;
                if (K.offx[d][comp] < -2 * n || K.offy[d][comp] < -2 * n)
                    return fail(FSIM_ERR_ARG, "channel: shift vector out of range");
            }
            // neg = conj(pos mirrored): delta = 1 per axis where the pos
            // fractional shift is non-zero (the neg one is then 1 - f)
            K.mdx[d][2 * b - 1] = K.fx[d][2 * b] != 0.0 ? 1 : 0;
            K.mdy[d][2 * b - 1] = K.fy[d][2 * b] != 0.0 ? 1 : 0;
        }
    }
    return FSIM_OK;
}

}  // namespace fsim

using namespace fsim;

struct fsim_channel_s {
    fsim_recon_params p;
    ReconConst K;
    ReconConst *dK = nullptr;
    float2 *wa = nullptr;
    float2 *otf2d = nullptr;
    float2 *ramp = nullptr;
    double *fade = nullptr;
    float2 *A = nullptr, *B = nullptr;
    int cap = 0;
};

const ReconConst *fsim::channel_device_const(const fsim_channel_s *ch) { return ch->dK; }
const ReconConst &fsim::channel_host_const(const fsim_channel_s *ch) { return ch->K; }

static void channel_free(fsim_channel_s *ch) {
    (void)hipFree(ch->dK);
    (void)hipFree(ch->wa);
    (void)hipFree(ch->otf2d);
    (void)hipFree(ch->ramp);
    (void)hipFree(ch->fade);
    (void)hipFree(ch->A);
    (void)hipFree(ch->B);
}

extern "C" {

int fsim_channel_create(const fsim_recon_params *p, fsim_otf oh, fsim_channel *out) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!p || !oh || !out) return fail(FSIM_ERR_ARG, "channel_create: null argument");
    const Otf &o = oh->o;
    fsim_channel_s *ch = new fsim_channel_s();
    ch->p = *p;
    // Hermitian pairing needs a real OTF (then conj(OTF) is symmetric in k)
    bool sym = p->mode == 0;
    if (sym) {
        std::vector<float2> hv((size_t)o.nband * o.ns);
        FSIM_HIP(hipMemcpy(hv.data(), p->use_att ? o.vals_att : o.vals, hv.size() * sizeof(float2),
                           hipMemcpyDeviceToHost));
        for (const float2 &z : hv)
            if (z.y != 0.f) sym = false;
    }
    rc = build_channel_const(*p, o, sym, ch->K);
    if (rc) {
        delete ch;
        return rc;
    }
    const int n = p->n, M = 2 * n;
    const size_t mm = (size_t)M * M;
    float *wtmp = nullptr, *atmp = nullptr;
    if (hipMalloc(&ch->dK, sizeof(ReconConst)) != hipSuccess || hipMalloc(&ch->wa, mm * sizeof(float2)) != hipSuccess ||
        hipMalloc(&ch->otf2d, (size_t)o.nband * n * n * sizeof(float2)) != hipSuccess ||
        hipMalloc(&ch->fade, 2 * n * sizeof(double)) != hipSuccess ||
        hipMalloc(&ch->ramp, (size_t)kMaxDir * kMaxComp * 2 * n * sizeof(float2)) != hipSuccess ||
        hipMalloc(&wtmp, mm * sizeof(float)) != hipSuccess || hipMalloc(&atmp, mm * sizeof(float)) != hipSuccess) {
        (void)hipFree(wtmp);
        (void)hipFree(atmp);
        channel_free(ch);
        delete ch;
        return fail(FSIM_ERR_OOM, "channel_create: device allocation failed");
    }
    std::vector<double> f(2 * n);
    fade_factors_host(n, p->fade_px, f.data(), f.data() + n);
    hipStream_t s = stream();
    FSIM_HIP(hipMemcpyAsync(ch->fade, f.data(), 2 * n * sizeof(double), hipMemcpyHostToDevice, s));
    ch->K.fade_lo = ch->fade;
    ch->K.fade_hi = ch->fade + n;
    ch->K.wa = ch->wa;
    {
        // per direction/component ramp tables exp(2 pi i f t / n), t < n, for
        // the x and y fractional shifts, from double precision
        std::vector<float2> r((size_t)kMaxDir * kMaxComp * 2 * n, make_float2(1.f, 0.f));
        for (int d = 0; d < p->ndir; d++)
            for (int c = 0; c < ch->K.C; c++)
                for (int a = 0; a < 2; a++) {
                    const double f = a == 0 ? ch->K.fx[d][c] : ch->K.fy[d][c];
                    float2 *row = r.data() + ((size_t)(d * kMaxComp + c) * 2 + a) * n;
                    for (int t = 0; t < n; t++) {
                        const double ph = 2 * M_PI * f * t / n;
                        row[t] = make_float2((float)std::cos(ph), (float)std::sin(ph));
                    }
                }
        FSIM_HIP(hipMemcpyAsync(ch->ramp, r.data(), r.size() * sizeof(float2), hipMemcpyHostToDevice, s));
        FSIM_HIP(hipStreamSynchronize(s));
        ch->K.ramp = ch->ramp;
    }
    {
        // otfToVector(band, 0, 0, write) per OTF band on the n^2 grid
        const int nb = o.nband;
        const size_t nn = (size_t)n * n;
        const float2 *tab = p->use_att ? o.vals_att : o.vals;
        for (int b = 0; b < nb; b++) {
            rc = otf_table_device(ch->otf2d + b * nn, n, n, tab + (size_t)b * o.ns, o.ns, o.cpm, o.cutoff,
                                  p->vec_cpm, s);
            if (rc) return rc;
        }
        ch->K.otf2d = ch->otf2d;
    }
    ShiftSet sh{};
    sh.ndir = p->ndir;
    sh.nband = p->nband;
    for (int d = 0; d < p->ndir; d++) {
        sh.px[d] = p->px[d];
        sh.py[d] = p->py[d];
    }
    rc = wiener_device(wtmp, M, M, o, p->vec_cpm, sh, p->use_att, p->wiener, s);
    if (rc) return rc;
    rc = write_apo_device(atmp, M, M, p->vec_cpm, o.cutoff, p->apo_bend, p->apo_cut, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pack_wa, dim3(1024), dim3(256), 0, s, ch->wa, wtmp, atmp, mm);
    FSIM_HIP(hipGetLastError());
    FSIM_HIP(hipMemcpyAsync(ch->dK, &ch->K, sizeof(ReconConst), hipMemcpyHostToDevice, s));
    FSIM_HIP(hipStreamSynchronize(s));
    (void)hipFree(wtmp);
    (void)hipFree(atmp);
    *out = ch;
    return FSIM_OK;
}

int fsim_channel_destroy(fsim_channel ch) {
    if (!ch) return FSIM_OK;
    FSIM_HIP(hipStreamSynchronize(stream()));
    channel_free(ch);
    delete ch;
    return FSIM_OK;
}

int fsim_reconstruct(fsim_channel ch, const uint16_t *raw, float *out, int count, void *strm) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!ch || !raw || !out || count < 1) return fail(FSIM_ERR_ARG, "reconstruct: bad arguments");
    const ReconConst &K = ch->K;
    const int n = K.n;
    hipStream_t s = strm ? (hipStream_t)strm : stream();
    if (count > ch->cap) {
        FSIM_HIP(hipStreamSynchronize(s));
        (void)hipFree(ch->A);
        (void)hipFree(ch->B);
        ch->A = ch->B = nullptr;
        ch->cap = 0;
        if (hipMalloc(&ch->A, workspace_a_elems(n, K.D, K.Cs) * count * sizeof(float2)) != hipSuccess ||
            hipMalloc(&ch->B, workspace_b_elems(n) * count * sizeof(float2)) != hipSuccess) {
            (void)hipFree(ch->A);
            ch->A = nullptr;
            (void)hipGetLastError();
            return fail(FSIM_ERR_OOM, "reconstruct: workspace allocation failed");
        }
        ch->cap = count;
    }
    JobList jl{};
    jl.jobs = nullptr;
    jl.base.c = ch->dK;
    jl.base.raw = raw;
    jl.base.A = ch->A;
    jl.base.B = ch->B;
    jl.base.out = out;
    jl.raw_stride = (size_t)K.D * K.P * n * n;
    jl.a_stride = workspace_a_elems(n, K.D, K.Cs);
    jl.b_stride = workspace_b_elems(n);
    jl.out_stride = (size_t)4 * n * n;
    return launch_pipeline(n, jl, count, K.D, K.Cs, s);
}

int fsim_stage_timing(int enable) {
    std::lock_guard<std::mutex> lk(g_timer.mu);
    g_timer.on = enable != 0;
    return FSIM_OK;
}

int fsim_stage_times(double *ms, int *launches, int nstage) {
    if (!ms || !launches || nstage < 5) return fail(FSIM_ERR_ARG, "stage_times: need 5 slots");
    std::lock_guard<std::mutex> lk(g_timer.mu);
    for (int i = 0; i < nstage; i++) {
        ms[i] = 0;
        launches[i] = 0;
    }
    std::set<hipEvent_t> owned;
    for (auto &e : g_timer.ev) {
        FSIM_HIP(hipEventSynchronize(e.second.second));
        float t = 0;
        FSIM_HIP(hipEventElapsedTime(&t, e.second.first, e.second.second));
        ms[e.first] += t;
        launches[e.first] += 1;
        owned.insert(e.second.first);
        owned.insert(e.second.second);
    }
    for (auto e : owned) (void)hipEventDestroy(e);
    g_timer.ev.clear();
    return FSIM_OK;
}

}  // extern "C"
