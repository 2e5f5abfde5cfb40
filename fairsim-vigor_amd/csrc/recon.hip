// Fused fast-shift SIM reconstruction.
//
// Dataflow of one stack (livemode/ReconstructionRunner.java:456-542 and
// tests/TestAccelSpeed.java:237-252, 563-703), regrouped into five HBM passes
// over the raw side n and the doubled side m = 2n:
//
//  K1 k_sep_rows    raw u16 -> fade -> band separation (linear, so it moves in
//                   front of the FFT) -> row FFT            writes A[d][c]
//  K2 k_col_shift   shifted components only: column FFT -> * conj(OTF) ->
//                   column IFFT -> column factor of the fast-shift phase ramp
//                   -> column FFT                             in place on A
//  K3 k_row_shift   shifted components only: row IFFT -> row factor of the
//                   ramp -> row FFT                           in place on A
//                   The ramp exp(2 pi i (fx x + fy y) / n) of SimUtils.java:
//                   74-81 is separable, so IFFT2 -> ramp -> FFT2 splits into
//                   a column part (K2) and a row part (K3); A then holds the
//                   full shifted spectrum of every shifted component.
//  K5 k_gather      per column tile of the m x m result: DC component column
//                   FFT * conj(OTF), pasteFreq of every component as a gather
//                   into a register accumulator (SimUtils.java:83-86), then
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

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <mutex>
#include <type_traits>
#include <map>
#include <set>
#include <utility>
#include <vector>

#ifndef FSIM_GATHER_CW
#define FSIM_GATHER_CW 8
#endif
#ifndef FSIM_JOBMAP
#define FSIM_JOBMAP 1
#endif
#ifndef FSIM_COL_CW
#define FSIM_COL_CW 8
#endif
#ifndef FSIM_COL_XCD
#define FSIM_COL_XCD 1
#endif
#ifndef FSIM_TILE_GROUPS
#define FSIM_TILE_GROUPS 1
#endif
#ifndef FSIM_COL_WPE
#define FSIM_COL_WPE 6
#endif
// Twiddles of the row kernels: 0 sincospi per pass, 1 a block-shared LDS
// table read per pass, 2 the thread's twiddles in registers, read once from
// that table.  Column pass: 1 LDS table per pass, 2 registers from the LDS
// table.  Defaults measured per kernel at 512^2 (64 stacks): the all-units
// separation runs four transforms per thread (registers: 0.166 ms against
// 0.178 with sincospi), the row shift takes the table only from n = 2048
// on (sincospi: 0.159 against 0.169 ms at 512^2, 0.169 against 0.187 at
// 1024^2 with 16 stacks, where the table costs a block per CU; 0.411
// against 0.398 ms at 2048^2 with 8), both separation kernels share one mode (the same
// arithmetic, so the two forms give bitwise equal spectra), the complex-to-real pass loses to a table fill
// (0.093 ms sincospi, 0.114 table), the column pass loses with register
// twiddles (0.195 ms table, 0.24 registers: occupancy).
// the row shift reads the LDS table from this n on (sincospi below)
#ifndef FSIM_ROW_TW_MIN
#define FSIM_ROW_TW_MIN 2048
#endif
#ifndef FSIM_ROW_TW
#define FSIM_ROW_TW 1
#endif
#ifndef FSIM_SEP_TW_MAX
#define FSIM_SEP_TW_MAX 1024
#endif
#ifndef FSIM_SEP_TW
#define FSIM_SEP_TW 2
#endif
#ifndef FSIM_C2R_TW
#define FSIM_C2R_TW 0
#endif
#ifndef FSIM_COL_TW
#define FSIM_COL_TW 1
#endif

namespace fsim {

using KRef = const FSIM_CST ReconConst &;

__device__ __forceinline__ Job get_job(const JobList &jl, int j) {
    Job J;
    if (jl.jobs) {
        const FSIM_CST Job &g = cst(jl.jobs + j);
        J.c = g.c;
        J.raw = g.raw;
        J.A = g.A;
        J.B = g.B;
        J.out = g.out;
        return J;
    }
    J = jl.base;
    J.raw += (size_t)j * jl.raw_stride;
    J.A += (size_t)j * jl.a_stride;
    J.B += (size_t)j * jl.b_stride;
    J.out += (size_t)j * jl.out_stride;
    return J;
}

// Work index w -> (unit, tile) in tile-group-major order: groups of TG
// adjacent tiles, and within a group every unit (component) with its TG
// tiles in turn.  With xcd_linear, an XCD walks one group for all units
// before the next, so the OTF tiles of that group stay in its L2 while the
// components stream past (otherwise each component re-reads the table), and
// neighbouring tiles, which share cache lines, still run back to back.
template <int TG>
__device__ __forceinline__ void tile_group_order(int w, int units, int &unit, int &tile) {
    const int g = w / (units * TG), r = w % (units * TG);
    unit = r / TG;
    tile = g * TG + r % TG;
}

// Fade factor of pixel (x, y): fadeBorderCos on a vector of ones
// (SimUtils.java:162-185, ReconstructionRunner.java:781-783): the products
// 1 * lo[y] * hi[y] * lo[x] * hi[x], each rounded to float.
// fade_rowv: (float)((float)(1 * lo[y]) * hi[y]).  In a row, lo[x] is 1.0
// except in the left border and hi[x] except in the right one (the borders
// do not overlap), so the column part is one rounded product with the
// factor that is not 1 (multiplying by 1.0 is exact): fade_factor returns
// that factor (1.0 inside), fade_colv applies it.
__device__ __forceinline__ float fade_rowv(double lo_y, double hi_y) {
    float f = 1.0f;
    f = (float)((double)f * lo_y);
    f = (float)((double)f * hi_y);
    return f;
}
__device__ __forceinline__ double fade_factor(const FSIM_GLB double *lo, const FSIM_GLB double *hi, int x, int n,
                                              int px) {
    const bool left = x < px, right = x >= n - px;
    const FSIM_GLB double *tab = left ? lo : hi;
    const double v = tab[(left || right) ? x : 0];
    return (left || right) ? v : 1.0;
}
__device__ __forceinline__ float fade_colv(float fy, double fx) { return (float)((double)fy * fx); }

// ---------------------------------------------------------------------------
// K1: band separation + row FFT.  One wavefront (T = n/E threads) per row of
// one unit: the DC component summed over all directions (the gather filters
// the DC components of every direction with the same band-0 OTF at the same
// offset, so by linearity their sum is all it needs; with a Hermitian result
// only its columns 0..n/2), or one stored shifted component of one
// direction.  The raw rows are read (the blocks of the other units of the
// row run on the same XCD and find them in L2), faded, combined as
// c = sum_f S[f / P][c][f % P] * raw_f over the unit's frames and
// transformed.  VEC: each lane reads 4 adjacent pixels per load (8-byte
// aligned rows) and the combined values reach the transform's element order
// through the row's LDS region.
// ---------------------------------------------------------------------------
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));

// Component spectra in A are tile-major: 8-column tiles, each one contiguous
// run over the n rows, element (x, y) at ((x / 8) * n + y) * 8 + x % 8.  The
// column kernels then read and write whole runs; the row kernels move the
// runs of their RW rows (RW * 8 elements per tile) through LDS.
constexpr int kAT = 8;
// Tile height: the whole column (row bands of 16-64 rows measured no
// different).
template <int N>
struct ATile {
    static constexpr int rh = N;
};
template <int N>
__device__ __forceinline__ size_t at_col(int x) {
    return (size_t)(x / kAT) * ATile<N>::rh * kAT + (x % kAT);
}
template <int N>
__device__ __forceinline__ size_t at_row(int y) {
    constexpr int RH = ATile<N>::rh;
    return (size_t)(y / RH) * N * RH + (size_t)(y % RH) * kAT;
}

// Rows [y0, y0 + RW) of a tile-major n x n array, tiles [0, ntiles), into
// row-major LDS rows (row r at rows + r * RP) / back from them.  Every thread
// of the block (NT = n * RW / E threads, E elements each) calls them; the
// caller places the block barriers.
template <int N, int RW, int NT>
__device__ __forceinline__ void tiled_rows_load(const FSIM_GLB float2 *a, int y0, float2 *rows, int RP, int ntiles,
                                                int tid) {
    constexpr int RUN = RW * kAT, PER = N * RW / NT;
    static_assert(ATile<N>::rh % RW == 0, "a block's rows lie in one tile row");
    const FSIM_GLB float2 *base = a + at_row<N>(y0);
    const int lim = ntiles * RUN;
    float2 u[PER];
#pragma unroll
    for (int r = 0; r < PER; r++) {
        const int e = tid + NT * r;
        if (e < lim) u[r] = ld(base + at_col<N>((e / RUN) * kAT) + e % RUN);
    }
#pragma unroll
    for (int r = 0; r < PER; r++) {
        const int e = tid + NT * r;
        if (e < lim) {
            const int q = e % RUN;
            rows[(q / kAT) * RP + (e / RUN) * kAT + q % kAT] = u[r];
        }
    }
}
// Row y of a tile-major n x n array from the registers of its transform
// (element k = t + T m): 8 lanes write one tile's 64 bytes of the row; the
// other half of each 128-byte line is the neighbouring row, written by the
// neighbouring wave, and the halves meet in L2.
template <int N, int E>
__device__ __forceinline__ void tiled_row_store_regs(FSIM_GLB float2 *a, int y, int t, const float2 (&v)[E], int kmax) {
    constexpr int T = N / E;
    FSIM_GLB float2 *base = a + at_row<N>(y);
#pragma unroll
    for (int m = 0; m < E; m++) {
        const int k = t + T * m;
        if (k <= kmax) st(base + at_col<N>(k), v[m]);
    }
}

template <int N, int RW, int NT>
__device__ __forceinline__ void tiled_rows_store(FSIM_GLB float2 *a, int y0, const float2 *rows, int RP, int ntiles,
                                                 int tid) {
    constexpr int RUN = RW * kAT, PER = N * RW / NT;
    static_assert(ATile<N>::rh % RW == 0, "a block's rows lie in one tile row");
    FSIM_GLB float2 *base = a + at_row<N>(y0);
    const int lim = ntiles * RUN;
#pragma unroll
    for (int r = 0; r < PER; r++) {
        const int e = tid + NT * r;
        if (e < lim) {
            const int q = e % RUN;
            st(base + at_col<N>((e / RUN) * kAT) + q, rows[(q / kAT) * RP + (e / RUN) * kAT + q % kAT]);
        }
    }
}

// LDS of the row kernels' items (float2 elements): the rows' transform
// regions and the twiddle table (tl elements) and, for the complex-to-real
// pass, the row-major copy of the block's rows.
// pitch: the rows' regions lie pitch elements apart, 8 more than a region
// where a block holds several rows, so that the tile-run staging (8 lanes
// per row of a run) puts a wave's rows on different LDS banks (a pitch that
// is a multiple of 32 banks would put all of them on the same ones).
template <int N>
struct RowItemLds {
    static constexpr int E = FftE<N>::value, T = N / E;
    static constexpr int RW = T >= 256 ? 1 : 256 / T;
    static constexpr int pitch = RW > 1 ? FftLds<N, E>::len + 8 : FftLds<N, E>::len;
    static constexpr int rows = RW * pitch;
};

// The separation kernels' twiddles (one choice per n for both forms, so
// that they give bitwise equal spectra): below FSIM_SEP_TW_MAX the LDS table
// (FSIM_SEP_TW 1) or each thread's twiddles loaded from it into registers
// (2, where they fit); from FSIM_SEP_TW_MAX on sincospi, which frees the
// table's LDS for another block per CU (one-unit kernel at 2048^2: 0.445
// against 0.477 ms per 8 stacks; 1024^2: 0.210 against 0.219 ms per 16).
template <int N>
struct SepTw {
    static constexpr bool table = FSIM_SEP_TW != 0 && N < FSIM_SEP_TW_MAX;
    static constexpr bool regs = FSIM_SEP_TW == 2 && table && tw_reg_fits<N, FftE<N>::value>();
};

// One (stack, row block, unit) item of the one-unit-per-block separation
// (the kernel k_sep_rows1, and the persistent work-list kernels): bx = the
// block's position L among the stack's row blocks and units, by = the stack.
template <int N, bool VEC>
__device__ __forceinline__ void sep_rows1_item(const JobList &jl, int nrb, int units, int half,
                                               const float2 *__restrict__ tw, unsigned bx, unsigned by, float2 *lds,
                                               float2 *twl, bool &twk) {
    constexpr int E = FftE<N>::value, T = N / E;
    constexpr int RW = T >= 256 ? 1 : 256 / T;
    constexpr int LP = FftLds<N, E>::len;
    const Job J = get_job(jl, by);
    KRef K = cst(J.c);
    const int C = K.Cs, P = K.P;
    // unit 0 of a row block: the DC component summed over the directions;
    // unit 1 + d (C - 1) + (c - 1): stored component c >= 1 of direction d.
    // The units of one row block run on one XCD (blocks L, L + 8, ...), so
    // the raw rows they share are read from L2.
    const int L = bx;
    const int rb = (L / 8 / units) * 8 + (L % 8);
    const int u = (L / 8) % units;
    if (rb >= nrb) return;
    int d = 0, c = 0, f0 = 0, nf = K.D * P;
    if (u > 0) {
        d = (u - 1) / (C - 1);
        c = 1 + (u - 1) % (C - 1);
        f0 = d * P;
        nf = P;
    }
    const int crow = K.srow[c];          // separation-matrix row
    using TW = TwLds<N, SepTw<N>::table>;
    const int tid = item_tid();
    const float2 *tws = tw_table<TW::tl, T * RW>(twl, tw, tid, twk);
    const int t = tid % T, rr = tid / T;
    const int y = rb * RW + rr;
    const size_t nn = (size_t)N * N;
    // frames are angle-major: frame f = d * P + p
    const FSIM_GLB uint16_t *raw = glb(J.raw) + (size_t)f0 * nn + (size_t)y * N;
    const FSIM_GLB double *flo = glb(K.fade_lo), *fhi = glb(K.fade_hi);
    const int fpx = K.fade_px;
    float2 v[E];
    // The loads of the row (fade factors, the pixels of up to G frames) are
    // issued before their first use, so a wave waits for memory once per G
    // frames.
    constexpr int G = 4;
    if constexpr (VEC) {
        static_assert(E % 4 == 0, "4 pixels per load");
        constexpr int Q = E / 4;
        using PL = FftLds<N, E>;
        // lane t holds pixels x = 4 t + 4 T k + j, k < Q, j < 4
        double fx[Q][4];
#pragma unroll
        for (int k = 0; k < Q; k++)
#pragma unroll
            for (int j = 0; j < 4; j++) fx[k][j] = fade_factor(flo, fhi, 4 * t + 4 * T * k + j, N, fpx);
        const double fly = flo[y], fhy = fhi[y];
        float2 w[Q][4];
        float fd[Q][4];
        for (int p0 = 0; p0 < nf; p0 += G) {
            u16x4 r[G][Q];
#pragma unroll
            for (int g = 0; g < G; g++)
                if (p0 + g < nf) {
                    const FSIM_GLB u16x4 *rp = (const FSIM_GLB u16x4 *)(raw + (size_t)(p0 + g) * nn);
#pragma unroll
                    for (int k = 0; k < Q; k++) r[g][k] = rp[t + T * k];
                }
            if (p0 == 0) {
                const float fy = fade_rowv(fly, fhy);
#pragma unroll
                for (int k = 0; k < Q; k++)
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        w[k][j] = make_float2(0.f, 0.f);
                        fd[k][j] = fade_colv(fy, fx[k][j]);
                    }
            }
#pragma unroll
            for (int g = 0; g < G; g++)
                if (p0 + g < nf) {
                    const int f = f0 + p0 + g;
                    const float2 s = ld(&K.S[f / P][crow][f % P]);
#pragma unroll
                    for (int k = 0; k < Q; k++) {
                        const float r4[4] = {(float)r[g][k].x, (float)r[g][k].y, (float)r[g][k].z,
                                             (float)r[g][k].w};
#pragma unroll
                        for (int j = 0; j < 4; j++) {
                            const float px = r4[j] * fd[k][j];
                            w[k][j].x += s.x * px;
                            w[k][j].y += s.y * px;
                        }
                    }
                }
        }
        float2 *row = lds + rr * LP;
#pragma unroll
        for (int k = 0; k < Q; k++)
#pragma unroll
            for (int j = 0; j < 4; j++) row[PL::pad(4 * t + 4 * T * k + j)] = w[k][j];
        pass_sync<RowWave<N>::value>();
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = row[PL::pad(t + T * m)];
        pass_sync<RowWave<N>::value>();
    } else {
        double fx[E];
#pragma unroll
        for (int m = 0; m < E; m++) fx[m] = fade_factor(flo, fhi, t + T * m, N, fpx);
        const double fly = flo[y], fhy = fhi[y];
        float fade[E];
        for (int p0 = 0; p0 < nf; p0 += G) {
            uint16_t r[G][E];
#pragma unroll
            for (int g = 0; g < G; g++)
                if (p0 + g < nf) {
                    const FSIM_GLB uint16_t *rp = raw + (size_t)(p0 + g) * nn;
#pragma unroll
                    for (int m = 0; m < E; m++) r[g][m] = rp[t + T * m];
                }
            if (p0 == 0) {
                const float fy = fade_rowv(fly, fhy);
#pragma unroll
                for (int m = 0; m < E; m++) {
                    v[m] = make_float2(0.f, 0.f);
                    fade[m] = fade_colv(fy, fx[m]);
                }
            }
#pragma unroll
            for (int g = 0; g < G; g++)
                if (p0 + g < nf) {
                    const int f = f0 + p0 + g;
                    const float2 s = ld(&K.S[f / P][crow][f % P]);
#pragma unroll
                    for (int m = 0; m < E; m++) {
                        const float px = (float)r[g][m] * fade[m];
                        v[m].x += s.x * px;
                        v[m].y += s.y * px;
                    }
                }
        }
    }
    if constexpr (SepTw<N>::regs) {
        TwReg<N, E> twr;
        tw_reg_load<N, E, TW::tl>(twr, t, tws);
        fft_regs_r<N, E, false, RowWave<N>::value>(v, lds + rr * LP, 1, t, twr);
    } else {
        fft_regs<N, E, false, RowWave<N>::value, TW::tl>(v, lds + rr * LP, 1, t, tws);
    }
    // with a Hermitian result the gather reads DC columns 0..n/2 only
    const int ntiles = (u == 0 && half) ? N / 2 / kAT + 1 : N / kAT;
    tiled_row_store_regs<N, E>(glb(J.A) + (size_t)(d * C + c) * nn, y, t, v, ntiles * kAT - 1);
}

template <int N, bool VEC>
__global__ void __launch_bounds__(256) k_sep_rows1(JobList jl, int nrb, int units, int half,
                                                   const float2 *__restrict__ tw) {
    __shared__ float2 lds[RowItemLds<N>::rows];
    __shared__ float2 twl[TwLds<N, SepTw<N>::table>::len];
    bool twk = false;
    sep_rows1_item<N, VEC>(jl, nrb, units, half, tw, blockIdx.x, blockIdx.y, lds, twl, twk);
}

// K1, all units of a row in one wavefront: the row's D*P raw frames are
// loaded once (4 pixels per load, all loads in flight together) and kept in
// registers as 16-bit pixels; the DC sum and every stored shifted component
// are then combined, transformed and stored in turn.  FR bounds the frames
// held (D*P <= FR: 9, or 15 for five phases in three directions); used for
// n <= 512 (8 pixels per lane), where the frames fit the register budget.
// One (stack, row block) item of the all-units separation: bx = the row
// block, by = the stack.
template <int N, int FR>
__device__ __forceinline__ void sep_rows_all_item(const JobList &jl, int half, const float2 *__restrict__ tw, unsigned bx,
                                                  unsigned by, float2 *lds, float2 *twl, bool &twk) {
    constexpr int E = FftE<N>::value, T = N / E;
    constexpr int RW = T >= 256 ? 1 : 256 / T;
    constexpr int LP = FftLds<N, E>::len;
    static_assert(E % 4 == 0, "4 pixels per load");
    constexpr int Q = E / 4;
    using PL = FftLds<N, E>;
    const Job J = get_job(jl, by);
    KRef K = cst(J.c);
    const int C = K.Cs, P = K.P, nf = K.D * P;
    using TW = TwLds<N, SepTw<N>::table>;
    const int tid = item_tid();
    const float2 *tws = tw_table<TW::tl, T * RW>(twl, tw, tid, twk);
    const int t = tid % T, rr = tid / T;
    const int y = bx * RW + rr;
    const size_t nn = (size_t)N * N;
    const FSIM_GLB uint16_t *raw = glb(J.raw) + (size_t)y * N;
    const FSIM_GLB double *flo = glb(K.fade_lo), *fhi = glb(K.fade_hi);
    const int fpx = K.fade_px;
    // lane t holds pixels x = 4 t + 4 T k + j, k < Q, j < 4
    u16x4 r[FR][Q];
#pragma unroll
    for (int f = 0; f < FR; f++)
        if (f < nf) {
            const FSIM_GLB u16x4 *rp = (const FSIM_GLB u16x4 *)(raw + (size_t)f * nn);
#pragma unroll
            for (int k = 0; k < Q; k++) r[f][k] = rp[t + T * k];
        }
    float fd[Q][4];
    {
        const float fy = fade_rowv(flo[y], fhi[y]);
#pragma unroll
        for (int k = 0; k < Q; k++)
#pragma unroll
            for (int j = 0; j < 4; j++) fd[k][j] = fade_colv(fy, fade_factor(flo, fhi, 4 * t + 4 * T * k + j, N, fpx));
    }
    float2 *row = lds + rr * LP;
    const int units = 1 + K.D * (C - 1);
    // one set of twiddles serves every unit's transform
    constexpr bool RT = SepTw<N>::regs;
    TwReg<N, E> twr;
    if constexpr (RT) tw_reg_load<N, E, TW::tl>(twr, t, tws);
    for (int u = 0; u < units; u++) {
        // unit 0: DC summed over the directions; then (d, c), c >= 1
        int d = 0, c = 0, f0 = 0, f1 = nf;
        if (u > 0) {
            d = (u - 1) / (C - 1);
            c = 1 + (u - 1) % (C - 1);
            f0 = d * P;
            f1 = f0 + P;
        }
        const int crow = K.srow[c];
        float2 w[Q][4];
#pragma unroll
        for (int k = 0; k < Q; k++)
#pragma unroll
            for (int j = 0; j < 4; j++) w[k][j] = make_float2(0.f, 0.f);
#pragma unroll
        for (int f = 0; f < FR; f++)
            if (f >= f0 && f < f1) {
                const float2 sv = ld(&K.S[f / P][crow][f % P]);
#pragma unroll
                for (int k = 0; k < Q; k++) {
                    const float r4[4] = {(float)r[f][k].x, (float)r[f][k].y, (float)r[f][k].z, (float)r[f][k].w};
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const float px = r4[j] * fd[k][j];
                        w[k][j].x += sv.x * px;
                        w[k][j].y += sv.y * px;
                    }
                }
            }
        float2 v[E];
#pragma unroll
        for (int k = 0; k < Q; k++)
#pragma unroll
            for (int j = 0; j < 4; j++) row[PL::pad(4 * t + 4 * T * k + j)] = w[k][j];
        pass_sync<RowWave<N>::value>();
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = row[PL::pad(t + T * m)];
        pass_sync<RowWave<N>::value>();
        if constexpr (RT)
            fft_regs_r<N, E, false, RowWave<N>::value>(v, row, 1, t, twr);
        else
            fft_regs<N, E, false, RowWave<N>::value, TW::tl>(v, row, 1, t, tws);
        // with a Hermitian result the gather reads DC columns 0..n/2 only
        const int ntiles = (u == 0 && half) ? N / 2 / kAT + 1 : N / kAT;
        tiled_row_store_regs<N, E>(glb(J.A) + (size_t)(d * C + c) * nn, y, t, v, ntiles * kAT - 1);
    }
}

template <int N, int FR>
__global__ void __launch_bounds__(256) k_sep_rows_all(JobList jl, int half, const float2 *__restrict__ tw) {
    __shared__ float2 lds[RowItemLds<N>::rows];
    __shared__ float2 twl[TwLds<N, SepTw<N>::table>::len];
    bool twk = false;
    sep_rows_all_item<N, FR>(jl, half, tw, blockIdx.x, blockIdx.y, lds, twl, twk);
}

// ---------------------------------------------------------------------------
// K2: shifted components: column FFT, * conj(OTF_band), column IFFT (1/n),
// column factor ry(y) of the sub-pixel ramp, column FFT.  Blocks own
// CW-column tiles; the tiles of one component run in order on one XCD.
// ---------------------------------------------------------------------------
// Waves per SIMD a kernel's registers are planned for: the blocks that the
// 160 KB of LDS per CU hold, clamped to [1, cap] (cap 4 = 128 VGPRs).
constexpr int lds_waves_per_simd(int threads, int lds_bytes, int cap) {
    const int w = (163840 / lds_bytes) * threads / 256;
    return w < 1 ? 1 : (w > cap ? cap : w);
}

// E0: elements per thread (the Stockham radix; 0: ColE), CW0: columns per
// block (capped by the 1024-thread block).  The pipeline picks one of the
// candidates kColE / kColCW per size (tuned or default).
// WC: one column's transform per wavefront (threads t-fastest, E chosen so
// that T = n/E <= 64): the transforms synchronise at wavefront level and
// the waves of a block run independently; the loads of a wave are 64-byte
// strided (one column of a tile run), the block's waves read the same lines.
template <int N>
struct WaveColE {
    static constexpr int value = N / 64 > 8 ? N / 64 : 8;
};
template <int N, int E0, int CW0, bool WC = false>
struct ColShiftTile {
    static constexpr int e = WC ? WaveColE<N>::value : ((E0 > 0 && E0 <= 16 && E0 < N / 2) ? E0 : ColE<N>::value);
    static constexpr int t = N / e;
    static_assert(!WC || t <= 64, "a column's transform inside one wavefront");
    static constexpr int cw0 = ColTile<N, t, e>::value;
    static constexpr int cw = cw0 < CW0 ? cw0 : CW0;
    static constexpr int value = t * cw;
    static constexpr int lds = FftLds<N, e>::len * cw;
    // the LDS twiddle table only where it costs no block per CU (at 2048^2 a
    // 4-column block would otherwise run alone on its CU, with no other
    // block loading while it transforms); sincospi twiddles otherwise
    static constexpr bool tw_fits =
        (lds + N) * 8 <= 163840 && 163840 / ((lds + N) * 8) == 163840 / (lds * 8);
    // register twiddles from that table: about 28 more VGPRs, at most 5
    // waves per SIMD keep them without spills
    static constexpr bool rt = FSIM_COL_TW == 2 && tw_fits && tw_reg_fits<N, e>();
    static constexpr int wpe = lds_waves_per_simd(value, (lds + (tw_fits ? N : 0)) * 8, rt ? 5 : FSIM_COL_WPE);
};

// One block's work of the column pass: b = the block's position in the
// pass's grid (blockIdx.x of k_col_shift).
template <int N, int E0, int CW0, bool WC>
__device__ __forceinline__ void col_shift_item(const JobList &jl, int nsh_all, int total,
                                               const float2 *__restrict__ tw, unsigned b, float2 *lds, float2 *twl,
                                               bool &twk) {
    using CT = ColShiftTile<N, E0, CW0, WC>;
    constexpr int E = CT::e, T = N / E;
    constexpr int CW = CT::cw;
    constexpr int tiles = N / CW;
    constexpr int LP = FftLds<N, E>::len;
    // nsh_all = D * (Cs - 1) shifted components per stack; total = stacks *
    // nsh_all * tiles
    int w;
    if (FSIM_COL_XCD) {
        w = xcd_linear_at(b, total);
        if (w >= total) return;
    } else {
        w = b;
    }
    // twiddles: in registers, loaded once from the global table, where
    // they fit (L <= 512); else the LDS table
    constexpr bool RT = CT::rt;
    using TW = TwLds<N, CT::tw_fits>;
    const int tid = item_tid();
    const float2 *tws = tw_table<TW::tl, CT::value>(twl, tw, tid, twk);
    int tile, u;
    if (FSIM_TILE_GROUPS) {
        constexpr int TG = tiles < 8 ? tiles : 8;
        tile_group_order<TG>(w, total / tiles, u, tile);
        // group g holds tiles g, g + G, g + 2G, ... (G groups), in runs of S
        // adjacent tiles that share A's 8-column runs (S = 8 / CW): each
        // XCD's share spans the whole column range, so the zero tiles beyond
        // the OTF cutoff (a contiguous range) spread evenly over the XCDs
        // instead of emptying one XCD's share
        constexpr int S = CW < kAT ? kAT / CW : 1;
        constexpr int TGB = TG / S > 0 ? TG / S : 1, NB = tiles / S;
        const int j = tile / S;
        tile = (j / TGB + (NB / TGB) * (j % TGB)) * S + tile % S;
    } else {
        tile = w % tiles;
        u = w / tiles;
    }
    const Job J = get_job(jl, u / nsh_all);
    KRef K = cst(J.c);
    const int C = K.Cs, nsh = C - 1, dc = u % nsh_all;
    const int d = dc / nsh, c = 1 + dc % nsh;   // stored shifted component
    const int cl = WC ? tid / T : tid % CW;
    const int t = WC ? tid % T : tid / CW;
    // LDS region of the column: interleaved with the tile's other columns,
    // or (WC) a contiguous one per column
    float2 *reg = WC ? lds + cl * LP : lds + cl;
    constexpr int LS = WC ? 1 : CW;
    const int x = tile * CW + cl;
    FSIM_GLB float2 *col = glb(J.A) + (size_t)(d * C + c) * N * N + at_col<N>(x);
    if (tile * CW >= K.zlo && tile * CW + CW - 1 <= K.zhi) {
        // every column of the tile lies beyond the OTF cutoff: the product
        // with conj(OTF) is zero, and so are its transforms
#pragma unroll
        for (int m = 0; m < E; m++) st(col + at_row<N>(t + T * m), make_float2(0.f, 0.f));
        return;
    }
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; m++) v[m] = ld(col + at_row<N>(t + T * m));
    TwReg<N, E> twr;
    if constexpr (RT) {
        tw_reg_load<N, E, TW::tl>(twr, t, tws);
        fft_regs_r<N, E, false, WC>(v, reg, LS, t, twr);
    } else {
        fft_regs<N, E, false, WC, TW::tl>(v, reg, LS, t, tws);
    }
    const int band = K.sband[c];
    const FSIM_GLB float2 *otf = glb(K.otf2d) + (size_t)(K.otf_nb > 1 ? band : 0) * N * N + x;
    if (K.otf_real) {
        // conj(OTF) real: the product's value without the zero terms
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = cscale(v[m], ld(otf + (size_t)(t + T * m) * N).x);
    } else {
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = times_conj_java(v[m], ld(otf + (size_t)(t + T * m) * N));
    }
    // column ramp factors (in flight during the inverse transform where the
    // registers allow)
    const FSIM_GLB float2 *ry = glb(K.ramp) + ((size_t)(d * kMaxComp + K.sramp[c]) * 2 + 1) * N;
    constexpr int ER = E <= 8 ? E : 1;
    float2 r[ER];
    if constexpr (E <= 8) {
#pragma unroll
        for (int m = 0; m < E; m++) r[m] = ld(ry + t + T * m);
    }
    if constexpr (RT)
        fft_regs_r<N, E, true, WC>(v, reg, LS, t, twr);
    else
        fft_regs<N, E, true, WC, TW::tl>(v, reg, LS, t, tws);
    // the ramp factors carry the inverse transform's 1/n (exact: a power of two)
#pragma unroll
    for (int m = 0; m < E; m++) {
        const float2 rm = E <= 8 ? r[m < ER ? m : 0] : ld(ry + t + T * m);
        v[m] = rot_java(v[m], rm.x, rm.y);
    }
    if constexpr (RT)
        fft_regs_r<N, E, false, WC>(v, reg, LS, t, twr);
    else
        fft_regs<N, E, false, WC, TW::tl>(v, reg, LS, t, tws);
#pragma unroll
    for (int m = 0; m < E; m++) st(col + at_row<N>(t + T * m), v[m]);
}

template <int N, int E0, int CW0, bool WC>
__global__ void __launch_bounds__((ColShiftTile<N, E0, CW0, WC>::value), (ColShiftTile<N, E0, CW0, WC>::wpe))
    k_col_shift(JobList jl, int nsh_all, int total, const float2 *__restrict__ tw) {
    using CT = ColShiftTile<N, E0, CW0, WC>;
    __shared__ float2 lds[CT::lds];
    __shared__ float2 twl[TwLds<N, CT::tw_fits>::len];
    bool twk = false;
    col_shift_item<N, E0, CW0, WC>(jl, nsh_all, total, tw, blockIdx.x, lds, twl, twk);
}

// ---------------------------------------------------------------------------
// K3: shifted components: row IFFT (1/n), row factor rx(x) of the sub-pixel
// ramp, row FFT.
// ---------------------------------------------------------------------------
// The row shift's twiddles: the LDS table (FSIM_ROW_TW) from n = 2048 on,
// sincospi below, where the table costs the kernel a block per CU.
template <int N>
struct RowShiftTw {
    static constexpr bool value = FSIM_ROW_TW != 0 && N >= FSIM_ROW_TW_MIN;
};

// One (stack, shifted component, row block) item: bx = component x row
// blocks + row block, by = the stack.
template <int N>
__device__ __forceinline__ void row_shift_item(const JobList &jl, const float2 *__restrict__ tw, unsigned bx,
                                               unsigned by, float2 *lds, float2 *twl, bool &twk) {
    constexpr int E = FftE<N>::value, T = N / E;
    constexpr int RW = T >= 256 ? 1 : 256 / T;
    constexpr int rblocks = N / RW;
    constexpr int LP = RowItemLds<N>::pitch;
    const Job J = get_job(jl, by);
    KRef K = cst(J.c);
    const int C = K.Cs, nsh = C - 1;
    using TW = TwLds<N, RowShiftTw<N>::value>;
    const int tid = item_tid();
    const float2 *tws = tw_table<TW::tl, T * RW>(twl, tw, tid, twk);
    const int rb = bx % rblocks, dc = bx / rblocks;
    const int d = dc / nsh, c = 1 + dc % nsh;   // stored shifted component
    const int t = tid % T, rr = tid / T;
    FSIM_GLB float2 *comp = glb(J.A) + (size_t)(d * C + c) * N * N;
    tiled_rows_load<N, RW, T * RW>(comp, rb * RW, lds, LP, N / kAT, tid);
    __syncthreads();
    constexpr bool RT = FSIM_ROW_TW == 2 && RowShiftTw<N>::value && tw_reg_fits<N, E>();
    TwReg<N, E> twr;
    if constexpr (RT) tw_reg_load<N, E, TW::tl>(twr, t, tws);
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; m++) v[m] = lds[rr * LP + t + T * m];
    // the transform's first exchange rewrites the row's region: every thread
    // of the row (two wavefronts from n = 2048 on) has read it first
    pass_sync<RowWave<N>::value>();
    if constexpr (RT)
        fft_regs_r<N, E, true, RowWave<N>::value>(v, lds + rr * LP, 1, t, twr);
    else
        fft_regs<N, E, true, RowWave<N>::value, TW::tl>(v, lds + rr * LP, 1, t, tws);
    // row factor rx(x) of the sub-pixel ramp (the column factor was applied
    // by the column pass); it carries the inverse transform's 1/n
    const FSIM_GLB float2 *rx = glb(K.ramp) + ((size_t)(d * kMaxComp + K.sramp[c]) * 2) * N;
#pragma unroll
    for (int m = 0; m < E; m++) {
        const float2 r = ld(rx + t + T * m);
        v[m] = rot_java(v[m], r.x, r.y);
    }
    if constexpr (RT)
        fft_regs_r<N, E, false, RowWave<N>::value>(v, lds + rr * LP, 1, t, twr);
    else
        fft_regs<N, E, false, RowWave<N>::value, TW::tl>(v, lds + rr * LP, 1, t, tws);
#pragma unroll
    for (int m = 0; m < E; m++) lds[rr * LP + t + T * m] = v[m];
    __syncthreads();
    tiled_rows_store<N, RW, T * RW>(comp, rb * RW, lds, LP, N / kAT, tid);
}

template <int N>
__global__ void __launch_bounds__(256) k_row_shift(JobList jl, const float2 *__restrict__ tw) {
    __shared__ float2 lds[RowItemLds<N>::rows];
    __shared__ float2 twl[TwLds<N, RowShiftTw<N>::value>::len];
    bool twk = false;
    row_shift_item<N>(jl, tw, blockIdx.x, blockIdx.y, lds, twl, twk);
}

// Block -> (stack, column tile) for the gather kernel.  With >= 8 stacks in a
// launch, XCD x takes stacks x, x+8, ... whole, tiles in order: the blocks
// that run together on an XCD read the same rows of the same components (DRAM
// row locality), and the cache lines that neighbouring tiles share (paste
// offsets are not tile aligned) stay in one L2.  With fewer stacks the tiles
// of a stack are spread over all XCDs, with adjacent tile pairs on the same
// XCD.  (Tile-group order, as in the column pass, keeps the Wiener x
// apodization tiles in L2 but spreads the concurrent blocks over 8 stacks;
// measured slower.)
__device__ __forceinline__ bool gather_block(unsigned b, int tiles, int njobs, int &job, int &tile) {
    const int L = (int)b;
    if (njobs >= 8 && FSIM_JOBMAP) {
        const int xcd = L % 8, k = L / 8;
        job = xcd + 8 * (k / tiles);
        tile = k % tiles;
        return job < njobs;
    }
    job = L / tiles;
    const int bx = L % tiles;
    tile = tiles % 16 ? bx : (bx / 16) * 16 + (bx % 8) * 2 + (bx / 8) % 2;
    return true;
}

static inline unsigned gather_grid(int tiles, int njobs) {
    return (unsigned)(njobs >= 8 && FSIM_JOBMAP ? tiles * ((njobs + 7) / 8) * 8 : tiles * njobs);
}

// ---------------------------------------------------------------------------
// K5: gather into the doubled grid, filter, column IFFT.
// Elements per thread for the n-point (en, 8) and 2n-point (em, FftE: 8 or
// 16) column transforms and the tile width: 8 columns up to n = 1024 (512
// threads at n = 512, 1024 at n = 1024), 4 at n = 2048, all within 128 VGPRs
// per thread.
// ---------------------------------------------------------------------------
template <int N>
struct GatherE {
    // 8 elements for the n-point DC transform: from n = 512 on its n/8
    // threads are exactly the 2n/16 of the 2n-point one, and the registers
    // stay free for the accumulator
    static constexpr int en = N >= 8 ? 8 : N;
    static constexpr int em = FftE<2 * N>::value;
    static constexpr int tn = N / en, tm = 2 * N / em;
    static constexpr int tk = tn > tm ? tn : tm;
    static constexpr int cw0 = ColTile<2 * N, tk, (en > em ? en : em)>::value;
    static constexpr int cw = cw0 < FSIM_GATHER_CW ? cw0 : FSIM_GATHER_CW;
    // LDS elements (float2): the larger padded transform, or the n x cw
    // staging tile.  split: the 2n-point transform exchanges its real and
    // imaginary parts in turn through a float region (half its LDS, twice
    // its barriers), chosen where that lets more blocks share a CU.
    static constexpr int lpn = FftLds<N, en>::len, lpm = FftLds<2 * N, em>::len;
    static constexpr int stage = (lpn > N ? lpn : N) * cw;
    static constexpr int full = (lpm * cw > stage ? lpm * cw : stage);
    static constexpr int half = ((lpm * cw + 1) / 2 > stage ? (lpm * cw + 1) / 2 : stage);
    // bytes of a block: the region, and the 2n-entry twiddle table where it fits
    static constexpr int bytes(int l) { return (l + 2 * N) * 8 <= 163840 ? (l + 2 * N) * 8 : l * 8; }
    static constexpr bool split =
        lds_waves_per_simd(tk * cw, bytes(half), 8) > lds_waves_per_simd(tk * cw, bytes(full), 8);
    static constexpr int lds = split ? half : full;
    // registers planned for at most 4 waves per SIMD (128 VGPRs), 6 (80
    // VGPRs) where the split exchange lets three 512-thread blocks share a CU
    static constexpr int wpe = lds_waves_per_simd(tk * cw, (lds + 2 * N) * 8, split ? 6 : 4);
};

// pasteFreq from the n grid into the 2n grid (BasicVectors.java:273-278), one
// axis: output index o, offset off.  p = (o - off) mod 2n; the output takes
// input index p mod n when p lies in [0, n/2) or [3n/2, 2n).
template <int N>
__device__ __forceinline__ int paste_p(int o, int off) { return (o - off) & (2 * N - 1); }
template <int N>
__device__ __forceinline__ bool paste_ok(int p) { return ((p + N / 2) & (2 * N - 1)) < N; }

// Does any column of [x0, x0 + cw) receive a pasted column (offset off)?
template <int N>
__device__ __forceinline__ bool tile_pasted(int x0, int cw, int off) {
    bool hit = false;
    for (int q = 0; q < cw; q++) hit |= paste_ok<N>(paste_p<N>(x0 + q, off));
    return hit;
}

// Hermitian result.  The reference keeps the real part of the inverse
// transform of the full spectrum S, i.e. transforms the Hermitian part
// (S(k) + conj S(-k)) / 2.  The contributions come in conjugate-mirror pairs
// (V_partner(k) = conj V_c(-k); the DC component is its own partner), so
// that part is sum_c w_c(k) V_c(k) with w_c(k) = ([k - K_c in W] +
// [-k - K_partner in W]) / 2, W the paste window [-n/2, n/2)^2: weight 1
// inside both windows, 1/2 on the edge pixels (the window is asymmetric, and
// the sub-pixel shift leaves values there), 0 outside.  Partner of
// contribution c: c itself for DC, else the other member of its band's
// (neg = 2b - 1, pos = 2b) pair.
__device__ __forceinline__ int partner_comp(int c) { return c == 0 ? 0 : ((c & 1) ? c + 1 : c - 1); }

// Does any column of [x0, x0 + cw) lie in the mirrored window (-x - off)?
template <int N>
__device__ __forceinline__ bool tile_pasted_mirror(int x0, int cw, int off) {
    bool hit = false;
    for (int q = 0; q < cw; q++) hit |= paste_ok<N>(paste_p<N>(-(x0 + q), off));
    return hit;
}

// Column tiles of the gather: all m / cw, or with a Hermitian result the
// n / cw tiles of columns [0, n) and one more whose first column is n (the
// Nyquist column; the rest of that tile is formed but not read).
template <int N, bool HALF>
struct GatherTiles {
    static constexpr int value = HALF ? N / GatherE<N>::cw + 1 : 2 * N / GatherE<N>::cw;
};

// One table of length m serves both of the gather's transforms (the n-point
// one at stride 2), where the LDS holds it beside the tile.
template <int N>
using GatherTw = TwLds<2 * N, (GatherE<N>::lds + 2 * N) * 8 <= 163840>;

// One block's work of the gather: b = the block's position in the gather's
// grid (blockIdx.x of k_gather).
template <int N, bool HALF>
__device__ __forceinline__ void gather_item(const JobList &jl, int njobs, const float2 *__restrict__ twn,
                                            const float2 *__restrict__ twm, unsigned b, float2 *lds, float2 *twl,
                                            bool &twk) {
    constexpr int M = 2 * N;
    constexpr int EN = GatherE<N>::en, TN = N / EN;
    constexpr int EM = GatherE<N>::em, TM = M / EM;
    constexpr int CW = GatherE<N>::cw;
    static_assert(!HALF || N + CW <= M, "the Nyquist tile must fit the row");
    constexpr int tiles = GatherTiles<N, HALF>::value;
    int job, tile;
    if (!gather_block(b, tiles, njobs, job, tile)) return;
    using TW = GatherTw<N>;
    const int tid = item_tid();
    const float2 *tws = tw_table<TW::tl, GatherE<N>::tk * GatherE<N>::cw>(twl, twm, tid, twk);
    const Job J = get_job(jl, job);
    KRef K = cst(J.c);
    const int D = K.D, C = K.C, Cs = K.Cs;
    const int cl = tid % CW, t = tid / CW;
    const int xo = tile * CW + cl;
    const FSIM_GLB float2 *A = glb(J.A);
    float2 acc[EM];
#pragma unroll
    for (int m = 0; m < EM; m++) acc[m] = make_float2(0.f, 0.f);
    const size_t nn = (size_t)N * N;
    // DC components of every direction: all are pasted at offset 0 and
    // filtered with the band-0 OTF, so their sum (formed by the separation
    // pass) takes one column FFT and one OTF product, staged through LDS for
    // the row gather.  Skipped as a whole when the tile lies outside the
    // paste window.
    if (tile_pasted<N>(tile * CW, CW, 0) || (HALF && tile_pasted_mirror<N>(tile * CW, CW, 0))) {
        const int px = paste_p<N>(xo, 0);
        const int xs = px & (N - 1);
        const bool okx = paste_ok<N>(px);
        const bool okx2 = HALF && paste_ok<N>(paste_p<N>(-xo, 0));
        const bool act = (t < TN) && (okx || okx2);
        float2 v[EN];
#pragma unroll
        for (int m = 0; m < EN; m++) v[m] = make_float2(0.f, 0.f);
        if (act) {
            // direction 0's DC slot holds the sum over the directions
            const FSIM_GLB float2 *col = A + at_col<N>(xs);
#pragma unroll
            for (int m = 0; m < EN; m++) v[m] = ld(col + at_row<N>(t + TN * m));
        }
        fft_regs<N, EN, false, false, TW::tl>(v, lds + cl, CW, t, tws, act);
        if (act) {
            const FSIM_GLB float2 *otf = glb(K.otf2d) + xs;
            const bool re = K.otf_real != 0;
#pragma unroll
            for (int m = 0; m < EN; m++) {
                const float2 o = ld(otf + (size_t)(t + TN * m) * N);
                v[m] = re ? cscale(v[m], o.x) : times_conj_java(v[m], o);
                lds[(t + TN * m) * CW + cl] = v[m];
            }
        }
        __syncthreads();
        if constexpr (HALF) {
            if (t < TM && (okx || okx2)) {
#pragma unroll
                for (int m = 0; m < EM; m++) {
                    const int y = t + TM * m;
                    const int py = paste_p<N>(y, 0);
                    const int r = (okx && paste_ok<N>(py)) + (okx2 && paste_ok<N>(paste_p<N>(-y, 0)));
                    if (r) acc[m] = cadd(acc[m], cscale(lds[(py & (N - 1)) * CW + cl], 0.5f * (float)r));
                }
            }
        } else if (t < TM && okx) {
#pragma unroll
            for (int m = 0; m < EM; m++) {
                const int py = paste_p<N>(t + TM * m, 0);
                if (paste_ok<N>(py)) acc[m] = cadd(acc[m], lds[(py & (N - 1)) * CW + cl]);
            }
        }
        __syncthreads();
    }
    // shifted contributions, directions then components
    for (int d = 0; d < D; d++) {
        // shifted components: full spectra in A, read in place; a mirrored
        // (neg) contribution reads its pos partner at the mirrored index and
        // conjugates.  Loads first, sums after, so the loads of a
        // contribution are in flight together.
        for (int c = 1; c < C; c++) {
            const int offx = K.offx[d][c];
            const int pc = partner_comp(c);
            const int offx2 = K.offx[d][pc], offy2 = K.offy[d][pc];
            if (!tile_pasted<N>(tile * CW, CW, offx) && !(HALF && tile_pasted_mirror<N>(tile * CW, CW, offx2)))
                continue;
            const int px = paste_p<N>(xo, offx);
            const bool colok = (t < TM) && paste_ok<N>(px);
            const bool colok2 = HALF && (t < TM) && paste_ok<N>(paste_p<N>(-xo, offx2));
            const int mir = K.mir[d][c];
            const int oy = K.offy[d][c], my = K.mdy[d][c];
            int xs = px & (N - 1);
            if (mir) xs = (K.mdx[d][c] - xs) & (N - 1);
            const FSIM_GLB float2 *comp = A + (size_t)(d * Cs + K.src[c]) * nn + at_col<N>(xs);
            const float sg = mir ? -1.0f : 1.0f;
            // Rows y and y + n of the doubled grid read the same component
            // row (the paste map is periodic in n), and each paste window
            // holds exactly one of the two (it spans n of the 2n rows).  So
            // a lane loads rows y = t + TM * i, i < EM / 2, once each,
            // unconditionally when its column is pasted (every such row
            // feeds element i or i + EM / 2), and each element takes the
            // value or zero by its own window test: the same bytes as
            // per-element conditional loads, without a branch per element.
            constexpr int EH = EM / 2;
            float2 sv[EH];
#pragma unroll
            for (int i = 0; i < EH; i++) sv[i] = make_float2(0.f, 0.f);
            if (colok || colok2) {
#pragma unroll
                for (int i = 0; i < EH; i++) {
                    int ys = paste_p<N>(t + TM * i, oy) & (N - 1);
                    if (mir) ys = (my - ys) & (N - 1);
                    sv[i] = ld(comp + at_row<N>(ys));
                }
            }
            // The window tests in closed form: with y = t + TM m and
            // TM EH = n, ((y - oy + n/2) mod 2n) < n holds exactly when bit
            // log2(EH) of ad + m is clear, ad = ((t - oy + n/2) mod 2n) / TM;
            // the mirror test ((-y - offy2 + n/2) mod 2n) < n when that bit
            // of am - m is clear (am offset by 2 EM to stay positive).
            // Elements i and i + EH fall on opposite sides of each window,
            // so their weights are r and (colok + colok2) - r.
            constexpr int LE = ilog2c(EH);
            const int ad = ((t - oy + N / 2) & (2 * N - 1)) / TM;
            const int am = ((-t - offy2 + N / 2) & (2 * N - 1)) / TM + 2 * EM;
            const int rs = (int)colok + (int)colok2;
            const float2 z = make_float2(0.f, 0.f);
#pragma unroll
            for (int i = 0; i < EH; i++) {
                const bool ind = (((ad + i) >> LE) & 1) == 0;
                const bool inm = (((am - i) >> LE) & 1) == 0;
                const int r0 = (int)(colok && ind) + (int)(HALF && colok2 && inm);
                const int rr[2] = {r0, rs - r0};
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    const int m = i + k * EH;
                    const float2 v = rr[k] ? sv[i] : z;
                    if constexpr (HALF) {
                        const float wt = 0.5f * (float)rr[k];
                        acc[m].x += v.x * wt;
                        acc[m].y += sg * v.y * wt;
                    } else {
                        acc[m].x += v.x;
                        acc[m].y += sg * v.y;
                    }
                }
            }
        }
    }
    // fullResult.times(wienerDenom); fullResult.times(apo), as one product
    // table (tile-major: the tile's cw columns of a row are adjacent)
    if (t < TM) {
        const FSIM_GLB float *wa = glb(K.wa) + (size_t)tile * M * CW + cl;
        float w[EM];
#pragma unroll
        for (int m = 0; m < EM; m++) w[m] = wa[(size_t)(t + TM * m) * CW];
#pragma unroll
        for (int m = 0; m < EM; m++) acc[m] = cscale(acc[m], w[m]);
    }
    if constexpr (GatherE<N>::split)
        fft_regs_split<M, EM, true, false, TW::tl>(acc, reinterpret_cast<float *>(lds) + cl, CW, t, tws, t < TM);
    else
        fft_regs<M, EM, true, false, TW::tl>(acc, lds + cl, CW, t, tws, t < TM);
    if (t < TM) {
        const float sc = 1.0f / M;
        if constexpr (HALF) {
            // tile-major: the tile's rows are one contiguous run
            FSIM_GLB float2 *B = glb(J.B) + (size_t)tile * M * CW + cl;
#pragma unroll
            for (int m = 0; m < EM; m++) st(B + (size_t)(t + TM * m) * CW, cscale(acc[m], sc));
        } else {
            FSIM_GLB float2 *B = glb(J.B) + xo;
#pragma unroll
            for (int m = 0; m < EM; m++) st(B + (size_t)(t + TM * m) * M, cscale(acc[m], sc));
        }
    }
}

template <int N, bool HALF>
__global__ void __launch_bounds__(GatherE<N>::tk * GatherE<N>::cw, GatherE<N>::wpe)
    k_gather(JobList jl, int njobs, const float2 *__restrict__ twn, const float2 *__restrict__ twm) {
    __shared__ float2 lds[GatherE<N>::lds];
    __shared__ float2 twl[GatherTw<N>::len];
    bool twk = false;
    gather_item<N, HALF>(jl, njobs, twn, twm, blockIdx.x, lds, twl, twk);
}

// ---------------------------------------------------------------------------
// K6: row IFFT of the doubled grid, real part out.
// ---------------------------------------------------------------------------
// One (stack, row block) item of the full-width last pass: bx = the row
// block, by = the stack.
template <int M>
__device__ __forceinline__ void row_real_item(const JobList &jl, const float2 *__restrict__ tw, unsigned bx,
                                              unsigned by, float2 *lds) {
    constexpr int E = FftE<M>::value, T = M / E;
    constexpr int RW = T >= 256 ? 1 : 256 / T;
    constexpr int LP = FftLds<M, E>::len;
    using TW = TwLds<M, false>;
    const float2 *tws = tw;
    const Job J = get_job(jl, by);
    const int tid = item_tid();
    const int t = tid % T, rr = tid / T;
    const int y = bx * RW + rr;
    const FSIM_GLB float2 *row = glb(J.B) + (size_t)y * M;
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; m++) v[m] = ld(row + t + T * m);
    fft_regs<M, E, true, RowWave<M>::value, TW::tl>(v, lds + rr * LP, 1, t, tws);
    const float sc = 1.0f / M;
    FSIM_GLB float *o = glb(J.out) + (size_t)y * M;
#pragma unroll
    for (int m = 0; m < E; m++) o[t + T * m] = v[m].x * sc;
}

template <int M>
__global__ void __launch_bounds__(256) k_row_real(JobList jl, const float2 *__restrict__ tw) {
    __shared__ float2 lds[RowItemLds<M>::rows];
    row_real_item<M>(jl, tw, blockIdx.x, blockIdx.y, lds);
}

// K6 for a Hermitian result: B holds columns 0..n of the m-point row
// spectra X (X[m - k] = conj(X[k])).  The m real outputs of a row come from
// one n-point inverse transform: with E[k] = X[k] + conj(X[n-k]) and
// O[k] = (X[k] - conj(X[n-k])) exp(2 pi i k / m), k < n, the inverse
// transform of E + iO is x[2j] + i x[2j+1] (the even and odd samples fold
// the m-point sum at k and k + n, and X[k + n] = conj(X[n - k])).  X[0] and
// X[n] enter with their real parts: the transform is that of the Hermitian
// extension, whose real part the full-width pass would keep.
// B is tile-major as the half gather writes it: element (k, y) at
// ((k / cw) * m + y) * cw + k % cw, so the RW rows of a block are, per tile,
// one contiguous run of RW * cw elements.  The block loads those runs with
// all its threads (whole cache lines per wave) into a row-major LDS copy,
// from which each row's threads take X[k] and X[n - k].  Blocks walk the
// rows in XCD-contiguous ranges (the grid's x extent is a multiple of 8),
// so rows that share cache lines are read on one XCD.
// Row-major LDS copy of columns [0, n + cw) of a complex-to-real block's rows.
template <int N>
struct C2rRows {
    static constexpr int CW = GatherE<N>::cw;
    static constexpr int RL = N + CW + 1;   // odd row pitch
    static constexpr int value = RowItemLds<N>::RW * RL;
};

// LDS of a complex-to-real block: the rows' transform regions, then the
// row copy (rows = its offset); from n = 1024 on the regions overlay the
// copy once it has been read (half the LDS, twice the blocks per CU: 0.100
// against 0.115 ms at 1024^2 with 16 stacks, 0.217 against 0.256 at 2048^2
// with 8; at 512^2 the extra blocks were 3 % slower).
template <int N>
struct C2rLds {
    static constexpr bool overlay = N >= 1024;
    static constexpr int rows = overlay ? 0 : RowItemLds<N>::rows;
    static constexpr int value = overlay ? (RowItemLds<N>::rows > C2rRows<N>::value ? RowItemLds<N>::rows : C2rRows<N>::value)
                                         : RowItemLds<N>::rows + C2rRows<N>::value;
};

// One (stack, row group) item of the complex-to-real pass: grp = the row
// group, by = the stack.
template <int N>
__device__ __forceinline__ void row_c2r_item(const JobList &jl, const float2 *__restrict__ twm, int grp, unsigned by,
                                             float2 *lds, float2 *twl, float2 *rows, bool &twk) {
    constexpr int M = 2 * N;
    constexpr int E = FftE<N>::value, T = N / E;
    constexpr int RW = T >= 256 ? 1 : 256 / T;
    constexpr int LP = FftLds<N, E>::len;
    constexpr int CW = GatherE<N>::cw;
    constexpr int NT = T * RW;
    // the twiddles of the n-point transform from the m-point table
    // (TL = m, a multiple of n)
    using TW = TwLds<M, FSIM_C2R_TW != 0>;
    const int tid = item_tid();
    const float2 *tws = tw_table<TW::tl, NT>(twl, twm, tid, twk);
    const Job J = get_job(jl, by);
    const int t = tid % T, rr = tid / T;
    const int y = grp * RW + rr;
    const FSIM_GLB float2 *tg = glb(twm);
    float2 a[E], b[E], w[E];
    // row-major LDS copy of columns [0, n + cw) of the block's rows
    constexpr int RL = C2rRows<N>::RL;
    constexpr int RUN = RW * CW;                // elements per tile run
    constexpr int TOT = (N / CW + 1) * RUN;
    constexpr int PER = (TOT + NT - 1) / NT;
    const FSIM_GLB float2 *bt = glb(J.B) + (size_t)grp * RUN;
    float2 u[PER];
#pragma unroll
    for (int r = 0; r < PER; r++) {
        const int e = tid + NT * r;
        if (TOT % NT == 0 || e < TOT) u[r] = ld(bt + (size_t)(e / RUN) * M * CW + e % RUN);
    }
#pragma unroll
    for (int r = 0; r < PER; r++) {
        const int e = tid + NT * r;
        if (TOT % NT == 0 || e < TOT) {
            const int q = e % RUN;
            rows[(q / CW) * RL + (e / RUN) * CW + q % CW] = u[r];
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < E; m++) {
        const int k = t + T * m;
        a[m] = rows[rr * RL + k];
        b[m] = rows[rr * RL + N - k];
        w[m] = ld(tg + k);   // exp(-2 pi i k / m)
    }
    // where the kernels overlay the transforms' region on the row copy,
    // every thread has read its values before the first exchange writes
    if constexpr (C2rLds<N>::overlay) __syncthreads();
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; m++) {
        float2 x = a[m], z = b[m];
        if (t + T * m == 0) {
            x.y = 0.f;
            z.y = 0.f;
        }
        const float2 e = make_float2(x.x + z.x, x.y - z.y);
        const float2 d = make_float2(x.x - z.x, x.y + z.y);
        const float2 o = make_float2(d.x * w[m].x + d.y * w[m].y, d.y * w[m].x - d.x * w[m].y);
        v[m] = make_float2(e.x - o.y, e.y + o.x);
    }
    if constexpr (FSIM_C2R_TW == 2 && tw_reg_fits<N, E>()) {
        TwReg<N, E> twr;
        tw_reg_load<N, E, TW::tl>(twr, t, tws);
        fft_regs_r<N, E, true, RowWave<N>::value>(v, lds + rr * LP, 1, t, twr);
    } else {
        fft_regs<N, E, true, RowWave<N>::value, TW::tl>(v, lds + rr * LP, 1, t, tws);
    }
    const float sc = 1.0f / M;
    FSIM_GLB float2 *o = reinterpret_cast<FSIM_GLB float2 *>(glb(J.out) + (size_t)y * M);
#pragma unroll
    for (int m = 0; m < E; m++) st(o + t + T * m, cscale(v[m], sc));
}

template <int N>
__global__ void __launch_bounds__(256) k_row_c2r(JobList jl, const float2 *__restrict__ twm) {
    constexpr int RW = RowItemLds<N>::RW;
    __shared__ __align__(16) float2 buf[C2rLds<N>::value];
    __shared__ float2 twl[TwLds<2 * N, FSIM_C2R_TW != 0>::len];
    const int grp = (2 * N / RW) % 8 ? (int)blockIdx.x : xcd_linear(2 * N / RW);
    bool twk = false;
    row_c2r_item<N>(jl, twm, grp, blockIdx.y, buf, twl, buf + C2rLds<N>::rows, twk);
}

// Tile-major position of the tables the gather reads: element j of an
// m x m table with tile width cw holds pixel (x, y).
__device__ __forceinline__ void tile_major_xy(size_t j, int M, int cw, int &x, int &y) {
    x = (int)(j / ((size_t)M * cw)) * cw + (int)(j % cw);
    y = (int)((j / cw) % M);
}

// (1/Wiener denominator) * apodization, and the apodization alone (for the
// per-stack tables), from row-major w and a into the tile-major layout
__global__ void k_pack_wa(float *__restrict__ wa, float *__restrict__ apo, const float *__restrict__ w,
                          const float *__restrict__ a, int M, int cw) {
#pragma clang fp contract(off)
    const size_t mm = (size_t)M * M;
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < mm; j += (size_t)gridDim.x * blockDim.x) {
        int x, y;
        tile_major_xy(j, M, cw, x, y);
        const size_t i = (size_t)y * M + x;
        wa[j] = w[i] * a[i];
        apo[j] = a[i];
    }
}

// Sub-pixel ramp tables exp(2 pi i f t / n) / n, t < n, of the x and y
// fractional shifts f of every direction and component
// (SimUtils.java:74-81), from double precision: ramp[k * n + t] for the
// shift f[k].
__global__ void k_ramps(float2 *__restrict__ ramp, const double *__restrict__ f, int n, size_t total) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int t = (int)(i % n);
        const double ph = 2 * M_PI * f[i / n] * t / n;
        double sn, cs;
        sincos(ph, &sn, &cs);
        // times the 1/n of the inverse transform it follows (exact: n is a
        // power of two), so the kernels apply one product
        const float inv_n = 1.0f / (float)n;
        ramp[i] = make_float2((float)cs * inv_n, (float)sn * inv_n);
    }
}

// Per-stack (1 / Wiener denominator) * apodization tables, tile-major like
// the channel's: stack s uses its own shifts sh[s] and the channel's
// apodization apo.  Only the first `len` elements (the tiles the gather
// reads) are formed.  Grid (pixels, stacks).
__global__ void k_wiener_multi(float *__restrict__ wa, const float *__restrict__ apo, int M, int cw, size_t len,
                               const ShiftSet *__restrict__ sh, const float2 *__restrict__ vals,
                               const float *__restrict__ only_att, int ns, int otf_nb, double cpm, double cutoff,
                               double vec_cpm, int use_att, float w2) {
#pragma clang fp contract(off)
    const size_t mm = (size_t)M * M;
    // squared distance beyond which every OTF term is zero (with margin);
    // a cutoff or scale that makes it not finite disables the skip
    double far2 = (cutoff / vec_cpm) * (cutoff / vec_cpm) * (1 + 1e-9);
    if (!(far2 < __builtin_inf())) far2 = __builtin_inf();
    const int st = blockIdx.y;
    // read in place (uniform loads): a local copy indexed by the direction
    // loop lives in scratch memory
    const ShiftSet &S = sh[st];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (size_t)gridDim.x * blockDim.x) {
        int x, y;
        tile_major_xy(i, M, cw, x, y);
        const double xh = xh_of(x, M), yh = yh_of(y, M);
        // |OTF|^2 (times the attenuation) at squared distance q from a band
        // centre: zero beyond the cutoff.  q > far2 puts the radius
        // sqrt(q) * vec_cpm above the cutoff by a relative 5e-10, far more
        // than the roundings of the exact path, which would return +0 there
        // too; so those terms skip the FP64 square root, division and table
        // lookups (most of the 2n x 2n grid lies outside a shifted band's
        // support).
        auto term = [&](double q, size_t tb) -> float {
            float v = 0.f;
            if (q <= far2) {
                const double r = sqrt(q) * vec_cpm;
                const float2 o = otf_val(vals + tb, ns, cpm, cutoff, r);
                v = o.x * o.x + o.y * o.y;
                if (use_att) v *= att_val(only_att + tb, cpm, cutoff, r);
            }
            return v;
        };
        // band 0 (px(0) = 0): both distances are |(xh, yh)| in every
        // direction, so one value serves all of them (the same operands as
        // wiener_sum's, hence the same float)
        const float v0 = term(xh * xh + yh * yh, 0);
        float acc = 0.f;
        for (int d = 0; d < S.ndir; d++) {
            acc = acc + v0 + v0;
            for (int b = 1; b < S.nband; b++) {
                const double bx = S.px[d] * b, by = S.py[d] * b;
                const size_t tb = (size_t)(otf_nb > 1 ? b : 0) * ns;
                const float v1 = term((xh - bx) * (xh - bx) + (yh - by) * (yh - by), tb);
                const float v2 = term((xh + bx) * (xh + bx) + (yh + by) * (yh + by), tb);
                acc = acc + v1 + v2;
            }
        }
        wa[(size_t)st * mm + i] = (1 / (acc + w2)) * apo[i];
    }
}

// Optional stage timing: HIP events recorded around each pipeline kernel on
// the launch stream, summed per stage when read back.  The events come from
// a pool created when timing is switched on (creating an event per launch
// would slow the launches being timed); marks beyond the pool are skipped.
struct StageTimer {
    std::mutex mu;
    bool on = false;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
};
static StageTimer g_timer;
static constexpr size_t kTimerEvents = 16384;

// Set while fsim_reconstruct_noalloc launches: the pipeline then only looks
// up existing twiddle tables and records no timing events.
static thread_local bool t_noalloc = false;

bool stage_timing_enabled() { return g_timer.on; }

static void stage_mark(int stage, hipStream_t s, hipEvent_t &start) {
    if (!g_timer.on || t_noalloc) return;
    std::lock_guard<std::mutex> lk(g_timer.mu);
    if (g_timer.used >= g_timer.pool.size()) {
        start = nullptr;
        return;
    }
    hipEvent_t e = g_timer.pool[g_timer.used++];
    (void)hipEventRecord(e, s);
    if (stage >= 0 && start) g_timer.ev.push_back({stage, {start, e}});
    start = e;
}

// Column-pass candidates (elements per thread E0, columns per block CW0; E0 =
// 0: the size's default element count): 0 (0, 8), 1 (16, 8), 2 (16, 16),
// 3 (8, 16), 4 (16, 4), 5 (8, 8); 6 and 7: one column per wavefront
// (WaveColE elements per thread), 8 and 4 columns per block.
// Default columns per block of the column pass: 8, and 4 at 2048^2, where
// two 4-column blocks (without the LDS twiddle table) share a CU and one loads
// while the other transforms (0.58 against 0.64 ms for 8 stacks).
template <int N>
struct DefColCW {
    static constexpr int value = N >= 2048 ? 4 : FSIM_COL_CW;
};

template <int N, int E0, int CW0, bool WC = false>
static void launch_col_shift(const JobList &jl, int njobs, int D, int C, const float2 *twn, hipStream_t s) {
    using CT = ColShiftTile<N, E0, CW0, WC>;
    const int nsh_all = D * (C - 1), total = njobs * nsh_all * (N / CT::cw);
    hipLaunchKernelGGL((k_col_shift<N, E0, CW0, WC>), dim3(FSIM_COL_XCD ? xcd_grid(total) : (unsigned)total),
                       dim3(CT::value), 0, s, jl, nsh_all, total, twn);
}

template <int N>
static void col_candidate_shape(int col, int &e, int &cw) {
    switch (col) {
        case 1: e = ColShiftTile<N, 16, 8>::e; cw = ColShiftTile<N, 16, 8>::cw; break;
        case 2: e = ColShiftTile<N, 16, 16>::e; cw = ColShiftTile<N, 16, 16>::cw; break;
        case 3: e = ColShiftTile<N, 8, 16>::e; cw = ColShiftTile<N, 8, 16>::cw; break;
        case 4: e = ColShiftTile<N, 16, 4>::e; cw = ColShiftTile<N, 16, 4>::cw; break;
        case 5: e = ColShiftTile<N, 8, 8>::e; cw = ColShiftTile<N, 8, 8>::cw; break;
        case 6: e = ColShiftTile<N, 0, 8, true>::e; cw = -ColShiftTile<N, 0, 8, true>::cw; break;
        case 7: e = ColShiftTile<N, 0, 4, true>::e; cw = -ColShiftTile<N, 0, 4, true>::cw; break;
        default: e = ColShiftTile<N, 0, DefColCW<N>::value>::e; cw = ColShiftTile<N, 0, DefColCW<N>::value>::cw; break;
    }
}

template <int N>
static int launch_pipeline_t(const JobList &jl, int njobs, int D, int P, int C, bool half, const PipeCfg &cfg,
                             hipStream_t s) {
    // C: stored components per direction
    constexpr int M = 2 * N;
    const float2 *twn = t_noalloc ? twiddles_cached(N) : twiddles(N);
    const float2 *twm = t_noalloc ? twiddles_cached(M) : twiddles(M);
    if (!twn || !twm)
        return fail(t_noalloc ? FSIM_ERR_STATE : FSIM_ERR_OOM,
                    t_noalloc ? "reconstruct_noalloc: twiddle tables missing (create the channel first)"
                              : "twiddle table allocation failed");
    hipEvent_t ev = nullptr;
    stage_mark(-1, s, ev);
    {
        constexpr int E = FftE<N>::value, T = N / E, RW = T >= 256 ? 1 : 256 / T;
        const int nrb = N / RW, units = 1 + D * (C - 1);
        const int gx = ((nrb + 7) / 8) * 8 * units;
        bool done = false;
        if constexpr (N <= 512) {
            if (cfg.sep != 0 && jl.raw_aligned && D * P <= 9) {
                hipLaunchKernelGGL((k_sep_rows_all<N, 9>), dim3(nrb, njobs), dim3(T * RW), 0, s, jl, (int)half, twn);
                done = true;
            } else if (cfg.sep != 0 && jl.raw_aligned && D * P <= 15) {
                // five phases in three directions
                hipLaunchKernelGGL((k_sep_rows_all<N, 15>), dim3(nrb, njobs), dim3(T * RW), 0, s, jl, (int)half,
                                   twn);
                done = true;
            }
        }
        if (done) {
        } else if (jl.raw_aligned)
            hipLaunchKernelGGL((k_sep_rows1<N, true>), dim3(gx, njobs), dim3(T * RW), 0, s, jl, nrb, units,
                               (int)half, twn);
        else
            hipLaunchKernelGGL((k_sep_rows1<N, false>), dim3(gx, njobs), dim3(T * RW), 0, s, jl, nrb, units,
                               (int)half, twn);
        stage_mark(0, s, ev);
    }
    if (C > 1) {
        switch (cfg.col) {
            case 1: launch_col_shift<N, 16, 8>(jl, njobs, D, C, twn, s); break;
            case 2: launch_col_shift<N, 16, 16>(jl, njobs, D, C, twn, s); break;
            case 3: launch_col_shift<N, 8, 16>(jl, njobs, D, C, twn, s); break;
            case 4: launch_col_shift<N, 16, 4>(jl, njobs, D, C, twn, s); break;
            case 5: launch_col_shift<N, 8, 8>(jl, njobs, D, C, twn, s); break;
            case 6: launch_col_shift<N, 0, 8, true>(jl, njobs, D, C, twn, s); break;
            case 7: launch_col_shift<N, 0, 4, true>(jl, njobs, D, C, twn, s); break;
            default: launch_col_shift<N, 0, DefColCW<N>::value>(jl, njobs, D, C, twn, s); break;
        }
        stage_mark(1, s, ev);
        constexpr int E = FftE<N>::value, T = N / E;
        constexpr int RW = T >= 256 ? 1 : 256 / T;
        hipLaunchKernelGGL(k_row_shift<N>, dim3(D * (C - 1) * (N / RW), njobs), dim3(T * RW), 0, s, jl, twn);
        stage_mark(2, s, ev);
    }
    {
        constexpr int TK = GatherE<N>::tk, CW = GatherE<N>::cw;
        if (jl.gather_after) FSIM_HIP(hipStreamWaitEvent(s, jl.gather_after, 0));
        if (half)
            hipLaunchKernelGGL((k_gather<N, true>), dim3(gather_grid(GatherTiles<N, true>::value, njobs)),
                               dim3(TK * CW), 0, s, jl, njobs, twn, twm);
        else
            hipLaunchKernelGGL((k_gather<N, false>), dim3(gather_grid(GatherTiles<N, false>::value, njobs)),
                               dim3(TK * CW), 0, s, jl, njobs, twn, twm);
        stage_mark(3, s, ev);
    }
    if (half) {
        constexpr int E = FftE<N>::value, T = N / E, RW = T >= 256 ? 1 : 256 / T;
        hipLaunchKernelGGL(k_row_c2r<N>, dim3(M / RW, njobs), dim3(T * RW), 0, s, jl, twm);
        stage_mark(4, s, ev);
    } else {
        constexpr int E = FftE<M>::value, T = M / E, RW = T >= 256 ? 1 : 256 / T;
        hipLaunchKernelGGL(k_row_real<M>, dim3(M / RW, njobs), dim3(T * RW), 0, s, jl, twm);
        stage_mark(4, s, ev);
    }
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

bool recon_size_ok(int n) { return n == 64 || n == 128 || n == 256 || n == 512 || n == 1024 || n == 2048; }

int gather_cw(int n) {
    switch (n) {
        case 64: return GatherE<64>::cw;
        case 128: return GatherE<128>::cw;
        case 256: return GatherE<256>::cw;
        case 512: return GatherE<512>::cw;
        case 1024: return GatherE<1024>::cw;
        case 2048: return GatherE<2048>::cw;
        default: return 1;
    }
}

int col_shape(int n, int col, int *e, int *cw) {
    switch (n) {
        case 64: col_candidate_shape<64>(col, *e, *cw); return FSIM_OK;
        case 128: col_candidate_shape<128>(col, *e, *cw); return FSIM_OK;
        case 256: col_candidate_shape<256>(col, *e, *cw); return FSIM_OK;
        case 512: col_candidate_shape<512>(col, *e, *cw); return FSIM_OK;
        case 1024: col_candidate_shape<1024>(col, *e, *cw); return FSIM_OK;
        case 2048: col_candidate_shape<2048>(col, *e, *cw); return FSIM_OK;
        default: return fail(FSIM_ERR_ARG, "col_shape: unsupported size");
    }
}

bool sep_all_eligible(int n, int D, int P) { return n <= 512 && D * P <= 15; }

int launch_pipeline(int n, const JobList &jl, int njobs, int D, int P, int C, bool half, hipStream_t s) {
    PipeCfg cfg;
    if (!pipe_cfg_lookup(n, D, P, C, half, cfg)) cfg = PipeCfg{};
    return launch_pipeline_cfg(n, jl, njobs, D, P, C, half, cfg, s);
}

// ---------------------------------------------------------------------------
// Persistent work-list kernels.  For each stage, one kernel per launch group
// runs a device list of items (class k, bx, by): item (k, bx, by) is the work
// of block (bx, by) of class k's own launch of that stage, run by the same
// block function.  The blocks (as many as stay resident) claim items
// through cursors and synchronise between items (the LDS is reused); every
// class's run of items starts at a multiple of 8, so an item's position
// modulo 8 is its block's index modulo 8 in the class's launch where that
// grid's x extent is a multiple of 8, and it runs on the same XCD.  The
// lists put the larger size first, so the small items fill the end of the
// launch.
// ---------------------------------------------------------------------------
template <int G>
struct PSizes;
template <>
struct PSizes<0> {
    static constexpr int a = 256, b = 512;
};
template <>
struct PSizes<1> {
    static constexpr int a = 1024, b = 2048;
};
constexpr int cmax(int x, int y) { return x > y ? x : y; }
constexpr int cmin(int x, int y) { return x < y ? x : y; }

int persistent_group(int n) {
    switch (n) {
        case 256: case 512: return 0;
        case 1024: case 2048: return 1;
        default: return -1;
    }
}

// Column pass tile of the persistent kernels: 512-thread blocks in both
// groups (16 columns at n = 256, the same 8-point radix as its 8-column
// default; the default tile elsewhere).
template <int N>
struct PCol {
    static constexpr int e0 = N == 256 ? 8 : 0;
    static constexpr int cw0 = N == 256 ? 16 : DefColCW<N>::value;
    using CT = ColShiftTile<N, e0, cw0, false>;
};

// The item at position i (uniform across the block); false for padding.
__device__ __forceinline__ bool p_item(const int4 *__restrict__ items, int i, int &k, unsigned &bx, unsigned &by) {
    const int4 v = items[i];
    k = __builtin_amdgcn_readfirstlane(v.x);
    bx = (unsigned)__builtin_amdgcn_readfirstlane(v.y);
    by = (unsigned)__builtin_amdgcn_readfirstlane(v.z);
    return k >= 0;
}

// List walking.  Positions p of a list with p % 8 == x belong to XCD x
// (blocks are dealt to the 8 XCDs round robin by block index): the blocks of
// XCD x claim them one at a time through cursor x of the list (claim j is
// position 8 j + x), so an item runs on the XCD of its block in the class's
// own launch, and blocks that start late, behind other work on their CUs,
// take less.  A block claims its next position while it runs the current
// item (thread 0 issues the atomic add, and the block reads the claim from
// LDS after the item's final barrier; two slots, so that a wave still
// reading the previous claim is never overwritten).
struct PWalk {
    int *cur;     // this XCD's cursor
    int *slot;    // two LDS ints
    int x;        // this block's XCD
    int par;      // slot of the next claim
    int p;        // current position (uniform)
};

__device__ __forceinline__ void p_start(PWalk &w, int *cur, int *slot) {
    w.x = (int)(blockIdx.x % 8);
    w.cur = cur + w.x * kPCursorStride;
    w.slot = slot;
    if (threadIdx.x == 0) slot[0] = atomicAdd(w.cur, 1);
    __syncthreads();
    w.p = __builtin_amdgcn_readfirstlane(8 * slot[0] + w.x);
    w.par = 1;
}

// Runs the block's items while its position lies in one list segment
// (positions below end; segments end on multiples of 8, and a block's
// claims grow, so the next segment starts where this one leaves the
// position).  Each segment runs one block function in a loop of its own, so
// the registers of the other segments' functions are not held beside it.
template <class F>
__device__ __forceinline__ void p_walk(const int4 *__restrict__ items, PWalk &w, int end, F &&f) {
    // the segment's twiddle table stays in LDS from its first item on
    bool twk = false;
    while (w.p < end) {
        int k;
        unsigned bx, by;
        const bool ok = p_item(items, w.p, k, bx, by);
        int next = 0;
        if (threadIdx.x == 0) next = atomicAdd(w.cur, 1);
        if (ok) f(k, bx, by, twk);
        if (threadIdx.x == 0) w.slot[w.par] = next;
        __syncthreads();
        w.p = __builtin_amdgcn_readfirstlane(8 * w.slot[w.par] + w.x);
        w.par ^= 1;
    }
}

// Zeroes the list cursors before a run (a kernel, so that a captured graph
// holds no memory set node).
__global__ void k_p_reset(int *cur, int count) {
    for (int i = threadIdx.x; i < count; i += blockDim.x) cur[i] = 0;
}

// The lists have four segments, ends e.x <= e.y <= e.z <= e.w: the larger
// size (variant 0, variant 1), then the smaller.  Variants: separation with
// all units of a row in one wavefront / one unit per block; gather and last
// pass of a Hermitian / full result; the column pass and row shift have
// variant 0 only.
template <int G>
__global__ void __launch_bounds__(256) k_p_sep(const PClass *__restrict__ cls, const int4 *__restrict__ items,
                                               int4 e, int *cur) {
    constexpr int A = PSizes<G>::a, B = PSizes<G>::b;
    __shared__ float2 lds[cmax(RowItemLds<A>::rows, RowItemLds<B>::rows)];
    __shared__ float2 twl[cmax(TwLds<A, SepTw<A>::table>::len, TwLds<B, SepTw<B>::table>::len)];
    __shared__ int slot[2];
    PWalk w;
    p_start(w, cur, slot);
    auto all = [&](auto n) {
        constexpr int N = decltype(n)::value;
        return [&](int k, unsigned bx, unsigned by, bool &twk) {
            const PClass &c = cls[k];
            sep_rows_all_item<N, 9>(c.jl, c.half, c.twn, bx, by, lds, twl, twk);
        };
    };
    auto one = [&](auto n) {
        constexpr int N = decltype(n)::value;
        return [&](int k, unsigned bx, unsigned by, bool &twk) {
            constexpr int E = FftE<N>::value, T = N / E, RW = T >= 256 ? 1 : 256 / T;
            const PClass &c = cls[k];
            sep_rows1_item<N, true>(c.jl, N / RW, 1 + c.D * (c.C - 1), c.half, c.twn, bx, by, lds, twl, twk);
        };
    };
    if constexpr (B <= 512) p_walk(items, w, e.x, all(std::integral_constant<int, B>{}));
    p_walk(items, w, e.y, one(std::integral_constant<int, B>{}));
    if constexpr (A <= 512) p_walk(items, w, e.z, all(std::integral_constant<int, A>{}));
    p_walk(items, w, e.w, one(std::integral_constant<int, A>{}));
}

template <int G>
__global__ void __launch_bounds__(PCol<PSizes<G>::a>::CT::value,
                                  cmin(PCol<PSizes<G>::a>::CT::wpe, PCol<PSizes<G>::b>::CT::wpe))
    k_p_col(const PClass *__restrict__ cls, const int4 *__restrict__ items, int4 e, int *cur) {
    constexpr int A = PSizes<G>::a, B = PSizes<G>::b;
    using CA = typename PCol<A>::CT;
    using CB = typename PCol<B>::CT;
    static_assert(CA::value == CB::value, "one block shape per group");
    __shared__ float2 lds[cmax(CA::lds, CB::lds)];
    __shared__ float2 twl[cmax(TwLds<A, CA::tw_fits>::len, TwLds<B, CB::tw_fits>::len)];
    __shared__ int slot[2];
    PWalk w;
    p_start(w, cur, slot);
    auto col = [&](auto n) {
        constexpr int N = decltype(n)::value;
        return [&](int k, unsigned bx, unsigned, bool &twk) {
            using P = PCol<N>;
            const PClass &c = cls[k];
            const int nsh_all = c.D * (c.C - 1);
            col_shift_item<N, P::e0, P::cw0, false>(c.jl, nsh_all, c.njobs * nsh_all * (N / P::CT::cw), c.twn, bx,
                                                    lds, twl, twk);
        };
    };
    p_walk(items, w, e.y, col(std::integral_constant<int, B>{}));
    p_walk(items, w, e.w, col(std::integral_constant<int, A>{}));
}

template <int G>
__global__ void __launch_bounds__(256) k_p_row_shift(const PClass *__restrict__ cls, const int4 *__restrict__ items,
                                                     int4 e, int *cur) {
    constexpr int A = PSizes<G>::a, B = PSizes<G>::b;
    __shared__ float2 lds[cmax(RowItemLds<A>::rows, RowItemLds<B>::rows)];
    __shared__ float2 twl[cmax(TwLds<A, RowShiftTw<A>::value>::len, TwLds<B, RowShiftTw<B>::value>::len)];
    __shared__ int slot[2];
    PWalk w;
    p_start(w, cur, slot);
    auto rs = [&](auto n) {
        constexpr int N = decltype(n)::value;
        return [&](int k, unsigned bx, unsigned by, bool &twk) { row_shift_item<N>(cls[k].jl, cls[k].twn, bx, by, lds, twl, twk); };
    };
    p_walk(items, w, e.y, rs(std::integral_constant<int, B>{}));
    p_walk(items, w, e.w, rs(std::integral_constant<int, A>{}));
}

template <int G>
__global__ void __launch_bounds__(GatherE<PSizes<G>::a>::tk * GatherE<PSizes<G>::a>::cw,
                                  cmin(GatherE<PSizes<G>::a>::wpe, GatherE<PSizes<G>::b>::wpe))
    k_p_gather(const PClass *__restrict__ cls, const int4 *__restrict__ items, int4 e, int *cur) {
    constexpr int A = PSizes<G>::a, B = PSizes<G>::b;
    static_assert(GatherE<A>::tk * GatherE<A>::cw == GatherE<B>::tk * GatherE<B>::cw, "one block shape per group");
    __shared__ float2 lds[cmax(GatherE<A>::lds, GatherE<B>::lds)];
    __shared__ float2 twl[cmax(GatherTw<A>::len, GatherTw<B>::len)];
    __shared__ int slot[2];
    PWalk w;
    p_start(w, cur, slot);
    auto ga = [&](auto n, auto h) {
        constexpr int N = decltype(n)::value;
        constexpr bool H = decltype(h)::value;
        return [&](int k, unsigned bx, unsigned, bool &twk) {
            const PClass &c = cls[k];
            gather_item<N, H>(c.jl, c.njobs, c.twn, c.twm, bx, lds, twl, twk);
        };
    };
    p_walk(items, w, e.x, ga(std::integral_constant<int, B>{}, std::true_type{}));
    p_walk(items, w, e.y, ga(std::integral_constant<int, B>{}, std::false_type{}));
    p_walk(items, w, e.z, ga(std::integral_constant<int, A>{}, std::true_type{}));
    p_walk(items, w, e.w, ga(std::integral_constant<int, A>{}, std::false_type{}));
}

// LDS of the last pass's items: a complex-to-real item's row regions,
// twiddle table and row-major copy, or a full-width item's row regions,
// carved from one array (an item uses one or the other).
template <int N>
struct PLastLds {
    // the complex-to-real items: the transform regions and the row copy
    // (C2rLds), then the twiddle table
    static constexpr int reg = C2rLds<N>::value, twl = TwLds<2 * N, FSIM_C2R_TW != 0>::len;
    static constexpr int value = cmax(reg + twl, RowItemLds<2 * N>::rows);
};

template <int G>
__global__ void __launch_bounds__(256) k_p_last(const PClass *__restrict__ cls, const int4 *__restrict__ items,
                                                int4 e, int *cur) {
    constexpr int A = PSizes<G>::a, B = PSizes<G>::b;
    __shared__ __align__(16) float2 buf[cmax(PLastLds<A>::value, PLastLds<B>::value)];
    __shared__ int slot[2];
    PWalk w;
    p_start(w, cur, slot);
    auto c2r = [&](auto n) {
        constexpr int N = decltype(n)::value;
        return [&](int k, unsigned bx, unsigned by, bool &twk) {
            using L = PLastLds<N>;
            constexpr int M = 2 * N, E = FftE<N>::value, T = N / E, RW = T >= 256 ? 1 : 256 / T;
            const int grp = (M / RW) % 8 ? (int)bx : xcd_linear_at(bx, M / RW);
            row_c2r_item<N>(cls[k].jl, cls[k].twm, grp, by, buf, buf + L::reg, buf + C2rLds<N>::rows, twk);
        };
    };
    auto real = [&](auto n) {
        constexpr int N = decltype(n)::value;
        return [&](int k, unsigned bx, unsigned by, bool &) { row_real_item<2 * N>(cls[k].jl, cls[k].twm, bx, by, buf); };
    };
    p_walk(items, w, e.x, c2r(std::integral_constant<int, B>{}));
    p_walk(items, w, e.y, real(std::integral_constant<int, B>{}));
    p_walk(items, w, e.z, c2r(std::integral_constant<int, A>{}));
    p_walk(items, w, e.w, real(std::integral_constant<int, A>{}));
}

// Grid (gx, gy) of class c's own launch of a stage (0 x 0: the stage does
// not run for the class).
template <int N>
static void pstage_grid(const PClass &c, int stage, int &gx, int &gy) {
    constexpr int M = 2 * N;
    constexpr int E = FftE<N>::value, T = N / E, RW = T >= 256 ? 1 : 256 / T;
    gx = gy = 0;
    switch (stage) {
        case 0: {
            const int nrb = N / RW, units = 1 + c.D * (c.C - 1);
            gx = c.sep_all ? nrb : ((nrb + 7) / 8) * 8 * units;
            gy = c.njobs;
            break;
        }
        case 1:
            if (c.C > 1) {
                const int total = c.njobs * c.D * (c.C - 1) * (N / PCol<N>::CT::cw);
                gx = FSIM_COL_XCD ? (int)xcd_grid(total) : total;
                gy = 1;
            }
            break;
        case 2:
            if (c.C > 1) {
                gx = c.D * (c.C - 1) * (N / RW);
                gy = c.njobs;
            }
            break;
        case 3:
            gx = (int)gather_grid(c.half ? GatherTiles<N, true>::value : GatherTiles<N, false>::value, c.njobs);
            gy = 1;
            break;
        default:
            if (c.half) {
                gx = M / RW;
            } else {
                constexpr int EM = FftE<M>::value, TM = M / EM, RWM = TM >= 256 ? 1 : 256 / TM;
                gx = M / RWM;
            }
            gy = c.njobs;
            break;
    }
}

// Segment of class c's items in a stage's list (0, 1: the group's larger
// size, variants 0 and 1; 2, 3: the smaller size).
static int pstage_segment(const PClass &c, int stage) {
    const int g = persistent_group(c.n);
    const int big = g == 0 ? 512 : 2048;
    int v = 0;
    if (stage == 0) v = c.sep_all ? 0 : 1;
    if (stage == 3 || stage == 4) v = c.half ? 0 : 1;
    return (c.n == big ? 0 : 2) + v;
}

void persistent_list(const PClass *cls, int ncls, int stage, std::vector<int4> &items, int ends[4]) {
    items.clear();
    for (int seg = 0; seg < 4; seg++) {
        for (int k = 0; k < ncls; k++) {
            const PClass &c = cls[k];
            if (persistent_group(c.n) < 0 || pstage_segment(c, stage) != seg) continue;
            int gx = 0, gy = 0;
            switch (c.n) {
                case 256: pstage_grid<256>(c, stage, gx, gy); break;
                case 512: pstage_grid<512>(c, stage, gx, gy); break;
                case 1024: pstage_grid<1024>(c, stage, gx, gy); break;
                default: pstage_grid<2048>(c, stage, gx, gy); break;
            }
            for (int y = 0; y < gy; y++)
                for (int x = 0; x < gx; x++) items.push_back(make_int4(k, x, y, 0));
            while (items.size() % 8) items.push_back(make_int4(-1, 0, 0, 0));
        }
        ends[seg] = (int)items.size();
    }
}

// Kernel and block size of a group's stage.
static void pstage_kernel(int group, int stage, const void *&fn, int &threads) {
    threads = 256;
    switch (group * kPStages + stage) {
        case 0: fn = (const void *)k_p_sep<0>; break;
        case 1: fn = (const void *)k_p_col<0>; threads = PCol<256>::CT::value; break;
        case 2: fn = (const void *)k_p_row_shift<0>; break;
        case 3: fn = (const void *)k_p_gather<0>; threads = GatherE<256>::tk * GatherE<256>::cw; break;
        case 4: fn = (const void *)k_p_last<0>; break;
        case 5: fn = (const void *)k_p_sep<1>; break;
        case 6: fn = (const void *)k_p_col<1>; threads = PCol<1024>::CT::value; break;
        case 7: fn = (const void *)k_p_row_shift<1>; break;
        case 8: fn = (const void *)k_p_gather<1>; threads = GatherE<1024>::tk * GatherE<1024>::cw; break;
        default: fn = (const void *)k_p_last<1>; break;
    }
}

int launch_persistent_reset(int *cur, int count, hipStream_t s) {
    hipLaunchKernelGGL(k_p_reset, dim3(1), dim3(256), 0, s, cur, count);
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

int launch_persistent(int group, int stage, const PClass *cls_d, const int4 *items_d, const int ends[4], int *cur,
                      hipStream_t s) {
    if (group < 0 || group >= kPGroups || stage < 0 || stage >= kPStages)
        return fail(FSIM_ERR_ARG, "launch_persistent: bad group or stage");
    const int nitems = ends[3];
    if (nitems <= 0) return FSIM_OK;
    const int4 e = make_int4(ends[0], ends[1], ends[2], ends[3]);
    // resident blocks per CU of each kernel, looked up once per device
    static std::mutex mu;
    static int resident[64][kPGroups * kPStages] = {};
    static int cus[64] = {};
    int dev = 0;
    FSIM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(FSIM_ERR_ARG, "launch_persistent: device index out of range");
    const void *fn = nullptr;
    int threads = 256;
    pstage_kernel(group, stage, fn, threads);
    int per_cu, ncu;
    {
        std::lock_guard<std::mutex> lk(mu);
        int &r = resident[dev][group * kPStages + stage];
        if (!cus[dev]) FSIM_HIP(hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev));
        if (!r) {
            FSIM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&r, fn, threads, 0));
            if (r < 1) return fail(FSIM_ERR_HIP, "launch_persistent: kernel does not fit a CU");
        }
        per_cu = r;
        ncu = cus[dev];
    }
    // a multiple of 8 blocks, an equal share per XCD
    const int grid = std::max(8, std::min(per_cu * ncu, nitems) / 8 * 8);
    switch (group * kPStages + stage) {
        case 0: hipLaunchKernelGGL(k_p_sep<0>, dim3(grid), dim3(threads), 0, s, cls_d, items_d, e, cur); break;
        case 1: hipLaunchKernelGGL(k_p_col<0>, dim3(grid), dim3(threads), 0, s, cls_d, items_d, e, cur); break;
        case 2: hipLaunchKernelGGL(k_p_row_shift<0>, dim3(grid), dim3(threads), 0, s, cls_d, items_d, e, cur); break;
        case 3: hipLaunchKernelGGL(k_p_gather<0>, dim3(grid), dim3(threads), 0, s, cls_d, items_d, e, cur); break;
        case 4: hipLaunchKernelGGL(k_p_last<0>, dim3(grid), dim3(threads), 0, s, cls_d, items_d, e, cur); break;
        case 5: hipLaunchKernelGGL(k_p_sep<1>, dim3(grid), dim3(threads), 0, s, cls_d, items_d, e, cur); break;
        case 6: hipLaunchKernelGGL(k_p_col<1>, dim3(grid), dim3(threads), 0, s, cls_d, items_d, e, cur); break;
        case 7: hipLaunchKernelGGL(k_p_row_shift<1>, dim3(grid), dim3(threads), 0, s, cls_d, items_d, e, cur); break;
        case 8: hipLaunchKernelGGL(k_p_gather<1>, dim3(grid), dim3(threads), 0, s, cls_d, items_d, e, cur); break;
        default: hipLaunchKernelGGL(k_p_last<1>, dim3(grid), dim3(threads), 0, s, cls_d, items_d, e, cur); break;
    }
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

int launch_pipeline_cfg(int n, const JobList &jl, int njobs, int D, int P, int C, bool half, const PipeCfg &cfg,
                        hipStream_t s) {
    // C: stored components per direction (ReconConst::Cs)
    if (cfg.col < 0 || cfg.col >= kColCands) return fail(FSIM_ERR_ARG, "pipeline: bad column-pass candidate");
    switch (n) {
        case 64: return launch_pipeline_t<64>(jl, njobs, D, P, C, half, cfg, s);
        case 128: return launch_pipeline_t<128>(jl, njobs, D, P, C, half, cfg, s);
        case 256: return launch_pipeline_t<256>(jl, njobs, D, P, C, half, cfg, s);
        case 512: return launch_pipeline_t<512>(jl, njobs, D, P, C, half, cfg, s);
        case 1024: return launch_pipeline_t<1024>(jl, njobs, D, P, C, half, cfg, s);
        case 2048: return launch_pipeline_t<2048>(jl, njobs, D, P, C, half, cfg, s);
        default: return fail(FSIM_ERR_ARG, "reconstruct: n must be a power of two in [64, 2048]");
    }
}

size_t workspace_a_elems(int n, int D, int C) { return (size_t)D * C * n * n; }
size_t workspace_b_elems(int n) { return (size_t)4 * n * n; }

void chunk_plan(int n, int D, int Cs, int count, const PipeCfg &cfg, int &K, int &S) {
    K = count;
    S = 1;
    int k = cfg.chunk;
    if (k < 0) {
        // default: up to 512^2, the largest power of two of stacks whose work
        // buffers take at most 160 MiB (8 stacks at 512^2, 3 x 3 frames)
        if (n > 512) return;
        const size_t per = (workspace_a_elems(n, D, Cs) + workspace_b_elems(n)) * sizeof(float2);
        k = 1;
        while ((size_t)(2 * k) * per <= ((size_t)160 << 20)) k *= 2;
    }
    if (k <= 0 || k >= count) return;
    const int chunks = (count + k - 1) / k;
    K = k;
    S = std::max(1, std::min({cfg.streams, kMaxChunkStreams, chunks}));
}

// Library streams of chunked launches and their fork / join events; `mu`
// is held while a call enqueues (the events are reused).
struct ChunkStreams {
    std::mutex mu;
    bool made = false, failed = false;
    hipStream_t s[kMaxChunkStreams] = {};
    hipEvent_t fork = nullptr, aux = nullptr;
    hipEvent_t join[kMaxChunkStreams] = {};
};
static ChunkStreams g_chunk;

static bool chunk_streams_ready() {
    if (g_chunk.made) return true;
    if (g_chunk.failed) return false;
    bool ok = hipEventCreateWithFlags(&g_chunk.fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&g_chunk.aux, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; i < kMaxChunkStreams && ok; i++)
        ok = hipStreamCreateWithFlags(&g_chunk.s[i], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&g_chunk.join[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        g_chunk.failed = true;
        return false;
    }
    g_chunk.made = true;
    return true;
}

std::mutex &chunk_mutex() { return g_chunk.mu; }

const hipStream_t *chunk_streams_locked() { return chunk_streams_ready() ? g_chunk.s : nullptr; }

int launch_chunked(int n, const JobList &jl, int count, int D, int P, int Cs, bool half, const PipeCfg &cfg, int K,
                   int S, hipStream_t s, float2 *const *ws_a, float2 *const *ws_b,
                   const std::function<int(hipStream_t)> *gather_input) {
    if (!chunk_streams_ready()) return fail(FSIM_ERR_HIP, "reconstruct: chunk streams could not be created");
    FSIM_HIP(hipEventRecord(g_chunk.fork, s));
    for (int i = 0; i < S; i++) FSIM_HIP(hipStreamWaitEvent(g_chunk.s[i], g_chunk.fork, 0));
    JobList base = jl;
    if (gather_input) {
        const int rc = (*gather_input)(s);
        if (rc) return rc;
        FSIM_HIP(hipEventRecord(g_chunk.aux, s));
        base.gather_after = g_chunk.aux;
    }
    for (int c = 0, first = 0; first < count; c++, first += K) {
        const int i = c % S, kc = std::min(K, count - first);
        JobList cj = base;
        if (jl.jobs) {
            // a device job list: its entries carry their chunk's work buffers
            cj.jobs = jl.jobs + first;
        } else {
            cj.base.raw += (size_t)first * jl.raw_stride;
            cj.base.out += (size_t)first * jl.out_stride;
            cj.base.A = ws_a[i];
            cj.base.B = ws_b[i];
        }
        const int rc = launch_pipeline_cfg(n, cj, kc, D, P, Cs, half, cfg, g_chunk.s[i]);
        if (rc) return rc;
    }
    for (int i = 0; i < S; i++) {
        FSIM_HIP(hipEventRecord(g_chunk.join[i], g_chunk.s[i]));
        FSIM_HIP(hipStreamWaitEvent(s, g_chunk.join[i], 0));
    }
    return FSIM_OK;
}

// ---------------------------------------------------------------------------
// channel set-up: the FilterUpdateThread work (ReconstructionRunner.java:
// 773-828) plus the per-direction constants of the reconstruction loop.
// ---------------------------------------------------------------------------
int build_channel_const(const fsim_recon_params &p, const Otf &o, bool sym, bool otf_real, ReconConst &K) {
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
    K.otf_nb = p.otf_band0 ? 1 : o.nband;   // bands of the N^2 OTF tables the components use
    K.otf_cpm = o.cpm;
    K.cutoff = o.cutoff;
    K.vec_cpm = p.vec_cpm;
    K.otf_tab = p.use_att ? o.vals_att : o.vals;
    K.Cs = sym ? B : C;
    K.half = sym ? 1 : 0;
    K.otf_real = otf_real ? 1 : 0;
    {
        // |xh| > R with R = cutoff / vec_cpm (a relative margin of 1e-9 keeps
        // columns whose radius rounds to the cutoff): otfToVector writes zero
        // for every row of such a column (OtfProvider.java:352)
        const double R = o.cutoff / p.vec_cpm * (1 + 1e-9);
        K.zlo = 1;
        K.zhi = 0;
        if (std::isfinite(R) && R >= 0 && R < n / 2) {
            K.zlo = (int)std::floor(R) + 1;
            K.zhi = n - K.zlo;
        }
    }
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

// Work buffers of one stream's calls on a channel: A/B for `cap` stacks and
// the per-parameter state of fsim_reconstruct_params.  Each stream that
// calls into a channel gets its own, so calls on different streams (the
// reference's concurrent ReconstructionThreads, each with its own vectors,
// ReconstructionRunner.java:243-251, 398-413) never share intermediates.
struct Workspace {
    float2 *A = nullptr, *B = nullptr;
    int cap = 0;
    MultiWork mw;
};

struct fsim_channel_s {
    fsim_recon_params p;
    Otf o;
    bool sym = false;
    bool otf_real = false;
    ReconConst K;
    ReconConst *dK = nullptr;
    float *wa = nullptr;    // (1/Wiener denominator) * apodization, tile-major
    float *apo = nullptr;   // apodization, tile-major
    float2 *otf2d = nullptr;
    float2 *ramp = nullptr;
    double *fade = nullptr;
    std::mutex ws_mu;
    std::map<hipStream_t, Workspace *> ws;
};

const ReconConst *fsim::channel_device_const(const fsim_channel_s *ch) { return ch->dK; }
const ReconConst &fsim::channel_host_const(const fsim_channel_s *ch) { return ch->K; }

static void workspace_release(Workspace *w) {
    dev_free(w->A);
    dev_free(w->B);
    w->A = w->B = nullptr;
    w->cap = 0;
    multiwork_release(w->mw);
}

// Frees every device buffer of the channel (tables and workspaces); the
// caller has synchronised the work that reads them.
static void channel_free(fsim_channel_s *ch) {
    dev_free(ch->dK);
    dev_free(ch->wa);
    dev_free(ch->apo);
    dev_free(ch->otf2d);
    dev_free(ch->ramp);
    dev_free(ch->fade);
    ch->dK = nullptr;
    ch->wa = ch->apo = nullptr;
    ch->otf2d = ch->ramp = nullptr;
    ch->fade = nullptr;
    for (auto &kv : ch->ws) {
        workspace_release(kv.second);
        delete kv.second;
    }
    ch->ws.clear();
}

// ramp tables of a set of stacks' constants: fx/fy of each direction and
// component, computed on the device
static int fill_ramps(float2 *ramp, double *fdev, std::vector<double> &hf, const ReconConst *K, int count, int n,
                      hipStream_t s) {
    const size_t per = (size_t)kMaxDir * kMaxComp * 2;
    hf.assign(per * count, 0.0);
    for (int st = 0; st < count; st++)
        for (int d = 0; d < kMaxDir; d++)
            for (int c = 0; c < kMaxComp; c++) {
                hf[st * per + (d * kMaxComp + c) * 2] = K[st].fx[d][c];
                hf[st * per + (d * kMaxComp + c) * 2 + 1] = K[st].fy[d][c];
            }
    FSIM_HIP(hipMemcpyAsync(fdev, hf.data(), hf.size() * sizeof(double), hipMemcpyHostToDevice, s));
    const size_t total = per * count * n;
    hipLaunchKernelGGL(k_ramps, dim3((unsigned)std::min<size_t>((total + 255) / 256, 65535 * 4)), dim3(256), 0, s,
                       ramp, fdev, n, total);
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

// The calling stream's A/B workspace, grown to `count` stacks.  Growing
// waits for that stream's earlier work (the only work that reads the old
// buffers).
static int ensure_workspace(fsim_channel_s *ch, int count, hipStream_t s, Workspace **out) {
    Workspace *w;
    {
        std::lock_guard<std::mutex> lk(ch->ws_mu);
        auto it = ch->ws.find(s);
        if (it == ch->ws.end()) it = ch->ws.emplace(s, new Workspace()).first;
        w = it->second;
    }
    *out = w;
    if (count <= w->cap) return FSIM_OK;
    const ReconConst &K = ch->K;
    FSIM_HIP(hipStreamSynchronize(s));
    dev_free(w->A);
    dev_free(w->B);
    w->A = w->B = nullptr;
    w->cap = 0;
    if (dev_alloc_n(&w->A, workspace_a_elems(K.n, K.D, K.Cs) * count) ||
        dev_alloc_n(&w->B, workspace_b_elems(K.n) * count)) {
        dev_free(w->A);
        w->A = nullptr;
        return fail(FSIM_ERR_OOM, "reconstruct: workspace allocation failed");
    }
    w->cap = count;
    return FSIM_OK;
}

namespace fsim {

void multiwork_release(MultiWork &w) {
    (void)hipFree(w.dK);
    (void)hipFree(w.wa);
    (void)hipFree(w.ramp);
    (void)hipFree(w.f);
    (void)hipFree(w.sh);
    (void)hipFree(w.jobs);
    if (w.uploaded) (void)hipEventDestroy(w.uploaded);
    w = MultiWork{};
}

size_t multiwork_bytes(int n, int count, int nsets) {
    const size_t M = 2 * (size_t)n;
    return (size_t)nsets * (sizeof(ReconConst) + M * M * sizeof(float) +
                            (size_t)kMaxDir * kMaxComp * 2 * n * sizeof(float2) +
                            (size_t)kMaxDir * kMaxComp * 2 * sizeof(double) + sizeof(ShiftSet)) +
           (size_t)count * sizeof(Job);
}

// Reconstruction of `count` stacks with their own SIM parameters: sp[st *
// sp_step] for stack st (sp_step 0: one set for all).  Constants, Wiener
// tables and ramps are built once per run of equal consecutive sets (the
// timepoints of one fit share them).  A and B: work buffers for `count`
// stacks; w: the state that grows with the sets and jobs.
int reconstruct_params_with(const fsim_channel_s *ch, MultiWork &w, float2 *A, float2 *B, const uint16_t *raw, float *out,
                            int count, const fsim_sim_params *sp, int sp_step, hipStream_t s, const ParamsChunks *pc) {
    const ReconConst &K0 = ch->K;
    const int n = K0.n, M = 2 * n;
    const size_t mm = (size_t)M * M, rper = (size_t)kMaxDir * kMaxComp * 2 * n;
    // the previous call's uploads read the host arrays rewritten below
    if (w.uploaded) FSIM_HIP(hipEventSynchronize(w.uploaded));
    std::vector<int> set_of(count);
    std::vector<int> first;
    for (int st = 0; st < count; st++) {
        const fsim_sim_params *cur = sp + (size_t)st * sp_step;
        const bool same = st > 0 && (sp_step == 0 || std::memcmp(cur, cur - sp_step, sizeof(fsim_sim_params)) == 0);
        if (!same) first.push_back(st);
        set_of[st] = (int)first.size() - 1;
    }
    const int nset = (int)first.size();
    // host constants of every set: the channel's settings with the set's
    // shifts, phases and modulations
    w.hK.resize(nset);
    w.hsh.resize(nset);
    for (int k = 0; k < nset; k++) {
        const fsim_sim_params &q = sp[(size_t)first[k] * sp_step];
        fsim_recon_params p = ch->p;
        for (int d = 0; d < p.ndir; d++) {
            p.px[d] = q.px[d];
            p.py[d] = q.py[d];
            for (int ph = 0; ph < p.nphase; ph++) p.phases[d][ph] = q.phases[d][ph];
            for (int b = 0; b < p.nband; b++) p.modul[d][b] = q.modul[d][b];
        }
        const int rc = build_channel_const(p, ch->o, ch->sym, ch->otf_real, w.hK[k]);
        if (rc) return rc;
        ShiftSet &S = w.hsh[k];
        S = ShiftSet{};
        S.ndir = p.ndir;
        S.nband = p.nband;
        for (int d = 0; d < p.ndir; d++) {
            S.px[d] = p.px[d];
            S.py[d] = p.py[d];
        }
    }
    if (nset > w.cap_sets || count > w.cap_jobs) {
        FSIM_HIP(hipStreamSynchronize(s));
        const int cs = std::max(nset, w.cap_sets), cj = std::max(count, w.cap_jobs);
        hipEvent_t ev = w.uploaded;
        w.uploaded = nullptr;
        std::vector<ReconConst> hK;
        std::vector<ShiftSet> hsh;
        hK.swap(w.hK);
        hsh.swap(w.hsh);
        multiwork_release(w);
        w.uploaded = ev;
        w.hK.swap(hK);
        w.hsh.swap(hsh);
        if (hipMalloc(&w.dK, cs * sizeof(ReconConst)) != hipSuccess ||
            hipMalloc(&w.wa, cs * mm * sizeof(float)) != hipSuccess ||
            hipMalloc(&w.ramp, cs * rper * sizeof(float2)) != hipSuccess ||
            hipMalloc(&w.f, (size_t)cs * kMaxDir * kMaxComp * 2 * sizeof(double)) != hipSuccess ||
            hipMalloc(&w.sh, cs * sizeof(ShiftSet)) != hipSuccess ||
            hipMalloc(&w.jobs, cj * sizeof(Job)) != hipSuccess) {
            (void)hipGetLastError();
            multiwork_release(w);
            return fail(FSIM_ERR_OOM, "reconstruct_params: device allocation failed");
        }
        w.cap_sets = cs;
        w.cap_jobs = cj;
    }
    for (int k = 0; k < nset; k++) {
        ReconConst &K = w.hK[k];
        K.fade_lo = K0.fade_lo;
        K.fade_hi = K0.fade_hi;
        K.fade_px = K0.fade_px;
        K.otf2d = K0.otf2d;
        K.wa = w.wa + k * mm;
        K.ramp = w.ramp + k * rper;
    }
    w.hjobs.resize(count);
    bool aligned = true;
    for (int st = 0; st < count; st++) {
        Job &j = w.hjobs[st];
        j.c = w.dK + set_of[st];
        j.raw = raw + (size_t)st * K0.D * K0.P * n * n;
        if (pc) {
            const int c = st / pc->k, r = st - c * pc->k;
            j.A = pc->a[c % pc->ns] + r * workspace_a_elems(n, K0.D, K0.Cs);
            j.B = pc->b[c % pc->ns] + r * workspace_b_elems(n);
        } else {
            j.A = A + st * workspace_a_elems(n, K0.D, K0.Cs);
            j.B = B + st * workspace_b_elems(n);
        }
        j.out = out + (size_t)st * 4 * n * n;
        aligned = aligned && ((uintptr_t)j.raw & 7) == 0;
    }
    FSIM_HIP(hipMemcpyAsync(w.dK, w.hK.data(), nset * sizeof(ReconConst), hipMemcpyHostToDevice, s));
    FSIM_HIP(hipMemcpyAsync(w.sh, w.hsh.data(), nset * sizeof(ShiftSet), hipMemcpyHostToDevice, s));
    FSIM_HIP(hipMemcpyAsync(w.jobs, w.hjobs.data(), count * sizeof(Job), hipMemcpyHostToDevice, s));
    int rc = fill_ramps(w.ramp, w.f, w.hf, w.hK.data(), nset, n, s);
    if (rc) return rc;
    if (!w.uploaded) FSIM_HIP(hipEventCreateWithFlags(&w.uploaded, hipEventDisableTiming));
    FSIM_HIP(hipEventRecord(w.uploaded, s));
    const Otf &o = ch->o;
    const bool half = K0.half && ((uintptr_t)out & 7) == 0;
    const int cw = gather_cw(n);
    // the half gather reads the tiles of columns [0, n + cw)
    const size_t len = half ? (size_t)(n + cw) * M : mm;
    const std::function<int(hipStream_t)> wiener = [&](hipStream_t ws) {
        hipLaunchKernelGGL(k_wiener_multi, dim3((unsigned)std::min<size_t>((len + 255) / 256, 2048), nset),
                           dim3(256), 0, ws, w.wa, ch->apo, M, cw, len, w.sh, o.vals, o.only_att, o.ns, o.nband,
                           o.cpm, o.cutoff, ch->p.vec_cpm, ch->p.use_att, (float)(ch->p.wiener * ch->p.wiener));
        FSIM_HIP(hipGetLastError());
        return FSIM_OK;
    };
    JobList jl{};
    jl.jobs = w.jobs;
    jl.raw_aligned = aligned;
    // chunked: the Wiener tables (FP64 work) on the caller's stream while
    // the chunks' separation and shift passes (memory work) run
    if (pc)
        return launch_chunked(n, jl, count, K0.D, K0.P, K0.Cs, half, pc->cfg, pc->k, pc->ns, s, nullptr, nullptr,
                              &wiener);
    const int rc2 = wiener(s);
    if (rc2) return rc2;
    return launch_pipeline(n, jl, count, K0.D, K0.P, K0.Cs, half, s);
}

}  // namespace fsim

// Device temporaries of the channel set-up, freed on every exit.
struct ChannelTmp {
    float *wtmp = nullptr, *atmp = nullptr;
    double *fdev = nullptr;
    ~ChannelTmp() {
        dev_free(wtmp);
        dev_free(atmp);
        dev_free(fdev);
    }
};

// Tables of a channel: the FilterUpdateThread work of
// ReconstructionRunner.java:773-828 on the device.  On failure the caller
// frees what was allocated (channel_free).
static int channel_setup(fsim_channel_s *ch, const fsim_recon_params *p, const Otf &o, ChannelTmp &tmp) {
    const int n = p->n, M = 2 * n;
    const size_t mm = (size_t)M * M;
    if (dev_alloc_n(&ch->dK, 1) || dev_alloc_n(&ch->wa, mm) || dev_alloc_n(&ch->apo, mm) ||
        dev_alloc_n(&ch->otf2d, (size_t)o.nband * n * n) || dev_alloc_n(&ch->fade, 2 * (size_t)n) ||
        dev_alloc_n(&ch->ramp, (size_t)kMaxDir * kMaxComp * 2 * n) ||
        dev_alloc_n(&tmp.fdev, (size_t)kMaxDir * kMaxComp * 2) || dev_alloc_n(&tmp.wtmp, mm) ||
        dev_alloc_n(&tmp.atmp, mm))
        return fail(FSIM_ERR_OOM, "channel_create: device allocation failed");
    // the transforms' twiddle tables exist before the first reconstruction
    // (fsim_reconstruct_noalloc only looks them up)
    if (!twiddles(n) || !twiddles(M)) return fail(FSIM_ERR_OOM, "channel_create: twiddle table allocation failed");
    std::vector<double> f(2 * n);
    fade_factors_host(n, p->fade_px, f.data(), f.data() + n);
    hipStream_t s = stream();
    FSIM_HIP(hipMemcpyAsync(ch->fade, f.data(), 2 * n * sizeof(double), hipMemcpyHostToDevice, s));
    ch->K.fade_lo = ch->fade;
    ch->K.fade_hi = ch->fade + n;
    ch->K.fade_px = p->fade_px;
    ch->K.wa = ch->wa;
    std::vector<double> hf;
    int rc = fill_ramps(ch->ramp, tmp.fdev, hf, &ch->K, 1, n, s);
    if (rc) return rc;
    ch->K.ramp = ch->ramp;
    // otfToVector(band, 0, 0, write) per OTF band on the n^2 grid
    const size_t nn = (size_t)n * n;
    const float2 *tab = p->use_att ? o.vals_att : o.vals;
    for (int b = 0; b < o.nband; b++) {
        rc = otf_table_device(ch->otf2d + b * nn, n, n, tab + (size_t)b * o.ns, o.ns, o.cpm, o.cutoff, p->vec_cpm,
                              s);
        if (rc) return rc;
    }
    ch->K.otf2d = ch->otf2d;
    ShiftSet sh{};
    sh.ndir = p->ndir;
    sh.nband = p->nband;
    for (int d = 0; d < p->ndir; d++) {
        sh.px[d] = p->px[d];
        sh.py[d] = p->py[d];
    }
    rc = wiener_device(tmp.wtmp, M, M, o, p->vec_cpm, sh, p->use_att, p->wiener, s);
    if (rc) return rc;
    rc = write_apo_device(tmp.atmp, M, M, p->vec_cpm, o.cutoff, p->apo_bend, p->apo_cut, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pack_wa, dim3(1024), dim3(256), 0, s, ch->wa, ch->apo, tmp.wtmp, tmp.atmp, M, gather_cw(n));
    FSIM_HIP(hipGetLastError());
    FSIM_HIP(hipMemcpyAsync(ch->dK, &ch->K, sizeof(ReconConst), hipMemcpyHostToDevice, s));
    FSIM_HIP(hipStreamSynchronize(s));
    return FSIM_OK;
}

extern "C" {

int fsim_channel_create(const fsim_recon_params *p, fsim_otf oh, fsim_channel *out) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!p || !oh || !out) return fail(FSIM_ERR_ARG, "channel_create: null argument");
    const Otf &o = oh->o;
    // Hermitian pairing needs a real OTF (then conj(OTF) is symmetric in k)
    bool otf_real = true;
    {
        std::vector<float2> hv((size_t)o.nband * o.ns);
        FSIM_HIP(hipMemcpy(hv.data(), p->use_att ? o.vals_att : o.vals, hv.size() * sizeof(float2),
                           hipMemcpyDeviceToHost));
        for (const float2 &z : hv)
            if (z.y != 0.f) otf_real = false;
    }
    const bool sym = p->mode == 0 && otf_real;
    fsim_channel_s *ch = new fsim_channel_s();
    ch->p = *p;
    ch->otf_real = otf_real;
    rc = build_channel_const(*p, o, sym, otf_real, ch->K);
    if (rc) {
        delete ch;
        return rc;
    }
    ch->o = o;
    ch->sym = sym;
    {
        ChannelTmp tmp;
        rc = channel_setup(ch, p, o, tmp);
        if (rc) {
            // the set-up's launches may still read the buffers
            (void)hipStreamSynchronize(stream());
            (void)hipGetLastError();
        }
    }
    if (rc) {
        channel_free(ch);
        delete ch;
        return rc;
    }
    *out = ch;
    return FSIM_OK;
}

int fsim_channel_class(fsim_channel ch, int *n, int *D, int *P, int *Cs, int *half) {
    if (!ch || !n || !D || !P || !Cs || !half) return fail(FSIM_ERR_ARG, "channel_class: bad arguments");
    const ReconConst &K = ch->K;
    *n = K.n;
    *D = K.D;
    *P = K.P;
    *Cs = K.Cs;
    *half = K.half;
    return FSIM_OK;
}

int fsim_channel_destroy(fsim_channel ch) {
    if (!ch) return FSIM_OK;
    // calls on any stream may still read the channel's tables and workspaces
    FSIM_HIP(hipDeviceSynchronize());
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
    JobList jl{};
    jl.jobs = nullptr;
    jl.base.c = ch->dK;
    jl.base.raw = raw;
    jl.base.out = out;
    jl.raw_stride = (size_t)K.D * K.P * n * n;
    jl.a_stride = workspace_a_elems(n, K.D, K.Cs);
    jl.b_stride = workspace_b_elems(n);
    jl.out_stride = (size_t)4 * n * n;
    jl.raw_aligned = ((uintptr_t)raw & 7) == 0;
    // the half-spectrum pass writes output pixel pairs (8-byte stores); an
    // output that is only 4-byte aligned takes the full-width pass
    const bool half = K.half && ((uintptr_t)out & 7) == 0;
    PipeCfg cfg;
    if (!pipe_cfg_lookup(n, K.D, K.P, K.Cs, half, cfg)) cfg = PipeCfg{};
    int kc = count, ns = 1;
    chunk_plan(n, K.D, K.Cs, count, cfg, kc, ns);
    if (kc < count) {
        // chunks on the library streams, each with its own work buffers
        std::lock_guard<std::mutex> lk(chunk_mutex());
        const hipStream_t *cs = chunk_streams_locked();
        if (!cs) return fail(FSIM_ERR_HIP, "reconstruct: chunk streams could not be created");
        float2 *wa[kMaxChunkStreams], *wb[kMaxChunkStreams];
        for (int i = 0; i < ns; i++) {
            Workspace *w = nullptr;
            rc = ensure_workspace(ch, kc, cs[i], &w);
            if (rc) return rc;
            wa[i] = w->A;
            wb[i] = w->B;
        }
        return launch_chunked(n, jl, count, K.D, K.P, K.Cs, half, cfg, kc, ns, s, wa, wb);
    }
    Workspace *w = nullptr;
    rc = ensure_workspace(ch, count, s, &w);
    if (rc) return rc;
    jl.base.A = w->A;
    jl.base.B = w->B;
    return launch_pipeline_cfg(n, jl, count, K.D, K.P, K.Cs, half, cfg, s);
}

int fsim_reconstruct_noalloc(fsim_channel ch, const uint16_t *raw, float *out, int count, void *strm) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!ch || !raw || !out || count < 1) return fail(FSIM_ERR_ARG, "reconstruct_noalloc: bad arguments");
    const ReconConst &K = ch->K;
    const int n = K.n;
    const size_t a_stride = workspace_a_elems(n, K.D, K.Cs), b_stride = workspace_b_elems(n);
    const size_t need = (a_stride + b_stride) * sizeof(float2) * (size_t)count;
    if (need > kNoAllocScratchBytes)
        return fail(FSIM_ERR_ARG, "reconstruct_noalloc: " + std::to_string(count) + " stacks of " +
                                      std::to_string(n) + "^2 need " + std::to_string(need) +
                                      " bytes of work buffers, more than the fixed " +
                                      std::to_string(kNoAllocScratchBytes) + " (use fewer stacks per call)");
    std::string why;
    float2 *ptr = static_cast<float2 *>(noalloc_scratch(&why));
    if (!ptr) return fail(FSIM_ERR_STATE, "reconstruct_noalloc: " + why);
    hipStream_t s = strm ? (hipStream_t)strm : stream();
    JobList jl{};
    jl.jobs = nullptr;
    jl.base.c = ch->dK;
    jl.base.raw = raw;
    jl.base.A = ptr;
    jl.base.B = ptr + a_stride * (size_t)count;
    jl.base.out = out;
    jl.raw_stride = (size_t)K.D * K.P * n * n;
    jl.a_stride = a_stride;
    jl.b_stride = b_stride;
    jl.out_stride = (size_t)4 * n * n;
    jl.raw_aligned = ((uintptr_t)raw & 7) == 0;
    const bool half = K.half && ((uintptr_t)out & 7) == 0;
    t_noalloc = true;
    rc = launch_pipeline(n, jl, count, K.D, K.P, K.Cs, half, s);
    t_noalloc = false;
    return rc;
}

int fsim_reconstruct_params(fsim_channel ch, const uint16_t *raw, float *out, int count, const fsim_sim_params *sp,
                            void *strm) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!ch || !raw || !out || !sp || count < 1) return fail(FSIM_ERR_ARG, "reconstruct_params: bad arguments");
    if (count > 65535) return fail(FSIM_ERR_ARG, "reconstruct_params: at most 65535 stacks per call");
    hipStream_t s = strm ? (hipStream_t)strm : stream();
    const ReconConst &K = ch->K;
    const bool half = K.half && ((uintptr_t)out & 7) == 0;
    PipeCfg cfg;
    if (!pipe_cfg_lookup(K.n, K.D, K.P, K.Cs, half, cfg)) cfg = PipeCfg{};
    int kc = count, ns = 1;
    chunk_plan(K.n, K.D, K.Cs, count, cfg, kc, ns);
    Workspace *w = nullptr;
    if (kc < count) {
        // chunks on the library streams as in fsim_reconstruct, each with its
        // own work buffers; the parameter state stays with the caller's stream
        std::lock_guard<std::mutex> lk(chunk_mutex());
        const hipStream_t *cs = chunk_streams_locked();
        if (!cs) return fail(FSIM_ERR_HIP, "reconstruct_params: chunk streams could not be created");
        ParamsChunks pc;
        pc.k = kc;
        pc.ns = ns;
        pc.cfg = cfg;
        for (int i = 0; i < ns; i++) {
            Workspace *wi = nullptr;
            rc = ensure_workspace(ch, kc, cs[i], &wi);
            if (rc) return rc;
            pc.a[i] = wi->A;
            pc.b[i] = wi->B;
        }
        rc = ensure_workspace(ch, 1, s, &w);
        if (rc) return rc;
        return reconstruct_params_with(ch, w->mw, nullptr, nullptr, raw, out, count, sp, 1, s, &pc);
    }
    rc = ensure_workspace(ch, count, s, &w);
    if (rc) return rc;
    return reconstruct_params_with(ch, w->mw, w->A, w->B, raw, out, count, sp, 1, s);
}

int fsim_stage_timing(int enable) {
    std::lock_guard<std::mutex> lk(g_timer.mu);
    if (enable && g_timer.pool.empty()) {
        g_timer.pool.resize(kTimerEvents, nullptr);
        for (auto &e : g_timer.pool)
            if (hipEventCreate(&e) != hipSuccess) {
                (void)hipGetLastError();
                for (auto &x : g_timer.pool)
                    if (x) (void)hipEventDestroy(x);
                g_timer.pool.clear();
                return fail(FSIM_ERR_HIP, "stage_timing: event creation failed");
            }
    }
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
    for (auto &e : g_timer.ev) {
        FSIM_HIP(hipEventSynchronize(e.second.second));
        float t = 0;
        FSIM_HIP(hipEventElapsedTime(&t, e.second.first, e.second.second));
        ms[e.first] += t;
        launches[e.first] += 1;
    }
    // the pool's events are reused by the next timed launches
    g_timer.ev.clear();
    g_timer.used = 0;
    return FSIM_OK;
}

}  // extern "C"
