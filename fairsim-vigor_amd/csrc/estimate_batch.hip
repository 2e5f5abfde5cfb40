// Batched SIM parameter estimation: SimAlgorithm.estimateParameters
// (SimAlgorithm.java:37-256) for many stacks in one stream-ordered sequence
// of launches, with no host round trip between its steps.
//
// Work items: a group is one (stack, direction); a pair is one correlation of
// band 0 with another band of a group (fitPeak: one pair per group; getPeak:
// one or two).  Per-group and per-pair state (the peak, the current shift and
// search radius, the correlation result) lives on the device, so the argmax
// that chooses the next fitPeak grid runs in a kernel and the host only
// launches.  Each step reuses the arithmetic of the per-call entry points
// (region_pixel, corr_candidate, the FFT kernels), so a batch gives the same
// results as the reference flow driven call by call.
//
// fitPeak's candidate search.  Evaluating all 100 candidates in the
// reference's arithmetic (float-rounded shift phase, MTool polynomial, float
// products) costs about 70 double-precision operations per pixel and
// candidate.  Each round instead first evaluates every candidate's
// correlation A(c) = sum q * e^{i phi_c} (q = b1 conj(b0)) separably, with
// accurate phasors: a 10-frequency DFT along y per column, then 10 along x —
// about 60 operations per pixel for all 100 candidates, fused with the
// inverse column transform that produces the pixels.  The reference value
// J(c) differs from the computed A(c) by at most
//   D = (sum |q| 2^-24 phi_max(x, y) + sum|q| (2e-6 + 64 * 2^-24)) * 1.5
// (2^-24 |phi|: rounding of the phase to float, with the pixel's largest
// phase magnitude over the grid phi_max(x, y) = 2 pi (ax x + ay y) / n,
// ax = |kx| + search, ay = |ky| + search; 2e-6 covers the polynomial and the
// reference's float products with margin; 64 * 2^-24 the float part of the
// approximate sums, kApxArith), so a candidate can
// hold the maximum of |J| only if |A(c)| >= max|A| - 2D.  Those candidates
// (typically a handful) are then evaluated in the reference's arithmetic and
// the first maximum in y-outer, x-inner order is taken among them, which is
// the candidate the full evaluation picks.  When the filter keeps one
// candidate it is the strict maximum and nothing is evaluated: fitPeak's
// phase and magnitude of a grid round are not used by estimateParameters
// (SimAlgorithm.java:114-159) or the live refit.
//
// Device layout for G = count * ndir groups, nn = n * n:
//   X  [G][P][nn] complex: input spectra; after separation slots 0, 1 (, 2)
//      hold components 0, 1 (, 3) in place
//   W  [2][Q][nn] complex: coarse correlation (Q = G: attenuated component 0
//      and fit-band component), then the commonRegion copies of Q pairs
//      (band 0 copies first, then band 1 copies) after the inverse row pass;
//      the column kernels transform their tiles in registers and write
//      nothing back
#include "fsim_internal.h"
#include "fft_core.h"
#include "fft_launch.h"
#include "filters.h"
#include "estimate_dev.h"

#include <cmath>
#include <mutex>
#include <type_traits>
#include <vector>

namespace fsim {

struct EstPair {
    double kx, ky;     // current shift
    double search;     // > 0: fitPeak's 10 x 10 grid +- search around (kx, ky); 0: (kx, ky) alone
    double re, im;     // correlation of the chosen candidate, scaled by 1 / |b0|^2
    int grp;           // group (stack * ndir + direction)
    int slot1;         // slot of band 1 in the group's X
    int bn1;           // OTF band of band 1
    int err;           // shift beyond the vector in commonRegion
    int ring_ok;       // ring_lo / ring_hi valid (|k| positive and finite)
    int ring_lo, ring_hi;   // commonRegion's DC ring keeps squared radii in [ring_lo, ring_hi]
    int nsel;          // candidates evaluated in the reference's arithmetic
    unsigned char sel[kMaxCand];   // their indices yi * 10 + xi, ascending
};

struct EstGroup {
    unsigned long long key;   // locatePeak: best (magnitude, index) key
};

struct EbSep {
    float2 S[3][kMaxPhase];   // rows of components 0, 1 (, 3)
    int nphase, ncomp, fb;    // fb: slot of the fit band's component
    double vec_cpm, att_fwhm;
};

struct EbRegion {
    const float2 *tabs;       // radial OTF tables, band-major
    int ns, otf_nband;
    double cpm, cutoff, vec_cpm;
};

constexpr int kNormSlot = kMaxCand;            // partial slot of |b0|^2 (exact), sum |q| (filter)
constexpr int kWqSlot = kMaxCand + 1;          // filter: sum |q| (ax x + ay y), the phase bound's weight
constexpr int kPairSlots = kMaxCand + 2;

// Clears the groups' locatePeak keys, and finds locatePeak's exclusion as
// an integer bound: sqrt((double)d2) > kmin (Correlation.java:150-152) is a
// non-decreasing test of the integer squared radius d2, true exactly from
// *d2min on (binary search with the same arithmetic; 2 (n/2)^2 + 1 when no
// pixel of the image qualifies).
__global__ void k_eb_init(EstGroup *__restrict__ g, int G, int n, double kmin, int *__restrict__ d2min) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < G) g[i].key = 0ull;
    if (i == 0) {
        int a = 0, b = 2 * (n / 2) * (n / 2) + 1;
        while (a < b) {
            const int m = a + (b - a) / 2;
            if (sqrt((double)m) > kmin)
                b = m;
            else
                a = m + 1;
        }
        *d2min = a;
    }
}

// Row transforms of the estimation's row kernels.  REG: each thread's
// twiddles in registers, loaded once from the exact-rounded table (tw,
// length L) where they fit (L <= 512) -- worth it for a kernel that runs
// several transforms per thread; else sincospi of the first twiddle of
// each butterfly and its powers (one transform per thread: the register
// loads would cost more L2 traffic than they save).
template <int L, bool REG>
struct EbRowTw {
    static constexpr int E = FftE<L>::value;
    static constexpr bool reg = REG && tw_reg_fits<L, E>();
    TwReg<L, E> r;
    __device__ __forceinline__ EbRowTw(const float2 *tw, int t) {
        if constexpr (reg) tw_reg_load<L, E, L>(r, t, glb(tw));
    }
    template <bool INV>
    __device__ __forceinline__ void fft(float2 (&v)[E], float2 *lds, int t, bool active) const {
        if constexpr (reg)
            fft_regs_r<L, E, INV, RowWave<L>::value>(v, lds, 1, t, r, active);
        else
            fft_regs<L, E, INV, RowWave<L>::value, 0>(v, lds, 1, t, nullptr, active);
    }
};

// Input preparation of every frame (sim_gui/ImageControl.java:819-840) and
// the band separation, fused with the forward row transform.  Per row of a
// group: each phase frame's fadeBorderCos (top, bottom, left, right; each
// `float *= double` rounds once, SimUtils.java:162-185), the components the
// flow reads accumulated in BandSeparation.separateBands' phase order
// (BandSeparation.java:131-165, 183-197; S complex, the frames real), then
// each component's row transform.  The
// separation is linear and moves in front of the transform, as in the
// fused reconstruction: the components equal the reference's up to float
// rounding, and NC transforms per row remain instead of P.  X slots 0, 1
// (, 2) of a group = components 0, 1 (, 3).  One row per T = n/E threads.
// LIVE: the live refit's preparation instead (ReconstructionRunner.java:
// 614-617, 661-664): the image times the float vector borderDampen, a vector
// of ones faded by fadeBorderCos -- per pixel
// (float)((float)((float)((float)(1 * lo[y]) * hi[y]) * lo[x]) * hi[x]),
// each factor applied as `float *= double` -- in one float product.
template <int L, bool LIVE, int NC>
__global__ void __launch_bounds__(256) k_eb_load_sep(const float *__restrict__ img, float2 *__restrict__ X, int G,
                                                     EbSep a, const double *__restrict__ lo,
                                                     const double *__restrict__ hi, const float2 *__restrict__ tw) {
    constexpr int E = FftE<L>::value, T = L / E;
    constexpr int RW = T >= 256 ? 1 : 256 / T;
    constexpr int LP = FftLds<L, E>::len;
    constexpr size_t nn = (size_t)L * L;
    __shared__ float2 lds[RW * LP];
    const int t = threadIdx.x % T, rr = threadIdx.x / T;
    const int row = blockIdx.x * RW + rr;   // row of G groups of L rows
    const bool active = row < G * L;
    const int g = row / L, y = row % L;
    const EbRowTw<L, false> rtw(tw, t);
    float2 acc[NC][E];
#pragma unroll
    for (int k = 0; k < NC; k++)
#pragma unroll
        for (int m = 0; m < E; m++) acc[k][m] = make_float2(0.f, 0.f);
    if (active) {
        const float *src = img + (size_t)g * a.nphase * nn + (size_t)y * L;
        const double ly = lo[y], hy = hi[y];
        double lx[E], hx[E];
        bool one[E];   // the pixel lies inside the faded border
#pragma unroll
        for (int m = 0; m < E; m++) {
            lx[m] = lo[t + T * m];
            hx[m] = hi[t + T * m];
            one[m] = ly == 1.0 && hy == 1.0 && lx[m] == 1.0 && hx[m] == 1.0;
        }
        // the next phase's pixels are in flight while this one is used
        float d[E], nxt[E];
#pragma unroll
        for (int m = 0; m < E; m++) nxt[m] = src[t + T * m];
        for (int p = 0; p < a.nphase; p++) {
#pragma unroll
            for (int m = 0; m < E; m++) d[m] = nxt[m];
            if (p + 1 < a.nphase)
#pragma unroll
                for (int m = 0; m < E; m++) nxt[m] = src[(size_t)(p + 1) * nn + t + T * m];
#pragma unroll
            for (int m = 0; m < E; m++) {
                float e;
                if (one[m]) {
                    // every factor is 1.0: each rounding returns the pixel
                    e = d[m];
                } else if constexpr (LIVE) {
                    float b = 1.0f;
                    b = (float)((double)b * ly);
                    b = (float)((double)b * hy);
                    b = (float)((double)b * lx[m]);
                    b = (float)((double)b * hx[m]);
                    e = d[m] * b;
                } else {
                    e = (float)((double)d[m] * ly);
                    e = (float)((double)e * hy);
                    e = (float)((double)e * lx[m]);
                    e = (float)((double)e * hx[m]);
                }
#pragma unroll
                for (int k = 0; k < NC; k++) acc[k][m] = sep_axpy(acc[k][m], a.S[k][p], make_float2(e, 0.f));
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NC; k++) {
        rtw.template fft<false>(acc[k], lds + rr * LP, t, active);
        if (active) {
            float2 *dst = X + ((size_t)g * a.nphase + k) * nn + (size_t)y * L;
#pragma unroll
            for (int m = 0; m < E; m++) dst[t + T * m] = acc[k][m];
        }
    }
}

template <int L, int NC>
static void launch_load_sep_tn(const float *img, float2 *X, int G, const EbSep &sep, const double *lo,
                               const double *hi, bool live, hipStream_t s) {
    constexpr int E = FftE<L>::value, T = L / E;
    constexpr int RW = T >= 256 ? 1 : 256 / T;
    const dim3 grid(cdiv((size_t)G * L, RW)), block(T * RW);
    const float2 *tw = twiddles(L);
    if (live)
        hipLaunchKernelGGL((k_eb_load_sep<L, true, NC>), grid, block, 0, s, img, X, G, sep, lo, hi, tw);
    else
        hipLaunchKernelGGL((k_eb_load_sep<L, false, NC>), grid, block, 0, s, img, X, G, sep, lo, hi, tw);
}

template <int L>
static int launch_load_sep_t(const float *img, float2 *X, int G, const EbSep &sep, const double *lo, const double *hi,
                             bool live, hipStream_t s) {
    if (!twiddles(L)) return fail(FSIM_ERR_OOM, "twiddle table allocation failed");
    if (sep.ncomp == 3)
        launch_load_sep_tn<L, 3>(img, X, G, sep, lo, hi, live, s);
    else
        launch_load_sep_tn<L, 2>(img, X, G, sep, lo, hi, live, s);
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

static int launch_load_sep(int L, const float *img, float2 *X, int G, const EbSep &sep, const double *lo,
                           const double *hi, bool live, hipStream_t s) {
    switch (L) {
        case 16: return launch_load_sep_t<16>(img, X, G, sep, lo, hi, live, s);
        case 32: return launch_load_sep_t<32>(img, X, G, sep, lo, hi, live, s);
        case 64: return launch_load_sep_t<64>(img, X, G, sep, lo, hi, live, s);
        case 128: return launch_load_sep_t<128>(img, X, G, sep, lo, hi, live, s);
        case 256: return launch_load_sep_t<256>(img, X, G, sep, lo, hi, live, s);
        case 512: return launch_load_sep_t<512>(img, X, G, sep, lo, hi, live, s);
        case 1024: return launch_load_sep_t<1024>(img, X, G, sep, lo, hi, live, s);
        case 2048: return launch_load_sep_t<2048>(img, X, G, sep, lo, hi, live, s);
        case 4096: return launch_load_sep_t<4096>(img, X, G, sep, lo, hi, live, s);
        default: return fail(FSIM_ERR_ARG, "estimate: unsupported size");
    }
}

// The fused column kernels (the components' column pass, the coarse
// correlation's peak, and each round's inverse column transform of both bands of a pair
// with the candidate filter or the exact evaluation) use the stand-alone
// column pass's tile width, capped at 512 threads per block so that a
// thread may hold up to 256 VGPRs, and its twiddle source, so that the
// transforms return its values bit for bit.
// NTMAX = 1024 for kernels whose registers fit 128 VGPRs: the full tile
// width (128-byte row segments) of the stand-alone pass.
template <int L, int NTMAX = 512>
struct EbCols {
    static constexpr int E = FftE<L>::value, T = L / E;
    static constexpr int CW0 = ColW<L>::value;
    static constexpr int CW = T * CW0 <= NTMAX ? CW0 : (NTMAX / T >= 1 ? NTMAX / T : 1);
    static constexpr int NT = T * CW, NW = (NT + 63) / 64;
    static constexpr int NTILE = L / CW;
    using TW = TwLds<L, (FftLds<L, E>::len * CW0 + L) * 8 <= 163840>;
    static constexpr int LDS = FftLds<L, E>::len * CW;
};

// Attenuation vector of the coarse correlation (SimAlgorithm.java:84-92:
// writeAttenuationVector(.99, 0.15 * cutoff), OtfProvider.java:503-529) as a
// table over the squared integer radius d2 = dx^2 + dy^2 in [0, n^2 / 2]:
// centred at 0 the per-pixel value depends on d2 alone (attenuation_pixel
// computes the same d2 exactly), so the column kernel reads one float per
// pixel instead of evaluating a double exp.
__global__ void k_eb_att_tab(float *__restrict__ tab, int nd2, double vec_cpm, double fwhm) {
    for (int d2 = blockIdx.x * blockDim.x + threadIdx.x; d2 < nd2; d2 += gridDim.x * blockDim.x)
        tab[d2] = attenuation_d2((double)d2, vec_cpm, .99, fwhm);
}

// The components' column pass fused with the first half of the coarse
// correlation.  Per group and column tile: each component's forward column
// transform (as the stand-alone column pass; k_eb_load_sep transformed the
// rows), written back in place; then the attenuated copies of component 0
// and of the fit band's component (SimAlgorithm.java:84-92: duplicate,
// times(otfAtt)) are taken to real space by their inverse column transforms
// here (the inverse rows follow in k_eb_rows_corr) and go to W[g] and
// W[G + g].  NC: stored components.
template <int L, int NC>
__global__ void __launch_bounds__((EbCols<L, 1024>::NT)) k_eb_cols_sep(float2 *__restrict__ X, float2 *__restrict__ W,
                                                                     int G, EbSep a, const float *__restrict__ att_tab,
                                                                     const float2 *__restrict__ tw) {
    using C = EbCols<L, 1024>;
    constexpr int E = C::E, T = C::T, CW = C::CW, NT = C::NT, NTILE = C::NTILE;
    constexpr size_t nn = (size_t)L * L;
    __shared__ float2 lds[C::LDS];
    __shared__ float2 twl[C::TW::len];
    const int w = xcd_linear(G * NTILE);
    if (w >= G * NTILE) return;   // block-uniform, before any barrier
    const int g = w / NTILE, tile = w % NTILE;
    const float2 *tws = tw_table<C::TW::tl, NT>(twl, tw);
    const int c = threadIdx.x % CW, t = threadIdx.x / CW;
    const int x = tile * CW + c;
    const int fslot = NC == 3 && a.fb == 2 ? 2 : 1;
    float2 *in = X + (size_t)g * a.nphase * nn + x;
    float2 u[E], f[E], nxt[E];
#pragma unroll
    for (int m = 0; m < E; m++) nxt[m] = in[(size_t)(t + T * m) * L];
#pragma unroll
    for (int k = 0; k < NC; k++) {
        float2 v[E];
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = nxt[m];
        // the next component's loads stay in flight across this transform
        if (k + 1 < NC)
#pragma unroll
            for (int m = 0; m < E; m++) nxt[m] = in[(size_t)(k + 1) * nn + (size_t)(t + T * m) * L];
        fft_regs<L, E, false, false, C::TW::tl>(v, lds + c, CW, t, tws);
#pragma unroll
        for (int m = 0; m < E; m++) in[(size_t)k * nn + (size_t)(t + T * m) * L] = v[m];
        if (k == 0)
#pragma unroll
            for (int m = 0; m < E; m++) u[m] = v[m];
        if (k == fslot)
#pragma unroll
            for (int m = 0; m < E; m++) f[m] = v[m];
    }
#pragma unroll
    for (int m = 0; m < E; m++) {
        const int y = t + T * m;
        const int dx = x < L / 2 ? x : x - L, dy = y < L / 2 ? y : y - L;
        const float att = att_tab[dx * dx + dy * dy];
        u[m] = make_float2(u[m].x * att, u[m].y * att);
        f[m] = make_float2(f[m].x * att, f[m].y * att);
    }
    fft_regs<L, E, true, false, C::TW::tl>(u, lds + c, CW, t, tws);
    fft_regs<L, E, true, false, C::TW::tl>(f, lds + c, CW, t, tws);
    constexpr float sc = 1.0f / L;
    float2 *w0 = W + (size_t)g * nn + x, *w1 = W + ((size_t)G + g) * nn + x;
#pragma unroll
    for (int m = 0; m < E; m++) {
        w0[(size_t)(t + T * m) * L] = cscale(u[m], sc);
        w1[(size_t)(t + T * m) * L] = cscale(f[m], sc);
    }
}

// The coarse correlation's row pass: the inverse row transforms of both
// attenuated copies (their columns were transformed by k_eb_cols_sep; scaled
// by 1 / L as the stand-alone row pass scales), c_fb.timesConj(c0)
// (AbstractVectorCplx.java:288-304), and the forward row transform of the
// product (SimAlgorithm.java:94-101), written over c_fb.  One row per
// T = L / E threads; grid (ceil(G L / RW)).
template <int L>
__global__ void __launch_bounds__(256) k_eb_rows_corr(float2 *__restrict__ W, int G, const float2 *__restrict__ tw) {
    constexpr int E = FftE<L>::value, T = L / E;
    constexpr int RW = T >= 256 ? 1 : 256 / T;
    constexpr int LP = FftLds<L, E>::len;
    __shared__ float2 lds[RW * LP];
    constexpr size_t nn = (size_t)L * L;
    const int t = threadIdx.x % T, rr = threadIdx.x / T;
    const int row = blockIdx.x * RW + rr;   // row of G images of L rows
    const bool active = row < G * L;
    float2 *p0 = W + (size_t)row * L, *p1 = W + (size_t)G * nn + (size_t)row * L;
    float2 v[E], u[E];
    if (active) {
#pragma unroll
        for (int m = 0; m < E; m++) {
            v[m] = p0[t + T * m];
            u[m] = p1[t + T * m];
        }
    }
    const EbRowTw<L, true> rtw(tw, t);
    rtw.template fft<true>(v, lds + rr * LP, t, active);
    rtw.template fft<true>(u, lds + rr * LP, t, active);
    constexpr float sc = 1.0f / L;
#pragma unroll
    for (int m = 0; m < E; m++) u[m] = times_conj_java(cscale(u[m], sc), cscale(v[m], sc));
    rtw.template fft<false>(u, lds + rr * LP, t, active);
    if (active) {
#pragma unroll
        for (int m = 0; m < E; m++) p1[t + T * m] = u[m];
    }
}

// The coarse correlation's forward column pass fused with
// Correlation.locatePeak (Correlation.java:144-167): each group's key (the
// magnitude bits and inverted index, peak_key) is the maximum over its
// tiles, one atomicMax per block; the correlation map itself is never
// written.
template <int L>
__global__ void __launch_bounds__((EbCols<L, 1024>::NT)) k_eb_cols_peak(const float2 *__restrict__ C1, int G,
                                                                      const int *__restrict__ d2min,
                                                                      EstGroup *__restrict__ grp,
                                                                      const float2 *__restrict__ tw) {
    using C = EbCols<L, 1024>;
    constexpr int E = C::E, T = C::T, CW = C::CW, NT = C::NT, NW = C::NW, NTILE = C::NTILE;
    constexpr size_t nn = (size_t)L * L;
    __shared__ float2 lds[C::LDS];
    __shared__ float2 twl[C::TW::len];
    __shared__ unsigned long long sh[NW];
    __shared__ double shs[NW];
    const int w = xcd_linear(G * NTILE);
    if (w >= G * NTILE) return;   // block-uniform, before any barrier
    const int g = w / NTILE, tile = w % NTILE;
    const float2 *tws = tw_table<C::TW::tl, NT>(twl, tw);
    const int c = threadIdx.x % CW, t = threadIdx.x / CW;
    const int x = tile * CW + c;
    const float2 *src = C1 + (size_t)g * nn + x;
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; m++) v[m] = src[(size_t)(t + T * m) * L];
    fft_regs<L, E, false, false, C::TW::tl>(v, lds + c, CW, t, tws);
    // peak_key's magnitude (float)sqrt(s), s = re^2 + im^2 in double, is a
    // non-decreasing function of s: the block's largest key has the largest
    // s* among the pixels beyond the exclusion, so the square roots are taken
    // only for pixels whose s lies within a float step of s* (the smallest
    // index among those with the same magnitude wins, as in peak_key)
    const int dm = *d2min;
    const int dx2 = (x < L / 2) ? (x * x) : ((x - L) * (x - L));
    double sv[E];
    bool okv[E];
    double smax = -1.0;
#pragma unroll
    for (int m = 0; m < E; m++) {
        const int y = t + T * m;
        const int dy2 = (y < L / 2) ? (y * y) : ((y - L) * (y - L));
        const float2 cv = v[m];
        sv[m] = (double)cv.x * cv.x + (double)cv.y * cv.y;
        okv[m] = dx2 + dy2 >= dm;
        if (okv[m] && sv[m] > smax) smax = sv[m];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        if (o < NT) {
            const double q = __shfl_down(smax, o, 64);
            if (q > smax) smax = q;
        }
    if (threadIdx.x % 64 == 0) shs[threadIdx.x / 64] = smax;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < NW; i++)
            if (shs[i] > smax) smax = shs[i];
        shs[0] = smax;
    }
    __syncthreads();
    smax = shs[0];
    unsigned long long k = 0;
    if (smax >= 0.0) {
        const float amax = (float)sqrt(smax);
        const double slo = smax * (1 - 0x1p-20);
#pragma unroll
        for (int m = 0; m < E; m++) {
            if (okv[m] && sv[m] >= slo && (float)sqrt(sv[m]) == amax) {
                const unsigned i = (unsigned)((t + T * m) * L + x);
                const unsigned long long key =
                    ((unsigned long long)__float_as_uint(amax) << 32) | (unsigned long long)(0xFFFFFFFFu - i);
                if (key > k) k = key;
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        if (o < NT) {
            const unsigned long long q = __shfl_down(k, o, 64);
            if (q > k) k = q;
        }
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = k;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < NW; i++)
            if (sh[i] > k) k = sh[i];
        if (k) atomicMax(&grp[g].key, k);
    }
}

template <int L>
static int launch_front_t(float2 *X, float2 *W, int G, const EbSep &sep, const float *att_tab, const int *d2min,
                          EstGroup *grp, hipStream_t s) {
    using C = EbCols<L, 1024>;
    const float2 *tw = twiddles(L);
    if (!tw) return fail(FSIM_ERR_OOM, "twiddle table allocation failed");
    const unsigned gc = xcd_grid(G * C::NTILE);
    if (sep.ncomp == 3)
        hipLaunchKernelGGL((k_eb_cols_sep<L, 3>), dim3(gc), dim3(C::NT), 0, s, X, W, G, sep, att_tab, tw);
    else
        hipLaunchKernelGGL((k_eb_cols_sep<L, 2>), dim3(gc), dim3(C::NT), 0, s, X, W, G, sep, att_tab, tw);
    constexpr int E = FftE<L>::value, T = L / E;
    constexpr int RW = T >= 256 ? 1 : 256 / T;
    hipLaunchKernelGGL(k_eb_rows_corr<L>, dim3(cdiv((size_t)G * L, RW)), dim3(T * RW), 0, s, W, G, tw);
    hipLaunchKernelGGL(k_eb_cols_peak<L>, dim3(gc), dim3(C::NT), 0, s, W + (size_t)G * L * L, G, d2min, grp, tw);
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

// The components' column pass, then the coarse correlation to its peak
// (SimAlgorithm.java:84-110): three launches.
static int launch_front(int n, float2 *X, float2 *W, int G, const EbSep &sep, const float *att_tab, const int *d2min,
                        EstGroup *grp, hipStream_t s) {
    switch (n) {
        case 16: return launch_front_t<16>(X, W, G, sep, att_tab, d2min, grp, s);
        case 32: return launch_front_t<32>(X, W, G, sep, att_tab, d2min, grp, s);
        case 64: return launch_front_t<64>(X, W, G, sep, att_tab, d2min, grp, s);
        case 128: return launch_front_t<128>(X, W, G, sep, att_tab, d2min, grp, s);
        case 256: return launch_front_t<256>(X, W, G, sep, att_tab, d2min, grp, s);
        case 512: return launch_front_t<512>(X, W, G, sep, att_tab, d2min, grp, s);
        case 1024: return launch_front_t<1024>(X, W, G, sep, att_tab, d2min, grp, s);
        case 2048: return launch_front_t<2048>(X, W, G, sep, att_tab, d2min, grp, s);
        case 4096: return launch_front_t<4096>(X, W, G, sep, att_tab, d2min, grp, s);
        default: return fail(FSIM_ERR_ARG, "estimate: unsupported size");
    }
}

// locatePeak's result in its coordinate convention (Correlation.java:145-166):
// (-0, -1) start values when no pixel qualifies, then x > w/2 -> x - w,
// y > h/2 -> h - y, else -y.
__device__ __forceinline__ void peak_coords(unsigned long long key, int n, double &xpos, double &ypos) {
    xpos = -0.0;
    ypos = -1;
    if (key) {
        const unsigned idx = 0xFFFFFFFFu - (unsigned)(key & 0xFFFFFFFFull);
        xpos = idx % n;
        ypos = idx / n;
    }
    if (xpos > n / 2) xpos = xpos - n;
    if (ypos > n / 2)
        ypos = n - ypos;
    else
        ypos *= -1;
}

// fitPeak(separate[0], separate[fb], 0, fitBand, otf, -peak[0], -peak[1],
// 0.05, 2.5) (SimAlgorithm.java:114-115): one pair per group.
__global__ void k_eb_fit_init(const EstGroup *__restrict__ grp, EstPair *__restrict__ pairs, int G, int n,
                              int fb_slot, int fit_band, double *__restrict__ peaks) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    double xpos, ypos;
    peak_coords(grp[g].key, n, xpos, ypos);
    peaks[2 * g] = xpos;
    peaks[2 * g + 1] = ypos;
    EstPair p;
    p.kx = -xpos;
    p.ky = -ypos;
    p.search = 2.5;
    p.re = p.im = 0.0;
    p.grp = g;
    p.slot1 = fb_slot;
    p.bn1 = fit_band;
    p.err = 0;
    p.nsel = 0;
    pairs[g] = p;
}

// The DC-centred half of commonRegion's per-pixel work, which is the same
// for every pair and round: for each OTF band, the weight w at (0, 0)
// (otfToVector / writeOtfVector) reduced to what the pixel needs -- NaN
// where |w| < 0.05 (the pixel is zeroed), else the reciprocal (rr, ri)
// that Cplx.div multiplies by (Cplx.java:311-319).  OTF tables are finite
// (fsim_otf_create checks), so a NaN entry only ever marks the cut.
__global__ void k_eb_dc_tab(float2 *__restrict__ dct, int n, EbRegion r, int nb) {
#pragma clang fp contract(off)
    const size_t nn = (size_t)n * n;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn * nb; i += (size_t)gridDim.x * blockDim.x) {
        const int b = (int)(i / nn);
        const size_t k = i % nn;
        const float2 w = otf_pixel(r.tabs + (size_t)b * r.ns, r.ns, r.cpm, r.cutoff, (int)(k % n), (int)(k / n), n, n,
                                   0.0, 0.0, r.vec_cpm);
        float2 e = make_float2(__builtin_nanf(""), __builtin_nanf(""));
        if (!(abs_java(w) < 0.05)) {
            const float s = w.x * w.x + w.y * w.y;
            e = make_float2(w.x / s, -w.y / s);
        }
        dct[i] = e;
    }
}

// region_band0 / region_band1 with the DC weight from the table above;
// `cut`: the pixel lies in the DC ring's cut (ratio < dist or > 1 - dist).
__device__ __forceinline__ float2 region_dc(float2 v, float2 dc, float2 wt, bool cut, double wlim) {
#pragma clang fp contract(off)
    if (dc.x != dc.x || abs_java(wt) < wlim)
        v = make_float2(0.f, 0.f);
    else
        v = make_float2(v.x * dc.x - v.y * dc.y, v.x * dc.y + v.y * dc.x);
    if (cut) v = make_float2(0.f, 0.f);
    return v;
}

// The ring ratio of commonRegion at squared radius d2 (an integer):
// sqrt(d2) / |k|, rounded as the reference rounds it.
__device__ __forceinline__ double ring_ratio(int d2, double kmax) {
    return div_rn(sqrt((double)d2), kmax, 1 / kmax);
}

// Per pair, the squared radii the DC ring keeps.  The ratio is a
// non-decreasing function of d2 (a rounded square root, then a rounded
// division by a positive |k|), so ratio < dist holds exactly for d2 < lo and
// ratio > 1 - dist exactly for d2 > hi; two binary searches over
// d2 in [0, n^2 / 2] find lo and hi with the same arithmetic the per-pixel
// test uses, and the pass then compares integers.  |k| zero or not finite
// keeps the per-pixel test.
__device__ __forceinline__ void eb_ring_pair(EstPair &P, int n, double dist) {
    const double kmax = k_len(P.kx, P.ky);
    if (!(kmax > 0.0 && kmax <= 1e300)) {
        P.ring_ok = 0;
        return;
    }
    const int dmax = 2 * (n / 2) * (n / 2);
    int a = 0, b = dmax + 1;   // first d2 with ratio >= dist
    while (a < b) {
        const int m = a + (b - a) / 2;
        if (ring_ratio(m, kmax) < dist)
            a = m + 1;
        else
            b = m;
    }
    const int lo = a;
    a = 0;
    b = dmax + 1;   // first d2 with ratio > 1 - dist
    while (a < b) {
        const int m = a + (b - a) / 2;
        if (ring_ratio(m, kmax) > 1 - dist)
            b = m;
        else
            a = m + 1;
    }
    P.ring_lo = lo;
    P.ring_hi = a - 1;
    P.ring_ok = 1;
}

// commonRegion's test on the OTF weight centred at the shifted position,
// |w| >= 0.05 (Correlation.java:262-264), depends on a pixel only through
// the squared distance d2 = (xh - kx)^2 + (yh - ky)^2 that the reference
// computes (d2_at): the radius is sqrt(d2) * vec_cpm, then the table index
// floor(radius / cpm), the weight zero at and beyond the cutoff or the
// table's end.  That index -- the key, with ns for zero -- is a
// non-decreasing function of d2 (rounded square root, rounded products and
// quotient by positive constants), so the test's outcome per table sample
// forms runs in d2.  Per OTF band, once per call, k_eb_runs classifies each
// sample (pass: |w| >= the limit whatever the interpolation fraction; fail:
// below it whatever the fraction, or zero; undecided: within a few rounding
// errors of the limit) and finds each run's first d2 by binary search with
// the same arithmetic; the region pass compares d2 with the run starts and
// forms the weight only in undecided runs (every pixel when a table has
// more than kMaxRuns runs).
constexpr int kMaxRuns = 16;
struct EbRuns {
    int n;                    // runs; 0: no classes, the weight at every pixel
    int cls[kMaxRuns];        // 0 fail, 1 pass, 2 undecided
    double start[kMaxRuns];   // first d2 of the run (start[0] = 0)
};

// Table key at squared distance d2 (see above): the sample index the weight
// interpolates from, or ns where the weight is zero.
__device__ __forceinline__ int otf_key(double d2, const EbRegion &r) {
    const double cycl = cycl_of_d2(d2, r.vec_cpm);
    if (cycl > r.cutoff) return r.ns;   // otf_pixel: zero
    float f;
    const int l = otf_index(r.ns, r.cpm, r.cutoff, cycl, f);
    return l < 0 ? r.ns : l;
}

// Class of sample v under the test abs_java(otf_weigh(v, f)) < wlim for
// every fraction f in [0, 1].  otf_weigh returns each part as
// v (1 - f) + v f in float: within 2.5 u of v (u = 2^-24: 1 - f rounds by at
// most u / 2, two products and a sum by u each, relative to |v|); abs_java
// adds u + 2^-52; so |w| lies within 3.6 u of |v|, and a margin of 8 u
// around the limit decides the sample for every f.  Zero samples fail.
__device__ __forceinline__ int sample_class(float2 v, double wlim) {
    const double m = sqrt((double)v.x * v.x + (double)v.y * v.y);
    if (m >= wlim * (1 + 8 * 0x1p-24)) return 1;
    if (m < wlim * (1 - 8 * 0x1p-24)) return 0;
    return 2;
}

// First d2 >= 0 with otf_key(d2) >= l: binary search over the bit patterns
// of non-negative doubles (ordered as the values) up to hi, which bounds
// every pixel's d2; +inf when no d2 up to hi reaches l.
__device__ __forceinline__ double first_d2(int l, const EbRegion &r, double hi) {
    unsigned long long a = 0, b = (unsigned long long)__double_as_longlong(hi) + 1;
    if (otf_key(hi, r) < l) return __longlong_as_double(0x7FF0000000000000ll);
    while (a < b) {
        const unsigned long long m = a + (b - a) / 2;
        if (otf_key(__longlong_as_double((long long)m), r) >= l)
            b = m;
        else
            a = m + 1;
    }
    return __longlong_as_double((long long)a);
}

// The run table of OTF band blockIdx.x (256 threads).  Run boundaries are
// the keys l whose class differs from that of l - 1, found chunk by chunk
// with a block prefix count; each run's first d2 by first_d2.
__global__ void __launch_bounds__(256) k_eb_runs(EbRegion r, int n, double wlim, EbRuns *__restrict__ out) {
    __shared__ int bl[kMaxRuns], bc[kMaxRuns], scan[256];
    __shared__ int cnt, overflow;
    const int b = blockIdx.x, tid = threadIdx.x, ns = r.ns;
    const float2 *tab = r.tabs + (size_t)b * ns;
    const int k0 = otf_key(0.0, r);
    auto cls = [&](int l) { return l >= ns ? 0 : sample_class(tab[l], wlim); };
    if (tid == 0) {
        bl[0] = k0;
        bc[0] = cls(k0);
        cnt = 1;
        overflow = 0;
    }
    __syncthreads();
    for (int l0 = k0 + 1; l0 <= ns; l0 += 256) {
        const int l = l0 + tid;
        int c = 0;
        const int flag = (l <= ns && (c = cls(l)) != cls(l - 1)) ? 1 : 0;
        scan[tid] = flag;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            const int v = tid >= o ? scan[tid - o] : 0;
            __syncthreads();
            scan[tid] += v;
            __syncthreads();
        }
        const int pos = cnt + scan[tid] - flag;
        if (flag) {
            if (pos < kMaxRuns) {
                bl[pos] = l;
                bc[pos] = c;
            } else {
                overflow = 1;
            }
        }
        __syncthreads();
        if (tid == 0) cnt += scan[255];
        __syncthreads();
    }
    const int nr = cnt < kMaxRuns ? cnt : kMaxRuns;
    const double hi = 64.0 * n * n;   // every pixel's d2 is below (1.5 n)^2 * 2
    if (tid < nr && !overflow) {
        out[b].cls[tid] = bc[tid];
        out[b].start[tid] = tid == 0 ? 0.0 : first_d2(bl[tid], r, hi);
    }
    if (tid == 0) out[b].n = overflow ? 0 : nr;
}

// The band value of commonRegion where the DC weight's reciprocal dc is
// valid and the shifted weight passes (region_dc without the cut).
__device__ __forceinline__ float2 region_pass(float2 v, float2 dc) {
#pragma clang fp contract(off)
    return make_float2(v.x * dc.x - v.y * dc.y, v.x * dc.y + v.y * dc.x);
}

// Correlation.commonRegion(b0, b1, 0, bn1, otf, kx, ky, 0.15, 0.05, true)
// on copies of each pair's bands (Correlation.java:60-63, 189-192), fused
// with the row pass of the inverse 2-D FFT that follows (fft2d runs rows
// first).  Grid
// (ceil(n / RW), 2 * npairs):
// blockIdx.y = 2 * pair + band; band 0 rows go to T[pair], band 1 rows to
// T[npairs + pair].
// RUNS: the shifted weight's test through the run tables (k_eb_runs);
// otherwise the weight at every pixel of the common region.
template <int L, bool RUNS>
__global__ void __launch_bounds__(256) k_eb_region_rows(const float2 *__restrict__ X, float2 *__restrict__ T,
                                                        EstPair *__restrict__ pairs, int npairs, int nphase,
                                                        EbRegion r, const float2 *__restrict__ dct,
                                                        const EbRuns *__restrict__ runs,
                                                        const float2 *__restrict__ tw) {
    constexpr int E = FftE<L>::value, Tn = L / E;
    constexpr int RW = Tn >= 256 ? 1 : 256 / Tn;
    constexpr int LP = FftLds<L, E>::len;
    __shared__ float2 lds[RW * LP];
    constexpr size_t nn = (size_t)L * L;
    const int t = threadIdx.x % Tn, rr = threadIdx.x / Tn;
    const int row = blockIdx.x * RW + rr;
    const bool active = row < L;   // small L: a block holds more row slots than rows
    const int p = blockIdx.y >> 1, band = blockIdx.y & 1;
    const double kx = pairs[p].kx, ky = pairs[p].ky;
    const int grp = pairs[p].grp, slot1 = pairs[p].slot1, bn1 = pairs[p].bn1;
    if (blockIdx.x == 0 && band == 0 && threadIdx.x == 0 && (fabs(kx) >= L || fabs(ky) >= L)) pairs[p].err = 1;
    RegionArgs a;
    a.tab0 = r.tabs;
    a.tab1 = r.tabs + (size_t)(r.otf_nband == 1 ? 0 : bn1) * r.ns;
    a.ns = r.ns;
    a.cpm = r.cpm;
    a.cutoff = r.cutoff;
    a.vec_cpm = r.vec_cpm;
    a.kx = kx;
    a.ky = ky;
    a.dist = 0.15;
    a.wlim = 0.05;
    a.ikx = (int)kx;
    a.iky = (int)ky;
    a.divide = 1;
    const double kmax = k_len(kx, ky);
    const bool ring_ok = pairs[p].ring_ok;
    const int ring_lo = pairs[p].ring_lo, ring_hi = pairs[p].ring_hi;
    float2 v[E];
    if (active) {
        const float2 *src = X + ((size_t)grp * nphase + (band ? slot1 : 0)) * nn + (size_t)row * L;
        const float2 *dc = dct + (size_t)(band && r.otf_nband > 1 ? bn1 : 0) * nn + (size_t)row * L;
        const float2 *tab = band ? a.tab1 : a.tab0;
        const double sx = band ? -kx : kx, sy = band ? -ky : ky;
        const EbRuns &R = runs[band ? (r.otf_nband == 1 ? 0 : bn1) : 0];
        const int nrun = RUNS ? R.n : 0;
        // ring: band 0 at its own pixel, band 1 at its source pixel
        // ((x + ikx) mod n, (y - iky) mod n).  The ring test needs integers
        // only, so the band values of the pixels inside the ring are loaded
        // with the DC weights, in one memory round trip.
        bool cuts[E];
        float2 dcs[E], vs[E];
#pragma unroll
        for (int m = 0; m < E; m++) {
            const int x = t + Tn * m;
            const int xq = band ? ((x + a.ikx) % L + L) % L : x, yq = band ? ((row - a.iky) % L + L) % L : row;
            const int dx = xq < L / 2 ? xq : xq - L, dy = yq < L / 2 ? yq : yq - L, d2 = dx * dx + dy * dy;
            cuts[m] = ring_ok ? (d2 < ring_lo || d2 > ring_hi)
                              : (ring_ratio(d2, kmax) < a.dist || ring_ratio(d2, kmax) > 1 - a.dist);
            dcs[m] = dc[x];
            vs[m] = cuts[m] ? make_float2(0.f, 0.f) : src[x];
        }
#pragma unroll
        for (int m = 0; m < E; m++) {
            const int x = t + Tn * m;
            const bool cut = cuts[m];
            // a pixel in the ring's cut or outside the DC weight's support is
            // zero whatever the shifted weight: the weight (FP64 radius and
            // table lookup) is formed only where it matters (a wavefront's
            // lanes hold 64 adjacent pixels, so whole runs outside the
            // common region skip it)
            const float2 dcv = dcs[m];
            float2 out = make_float2(0.f, 0.f);
            if (!cut && dcv.x == dcv.x) {
                const double d2 = d2_at(x, row, L, L, sx, sy);
                int cls = 2;
                if (nrun > 0) {
                    cls = R.cls[0];
                    for (int k = 1; k < nrun; k++)
                        if (d2 >= R.start[k]) cls = R.cls[k];
                }
                if (cls == 2) {
                    const double cycl = cycl_of_d2(d2, a.vec_cpm);
                    const float2 wt = cycl > a.cutoff ? make_float2(0.f, 0.f)
                                                      : otf_val(tab, a.ns, a.cpm, a.cutoff, cycl);
                    out = region_dc(vs[m], dcv, wt, false, a.wlim);
                } else if (cls == 1) {
                    out = region_pass(vs[m], dcv);
                }
            }
            v[m] = out;
        }
    }
    const EbRowTw<L, false> rtw(tw, t);
    rtw.template fft<true>(v, lds + rr * LP, t, active);
    if (active) {
        float2 *dst = T + (size_t)(band ? npairs + p : p) * nn + (size_t)row * L;
        constexpr float scale = 1.0f / L;
#pragma unroll
        for (int m = 0; m < E; m++) dst[t + Tn * m] = cscale(v[m], scale);
    }
}

// Test hook: the weight at every pixel instead of the run tables.
static int g_region_exact = 0;

template <int L>
static int launch_region_rows_t(const EbRegion &r, const float2 *dct, const EbRuns *runs, EstPair *pairs,
                                int npairs, int nphase, float2 *X, float2 *W, hipStream_t s) {
    constexpr int E = FftE<L>::value, Tn = L / E;
    constexpr int RW = Tn >= 256 ? 1 : 256 / Tn;
    const float2 *tw = twiddles(L);
    if (!tw) return fail(FSIM_ERR_OOM, "twiddle table allocation failed");
    if (g_region_exact)
        hipLaunchKernelGGL((k_eb_region_rows<L, false>), dim3(cdiv(L, RW), 2 * npairs), dim3(Tn * RW), 0, s, X, W,
                           pairs, npairs, nphase, r, dct, runs, tw);
    else
        hipLaunchKernelGGL((k_eb_region_rows<L, true>), dim3(cdiv(L, RW), 2 * npairs), dim3(Tn * RW), 0, s, X, W,
                           pairs, npairs, nphase, r, dct, runs, tw);
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

static int launch_region_rows(int n, const EbRegion &r, const float2 *dct, const EbRuns *runs, EstPair *pairs,
                              int npairs, int nphase, float2 *X, float2 *W, hipStream_t s) {
    switch (n) {
        case 16: return launch_region_rows_t<16>(r, dct, runs, pairs, npairs, nphase, X, W, s);
        case 32: return launch_region_rows_t<32>(r, dct, runs, pairs, npairs, nphase, X, W, s);
        case 64: return launch_region_rows_t<64>(r, dct, runs, pairs, npairs, nphase, X, W, s);
        case 128: return launch_region_rows_t<128>(r, dct, runs, pairs, npairs, nphase, X, W, s);
        case 256: return launch_region_rows_t<256>(r, dct, runs, pairs, npairs, nphase, X, W, s);
        case 512: return launch_region_rows_t<512>(r, dct, runs, pairs, npairs, nphase, X, W, s);
        case 1024: return launch_region_rows_t<1024>(r, dct, runs, pairs, npairs, nphase, X, W, s);
        case 2048: return launch_region_rows_t<2048>(r, dct, runs, pairs, npairs, nphase, X, W, s);
        case 4096: return launch_region_rows_t<4096>(r, dct, runs, pairs, npairs, nphase, X, W, s);
        default: return fail(FSIM_ERR_ARG, "estimate: unsupported size");
    }
}

// Inverse column transform of column tile `tile` of both bands of pair p
// (their rows transformed by the region pass; band 0 of pair p at
// W[p], band 1 at W[npairs + p]), as the stand-alone column pass computes
// it, scaled by 1 / L: a[m], b[m] = band 0, band 1 at column tile * CW + c,
// row t + T m.  Every thread of the block calls it (barriers inside).
template <int L>
__device__ __forceinline__ void eb_cols_load(const float2 *__restrict__ W, int npairs, int p, int tile, int c, int t,
                                             float2 (&a)[EbCols<L>::E], float2 (&b)[EbCols<L>::E], float2 *lds,
                                             const float2 *tws) {
    using C = EbCols<L>;
    constexpr size_t nn = (size_t)L * L;
    const float2 *pa = W + (size_t)p * nn + tile * C::CW + c;
    const float2 *pb = W + ((size_t)npairs + p) * nn + tile * C::CW + c;
#pragma unroll
    for (int m = 0; m < C::E; m++) {
        a[m] = pa[(size_t)(t + C::T * m) * L];
        b[m] = pb[(size_t)(t + C::T * m) * L];
    }
    // fft_regs ends with a barrier after its last LDS reads, so the second
    // transform may reuse the region at once
    fft_regs<L, C::E, true, false, C::TW::tl>(a, lds + c, C::CW, t, tws);
    fft_regs<L, C::E, true, false, C::TW::tl>(b, lds + c, C::CW, t, tws);
    constexpr float sc = 1.0f / L;
#pragma unroll
    for (int m = 0; m < C::E; m++) {
        a[m] = cscale(a[m], sc);
        b[m] = cscale(b[m], sc);
    }
}

// Sum of v over the block's NT threads, valid in thread 0; sh holds a
// double2 per wavefront.  Called by all threads (barriers inside).
template <int NT>
__device__ __forceinline__ double2 eb_block_sum(double2 v, double2 *sh) {
    constexpr int NW = (NT + 63) / 64;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        if (o < NT) {
            v.x += __shfl_down(v.x, o, 64);
            v.y += __shfl_down(v.y, o, 64);
        }
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    double2 r = make_double2(0.0, 0.0);
    if (threadIdx.x == 0)
        for (int i = 0; i < NW; i++) {
            r.x += sh[i].x;
            r.y += sh[i].y;
        }
    __syncthreads();
    return r;
}

// Phasor tables of the candidate filter, per pair (kApxTab n float2):
// TY[y][j] = e^{-2 pi i ypos_j y / n} rounded to float (10 n float2), then
// as double2 TXB[x] = e^{+2 pi i xpos_0 x / n} and TXS[x] =
// e^{+2 pi i (search / 4.5) x / n}: the grid's positions are
// pos_0 + i search / 4.5, so TX_i = TXB TXS^i (the shift of b1 by
// (xpos, -ypos)).  Phases reduced mod 1 and evaluated in double.
constexpr int kApxTab = 14;
__device__ __forceinline__ const double2 *apx_txb(const float2 *tabs, int p, int n) {
    return reinterpret_cast<const double2 *>(tabs + (size_t)p * kApxTab * n + 10 * n);
}
__device__ __forceinline__ void eb_tabs_pair(double kx, double ky, double search, int p, int n,
                                             float2 *__restrict__ tabs, int tid, int nt) {
    float2 *t = tabs + (size_t)p * kApxTab * n;
    for (int k = tid; k < 12 * n; k += nt) {
        double pos;
        int v;
        if (k < 10 * n) {
            pos = grid_pos(ky, k / n, search);
            v = k % n;
        } else {
            v = (k - 10 * n) % n;
            pos = k < 11 * n ? grid_pos(kx, 0, search) : search / 4.5;
        }
        double u = pos * v / n;
        u -= floor(u);
        double sn, cs;
        sincospi(2 * u, &sn, &cs);
        if (k < 10 * n)
            t[v * 10 + k / n] = make_float2((float)cs, (float)-sn);
        else
            reinterpret_cast<double2 *>(t + 10 * n)[k - 10 * n] = make_double2(cs, sn);
    }
}

// The ring bounds (thread 0) and, with `tables`, the filter's phasor
// tables of pair p for the round about to run, from its current shift and
// search radius.  Called by the nt threads of a block.
__device__ __forceinline__ void eb_prep_pair(EstPair *__restrict__ pairs, int p, int n, float2 *__restrict__ tabs,
                                             bool tables, double kx, double ky, double search, int tid, int nt) {
    if (tid == 0) eb_ring_pair(pairs[p], n, 0.15);
    if (tables) eb_tabs_pair(kx, ky, search, p, n, tabs, tid, nt);
}

// A round's set-up where no earlier kernel of the flow did it (the first
// fitPeak round and the getPeak round): one block per pair.
__global__ void __launch_bounds__(256) k_eb_prep(EstPair *__restrict__ pairs, int n, float2 *__restrict__ tabs,
                                                 int tables) {
    const int p = blockIdx.x;
    eb_prep_pair(pairs, p, n, tabs, tables != 0, pairs[p].kx, pairs[p].ky, pairs[p].search, threadIdx.x, 256);
}

// Distance of the approximate sums from the exact ones, relative to sum|q|,
// from their float part (u = 2^-24): q = b1 conj(b0) in float (<= 2 sqrt 2 u
// |q|), the float phasor (<= u), a thread's 2 E <= 32 fused multiply-adds per
// part, each rounding the running sum of terms bounded by sqrt 2 |q|
// (<= 2 sqrt 2 E u |q|), then at most 5 float additions across the lanes of
// a column in its wavefront, each rounding a partial sum bounded by
// sqrt 2 sum|q| (<= 5 sqrt 2 u); the rest is double.  Under 57 u for
// E = 16; 64 u.
constexpr double kApxArith = 64 * 0x1p-24;

// Candidates that can hold the maximum (see the top of this file); all 100
// when the bound cannot separate (no positive threshold, or NaN).
// Called by the nt threads of one block per pair; mag: kPairSlots doubles
// of shared memory.
__device__ __forceinline__ void eb_select_pair(EstPair *__restrict__ pairs, int p, const double2 *__restrict__ apart,
                                               int nblk, int n, double *mag, int tid, int nt) {
    const double2 *in = apart + (size_t)p * kPairSlots * nblk;
    for (int t = tid; t < kPairSlots; t += nt) {
        double2 s = make_double2(0.0, 0.0);
        for (int b = 0; b < nblk; b++) {
            s.x += in[(size_t)t * nblk + b].x;
            s.y += in[(size_t)t * nblk + b].y;
        }
        mag[t] = t >= kNormSlot ? s.x : hypot(s.x, s.y);
    }
    __syncthreads();
    if (tid != 0) return;
    double M = 0.0;
    bool bad = false;
    for (int c = 0; c < kMaxCand; c++) {
        if (!(mag[c] >= 0.0)) bad = true;
        if (mag[c] > M) M = mag[c];
    }
    // the phase of pixel (x, y) is at most 2 pi (ax x + ay y) / n in magnitude
    const double phisum = 2 * M_PI * mag[kWqSlot] / n;
    const double delta = 1.5 * (phisum * 0x1p-24 * (1 + 1e-6) + mag[kNormSlot] * (2e-6 + kApxArith));
    const double thr = M - 2 * delta - 1e-12 * M;
    const bool all = bad || !(thr > 0.0) || !(delta == delta);
    int ns = 0;
    for (int c = 0; c < kMaxCand; c++)
        if (all || mag[c] >= thr) pairs[p].sel[ns++] = (unsigned char)c;
    pairs[p].nsel = ns;
}

__global__ void __launch_bounds__(128) k_eb_select(EstPair *__restrict__ pairs, const double2 *__restrict__ apart,
                                                   int nblk, int n) {
    __shared__ double mag[kPairSlots];
    eb_select_pair(pairs, blockIdx.x, apart, nblk, n, mag, threadIdx.x, 128);
}

// A round's inverse column pass fused with the approximate correlation of
// all 100 candidates (the filter described at the top of this file): per
// column x, C_j(x) = sum_y TY_j(y) q(x, y), q = b1 conj(b0), each thread's
// terms in float and then in double; A_tile(i, j) = sum_x TX_i(x) C_j(x)
// over the tile's columns in double, and the tile's sum |q|.  The float part
// stays within kApxArith sum|q| of the exact sums (bound in eb_select_pair).
// Nothing is written back: the exact evaluation transforms the columns
// again.  Grid xcd_grid(npairs * NTILE): neighbouring tiles of a pair run
// on one XCD and share the cache lines of their rows.
// apart[(pair * kPairSlots + c) * NTILE + tile], sum |q| in slot kNormSlot,
// sum |q| (ax x + ay y) in slot kWqSlot.
template <int L>
__global__ void __launch_bounds__(EbCols<L>::NT) k_eb_cols_apx(const float2 *__restrict__ W, int npairs,
                                                               const EstPair *__restrict__ pairs,
                                                               const float2 *__restrict__ tabs,
                                                               const float2 *__restrict__ tw,
                                                               double2 *__restrict__ apart) {
    using C = EbCols<L>;
    constexpr int E = C::E, T = C::T, CW = C::CW, NT = C::NT, NW = C::NW, NTILE = C::NTILE;
    // the pair's phasor table TY (10 L float2) moves into the transforms'
    // LDS region once they are done, where it fits (L <= 512): read from
    // global memory, its ten entries per pixel were the kernel's bottleneck
    constexpr bool TY_LDS = 10 * L * (int)sizeof(float2) <= 65536;
    constexpr int LDSN = TY_LDS && 10 * L > C::LDS ? 10 * L : C::LDS;
    __shared__ __align__(16) float2 lds[LDSN];
    __shared__ float2 twl[C::TW::len];
    // the row sums per wavefront reuse the same region where it fits
    constexpr bool red_in_lds = NW * CW * 10 * sizeof(double2) <= LDSN * sizeof(float2);
    __shared__ double2 red_own[red_in_lds ? 1 : NW * CW * 10];
    double2 *red = red_in_lds ? reinterpret_cast<double2 *>(lds) : red_own;
    __shared__ double2 col[CW][10], txi[CW][10];
    __shared__ double2 ared[NW];
    const int w = xcd_linear(npairs * NTILE);
    if (w >= npairs * NTILE) return;   // block-uniform, before any barrier
    const int p = w / NTILE, tile = w % NTILE;
    // the phase bound's weights: |xpos| <= ax, |ypos| <= ay on the grid
    const float ax = (float)(fabs(pairs[p].kx) + pairs[p].search) * (1 + 0x1p-20f);
    const float ay = (float)(fabs(pairs[p].ky) + pairs[p].search) * (1 + 0x1p-20f);
    const float2 *TY = tabs + (size_t)p * kApxTab * L;
    const double2 *TXB = apx_txb(tabs, p, L), *TXS = TXB + L;
    // the table's loads are issued before the transforms and held in registers
    constexpr int RP = TY_LDS ? (10 * L + NT - 1) / NT : 1;
    float2 typ[RP];
    if constexpr (TY_LDS) {
#pragma unroll
        for (int r = 0; r < RP; r++) {
            const int k = threadIdx.x + r * NT;
            if (k < 10 * L) typ[r] = TY[k];
        }
    }
    const float2 *tws = tw_table<C::TW::tl, NT>(twl, tw);
    const int c = threadIdx.x % CW, t = threadIdx.x / CW;
    float2 a[E], b[E];
    eb_cols_load<L>(W, npairs, p, tile, c, t, a, b, lds, tws);
    if constexpr (TY_LDS) {
        // eb_cols_load ends with a barrier after the transforms' last LDS reads
#pragma unroll
        for (int r = 0; r < RP; r++) {
            const int k = threadIdx.x + r * NT;
            if (k < 10 * L) lds[k] = typ[r];
        }
        __syncthreads();
    }
    auto ty_at = [&](int i) -> float2 {
        if constexpr (TY_LDS)
            return lds[i];
        else
            return TY[i];
    };
    float2 Sf[10];
#pragma unroll
    for (int j = 0; j < 10; j++) Sf[j] = make_float2(0.f, 0.f);
    float aqf = 0.f, wqf = 0.f;
    const float bx = ax * (float)(tile * CW + c);
#pragma unroll
    for (int m = 0; m < E; m++) {
        const int y = t + T * m;
        const float2 u = a[m], v = b[m];
        const float qr = v.x * u.x + v.y * u.y;
        const float qi = v.y * u.x - v.x * u.y;
        // |q| for the bound's sums: a few ulps, far inside its factor 1.5
        const float aqp = __builtin_amdgcn_sqrtf(qr * qr + qi * qi);
        aqf += aqp;
        wqf = fmaf(aqp, fmaf(ay, (float)y, bx), wqf);
#pragma unroll
        for (int j = 0; j < 10; j++) {
            // q e accumulated by four fused multiply-adds
            const float2 e = ty_at(y * 10 + j);
            Sf[j].x = fmaf(qr, e.x, Sf[j].x);
            Sf[j].x = fmaf(-qi, e.y, Sf[j].x);
            Sf[j].y = fmaf(qr, e.y, Sf[j].y);
            Sf[j].y = fmaf(qi, e.x, Sf[j].y);
        }
    }
    // the table's last reads come before the row sums overwrite the region
    if constexpr (TY_LDS && red_in_lds) __syncthreads();
    // sum over the rows: the lanes of one column within the wavefront
    // (lane = t' CW + c) in float (half the cross-lane moves of double
    // sums; the bound's float part covers them), then over the wavefronts
    // in double
#pragma unroll
    for (int o = CW; o < 64; o <<= 1)
        if (o < NT)
#pragma unroll
            for (int j = 0; j < 10; j++) {
                Sf[j].x += __shfl_xor(Sf[j].x, o, 64);
                Sf[j].y += __shfl_xor(Sf[j].y, o, 64);
            }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        if (o < NT) {
            aqf += __shfl_down(aqf, o, 64);
            wqf += __shfl_down(wqf, o, 64);
        }
    double2 S[10];
#pragma unroll
    for (int j = 0; j < 10; j++) S[j] = make_double2((double)Sf[j].x, (double)Sf[j].y);
    const double aq = (double)aqf, wq = (double)wqf;
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
    if (lane < CW)
#pragma unroll
        for (int j = 0; j < 10; j++) red[(wv * CW + lane) * 10 + j] = S[j];
    if (lane == 0) ared[wv] = make_double2(aq, wq);
    __syncthreads();
    for (int k = threadIdx.x; k < CW * 10; k += NT) {
        const int cc = k / 10, j = k % 10;
        double2 s = make_double2(0.0, 0.0);
        for (int i = 0; i < NW; i++) {
            s.x += red[(i * CW + cc) * 10 + j].x;
            s.y += red[(i * CW + cc) * 10 + j].y;
        }
        col[cc][j] = s;
        // TX_j at this column (here j counts x candidates): TXB TXS^j
        const int x = tile * CW + cc;
        const double2 e = TXB[x], st = TXS[x];
        double zr = e.x, zi = e.y;
        for (int r = 0; r < j; r++) {
            const double nr = zr * st.x - zi * st.y;
            zi = zr * st.y + zi * st.x;
            zr = nr;
        }
        txi[cc][j] = make_double2(zr, zi);
    }
    __syncthreads();
    double2 *out = apart + (size_t)p * kPairSlots * NTILE;
    for (int k = threadIdx.x; k < kMaxCand; k += NT) {
        const int i = k % 10, j = k / 10;
        double2 s = make_double2(0.0, 0.0);
        for (int cc = 0; cc < CW; cc++) {
            const double2 e = txi[cc][i], R = col[cc][j];
            s.x += R.x * e.x - R.y * e.y;
            s.y += R.x * e.y + R.y * e.x;
        }
        out[(size_t)k * NTILE + tile] = s;
    }
    if (threadIdx.x == 0) {
        double s = 0.0, sw = 0.0;
        for (int i = 0; i < NW; i++) {
            s += ared[i].x;
            sw += ared[i].y;
        }
        out[(size_t)kNormSlot * NTILE + tile] = make_double2(s, 0.0);
        out[(size_t)kWqSlot * NTILE + tile] = make_double2(sw, 0.0);
    }
}

// Per pair: selected candidate sums in tile order, corr = sum * (1 / |b0|^2)
// as Cplx.mult with a real Cplx.Double; for a grid, the first maximum of
// hypot(corr) in y-outer, x-inner order becomes the new shift and the search
// radius shrinks by 5 (Correlation.java:99-131).
// Called by the nt threads of one block per pair (uniform branches only);
// sums: kPairSlots double2 of shared memory.  The pair's new state is left
// in pairs[p], written by thread 0.
__device__ __forceinline__ void eb_step_pair(EstPair *__restrict__ pairs, int p, const double2 *__restrict__ part,
                                             int nchunk, double2 *sums, int tid, int nt) {
    const double search = pairs[p].search;
    const int nsel = pairs[p].nsel;
    if (search > 0 && nsel == 1) {
        // the filter kept one candidate: it is the maximum (no sums made)
        if (tid == 0) {
            const int cc = pairs[p].sel[0];
            pairs[p].kx = grid_pos(pairs[p].kx, cc % 10, search);
            pairs[p].ky = grid_pos(pairs[p].ky, cc / 10, search);
            pairs[p].search = search / 5;
            pairs[p].re = pairs[p].im = 0.0;
        }
        return;
    }
    const double2 *in = part + (size_t)p * kPairSlots * nchunk;
    for (int t = tid; t <= nsel; t += nt) {
        const int c = t < nsel ? pairs[p].sel[t] : kNormSlot;
        double2 s = make_double2(0.0, 0.0);
        for (int k = 0; k < nchunk; k++) {
            s.x += in[(size_t)c * nchunk + k].x;
            s.y += in[(size_t)c * nchunk + k].y;
        }
        sums[c] = s;
    }
    __syncthreads();
    if (tid != 0) return;
    const double scal = 1. / sums[kNormSlot].x;
    if (search <= 0) {
        pairs[p].re = sums[0].x * scal - sums[0].y * 0.0;
        pairs[p].im = sums[0].x * 0.0 + sums[0].y * scal;
        return;
    }
    const double kx = pairs[p].kx, ky = pairs[p].ky;
    double max = 0, newKx = 0, newKy = 0, bre = 0, bim = 0;
    for (int k = 0; k < nsel; k++) {
        const int cc = pairs[p].sel[k], xi = cc % 10, yi = cc / 10;
        const double2 s = sums[cc];
        const double re = s.x * scal - s.y * 0.0, im = s.x * 0.0 + s.y * scal;
        const double mag = hypot(re, im);
        if (mag > max) {
            max = mag;
            newKx = grid_pos(kx, xi, search);
            newKy = grid_pos(ky, yi, search);
            bre = re;
            bim = im;
        }
    }
    pairs[p].kx = newKx;
    pairs[p].ky = newKy;
    pairs[p].search = search / 5;
    pairs[p].re = bre;
    pairs[p].im = bim;
}

// The step of a round, one block per pair; with prep_next also the next
// fitPeak round's set-up at the new shift (ring bounds, phasor tables).
__global__ void __launch_bounds__(256) k_eb_step(EstPair *__restrict__ pairs, const double2 *__restrict__ part,
                                                 int nchunk, int n, float2 *__restrict__ tabs, int prep_next) {
    __shared__ double2 sums[kPairSlots];
    __shared__ double nk[3];
    const int p = blockIdx.x;
    eb_step_pair(pairs, p, part, nchunk, sums, threadIdx.x, 256);
    if (!prep_next) return;
    __syncthreads();
    if (threadIdx.x == 0) {
        nk[0] = pairs[p].kx;
        nk[1] = pairs[p].ky;
        nk[2] = pairs[p].search;
    }
    __syncthreads();
    eb_prep_pair(pairs, p, n, tabs, true, nk[0], nk[1], nk[2], threadIdx.x, 256);
}

// The exact evaluation of a round: the inverse column pass again (the same
// values), then |b0|^2 and each selected candidate in the reference's
// arithmetic (corr_pixel), summed per tile;
// part[(pair * kPairSlots + c) * NTILE + tile].  Candidate c = yi * 10 + xi
// of the grid, or (kx, ky) alone.  A fitPeak round whose filter kept a
// single candidate needs no evaluation -- that candidate is the strict
// maximum, and a grid round's phase and magnitude are not used -- so such
// pairs return at once (eb_step_pair takes the candidate's position).
template <int L>
__global__ void __launch_bounds__(EbCols<L>::NT) k_eb_cols_corr(const float2 *__restrict__ W, int npairs,
                                                                const EstPair *__restrict__ pairs,
                                                                const float2 *__restrict__ tw,
                                                                double2 *__restrict__ part) {
    using C = EbCols<L>;
    constexpr int E = C::E, T = C::T, CW = C::CW, NT = C::NT, NTILE = C::NTILE;
    __shared__ float2 lds[C::LDS];
    __shared__ float2 twl[C::TW::len];
    __shared__ double2 sh[C::NW];
    const int w = xcd_linear(npairs * NTILE);
    if (w >= npairs * NTILE) return;   // block-uniform, before any barrier
    const int p = w / NTILE, tile = w % NTILE;
    const double kx = pairs[p].kx, ky = pairs[p].ky, search = pairs[p].search;
    const int nsel = pairs[p].nsel;
    if (search > 0 && nsel == 1) return;   // block-uniform
    {
        const float2 *tws = tw_table<C::TW::tl, NT>(twl, tw);
        const int c = threadIdx.x % CW, t = threadIdx.x / CW;
        const int x = tile * CW + c;
        float2 a[E], b[E];
        eb_cols_load<L>(W, npairs, p, tile, c, t, a, b, lds, tws);
        double2 *out = part + (size_t)p * kPairSlots * NTILE;
        double n2 = 0.0;
#pragma unroll
        for (int m = 0; m < E; m++) n2 = corr_norm2_add(n2, a[m]);
        const double2 s0 = eb_block_sum<NT>(make_double2(n2, 0.0), sh);
        if (threadIdx.x == 0) out[(size_t)kNormSlot * NTILE + tile] = s0;
        constexpr double inv_n = 1.0 / L;
        for (int k = 0; k < nsel; k++) {
            const int cc = pairs[p].sel[k];
            const double xpos = search > 0 ? grid_pos(kx, cc % 10, search) : kx;
            const double ypos = search > 0 ? grid_pos(ky, cc / 10, search) : ky;
            double2 acc = make_double2(0.0, 0.0);
#pragma unroll
            for (int m = 0; m < E; m++) {
                const float2 q = corr_pixel(a[m], b[m], x, t + T * m, xpos, ypos, inv_n);
                acc.x += (double)q.x;
                acc.y += (double)q.y;
            }
            const double2 s = eb_block_sum<NT>(acc, sh);
            if (threadIdx.x == 0) out[(size_t)cc * NTILE + tile] = s;
        }
    }
}

#define FSIM_EB_SWITCH(n, FN, ...)                                            \
    switch (n) {                                                              \
        case 16: return FN<16>(__VA_ARGS__);                                  \
        case 32: return FN<32>(__VA_ARGS__);                                  \
        case 64: return FN<64>(__VA_ARGS__);                                  \
        case 128: return FN<128>(__VA_ARGS__);                                \
        case 256: return FN<256>(__VA_ARGS__);                                \
        case 512: return FN<512>(__VA_ARGS__);                                \
        case 1024: return FN<1024>(__VA_ARGS__);                              \
        case 2048: return FN<2048>(__VA_ARGS__);                              \
        case 4096: return FN<4096>(__VA_ARGS__);                              \
        default: return fail(FSIM_ERR_ARG, "estimate: unsupported size");     \
    }

// Column tiles per image of the fused column kernels (partial sums per pair
// and candidate); 0 for an unsupported size.
static int eb_ntile(int n) {
    switch (n) {
        case 16: return EbCols<16>::NTILE;
        case 32: return EbCols<32>::NTILE;
        case 64: return EbCols<64>::NTILE;
        case 128: return EbCols<128>::NTILE;
        case 256: return EbCols<256>::NTILE;
        case 512: return EbCols<512>::NTILE;
        case 1024: return EbCols<1024>::NTILE;
        case 2048: return EbCols<2048>::NTILE;
        case 4096: return EbCols<4096>::NTILE;
        default: return 0;
    }
}

template <int L>
static int launch_cols_apx_t(const float2 *W, int npairs, const EstPair *pairs, const float2 *tabs, double2 *apart,
                             hipStream_t s) {
    using C = EbCols<L>;
    const float2 *tw = twiddles(L);
    if (!tw) return fail(FSIM_ERR_OOM, "twiddle table allocation failed");
    hipLaunchKernelGGL(k_eb_cols_apx<L>, dim3(xcd_grid(npairs * C::NTILE)), dim3(C::NT), 0, s, W, npairs, pairs, tabs,
                       tw, apart);
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}
static int launch_cols_apx(int n, const float2 *W, int npairs, const EstPair *pairs, const float2 *tabs,
                           double2 *apart, hipStream_t s) {
    FSIM_EB_SWITCH(n, launch_cols_apx_t, W, npairs, pairs, tabs, apart, s)
}

template <int L>
static int launch_cols_corr_t(const float2 *W, int npairs, const EstPair *pairs, double2 *part, hipStream_t s) {
    using C = EbCols<L>;
    const float2 *tw = twiddles(L);
    if (!tw) return fail(FSIM_ERR_OOM, "twiddle table allocation failed");
    hipLaunchKernelGGL(k_eb_cols_corr<L>, dim3(xcd_grid(npairs * C::NTILE)), dim3(C::NT), 0, s, W, npairs, pairs,
                       tw, part);
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}
static int launch_cols_corr(int n, const float2 *W, int npairs, const EstPair *pairs, double2 *part, hipStream_t s) {
    FSIM_EB_SWITCH(n, launch_cols_corr_t, W, npairs, pairs, part, s)
}


// getPeak pairs from the fitted peak (SimAlgorithm.java:118-159): two bands,
// getPeak(sep0, sep1, 0, 1, peak); three bands, the peak doubled when fitted
// on band 1, then getPeak(sep0, sep_lb, 0, 1, peak / 2) and
// getPeak(sep0, sep_hb, 0, 2, peak).  live: the refit's single getPeak of
// band 1 at peak / (bands - 1) (ReconstructionRunner.java:722-724; its
// second getPeak only feeds a trace line).
__global__ void k_eb_gp_init(const EstPair *__restrict__ fit, EstPair *__restrict__ gp, int G, int nband,
                             int fit_band, int live, double *__restrict__ shift) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    double kx = fit[g].kx, ky = fit[g].ky;
    EstPair p;
    p.search = 0;
    p.re = p.im = 0.0;
    p.grp = g;
    p.err = 0;
    p.nsel = 1;
    p.sel[0] = 0;
    if (live) {
        p.kx = kx / (nband - 1);
        p.ky = ky / (nband - 1);
        p.slot1 = 1;
        p.bn1 = 1;
        gp[g] = p;
    } else if (nband == 3) {
        if (fit_band == 1) {
            kx *= 2;
            ky *= 2;
        }
        p.kx = kx / 2;
        p.ky = ky / 2;
        p.slot1 = 1;
        p.bn1 = 1;
        gp[2 * g] = p;
        p.kx = kx;
        p.ky = ky;
        p.slot1 = 2;
        p.bn1 = 2;
        gp[2 * g + 1] = p;
    } else {
        p.kx = kx;
        p.ky = ky;
        p.slot1 = 1;
        p.bn1 = 1;
        gp[g] = p;
    }
    shift[2 * g] = -kx;
    shift[2 * g + 1] = -ky;
}

// Per-group results: setPxPy(-peak), setPhaOff(phase of p1), modulations |p1|
// (, |p2|) as Cplx.Double.abs = hypot (Cplx.java:49).
__global__ void k_eb_finish(const EstPair *__restrict__ fit, const EstPair *__restrict__ gp, int G, int npk,
                            int live, const double *__restrict__ peaks, const double *__restrict__ shift,
                            fsim_estimate_result *__restrict__ out, int *__restrict__ err) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    const EstPair &p1 = gp[(size_t)g * npk];
    fsim_estimate_result r;
    r.px = shift[2 * g];
    r.py = shift[2 * g + 1];
    r.pha_off = atan2(p1.im, p1.re);
    r.mod1 = hypot(p1.re, p1.im);
    r.mod2 = live ? __builtin_nan("") : 0.0;
    int e = fit[g].err | p1.err;
    if (npk == 2) {
        const EstPair &p2 = gp[(size_t)g * npk + 1];
        r.mod2 = hypot(p2.re, p2.im);
        e |= p2.err;
    }
    r.peak_x = peaks[2 * g];
    r.peak_y = peaks[2 * g + 1];
    out[g] = r;
    err[g] = e;
}

// Device workspace of the batched estimation, grown on demand; one batch at
// a time.
struct EbWork {
    std::mutex mu;
    float2 *X = nullptr, *W = nullptr;
    size_t x_bytes = 0, w_bytes = 0;
    void *small = nullptr;   // pairs, groups, partial sums, results
    size_t small_bytes = 0;
    std::vector<double> fade_h;
};
static EbWork g_eb;

static int eb_grow(void **p, size_t &have, size_t need) {
    if (need <= have) return FSIM_OK;
    (void)hipFree(*p);
    *p = nullptr;
    have = 0;
    if (hipMalloc(p, need) != hipSuccess) return fail(FSIM_ERR_OOM, "estimate: workspace allocation failed");
    have = need;
    return FSIM_OK;
}

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// One commonRegion / inverse FFT / correlation round for npairs pairs:
// (prep: the ring bounds and, with `grid`, the filter's phasor tables --
// otherwise the previous round's step made them) commonRegion fused with
// the inverse row pass, then with `grid` (fitPeak's 10 x 10 candidates) the
// inverse column pass fused with the candidate filter, and the selection;
// the exact evaluation of what the filter kept (the inverse column pass
// once more); the step, with prep_next also the next round's set-up.
// tabs: kApxTab n float2 of phasors per pair; apart, part: partial sums of
// eb_ntile(n) tiles.
static int eb_round(const EbRegion &r, const float2 *dct, const EbRuns *runs, EstPair *pairs, int npairs, bool grid,
                    int nphase, int n, double2 *part, float2 *tabs, double2 *apart, bool prep, bool prep_next,
                    hipStream_t s) {
    const int ntile = eb_ntile(n);
    if (prep) {
        hipLaunchKernelGGL(k_eb_prep, dim3(npairs), dim3(256), 0, s, pairs, n, tabs, grid ? 1 : 0);
        FSIM_HIP(hipGetLastError());
    }
    int rc = launch_region_rows(n, r, dct, runs, pairs, npairs, nphase, g_eb.X, g_eb.W, s);
    if (rc) return rc;
    if (grid) {
        rc = launch_cols_apx(n, g_eb.W, npairs, pairs, tabs, apart, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_eb_select, dim3(npairs), dim3(128), 0, s, pairs, apart, ntile, n);
        FSIM_HIP(hipGetLastError());
    }
    rc = launch_cols_corr(n, g_eb.W, npairs, pairs, part, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_eb_step, dim3(npairs), dim3(256), 0, s, pairs, part, ntile, n, tabs, prep_next ? 1 : 0);
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

}  // namespace fsim

using namespace fsim;

extern "C" int fsim_estimate(const float *images, int count, const fsim_estimate_params *p, fsim_otf otf,
                             fsim_estimate_result *res) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!images || !p || !otf || !res || count < 1) return fail(FSIM_ERR_ARG, "estimate: null argument or count < 1");
    const int n = p->n, D = p->ndir, P = p->nphase, B = p->nband;
    if (!fft_size_ok(n) || n < 16) return fail(FSIM_ERR_ARG, "estimate: n must be a power of two >= 16");
    if (reinterpret_cast<uintptr_t>(images) % 16) return fail(FSIM_ERR_ARG, "estimate: images must be 16-byte aligned");
    if (B != 2 && B != 3) return fail(FSIM_ERR_ARG, "estimate: 2 or 3 bands");
    const bool live = p->live != 0;
    if (p->live != 0 && p->live != 1) return fail(FSIM_ERR_ARG, "estimate: live must be 0 or 1");
    if (!live && p->fit_band != 1 && p->fit_band != 2) return fail(FSIM_ERR_ARG, "estimate: Fitband neither 1 nor 2");
    // the live refit fits the highest band with OTF band number 1
    const int fit_band = live ? (B == 3 ? 2 : 1) : p->fit_band;
    const int fit_otf_band = live ? 1 : p->fit_band;
    if (D < 1 || P < 2 * B - 1 || P > kMaxPhase) return fail(FSIM_ERR_ARG, "estimate: bad direction/phase count");
    if (!(p->vec_cpm > 0)) return fail(FSIM_ERR_STATE, "estimate: vector pixel size not set");
    if (p->fade_px < 0 || 2 * p->fade_px > n) return fail(FSIM_ERR_ARG, "estimate: fade border too wide");
    const Otf &o = otf->o;
    if (o.nband != 1 && (fit_otf_band >= o.nband || B - 1 >= o.nband))
        return fail(FSIM_ERR_ARG, "estimate: band idx too high for the OTF");
    const int G = count * D, npk = (B == 3 && !live) ? 2 : 1;
    const size_t nn = (size_t)n * n;
    if ((size_t)G * P > 65535 || (size_t)2 * npk * G > 65535)
        return fail(FSIM_ERR_ARG, "estimate: batch too large (at most 65535 images per transform)");

    // separation matrix for phases 2 pi p / P, default factors
    std::vector<double> ph(P), Sd((size_t)2 * (2 * B - 1) * P);
    for (int i = 0; i < P; i++) ph[i] = (2 * M_PI * i) / P + 0.0;
    rc = separation_matrix_host(ph.data(), P, B, nullptr, Sd.data());
    if (rc) return rc;
    EbSep sep{};
    sep.nphase = P;
    sep.ncomp = B == 3 ? 3 : 2;
    sep.fb = (B == 3 && fit_band == 2) ? 2 : 1;
    const int comps[3] = {0, 1, 3};
    for (int c = 0; c < sep.ncomp; c++)
        for (int q = 0; q < P; q++) {
            const size_t k = (size_t)comps[c] * P + q;
            sep.S[c][q] = make_float2((float)Sd[2 * k], (float)Sd[2 * k + 1]);
        }
    sep.vec_cpm = p->vec_cpm;
    sep.att_fwhm = 0.15 * o.cutoff;
    EbRegion reg{o.vals, o.ns, o.nband, o.cpm, o.cutoff, p->vec_cpm};

    const int ntile = eb_ntile(n);
    const int npair_max = npk * G;
    std::lock_guard<std::mutex> lk(g_eb.mu);
    rc = eb_grow((void **)&g_eb.X, g_eb.x_bytes, (size_t)G * P * nn * sizeof(float2));
    if (!rc) rc = eb_grow((void **)&g_eb.W, g_eb.w_bytes, (size_t)2 * npair_max * nn * sizeof(float2));
    // small buffers: fade table, groups, fit pairs, getPeak pairs, partials, peaks, shifts, results, errors
    const size_t o_fade = 0, o_grp = align256(o_fade + 2 * n * sizeof(double));
    const size_t o_fit = align256(o_grp + G * sizeof(EstGroup));
    const size_t o_gp = align256(o_fit + G * sizeof(EstPair));
    const size_t o_part = align256(o_gp + (size_t)npair_max * sizeof(EstPair));
    const size_t o_peak = align256(o_part + (size_t)npair_max * kPairSlots * ntile * sizeof(double2));
    const size_t o_shift = align256(o_peak + 2 * G * sizeof(double));
    const size_t o_res = align256(o_shift + 2 * G * sizeof(double));
    const size_t o_err = align256(o_res + G * sizeof(fsim_estimate_result));
    const size_t o_tabs = align256(o_err + G * sizeof(int));
    const size_t o_apart = align256(o_tabs + (size_t)G * kApxTab * n * sizeof(float2));
    const int ndc = o.nband == 1 ? 1 : (o.nband < 3 ? o.nband : 3);
    const size_t o_dct = align256(o_apart + (size_t)G * kPairSlots * ntile * sizeof(double2));
    const size_t o_runs = align256(o_dct + (size_t)ndc * nn * sizeof(float2));
    const size_t o_att = align256(o_runs + (size_t)ndc * sizeof(EbRuns));
    const size_t o_d2 = align256(o_att + (2 * (size_t)(n / 2) * (n / 2) + 1) * sizeof(float));
    const size_t small = align256(o_d2 + sizeof(int));
    if (!rc) rc = eb_grow(&g_eb.small, g_eb.small_bytes, small);
    if (rc) return rc;
    char *sb = reinterpret_cast<char *>(g_eb.small);
    double *fade = reinterpret_cast<double *>(sb + o_fade);
    EstGroup *grp = reinterpret_cast<EstGroup *>(sb + o_grp);
    EstPair *fit = reinterpret_cast<EstPair *>(sb + o_fit);
    EstPair *gp = reinterpret_cast<EstPair *>(sb + o_gp);
    double2 *part = reinterpret_cast<double2 *>(sb + o_part);
    double *peaks = reinterpret_cast<double *>(sb + o_peak);
    double *shift = reinterpret_cast<double *>(sb + o_shift);
    fsim_estimate_result *resd = reinterpret_cast<fsim_estimate_result *>(sb + o_res);
    int *errd = reinterpret_cast<int *>(sb + o_err);
    float2 *tabs = reinterpret_cast<float2 *>(sb + o_tabs);
    double2 *apart = reinterpret_cast<double2 *>(sb + o_apart);
    float2 *dct = reinterpret_cast<float2 *>(sb + o_dct);
    EbRuns *runs = reinterpret_cast<EbRuns *>(sb + o_runs);
    float *att_tab = reinterpret_cast<float *>(sb + o_att);
    int *d2min = reinterpret_cast<int *>(sb + o_d2);

    hipStream_t s = stream();
    g_eb.fade_h.resize(2 * (size_t)n);
    fade_factors_host(n, p->fade_px, g_eb.fade_h.data(), g_eb.fade_h.data() + n);
    FSIM_HIP(hipMemcpyAsync(fade, g_eb.fade_h.data(), 2 * n * sizeof(double), hipMemcpyHostToDevice, s));
    const unsigned gb = cdiv(G, 64);
    const double kmin = p->fit_exclude * o.cutoff / p->vec_cpm;
    hipLaunchKernelGGL(k_eb_init, dim3(gb), dim3(64), 0, s, grp, G, n, kmin, d2min);
    hipLaunchKernelGGL(k_eb_dc_tab, dim3((unsigned)std::min<size_t>((ndc * nn + 255) / 256, 4096)), dim3(256), 0, s,
                       dct, n, reg, ndc);
    hipLaunchKernelGGL(k_eb_runs, dim3(ndc), dim3(256), 0, s, reg, n, 0.05, runs);
    rc = launch_load_sep(n, images, g_eb.X, G, sep, fade, fade + n, live, s);
    if (rc) return rc;
    // the components' column pass and the attenuated copies' inverse
    // columns; their inverse rows, product and forward rows; the forward
    // columns with locatePeak
    const int nd2 = 2 * (n / 2) * (n / 2) + 1;
    hipLaunchKernelGGL(k_eb_att_tab, dim3((unsigned)std::min(cdiv(nd2, 256), 4096u)), dim3(256), 0, s, att_tab, nd2,
                       sep.vec_cpm, sep.att_fwhm);
    rc = launch_front(n, g_eb.X, g_eb.W, G, sep, att_tab, d2min, grp, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_eb_fit_init, dim3(gb), dim3(64), 0, s, grp, fit, G, n, sep.fb, fit_otf_band, peaks);
    FSIM_HIP(hipGetLastError());
    // fitPeak: three rounds of commonRegion, inverse FFTs, candidate filter, argmax
    for (int iter = 0; iter < 3; iter++) {
        rc = eb_round(reg, dct, runs, fit, G, true, P, n, part, tabs, apart, iter == 0, iter < 2, s);
        if (rc) return rc;
    }
    // getPeak of band 1 (and band 2)
    hipLaunchKernelGGL(k_eb_gp_init, dim3(gb), dim3(64), 0, s, fit, gp, G, B, fit_band, live ? 1 : 0, shift);
    FSIM_HIP(hipGetLastError());
    rc = eb_round(reg, dct, runs, gp, npk * G, false, P, n, part, tabs, apart, true, false, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_eb_finish, dim3(gb), dim3(64), 0, s, fit, gp, G, npk, live ? 1 : 0, peaks, shift, resd,
                       errd);
    FSIM_HIP(hipGetLastError());
    std::vector<int> err(G);
    FSIM_HIP(hipMemcpyAsync(res, resd, G * sizeof(fsim_estimate_result), hipMemcpyDeviceToHost, s));
    FSIM_HIP(hipMemcpyAsync(err.data(), errd, G * sizeof(int), hipMemcpyDeviceToHost, s));
    FSIM_HIP(hipStreamSynchronize(s));
    for (int g = 0; g < G; g++)
        if (err[g]) return fail(FSIM_ERR_ARG, "estimate: commonRegion shift larger than the vector (stack " +
                                                  std::to_string(g / D) + ", direction " + std::to_string(g % D) + ")");
    return FSIM_OK;
}

extern "C" int fsim_debug_region_exact(int on) {
    g_region_exact = on ? 1 : 0;
    return FSIM_OK;
}

extern "C" int fsim_params_from_estimate(const fsim_estimate_result *res, int count, int ndir, int nphase, int nband,
                                         const fsim_sim_params *base, fsim_sim_params *out) {
    if (!res || !base || !out || count < 0) return fail(FSIM_ERR_ARG, "params_from_estimate: null argument");
    if (ndir < 1 || ndir > 5 || nphase < 1 || nphase > 7 || (nband != 2 && nband != 3))
        return fail(FSIM_ERR_ARG, "params_from_estimate: 1-5 directions, 1-7 phases, 2 or 3 bands");
    for (int s = 0; s < count; s++) {
        fsim_sim_params p = *base;
        for (int d = 0; d < ndir; d++) {
            const fsim_estimate_result &r = res[(size_t)s * ndir + d];
            p.px[d] = r.px / (nband - 1);
            p.py[d] = r.py / (nband - 1);
            for (int ph = 0; ph < nphase; ph++) p.phases[d][ph] = base->phases[d][ph] + r.pha_off;
            p.modul[d][1] = r.mod1;
            if (nband == 3 && !std::isnan(r.mod2)) p.modul[d][2] = r.mod2;
        }
        out[s] = p;
    }
    return FSIM_OK;
}
