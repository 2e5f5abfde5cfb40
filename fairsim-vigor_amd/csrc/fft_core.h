// Device-side FFT building blocks for gfx950.
//
// An L-point complex transform is computed by T = L/E threads; thread t holds
// the E elements x[t + T*m] (m < E) in registers.  The transform runs as
// Stockham autosort passes of radix R <= E (all powers of two): the first pass
// works on the registers as loaded, middle passes exchange through LDS, the
// last pass leaves X[t + T*m] in the same registers.  So a kernel loads with
// a coalesced pattern, runs its prologue per element, calls fft_regs, runs its
// epilogue per element and stores with the same coalesced pattern; the only
// LDS traffic is (passes - 1) exchanges.
//
// Conventions follow the reference's FFT contract (JTransforms FloatFFT_2D as
// called from linalg/JTransformsConnector.java:76-90): forward e^{-2 pi i jk/L}
// unscaled; inverse e^{+2 pi i jk/L}; callers apply the 1/L scale.
#pragma once
#include <hip/hip_runtime.h>

// Default twiddle source of the kernels: 3 = a block-shared LDS table
// exp(-2 pi i s / TL) read per butterfly; 2 = sincospi of the first twiddle of
// each butterfly and its powers by complex products.
#ifndef FSIM_TW_MODE
#define FSIM_TW_MODE 3
#endif
#ifndef FSIM_LDS_PAD
#define FSIM_LDS_PAD 1
#endif

// Global (address space 1) and constant (4) views of device pointers that
// arrive inside structures.  With them the compiler emits global_* and s_load
// accesses instead of flat_* ones, whose completion also counts against the
// LDS counter: every LDS wait of an FFT exchange would otherwise wait for the
// outstanding global loads as well.
#define FSIM_GLB __attribute__((address_space(1)))
#define FSIM_CST __attribute__((address_space(4)))

namespace fsim {

template <typename T>
__device__ __forceinline__ FSIM_GLB T *glb(T *p) {
    return (FSIM_GLB T *)p;
}
template <typename T>
__device__ __forceinline__ const FSIM_CST T &cst(const T *p) {
    return *(const FSIM_CST T *)p;
}

// float2 moves through a native vector type (the float2 class only binds
// generic references)
typedef float f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ld(const FSIM_GLB float2 *p) {
    const f2v x = *(const FSIM_GLB f2v *)p;
    return make_float2(x.x, x.y);
}
__device__ __forceinline__ float2 ld(const FSIM_CST float2 *p) {
    const f2v x = *(const FSIM_CST f2v *)p;
    return make_float2(x.x, x.y);
}
__device__ __forceinline__ void st(FSIM_GLB float2 *p, float2 v) {
    f2v x;
    x.x = v.x;
    x.y = v.y;
    *(FSIM_GLB f2v *)p = x;
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// cos / sin of 2 pi m / 16, m = 0..7 (exact decimal expansions of the
// constants; the compiler rounds them to float once).
__device__ constexpr float kC16[8] = {1.0f,
                                      0.92387953251128675613f,
                                      0.70710678118654752440f,
                                      0.38268343236508977173f,
                                      0.0f,
                                      -0.38268343236508977173f,
                                      -0.70710678118654752440f,
                                      -0.92387953251128675613f};
__device__ constexpr float kS16[8] = {0.0f,
                                      0.38268343236508977173f,
                                      0.70710678118654752440f,
                                      0.92387953251128675613f,
                                      1.0f,
                                      0.92387953251128675613f,
                                      0.70710678118654752440f,
                                      0.38268343236508977173f};

// d * exp(-+ 2 pi i e / 16) with e in [0, 8), e known at compile time.
template <bool INV, int e>
__device__ __forceinline__ float2 tw16(float2 d) {
    if constexpr (e == 0) {
        return d;
    } else if constexpr (e == 4) {
        // forward: * (-i); inverse: * (+i)
        return INV ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);
    } else {
        constexpr float c = kC16[e];
        constexpr float s = INV ? kS16[e] : -kS16[e];
        return make_float2(d.x * c - d.y * s, d.x * s + d.y * c);
    }
}

// Synchronisation between Stockham passes.  When one transform lives inside a
// single wavefront (T <= 64, transforms packed along the wave), the LDS
// exchange only needs the wave's own LDS operations to be complete and
// ordered; otherwise a block barrier.
template <bool WAVE>
__device__ __forceinline__ void pass_sync() {
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v >> 1); }
constexpr int bitrevc(int v, int bits) {
    int r = 0;
    for (int b = 0; b < bits; b++) r |= ((v >> b) & 1) << (bits - 1 - b);
    return r;
}

// One radix-2 DIF stage of an R-point in-register DFT (span = R/2, R/4, ...).
template <int R, int SPAN, bool INV, int S = 0, int M = 0>
__device__ __forceinline__ void dif_stage(float2 (&v)[R]) {
    if constexpr (S < R) {
        if constexpr (M < SPAN) {
            float2 a = v[S + M], b = v[S + M + SPAN];
            v[S + M] = cadd(a, b);
            v[S + M + SPAN] = tw16<INV, M * (16 / (2 * SPAN))>(csub(a, b));
            dif_stage<R, SPAN, INV, S, M + 1>(v);
        } else {
            dif_stage<R, SPAN, INV, S + 2 * SPAN, 0>(v);
        }
    }
}

template <int R, int SPAN, bool INV>
__device__ __forceinline__ void dif_all(float2 (&v)[R]) {
    if constexpr (SPAN >= 1) {
        dif_stage<R, SPAN, INV>(v);
        dif_all<R, SPAN / 2, INV>(v);
    }
}

// R-point DFT in registers, R in {1,2,4,8,16}; natural order in and out.
template <int R, bool INV>
__device__ __forceinline__ void dft_reg(float2 (&v)[R]) {
    if constexpr (R > 1) {
        dif_all<R, R / 2, INV>(v);
        float2 t[R];
#pragma unroll
        for (int i = 0; i < R; i++) t[i] = v[i];
        constexpr int bits = ilog2c(R);
#pragma unroll
        for (int i = 0; i < R; i++) v[i] = t[bitrevc(i, bits)];
    }
}

// Elements per thread used for an L-point transform.
template <int L>
struct FftE {
    static constexpr int value = L >= 1024 ? 16 : (L >= 8 ? 8 : L);
};

// Elements per thread in the fused column kernels: as FftE (16 from
// L = 1024 up), which keeps a column's registers within 128 VGPRs.
template <int L>
struct ColE {
    static constexpr int value = FftE<L>::value;
};

// LDS layout of a transform: element i lives at lds[pad(i) * ls] with one
// spare slot every RE = min(E, 16) elements.  Stockham writes go out at stride
// R between threads; the padding makes consecutive threads land in different
// banks (stride RE + 1, odd), while the unit-stride reads stay conflict free.
template <int L, int E>
struct FftLds {
    static constexpr int re = E > 16 ? 16 : E;
    static constexpr int shift = !FSIM_LDS_PAD ? 0 : re >= 16 ? 4 : re >= 8 ? 3 : re >= 4 ? 2 : re >= 2 ? 1 : 0;
    static constexpr int len = L + (shift ? (L >> shift) : 0);   // padded length
    static __device__ __forceinline__ int pad(int i) { return shift ? i + (i >> shift) : i; }
};

// Radix of the Stockham pass that follows sub-transforms of size NS.
constexpr int tw_radix(int L, int E, int NS) {
    return (L / NS) >= (E > 16 ? 16 : E) ? (E > 16 ? 16 : E) : L / NS;
}
// Twiddle factors one thread applies in the passes from NS on: every pass
// after the first has Q = E/R butterflies per thread with R-1 twiddles each.
constexpr int tw_count_from(int L, int E, int NS) {
    return NS >= L ? 0
                   : ((NS > 1 ? (E / tw_radix(L, E, NS)) * (tw_radix(L, E, NS) - 1) : 0) +
                      tw_count_from(L, E, NS * tw_radix(L, E, NS)));
}

// A thread's twiddle factors for every pass of an L-point transform with E
// elements per thread (thread t, element order of fft_regs), held in
// registers: loaded once per kernel from the exact-rounded table
// exp(-2 pi i s / TL) and reused by all of the kernel's transforms of that
// length (an inverse transform conjugates them), so the passes do no table
// reads, no index arithmetic and no sincos.  Worth it where the count is
// small (<= 14 for L <= 512 with E = 8); the values equal the table entries,
// so results are bitwise those of the table path.
template <int L, int E>
struct TwReg {
    static constexpr int count = tw_count_from(L, E, 1);
    float2 w[count > 0 ? count : 1];
};
template <int L, int E>
constexpr bool tw_reg_fits() {
    return E <= 8 && TwReg<L, E>::count <= 16;
}

__device__ __forceinline__ float2 tw_fetch(const float2 *p, int i) { return p[i]; }
__device__ __forceinline__ float2 tw_fetch(const FSIM_GLB float2 *p, int i) { return ld(p + i); }

template <int L, int E, int TL, int NS, typename P>
__device__ __forceinline__ void tw_reg_load_pass(TwReg<L, E> &tw, int t, P tab) {
    if constexpr (NS < L) {
        constexpr int R = tw_radix(L, E, NS), Q = E / R, T = L / E;
        if constexpr (NS > 1) {
            constexpr int off = tw_count_from(L, E, 1) - tw_count_from(L, E, NS);
            constexpr int SC = TL / (NS * R);
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const int k = (t + T * q) & (NS - 1);
#pragma unroll
                for (int r = 1; r < R; r++) tw.w[off + q * (R - 1) + r - 1] = tw_fetch(tab, SC * k * r);
            }
        }
        tw_reg_load_pass<L, E, TL, NS * R>(tw, t, tab);
    }
}

// Fills tw for thread t from the table tab (LDS or global) of length TL, a
// multiple of L: entry s = exp(-2 pi i s / TL).
template <int L, int E, int TL, typename P>
__device__ __forceinline__ void tw_reg_load(TwReg<L, E> &tw, int t, P tab) {
    static_assert(TL % L == 0, "twiddle table length must be a multiple of L");
    tw_reg_load_pass<L, E, TL, 1>(tw, t, tab);
}

// Twiddle sources of fft_pass: a table (TL > 0: LDS table of length TL; TL =
// 0: sincospi of the first twiddle and its powers) or TwReg registers.
template <int TL>
struct TwTab {
    const float2 *tw;
    template <int L, int E, int NS, int R, bool INV>
    __device__ __forceinline__ void apply(float2 (&u)[R], int k, int q) const {
        if constexpr (TL > 0) {
            // w^r = exp(-+2 pi i k r / (NS R)) = table[(TL / (NS R)) k r];
            // the index advances by additions (k is opaque to the compiler,
            // which would otherwise form k r with quarter-rate multiplies)
            constexpr int SC = TL / (NS * R);
            const int step = SC * k;
            int idx = step;
#pragma unroll
            for (int r = 1; r < R; r++) {
                float2 w = tw[idx];
                idx += step;
                asm volatile("" : "+v"(idx));   // keeps the addition (no multiply by r)
                if (INV) w.y = -w.y;
                u[r] = cmul(u[r], w);
            }
        } else {
            // exp(-+2 pi i k / (NS R)): the argument k/(NS R) is exact
            // in float (power-of-two denominator), sincospi is accurate
            float sw, cw;
            sincospif((INV ? 2.0f : -2.0f) * (float)k / (float)(NS * R), &sw, &cw);
            const float2 w1 = make_float2(cw, sw);
            float2 w = w1;
#pragma unroll
            for (int r = 1; r < R; r++) {
                u[r] = cmul(u[r], w);
                w = cmul(w, w1);
            }
        }
    }
};
template <int L0, int E0>
struct TwRegRef {
    const TwReg<L0, E0> &tw;
    template <int L, int E, int NS, int R, bool INV>
    __device__ __forceinline__ void apply(float2 (&u)[R], int, int q) const {
        constexpr int off = tw_count_from(L, E, 1) - tw_count_from(L, E, NS);
#pragma unroll
        for (int r = 1; r < R; r++) {
            float2 w = tw.w[off + q * (R - 1) + r - 1];
            if (INV) w.y = -w.y;
            u[r] = cmul(u[r], w);
        }
    }
};

// Stockham pass NS (sub-transform size so far) of radix R = min(E, 16, L/NS);
// recursion walks the passes at compile time.  TS supplies the twiddles
// (TwTab or TwRegRef).
template <int L, int E, int NS, bool INV, bool FIRST, bool WAVE, typename TS>
__device__ __forceinline__ void fft_pass(float2 (&v)[E], float2 *lds, int ls, int t, const TS &ts, bool active) {
    constexpr int REM = L / NS;
    constexpr int RE = E > 16 ? 16 : E;  // in-register DFTs go up to radix 16
    constexpr int R = REM >= RE ? RE : REM;
    constexpr bool LAST = (NS * R == L);
    constexpr int T = L / E;
    constexpr int Q = E / R;
    using PL = FftLds<L, E>;
    if constexpr (!FIRST) {
        if (active) {
            if constexpr (PL::shift == 0 || (T % PL::re) == 0) {
                // pad(t + T m) = pad(t) + m (T + T / RE) when RE divides T
                constexpr int TP = T + (PL::shift ? (T >> PL::shift) : 0);
                const int pt = PL::pad(t);
#pragma unroll
                for (int m = 0; m < E; m++) v[m] = lds[(pt + TP * m) * ls];
            } else {
#pragma unroll
                for (int m = 0; m < E; m++) v[m] = lds[PL::pad(t + T * m) * ls];
            }
        }
        pass_sync<WAVE>();
    }
    if (active) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            float2 u[R];
#pragma unroll
            for (int r = 0; r < R; r++) u[r] = v[q + r * Q];
            const int j = t + T * q;
            int k = j & (NS - 1);
            // keep the twiddles computed here: hoisted out of a caller's
            // loop they stay live across it and push kernels into spills
            // (the empty asm emits no instruction)
            asm volatile("" : "+v"(k));
            if constexpr (NS > 1) ts.template apply<L, E, NS, R, INV>(u, k, q);
            dft_reg<R, INV>(u);
            if constexpr (LAST) {
#pragma unroll
                for (int r = 0; r < R; r++) v[q + r * Q] = u[r];
            } else {
                // pad(base + r NS) = pad(base) + r (NS + NS / RE): NS is 1 or a
                // multiple of RE, and for NS = 1 base is a multiple of R = RE
                const int pb = PL::pad((j - k) * R + k);
                constexpr int NSP = NS + (PL::shift ? (NS >> PL::shift) : 0);
#pragma unroll
                for (int r = 0; r < R; r++) lds[(pb + r * NSP) * ls] = u[r];
            }
        }
    }
    if constexpr (!LAST) {
        pass_sync<WAVE>();
        fft_pass<L, E, NS * R, INV, false, WAVE, TS>(v, lds, ls, t, ts, active);
    }
}

// Full L-point FFT on registers.  Every thread of the block must call it
// (it contains barriers); threads with active == false only join the
// barriers.  lds: this transform's region of FftLds<L, E>::len elements
// (times ls), element i at lds[FftLds<L, E>::pad(i) * ls].
// WAVE selects wavefront-level synchronisation, valid when every transform's
// T = L/E threads lie inside one wavefront; the caller's own later LDS use
// of the region is then ordered by the same wave.  On return the region may
// be reused (the last pass ends with a synchronisation after its reads).
// TL > 0: tw is the twiddle table returned by tw_table<TL> (TL a multiple of
// L); TL = 0: no table, twiddles from sincospi (tw unused).
template <int L, int E, bool INV, bool WAVE, int TL>
__device__ __forceinline__ void fft_regs(float2 (&v)[E], float2 *lds, int ls, int t,
                                         const float2 *__restrict__ tw, bool active = true) {
    static_assert(L % E == 0, "E must divide L");
    static_assert(TL == 0 || TL % L == 0, "twiddle table length must be a multiple of L");
    static_assert(!WAVE || (L / E) <= 64, "wave synchronisation needs the transform inside one wavefront");
    const TwTab<TL> ts{tw};
    fft_pass<L, E, 1, INV, true, WAVE, TwTab<TL>>(v, lds, ls, t, ts, active);
}

// fft_regs with the thread's twiddles in registers (tw_reg_load).
template <int L, int E, bool INV, bool WAVE>
__device__ __forceinline__ void fft_regs_r(float2 (&v)[E], float2 *lds, int ls, int t, const TwReg<L, E> &tw,
                                           bool active = true) {
    static_assert(L % E == 0, "E must divide L");
    static_assert(!WAVE || (L / E) <= 64, "wave synchronisation needs the transform inside one wavefront");
    const TwRegRef<L, E> ts{tw};
    fft_pass<L, E, 1, INV, true, WAVE, TwRegRef<L, E>>(v, lds, ls, t, ts, active);
}

// fft_pass with each LDS exchange in two halves, the real parts and then the
// imaginary parts, through a float region of FftLds<L, E>::len elements
// (times ls): half the LDS of fft_pass for twice the barriers.  The
// butterflies, twiddles and their order are those of fft_pass, and the
// exchange only moves values, so the results are bitwise equal.
template <int L, int E, int NS, bool INV, bool WAVE, typename TS>
__device__ __forceinline__ void fft_pass_split(float2 (&v)[E], float *lds, int ls, int t, const TS &ts,
                                               bool active) {
    constexpr int REM = L / NS;
    constexpr int RE = E > 16 ? 16 : E;
    constexpr int R = REM >= RE ? RE : REM;
    constexpr bool LAST = (NS * R == L);
    constexpr int T = L / E;
    constexpr int Q = E / R;
    using PL = FftLds<L, E>;
    int pb[Q];
    if (active) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            float2 u[R];
#pragma unroll
            for (int r = 0; r < R; r++) u[r] = v[q + r * Q];
            const int j = t + T * q;
            int k = j & (NS - 1);
            asm volatile("" : "+v"(k));
            if constexpr (NS > 1) ts.template apply<L, E, NS, R, INV>(u, k, q);
            dft_reg<R, INV>(u);
#pragma unroll
            for (int r = 0; r < R; r++) v[q + r * Q] = u[r];
            pb[q] = PL::pad((j - k) * R + k);
        }
    }
    if constexpr (!LAST) {
        constexpr int NSP = NS + (PL::shift ? (NS >> PL::shift) : 0);
#pragma unroll
        for (int h = 0; h < 2; h++) {
            if (active) {
#pragma unroll
                for (int q = 0; q < Q; q++)
#pragma unroll
                    for (int r = 0; r < R; r++) lds[(pb[q] + r * NSP) * ls] = h ? v[q + r * Q].y : v[q + r * Q].x;
            }
            pass_sync<WAVE>();
            if (active) {
                if constexpr (PL::shift == 0 || (T % PL::re) == 0) {
                    constexpr int TP = T + (PL::shift ? (T >> PL::shift) : 0);
                    const int pt = PL::pad(t);
#pragma unroll
                    for (int m = 0; m < E; m++) {
                        const float x = lds[(pt + TP * m) * ls];
                        if (h) v[m].y = x; else v[m].x = x;
                    }
                } else {
#pragma unroll
                    for (int m = 0; m < E; m++) {
                        const float x = lds[PL::pad(t + T * m) * ls];
                        if (h) v[m].y = x; else v[m].x = x;
                    }
                }
            }
            pass_sync<WAVE>();
        }
        fft_pass_split<L, E, NS * R, INV, WAVE, TS>(v, lds, ls, t, ts, active);
    }
}

// fft_regs through a float LDS region (fft_pass_split).
template <int L, int E, bool INV, bool WAVE, int TL>
__device__ __forceinline__ void fft_regs_split(float2 (&v)[E], float *lds, int ls, int t,
                                               const float2 *__restrict__ tw, bool active = true) {
    static_assert(L % E == 0, "E must divide L");
    static_assert(TL == 0 || TL % L == 0, "twiddle table length must be a multiple of L");
    static_assert(!WAVE || (L / E) <= 64, "wave synchronisation needs the transform inside one wavefront");
    const TwTab<TL> ts{tw};
    fft_pass_split<L, E, 1, INV, WAVE, TwTab<TL>>(v, lds, ls, t, ts, active);
}

// Twiddle table of a kernel whose largest transform has L points: tl is the
// table length passed to fft_regs (0: no table), len the LDS elements the
// kernel reserves for it.  use = false forces the sincospi form (a kernel
// whose LDS cannot hold the table).
template <int L, bool use = true>
struct TwLds {
    static constexpr int tl = (FSIM_TW_MODE == 3 && use) ? L : 0;
    static constexpr int len = tl > 0 ? tl : 1;
};

// The thread's index in its block, passed through an empty asm statement so
// that the optimizer keeps the index arithmetic at its uses: the persistent
// work-list kernels run block functions in loops, and arithmetic that
// depends on the thread alone would otherwise be hoisted out of every loop
// and held in registers for all of the functions at once.
__device__ __forceinline__ int item_tid() {
    unsigned v = threadIdx.x;
    asm volatile("" : "+v"(v));
    __builtin_assume(v < 1024u);
    return (int)v;
}

// Fills the LDS table twl of length TL (> 0) from the global table
// g[s] = exp(-2 pi i s / TL) (rounded from double) and synchronises the
// block; every one of the NT threads must call it.  All loads are issued
// before the first LDS write.  With TL = 0 it does nothing.
template <int TL, int NT>
__device__ __forceinline__ const float2 *tw_table(float2 *twl, const float2 *g, int tid) {
    if constexpr (TL > 0) {
        constexpr int R = (TL + NT - 1) / NT;
        const FSIM_GLB float2 *gg = glb(g);
        float2 tmp[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = tid + r * NT;
            if (TL % NT == 0 || i < TL) tmp[r] = ld(gg + i);
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = tid + r * NT;
            if (TL % NT == 0 || i < TL) twl[i] = tmp[r];
        }
        __syncthreads();
        return twl;
    } else {
        return g;
    }
}
template <int TL, int NT>
__device__ __forceinline__ const float2 *tw_table(float2 *twl, const float2 *g) {
    return tw_table<TL, NT>(twl, g, (int)threadIdx.x);
}
// tw_table for blocks that run several items with the same table: loads it
// (and sets loaded) only when loaded is false.  loaded must be uniform.
template <int TL, int NT>
__device__ __forceinline__ const float2 *tw_table(float2 *twl, const float2 *g, int tid, bool &loaded) {
    if constexpr (TL > 0) {
        if (!loaded) {
            tw_table<TL, NT>(twl, g, tid);
            loaded = true;
        }
        return twl;
    } else {
        return g;
    }
}

// Row layout (t fastest): one transform per T consecutive threads.
template <int L>
struct RowWave {
    static constexpr bool value = (L / FftE<L>::value) <= 64;
};



}  // namespace fsim
