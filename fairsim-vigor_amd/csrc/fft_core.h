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

#ifndef FSIM_TW_MODE
#define FSIM_TW_MODE 2
#endif
#ifndef FSIM_LDS_PAD
#define FSIM_LDS_PAD 1
#endif

namespace fsim {

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

// Elements per thread in column kernels: 32 from L = 1024 up, so that the
// 16-column tiles fit 512-thread blocks (256 VGPRs per thread available).
template <int L>
struct ColE {
    static constexpr int value = L >= 1024 ? 32 : FftE<L>::value;
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

// Stockham pass NS (sub-transform size so far) of radix R = min(E, 16, L/NS);
// recursion walks the passes at compile time.
template <int L, int E, int NS, bool INV, bool FIRST, bool WAVE>
__device__ __forceinline__ void fft_pass(float2 (&v)[E], float2 *lds, int ls, int t,
                                         const float2 *__restrict__ tw, bool active) {
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
            if constexpr (NS > 1) {
                const int st = (L / (NS * R)) * k;
#if FSIM_TW_MODE == 1 || FSIM_TW_MODE == 2
                // first twiddle of the butterfly, powers by complex products
#if FSIM_TW_MODE == 2
                // exp(-+2 pi i k / (NS R)): the argument k/(NS R) is exact in
                // float (power-of-two denominator), sincospi is accurate
                float sw, cw;
                sincospif((INV ? 2.0f : -2.0f) * (float)k / (float)(NS * R), &sw, &cw);
                float2 w1 = make_float2(cw, sw);
#else
                float2 w1 = tw[st];
                if (INV) w1.y = -w1.y;
#endif
                float2 w = w1;
#pragma unroll
                for (int r = 1; r < R; r++) {
                    u[r] = cmul(u[r], w);
                    w = cmul(w, w1);
                }
#else
#pragma unroll
                for (int r = 1; r < R; r++) {
                    float2 w = tw[r * st];
                    if (INV) w.y = -w.y;
                    u[r] = cmul(u[r], w);
                }
#endif
            }
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
        fft_pass<L, E, NS * R, INV, false, WAVE>(v, lds, ls, t, tw, active);
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
template <int L, int E, bool INV, bool WAVE = false>
__device__ __forceinline__ void fft_regs(float2 (&v)[E], float2 *lds, int ls, int t,
                                         const float2 *__restrict__ tw, bool active = true) {
    static_assert(L % E == 0, "E must divide L");
    static_assert(!WAVE || (L / E) <= 64, "wave synchronisation needs the transform inside one wavefront");
    fft_pass<L, E, 1, INV, true, WAVE>(v, lds, ls, t, tw, active);
}

// Row layout (t fastest): one transform per T consecutive threads.
template <int L>
struct RowWave {
    static constexpr bool value = (L / FftE<L>::value) <= 64;
};



}  // namespace fsim
