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

// Stockham pass NS (sub-transform size so far) of radix R = min(E, L/NS);
// recursion walks the passes at compile time.
template <int L, int E, int NS, bool INV, bool FIRST>
__device__ __forceinline__ void fft_pass(float2 (&v)[E], float2 *lds, int ls, int t,
                                         const float2 *__restrict__ tw, bool active) {
    constexpr int REM = L / NS;
    constexpr int R = REM >= E ? E : REM;
    constexpr bool LAST = (NS * R == L);
    constexpr int T = L / E;
    constexpr int Q = E / R;
    if constexpr (!FIRST) {
        if (active) {
#pragma unroll
            for (int m = 0; m < E; m++) v[m] = lds[(t + T * m) * ls];
        }
        __syncthreads();
    }
    if (active) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            float2 u[R];
#pragma unroll
            for (int r = 0; r < R; r++) u[r] = v[q + r * Q];
            const int j = t + T * q;
            const int k = j & (NS - 1);
            if constexpr (NS > 1) {
                const int st = (L / (NS * R)) * k;
#pragma unroll
                for (int r = 1; r < R; r++) {
                    float2 w = tw[r * st];
                    if (INV) w.y = -w.y;
                    u[r] = cmul(u[r], w);
                }
            }
            dft_reg<R, INV>(u);
            if constexpr (LAST) {
#pragma unroll
                for (int r = 0; r < R; r++) v[q + r * Q] = u[r];
            } else {
                const int base = (j - k) * R + k;
#pragma unroll
                for (int r = 0; r < R; r++) lds[(base + r * NS) * ls] = u[r];
            }
        }
    }
    if constexpr (!LAST) {
        __syncthreads();
        fft_pass<L, E, NS * R, INV, false>(v, lds, ls, t, tw, active);
    }
}

// Full L-point FFT on registers.  Every thread of the block must call it
// (it contains block barriers); threads with active == false only join the
// barriers.  lds: this transform's region, element i at lds[i * ls].
// On return the LDS region may be reused after a __syncthreads().
template <int L, int E, bool INV>
__device__ __forceinline__ void fft_regs(float2 (&v)[E], float2 *lds, int ls, int t,
                                         const float2 *__restrict__ tw, bool active = true) {
    static_assert(L % E == 0, "E must divide L");
    fft_pass<L, E, 1, INV, true>(v, lds, ls, t, tw, active);
}

// Elements per thread used for an L-point transform.
template <int L>
struct FftE {
    static constexpr int value = L >= 1024 ? 16 : (L >= 8 ? 8 : L);
};

}  // namespace fsim
