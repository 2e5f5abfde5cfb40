// Device pieces of the parameter estimation (sim_algorithm/Correlation.java)
// shared by the per-call entry points (estimate.hip) and the batched
// estimation (estimate_batch.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "simmath.h"

namespace fsim {

// |v| as Cplx.Float.abs (Cplx.java:159-161): sqrt in double, rounded to float.
__device__ __forceinline__ float abs_java(float2 v) {
    return (float)sqrt((double)v.x * (double)v.x + (double)v.y * (double)v.y);
}

// a / b as Cplx.Float.div = mult(a, reciproc(b)) (Cplx.java:311-319).
__device__ __forceinline__ float2 div_java(float2 a, float2 b) {
#pragma clang fp contract(off)
    const float s = b.x * b.x + b.y * b.y;
    const float rr = b.x / s, ri = -b.y / s;
    return make_float2(a.x * rr - a.y * ri, a.x * ri + a.y * rr);
}

// Distance of pixel (x, y) to the DC component, in pixels, as the reference
// computes it (integer squares, then sqrt in double).
__device__ __forceinline__ double dc_rad(int x, int y, int w, int h) {
    const int dx = x < w / 2 ? x : x - w, dy = y < h / 2 ? y : y - h;
    return sqrt((double)(dx * dx + dy * dy));
}

// |(kx, ky)| as Math.sqrt(kx * kx + ky * ky), without fused multiply-adds.
__device__ __forceinline__ double k_len(double kx, double ky) {
#pragma clang fp contract(off)
    return sqrt(kx * kx + ky * ky);
}

struct RegionArgs {
    const float2 *tab0, *tab1;   // radial OTF tables of bands bn0 and bn1 (no attenuation)
    int ns;
    double cpm, cutoff, vec_cpm;
    double kx, ky, dist, wlim;
    int ikx, iky;                // (int)kx, (int)ky: truncated toward zero
    int divide;
};

// Correlation.commonRegion (Correlation.java:227-288) at pixel (x, y).
// band0: zero where either OTF weight (centred at 0 and at (kx, ky)) is below
// the limit, else divided by the weight; zero in the ring ratio < dist or
// ratio > 1 - dist around DC.  band1: the same with the weights centred at 0
// and (-kx, -ky); the ring pixels (xq, yq) also zero band1 at
// ((xq - ikx) mod w, (yq + iky) mod h), so band1(x, y) is zero when its
// source pixel ((x + ikx) mod w, (y - iky) mod h) lies in the ring.  The
// reference's loop order decides only the sign of such zeros.
// kmax = |(kx, ky)|.
__device__ __forceinline__ float2 region_band0(const RegionArgs &a, int x, int y, int w, int h, double kmax,
                                               float2 v0) {
    const float2 w0 = otf_pixel(a.tab0, a.ns, a.cpm, a.cutoff, x, y, w, h, 0.0, 0.0, a.vec_cpm);
    const float2 wt0 = otf_pixel(a.tab0, a.ns, a.cpm, a.cutoff, x, y, w, h, a.kx, a.ky, a.vec_cpm);
    if (abs_java(w0) < a.wlim || abs_java(wt0) < a.wlim)
        v0 = make_float2(0.f, 0.f);
    else if (a.divide)
        v0 = div_java(v0, w0);
    const double r0 = dc_rad(x, y, w, h) / kmax;
    if (r0 < a.dist || r0 > 1 - a.dist) v0 = make_float2(0.f, 0.f);
    return v0;
}

__device__ __forceinline__ float2 region_band1(const RegionArgs &a, int x, int y, int w, int h, double kmax,
                                               float2 v1) {
    const float2 w1 = otf_pixel(a.tab1, a.ns, a.cpm, a.cutoff, x, y, w, h, 0.0, 0.0, a.vec_cpm);
    const float2 wt1 = otf_pixel(a.tab1, a.ns, a.cpm, a.cutoff, x, y, w, h, -a.kx, -a.ky, a.vec_cpm);
    if (abs_java(w1) < a.wlim || abs_java(wt1) < a.wlim)
        v1 = make_float2(0.f, 0.f);
    else if (a.divide)
        v1 = div_java(v1, w1);
    const int xq = ((x + a.ikx) % w + w) % w, yq = ((y - a.iky) % h + h) % h;
    const double rq = dc_rad(xq, yq, w, h) / kmax;
    if (rq < a.dist || rq > 1 - a.dist) v1 = make_float2(0.f, 0.f);
    return v1;
}

__device__ __forceinline__ void region_pixel(const RegionArgs &a, int x, int y, int w, int h, double kmax,
                                             float2 &v0, float2 &v1) {
    v0 = region_band0(a, x, y, w, h, kmax, v0);
    v1 = region_band1(a, x, y, w, h, kmax, v1);
}

// Candidate correlations: b1 shifted by (xpos, -ypos) (Vec2d.Cplx.fourierShift,
// BasicVectors.java:309-339), times conj(b0) (AbstractVectorCplx.timesConj),
// summed over all pixels in double (sumElements).  A block holds kCorrPix
// pixels of both bands and evaluates up to kCandPerBlock candidates.
constexpr int kMaxCand = 100;
constexpr int kCorrThreads = 256, kCorrPpt = 8, kCorrPix = kCorrThreads * kCorrPpt;
constexpr int kCandPerBlock = 25;

// Block-wide double2 sum, result valid in thread 0.  Called by all threads.
__device__ __forceinline__ double2 block_sum(double2 v, double2 *sh) {
    for (int o = 32; o > 0; o >>= 1) {
        v.x += __shfl_down(v.x, o, 64);
        v.y += __shfl_down(v.y, o, 64);
    }
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    double2 r = make_double2(0.0, 0.0);
    if (threadIdx.x == 0)
        for (int i = 0; i < kCorrThreads / 64; i++) {
            r.x += sh[i].x;
            r.y += sh[i].y;
        }
    __syncthreads();
    return r;
}

// The pixels of chunk `chunk` of both N x N bands held by this thread.
struct CorrPix {
    float2 p0[kCorrPpt], p1[kCorrPpt];
    int px[kCorrPpt], py[kCorrPpt];
    bool ok[kCorrPpt];
};

__device__ __forceinline__ void corr_load(CorrPix &c, const float2 *__restrict__ b0, const float2 *__restrict__ b1,
                                          int N, int chunk) {
    const size_t n = (size_t)N * N;
#pragma unroll
    for (int j = 0; j < kCorrPpt; j++) {
        const size_t i = (size_t)chunk * kCorrPix + (size_t)j * kCorrThreads + threadIdx.x;
        c.ok[j] = i < n;
        const size_t ic = c.ok[j] ? i : 0;
        c.p0[j] = b0[ic];
        c.p1[j] = b1[ic];
        c.px[j] = (int)(ic % N);
        c.py[j] = (int)(ic / N);
    }
}

// One pixel's term of sum(shift(b1, xpos, -ypos) * conj(b0)) for a
// power-of-two size N (inv_n = 1 / N): fourierShift's phase and rotation,
// then timesConj, in the reference's float operations.
__device__ __forceinline__ float2 corr_pixel(float2 p0, float2 p1, int x, int y, double xpos, double ypos,
                                             double inv_n) {
    float co, si;
    shift_cos_sin_pow2(xpos, -ypos, x, y, inv_n, co, si);
    return times_conj_java(rot_java(p1, co, si), p0);
}

// This thread's part of sum(shift(b1, xpos, -ypos) * conj(b0)).
__device__ __forceinline__ double2 corr_candidate(const CorrPix &c, double xpos, double ypos, int N) {
    double2 acc = make_double2(0.0, 0.0);
    const bool pow2 = (N & (N - 1)) == 0;
#pragma unroll
    for (int j = 0; j < kCorrPpt; j++) {
        float2 q;
        if (pow2) {
            q = corr_pixel(c.p0[j], c.p1[j], c.px[j], c.py[j], xpos, ypos, 1.0 / N);
        } else {
            float co, si;
            shift_cos_sin(xpos, -ypos, c.px[j], c.py[j], N, co, si);
            q = times_conj_java(rot_java(c.p1[j], co, si), c.p0[j]);
        }
        if (c.ok[j]) {
            acc.x += (double)q.x;
            acc.y += (double)q.y;
        }
    }
    return acc;
}

// b0.norm2() (AbstractVectorCplx.java:241-247) accumulation of one pixel:
// float squares added in double.
__device__ __forceinline__ double corr_norm2_add(double a, float2 v) {
    const float re2 = v.x * v.x, im2 = v.y * v.y;
    a += (double)re2;
    a += (double)im2;
    return a;
}

// This thread's part of b0.norm2().
__device__ __forceinline__ double corr_norm2(const CorrPix &c) {
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < kCorrPpt; j++)
        if (c.ok[j]) a = corr_norm2_add(a, c.p0[j]);
    return a;
}

// fitPeak's candidate grid (Correlation.java:84-88): kx + ((i - 4.5) / 4.5) * search
__device__ __host__ __forceinline__ double grid_pos(double k, int i, double search) {
    return k + ((i - 4.5) / 4.5) * search;
}

}  // namespace fsim
