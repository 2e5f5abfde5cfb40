// Per-element SIM math shared by the vector-API kernels and the fused
// reconstruction kernels.  Each function reproduces the Java arithmetic it
// cites operation by operation (FMA contraction is switched off inside), so
// that tables and integer maps match the reference bit for bit.
#pragma once
#include <hip/hip_runtime.h>

namespace fsim {

// a / b rounded to nearest, from ib = 1 / b rounded to nearest: q = a * ib is
// within an ulp of a / b, the residual a - b * q is exact in one FMA, and
// RN(q + (a - b * q) * ib) = RN(a / b) (Markstein's theorem), so the result is
// the IEEE quotient bit for bit.  Two FMAs and a product instead of the
// division sequence; where a or q is zero, NaN, infinite, huge or tiny (the
// residual could overflow or underflow; b zero, infinite or NaN lands here
// too) it divides.  Callers pass b's reciprocal as an expression of
// loop-invariant values (a constant, a kernel argument), which the compiler
// folds or computes once.
__device__ __forceinline__ double div_rn(double a, double b, double ib) {
    const double q = a * ib;
    const double m = fabs(a), mq = fabs(q);
    if (!(m >= 0x1p-900 && m <= 0x1p900 && mq >= 0x1p-900 && mq <= 0x1p900)) return a / b;
    const double r = fma(-q, b, a);
    return fma(r, ib, q);
}

// MTool.aproxSin01 / fsin / fcos (linalg/MTool.java:34-59), in double.
__device__ __forceinline__ double aprox_sin01(double x) {
#pragma clang fp contract(off)
    double x2 = x * x;
    return ((((.00015148419 * x2 - .00467376557) * x2 + .07968967928) * x2 - .64596371106) * x2 +
            1.57079631847) *
           x;
}
__device__ __forceinline__ double fsin_java(double x) {
#pragma clang fp contract(off)
    x = div_rn(x, M_PI / 2, 1 / (M_PI / 2));
    x -= floor(0.25 * x) * 4;
    if (x > 3) return -aprox_sin01(4 - x);
    if (x > 1) return aprox_sin01(2 - x);
    return aprox_sin01(x);
}

// Phase of Vec2d.Cplx.fourierShift (BasicVectors.java:320-323):
// phaVal = (float)(2*PI*(kx*x+ky*y)/N); co/si = (float)fcos/fsin(phaVal).
__device__ __forceinline__ void shift_cos_sin(double kx, double ky, int x, int y, int N, float &co,
                                              float &si) {
#pragma clang fp contract(off)
    float pha = (float)div_rn(2 * M_PI * (kx * x + ky * y), (double)N, 1.0 / N);
    double p = (double)pha;
    co = (float)fsin_java(p + M_PI / 2);
    si = (float)fsin_java(p);
}

// The same for a power-of-two N, from inv_n = 1 / N: the product with the
// exact reciprocal of a power of two is the quotient, bit for bit.
__device__ __forceinline__ void shift_cos_sin_pow2(double kx, double ky, int x, int y, double inv_n, float &co,
                                                   float &si) {
#pragma clang fp contract(off)
    float pha = (float)(2 * M_PI * (kx * x + ky * y) * inv_n);
    double p = (double)pha;
    co = (float)fsin_java(p + M_PI / 2);
    si = (float)fsin_java(p);
}

// Cplx.multReal / multImag of (re, im) and (co, si) (linalg/Cplx.java:325-332)
__device__ __forceinline__ float2 rot_java(float2 v, float co, float si) {
#pragma clang fp contract(off)
    return make_float2((v.x * co) - (v.y * si), (v.y * co) + (v.x * si));
}

// Wrapped, y-flipped frequency coordinates (OtfProvider.java:345-346).
__device__ __forceinline__ double xh_of(int x, int w) { return (double)((x < w / 2) ? x : x - w); }
__device__ __forceinline__ double yh_of(int y, int h) { return (double)((y < h / 2) ? -y : h - y); }

// OtfProvider.getOtfVal (OtfProvider.java:189-222) on a device table; both
// interpolation weights use the lower sample, as the reference does.
// otf_index: the table index l and fraction f at radius cycl, or -1 where the
// value is zero; otf_weigh: the value from the sample tab[l].
__device__ __forceinline__ int otf_index(int ns, double cpm, double cutoff, double cycl, float &f) {
#pragma clang fp contract(off)
    f = 0.f;
    if (cycl >= cutoff) return -1;
    const double pos = div_rn(cycl, cpm, 1 / cpm);
    if (ceil(pos) >= (double)ns) return -1;
    const int l = (int)floor(pos);
    f = (float)(pos - l);
    return l;
}
__device__ __forceinline__ float2 otf_weigh(float2 v, float f) {
#pragma clang fp contract(off)
    const float omf = 1 - f;
    return make_float2(v.x * omf + v.x * f, v.y * omf + v.y * f);
}
__device__ __forceinline__ float2 otf_val(const float2 *__restrict__ tab, int ns, double cpm, double cutoff,
                                          double cycl) {
    float f;
    const int l = otf_index(ns, cpm, cutoff, cycl, f);
    if (l < 0) return make_float2(0.f, 0.f);
    return otf_weigh(tab[l], f);
}

// OtfProvider.getAttVal (OtfProvider.java:246-269)
__device__ __forceinline__ float att_val(const float *__restrict__ only_att, double cpm, double cutoff,
                                         double cycl) {
#pragma clang fp contract(off)
    if (cycl >= cutoff) return 1.f;
    const double pos = div_rn(cycl, cpm, 1 / cpm);
    const int l = (int)floor(pos);
    const float f = (float)(pos - l);
    const float retl = only_att[l] * f;
    const float reth = only_att[l] * (1 - f);
    return retl + reth;
}

// Squared distance of pixel (x, y) from centre (kx, ky) in pixels, and the
// radius in cycles/micron from it: MTool.fhypot(xh-kx, yh-ky) *
// vecCyclesPerMicron (OtfProvider.java:347-349).
__device__ __forceinline__ double d2_at(int x, int y, int w, int h, double kx, double ky) {
#pragma clang fp contract(off)
    const double dx = xh_of(x, w) - kx, dy = yh_of(y, h) - ky;
    return dx * dx + dy * dy;
}
__device__ __forceinline__ double cycl_of_d2(double d2, double vec_cpm) {
#pragma clang fp contract(off)
    return sqrt(d2) * vec_cpm;
}
__device__ __forceinline__ double cycl_at(int x, int y, int w, int h, double kx, double ky, double vec_cpm) {
    return cycl_of_d2(d2_at(x, y, w, h, kx, ky), vec_cpm);
}

// otfToVector value at a pixel (OtfProvider.java:351-365): 0 above cutoff.
__device__ __forceinline__ float2 otf_pixel(const float2 *__restrict__ tab, int ns, double cpm, double cutoff,
                                            int x, int y, int w, int h, double kx, double ky, double vec_cpm) {
    const double cycl = cycl_at(x, y, w, h, kx, ky, vec_cpm);
    if (cycl > cutoff) return make_float2(0.f, 0.f);
    return otf_val(tab, ns, cpm, cutoff, cycl);
}

// OtfProvider.writeAttenuationVector value at pixel (x, y)
// (OtfProvider.java:503-529): 1 - str * exp(-cycl^2 / (2 (fwhm / 2.355)^2)).
// attenuation_d2: the same from the squared distance d2_at returns.
__device__ __forceinline__ float attenuation_d2(double d2, double vec_cpm, double str, double fwhm) {
#pragma clang fp contract(off)
    const double cycl = cycl_of_d2(d2, vec_cpm);
    return (float)(1 - str * exp(-(cycl * cycl) / (2 * ((fwhm / 2.355) * (fwhm / 2.355)))));
}
__device__ __forceinline__ float attenuation_pixel(int x, int y, int w, int h, double vec_cpm, double str,
                                                   double fwhm, double kx, double ky) {
    return attenuation_d2(d2_at(x, y, w, h, kx, ky), vec_cpm, str, fwhm);
}

// Correlation.locatePeak (Correlation.java:144-167) ranking key of pixel i =
// (x, y): |c| as (float)sqrt((double)re*re + (double)im*im) in the high
// word, ~i in the low word, so that the largest key is the first maximum in
// y-major order; 0 for pixels within kmin of DC (and NaN magnitudes).
__device__ __forceinline__ unsigned long long peak_key(float2 c, int x, int y, int w, int h, double kmin,
                                                       unsigned i) {
    const int dx2 = (x < w / 2) ? (x * x) : ((x - w) * (x - w));
    const int dy2 = (y < h / 2) ? (y * y) : ((y - h) * (y - h));
    const double rad = sqrt((double)(dx2 + dy2));
    const float a = (float)sqrt((double)c.x * c.x + (double)c.y * c.y);
    if (!(rad > kmin) || a != a) return 0ull;
    return ((unsigned long long)__float_as_uint(a) << 32) | (unsigned long long)(0xFFFFFFFFu - i);
}

// BandSeparation.separateBands accumulation (BandSeparation.java:157-163,
// AbstractVectorCplx.axpy): acc += s * x in the reference's float order.
__device__ __forceinline__ float2 sep_axpy(float2 acc, float2 s, float2 x) {
#pragma clang fp contract(off)
    return make_float2(acc.x + ((s.x * x.x) - (s.y * x.y)), acc.y + ((s.y * x.x) + (s.x * x.y)));
}

// Multiply by the conjugate: Cplx.Float.mult(val.conj()) (Cplx.java:298-300)
__device__ __forceinline__ float2 mul_conj_java(float2 v, float2 o) {
#pragma clang fp contract(off)
    const float cr = o.x, ci = -o.y;
    return make_float2(v.x * cr - v.y * ci, v.x * ci + v.y * cr);
}
// AbstractVectorCplx.times(in, conj=true) (AbstractVectorCplx.java:281-295):
// this = multReal(in.re, -in.im, this.re, this.im), ...
__device__ __forceinline__ float2 times_conj_java(float2 y, float2 x) {
#pragma clang fp contract(off)
    const float xr = x.x, xi = -x.y;
    return make_float2((xr * y.x) - (xi * y.y), (xi * y.x) + (xr * y.y));
}

// Inverse of the pasteFreq index map (BasicVectors.java:273-278) along one
// axis: output index -> input index, or -1 where nothing is pasted.
// Valid for wo >= 2*wi, where the forward map is injective.
__device__ __forceinline__ int paste_src(int o, int wo, int wi, int off) {
    int p = (o - off) % wo;
    if (p < 0) p += wo;
    if (p < wi / 2) return p;
    if (p >= wi / 2 + wo / 2 && p < wi + wo / 2) return p - wo / 2;
    return -1;
}

}  // namespace fsim
