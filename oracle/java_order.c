/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * FP32 "Java-order" CPU restatement of fairSIM's 2D SR-SIM reconstruction
 * path.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load this library, and only as the checker / CPU baseline; the product
 * (libfsim.so) never calls it.
 *
 * Every function follows the Java text it cites (paths relative to the
 * reference tree).  Java float semantics are reproduced: binary32 per
 * operation, no FMA contraction (build with -ffp-contract=off), transcendental
 * functions and radii in double, `float *= double` evaluated in double and
 * rounded once.
 *
 * Parity status per function (how each is pinned is listed in tests/):
 *   ref_fsin/ref_fcos            pinned: max error vs libm sin/cos (MTool.main claim)
 *   ref_fade_border              pinned: closed form at border/interior pixels
 *   ref_otf_*                    pinned: scalar evaluation of the Java formulas
 *   ref_wiener_denominator       pinned: scalar re-evaluation at random pixels
 *   ref_separation_matrix        pinned: S*M = I, pseudo-inverse identity, KAT
 *   ref_separate_bands           pinned: known-answer synthetic bands
 *   ref_paste_freq               pinned: brute force vs an independent map
 *   ref_fourier_shift            pinned: integer shift == circular roll
 *   ref_fft2d                    pinned: delta/constant/cosine closed forms,
 *                                Parseval, numpy FP64.  The bit-level rounding
 *                                of JTransforms 3.1 (not vendored) is
 *                                "parity unpinned".
 *   ref_locate_peak              pinned: brute force on tiny inputs
 *   ref_recon_fast               pinned end to end against the FP64 numpy
 *                                restatement (rel-L2) and a noise-free KAT.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { float re, im; } cf;
typedef struct { double re, im; } cd;

static int g_threads = 1;

void ref_set_threads(int n) {
    g_threads = n < 1 ? 1 : n;
#ifdef _OPENMP
    omp_set_num_threads(g_threads);
#endif
}

int ref_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------ */
/* MTool.fsin / fcos / aproxSin01  (linalg/MTool.java:34-59)           */
/* ------------------------------------------------------------------ */
static double aprox_sin01(double x) {
    double x2 = x * x;
    return ((((.00015148419 * x2 - .00467376557) * x2 + .07968967928) * x2
             - .64596371106) * x2 + 1.57079631847) * x;
}

double ref_fsin(double x) {
    x = x / (M_PI / 2);
    x -= floor(0.25 * x) * 4;
    if (x > 3) return -aprox_sin01(4 - x);
    if (x > 1) return aprox_sin01(2 - x);
    return aprox_sin01(x);
}

double ref_fcos(double x) { return ref_fsin(x + M_PI / 2); }

/* ------------------------------------------------------------------ */
/* SimUtils.fadeBorderCos  (sim_algorithm/SimUtils.java:162-185)       */
/* applied to a vector pre-set to 1 (ReconstructionRunner.java:781-783)*/
/* ------------------------------------------------------------------ */
void ref_fade_border(float *dat, int w, int h, int px) {
    const double fac = 1. / px * M_PI / 2.;
    for (int y = 0; y < px; y++)
        for (int x = 0; x < w; x++)
            dat[x + y * w] = (float)((double)dat[x + y * w] * pow(sin(y * fac), 2));
    for (int y = h - px; y < h; y++)
        for (int x = 0; x < w; x++)
            dat[x + y * w] = (float)((double)dat[x + y * w] * pow(sin((h - y - 1) * fac), 2));
    for (int y = 0; y < h; y++)
        for (int x = 0; x < px; x++)
            dat[x + y * w] = (float)((double)dat[x + y * w] * pow(sin(x * fac), 2));
    for (int y = 0; y < h; y++)
        for (int x = w - px; x < w; x++)
            dat[x + y * w] = (float)((double)dat[x + y * w] * pow(sin((w - x - 1) * fac), 2));
}

/* ------------------------------------------------------------------ */
/* OtfProvider (sim_algorithm/OtfProvider.java)                        */
/* Tables are complex (vals, valsAtt) and real (valsOnlyAtt), 1 band.  */
/* ------------------------------------------------------------------ */

/* OtfProvider.valIdealOTF  (OtfProvider.java:74-78) */
double ref_val_ideal_otf(double dist) {
    if ((dist < 0) || (dist > 1)) return 0;
    return (2 / M_PI) * (acos(dist) - dist * sqrt(1 - dist * dist));
}

/* OtfProvider.valAttenuation  (OtfProvider.java:238-240) */
float ref_val_attenuation(double dist, double str, double fwhm) {
    return (float)(1 - str * exp(-(pow(dist, 2)) / (2 * pow(fwhm / 2.355, 2))));
}

/* OtfProvider.fromEstimate  (OtfProvider.java:96-127): fills vals (complex,
 * n samples), returns cutoff and the table's cycles/micron per sample. */
void ref_otf_from_estimate(double na, double lambda, double a, int n,
                           cf *vals, double *cutoff, double *cpm) {
    double cut = 1000 / (lambda / na / 2);
    *cutoff = cut;
    *cpm = cut / n;
    for (int i = 0; i < n; i++) {
        double v = i / (double)n;
        float r = (float)(ref_val_ideal_otf(v) * (pow(a, v)));
        vals[i].re = r;
        vals[i].im = 0;
    }
}

/* OtfProvider.setAttenuation  (OtfProvider.java:276-287) */
void ref_otf_set_attenuation(const cf *vals, int n, double cpm, double str,
                             double fwhm, cf *vals_att, float *vals_only_att) {
    for (int v = 0; v < n; v++) {
        double dist = v * cpm;
        vals_only_att[v] = ref_val_attenuation(dist, str, fwhm);
        vals_att[v].re = vals[v].re * vals_only_att[v];
        vals_att[v].im = vals[v].im * vals_only_att[v];
    }
}

/* OtfProvider.getOtfVal  (OtfProvider.java:189-222), including the lPos
 * quirk: both interpolation weights multiply the lower sample. */
cf ref_get_otf_val(const cf *tab, int n, double cpm, double cutoff, double cycl) {
    cf z = {0.f, 0.f};
    if (cycl >= cutoff) return z;
    const double pos = cycl / cpm;
    if (ceil(pos) >= n) return z;
    int lpos = (int)floor(pos);
    float f = (float)(pos - lpos);
    float omf = 1 - f;
    cf l = {tab[lpos].re * omf, tab[lpos].im * omf};
    cf hh = {tab[lpos].re * f, tab[lpos].im * f};
    cf r = {l.re + hh.re, l.im + hh.im};
    return r;
}

/* OtfProvider.getAttVal  (OtfProvider.java:246-269), same quirk */
float ref_get_att_val(const float *only_att, double cpm, double cutoff, double cycl) {
    if (cycl >= cutoff) return 1.f;
    double pos = cycl / cpm;
    int lpos = (int)floor(pos);
    float f = (float)(pos - lpos);
    float retl = only_att[lpos] * f;
    float reth = only_att[lpos] * (1 - f);
    return retl + reth;
}

/* OtfProvider.otfToVector  (OtfProvider.java:331-373).  write!=0 writes the
 * OTF, otherwise multiplies the vector by conj(OTF). */
void ref_otf_to_vector(cf *vec, int w, int h, const cf *tab, int n, double cpm,
                       double cutoff, double vec_cpm, double kx, double ky, int write) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            double xh = (x < w / 2) ? (x) : (x - w);
            double yh = (y < h / 2) ? (-y) : (h - y);
            double rad = sqrt((xh - kx) * (xh - kx) + (yh - ky) * (yh - ky));
            double cycl = rad * vec_cpm;
            cf *p = &vec[x + y * w];
            if (cycl > cutoff) { p->re = 0; p->im = 0; }
            if (cycl <= cutoff) {
                cf val = ref_get_otf_val(tab, n, cpm, cutoff, cycl);
                if (!write) {
                    /* Cplx.Float.mult(val.conj()) (linalg/Cplx.java:298-300) */
                    float cr = val.re, ci = -val.im;
                    float r = p->re * cr - p->im * ci;
                    float i = p->re * ci + p->im * cr;
                    p->re = r; p->im = i;
                } else {
                    *p = val;
                }
            }
        }
    }
}

/* OtfProvider.writeApoVector (real)  (OtfProvider.java:469-492) */
void ref_write_apo(float *vec, int w, int h, double vec_cpm, double cutoff,
                   double bend, double apo_cut) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            double xh = (x < w / 2) ? (x) : (x - w);
            double yh = (y < h / 2) ? (-y) : (h - y);
            double rad = sqrt(xh * xh + yh * yh);
            double cycl = rad * vec_cpm;
            double frac = cycl / (cutoff * apo_cut);
            double val = pow(ref_val_ideal_otf(frac), bend);
            vec[x + y * w] = (float)val;
        }
}

/* OtfProvider.writeAttenuationVector  (OtfProvider.java:503-529) */
void ref_write_attenuation(float *vec, int w, int h, double vec_cpm, double str,
                           double fwhm, double kx, double ky) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            double xh = (x < w / 2) ? (x) : (x - w);
            double yh = (y < h / 2) ? (-y) : (h - y);
            double rad = sqrt((xh - kx) * (xh - kx) + (yh - ky) * (yh - ky));
            double cycl = rad * vec_cpm;
            vec[x + y * w] = ref_val_attenuation(cycl, str, fwhm);
        }
}

/* OtfProvider.maskOtf  (OtfProvider.java:538-559) */
void ref_mask_otf(cf *vec, int w, int h, double vec_cpm, double cutoff, double kx, double ky) {
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            double xh = (x < w / 2) ? (x) : (x - w);
            double yh = (y < h / 2) ? (-y) : (h - y);
            double rad = sqrt((xh - kx) * (xh - kx) + (yh - ky) * (yh - ky)) * vec_cpm;
            if (rad > cutoff) { vec[x + y * w].re = 0; vec[x + y * w].im = 0; }
        }
}

/* WienerFilter.updateCache + addWienerDenominator + getDenominator
 * (sim_algorithm/WienerFilter.java:46-122).  px/py hold the band-1 shift per
 * direction; band b uses b*px (SimParam.Dir.px, SimParam.java:552-562).
 * Output is the reciprocal denominator 1/(sum + (float)(w*w)). */
void ref_wiener_denominator(float *out, int w, int h, const cf *vals,
                            const float *only_att, int n, double cpm, double cutoff,
                            double vec_cpm, int ndir, int nband, const double *px,
                            const double *py, int use_att, double wparam) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            float acc = 0.f;
            double xh = (x < w / 2) ? (x) : (x - w);
            double yh = (y < h / 2) ? (-y) : (h - y);
            for (int d = 0; d < ndir; d++)
                for (int b = 0; b < nband; b++) {
                    double bx = px[d] * b, by = py[d] * b;
                    double rad1 = sqrt((xh - bx) * (xh - bx) + (yh - by) * (yh - by)) * vec_cpm;
                    double rad2 = sqrt((xh + bx) * (xh + bx) + (yh + by) * (yh + by)) * vec_cpm;
                    cf o1 = ref_get_otf_val(vals, n, cpm, cutoff, rad1);
                    cf o2 = ref_get_otf_val(vals, n, cpm, cutoff, rad2);
                    float v1 = o1.re * o1.re + o1.im * o1.im;
                    float v2 = o2.re * o2.re + o2.im * o2.im;
                    if (use_att) {
                        v1 *= ref_get_att_val(only_att, cpm, cutoff, rad1);
                        v2 *= ref_get_att_val(only_att, cpm, cutoff, rad2);
                    }
                    acc = acc + v1 + v2;
                }
            out[x + y * w] = 1 / (acc + (float)(wparam * wparam));
        }
}

/* ------------------------------------------------------------------ */
/* Complex double helpers as in linalg/Cplx.java:254-286               */
/* ------------------------------------------------------------------ */
static cd cd_add(cd a, cd b) { cd r = {a.re + b.re, a.im + b.im}; return r; }
static cd cd_sub(cd a, cd b) { cd r = {a.re - b.re, a.im - b.im}; return r; }
static cd cd_mult(cd a, cd b) {
    cd r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static cd cd_reciproc(cd a) {
    double s = a.re * a.re + a.im * a.im;
    cd r = {a.re / s, -a.im / s};
    return r;
}
static cd cd_div(cd a, cd b) { return cd_mult(a, cd_reciproc(b)); }
static double cd_hypot(cd a) { return hypot(a.re, a.im); }

/* MatrixComplex (linalg/MatrixComplex.java), row-major m x n arrays */
static void mat_mult(const cd *a, int am, int an, const cd *b, int bn, cd *c) {
    for (int i = 0; i < am * bn; i++) { c[i].re = 0; c[i].im = 0; }
    for (int i = 0; i < am; i++)
        for (int k = 0; k < an; k++)
            for (int j = 0; j < bn; j++)
                c[i * bn + j] = cd_add(c[i * bn + j], cd_mult(a[i * an + k], b[k * bn + j]));
}

/* MatrixComplex.pivot  (MatrixComplex.java:56-77): compares magnitudes of the
 * unpermuted matrix, swaps rows of the permutation only. */
static void mat_pivot(const cd *v, int n, cd *p) {
    for (int i = 0; i < n * n; i++) { p[i].re = 0; p[i].im = 0; }
    for (int i = 0; i < n; i++) p[i * n + i].re = 1;
    for (int i = 0; i < n; i++) {
        int max_j = i;
        for (int j = i; j < n; j++)
            if (cd_hypot(v[j * n + i]) > cd_hypot(v[max_j * n + i])) max_j = j;
        if (max_j != i)
            for (int k = 0; k < n; k++) {
                cd t = p[i * n + k];
                p[i * n + k] = p[max_j * n + k];
                p[max_j * n + k] = t;
            }
    }
}

/* MatrixComplex.decomposeLU + inverse  (MatrixComplex.java:81-104, 212-237) */
static void mat_inverse(const cd *a, int n, cd *x) {
    cd *U = calloc(n * n, sizeof(cd)), *L = calloc(n * n, sizeof(cd));
    cd *P = calloc(n * n, sizeof(cd)), *Ap = calloc(n * n, sizeof(cd));
    cd *y = calloc(n * n, sizeof(cd));
    for (int i = 0; i < n; i++) L[i * n + i].re = 1;
    mat_pivot(a, n, P);
    mat_mult(P, n, n, a, n, Ap);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            if (j <= i) {
                cd s = {0, 0};
                for (int k = 0; k < j; k++) s = cd_add(s, cd_mult(L[j * n + k], U[k * n + i]));
                U[j * n + i] = cd_sub(Ap[j * n + i], s);
            }
            if (j >= i) {
                cd s = {0, 0};
                for (int k = 0; k < i; k++) s = cd_add(s, cd_mult(L[j * n + k], U[k * n + i]));
                L[j * n + i] = cd_div(cd_sub(Ap[j * n + i], s), U[i * n + i]);
            }
        }
    for (int k = 0; k < n; k++)
        for (int i = 0; i < n; i++) {
            cd res = P[i * n + k];
            for (int j = 0; j < i; j++) res = cd_sub(res, cd_mult(L[i * n + j], y[j * n + k]));
            y[i * n + k] = cd_div(res, L[i * n + i]);
        }
    for (int k = 0; k < n; k++)
        for (int i = n - 1; i >= 0; i--) {
            cd res = y[i * n + k];
            for (int j = i + 1; j < n; j++) res = cd_sub(res, cd_mult(U[i * n + j], x[j * n + k]));
            x[i * n + k] = cd_div(res, U[i * n + i]);
        }
    free(U); free(L); free(P); free(Ap); free(y);
}

/* BandSeparation.createSeparationMatrix (BandSeparation.java:40-114) with the
 * phase expansion of separateBands (:143-146).
 * phases: P absolute band-1 phases; fac: nband factors or NULL (defaults
 * {1, .5, ...}; otherwise fac[b>0] is multiplied by 0.5).
 * out: C x P complex (C = 2*nband-1), row-major, interleaved (re, im). */
int ref_separation_matrix(const double *phases, int nphase, int nband,
                          const double *fac_in, double *out) {
    int C = 2 * nband - 1;
    if (nband < 2 || nphase < C) return -1;
    double fac[16];
    for (int i = 0; i < nband; i++) {
        if (!fac_in) fac[i] = (i == 0) ? 1.0 : .5;
        else fac[i] = (i == 0) ? fac_in[0] : fac_in[i] * 0.5;
    }
    cd *M = calloc(nphase * C, sizeof(cd));
    for (int i = 0; i < nphase; i++) {
        M[i * C + 0].re = fac[0];
        M[i * C + 0].im = 0;
        for (int j = 0; j < nband - 1; j++) {
            double ph = phases[i] * (j + 1);
            M[i * C + 2 * j + 1].re = fac[j + 1] * cos(ph);
            M[i * C + 2 * j + 1].im = fac[j + 1] * sin(ph);
            M[i * C + 2 * j + 2].re = fac[j + 1] * cos(-ph);
            M[i * C + 2 * j + 2].im = fac[j + 1] * sin(-ph);
        }
    }
    cd *inv = calloc(C * nphase, sizeof(cd));
    if (nphase == C) {
        mat_inverse(M, C, inv);
    } else {
        /* MatrixComplex.pseudoInverse (MatrixComplex.java:256-264) */
        cd *adj = calloc(C * nphase, sizeof(cd)), *axa = calloc(C * C, sizeof(cd));
        cd *ai = calloc(C * C, sizeof(cd));
        for (int i = 0; i < nphase; i++)
            for (int j = 0; j < C; j++) {
                adj[j * nphase + i].re = M[i * C + j].re;
                adj[j * nphase + i].im = -M[i * C + j].im;
            }
        mat_mult(adj, C, nphase, M, C, axa);
        mat_inverse(axa, C, ai);
        mat_mult(ai, C, C, adj, nphase, inv);
        free(adj); free(axa); free(ai);
    }
    for (int i = 0; i < C * nphase; i++) {
        out[2 * i] = inv[i].re;
        out[2 * i + 1] = inv[i].im;
    }
    free(M); free(inv);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Vector operations (linalg/AbstractVectorCplx.java, BasicVectors.java)*/
/* ------------------------------------------------------------------ */

/* BandSeparation.separateBands inner loop (BandSeparation.java:157-163) with
 * AbstractVectorCplx.axpy(Cplx.Float) (AbstractVectorCplx.java:185-195).
 * S: C x P complex float (toFlt of the matrix). in: P vectors, out: C. */
void ref_separate_bands(const cf *const *in, cf *const *out, long n, const float *S,
                        int nphase, int ncomp) {
    for (int c = 0; c < ncomp; c++) memset(out[c], 0, sizeof(cf) * n);
    for (int p = 0; p < nphase; p++)
        for (int c = 0; c < ncomp; c++) {
            float are = S[2 * (c * nphase + p)], aim = S[2 * (c * nphase + p) + 1];
            cf *o = out[c];
            const cf *x = in[p];
#pragma omp parallel for schedule(static)
            for (long i = 0; i < n; i++) {
                o[i].re += (are * x[i].re) - (aim * x[i].im);
                o[i].im += (aim * x[i].re) + (are * x[i].im);
            }
        }
}

/* Vec2d.Cplx.pasteFreq  (BasicVectors.java:259-285) */
void ref_paste_freq(cf *out, int wo, int ho, const cf *in, int wi, int hi, int xoff, int yoff) {
    memset(out, 0, sizeof(cf) * (size_t)wo * ho);
    for (int y = 0; y < hi; y++)
        for (int x = 0; x < wi; x++) {
            int xo = (x < wi / 2) ? (x) : (x + wo / 2);
            int yo = (y < hi / 2) ? (y) : (y + ho / 2);
            xo = (xo + xoff + wo) % wo;
            yo = (yo + yoff + ho) % ho;
            out[xo + wo * yo] = in[x + wi * y];
        }
}

/* Vec2d.Cplx.fourierShift  (BasicVectors.java:309-339), fast sin/cos */
void ref_fourier_shift(cf *v, int N, double kx, double ky) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < N; y++)
        for (int x = 0; x < N; x++) {
            float pha = (float)(2 * M_PI * (kx * x + ky * y) / N);
            float co = (float)ref_fcos(pha);
            float si = (float)ref_fsin(pha);
            float re = v[y * N + x].re, im = v[y * N + x].im;
            v[y * N + x].re = (re * co) - (im * si);
            v[y * N + x].im = (im * co) + (re * si);
        }
}

/* ------------------------------------------------------------------ */
/* 2D complex FFT, FP32 (stands in for JTransforms FloatFFT_2D,        */
/* linalg/JTransformsConnector.java:76-90): forward unscaled,          */
/* e^{-2 pi i}; inverse e^{+2 pi i} scaled 1/n per 1-D pass.           */
/* Radix-2 decimation in time with twiddles from double precision.     */
/* ------------------------------------------------------------------ */
static void fft1d(cf *a, int n, long stride, int inverse, const cf *tw, cf *tmp) {
    /* gather into tmp in bit-reversed order */
    int lg = 0;
    while ((1 << lg) < n) lg++;
    for (int i = 0; i < n; i++) {
        int r = 0;
        for (int b = 0; b < lg; b++) r |= ((i >> b) & 1) << (lg - 1 - b);
        tmp[r] = a[(long)i * stride];
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int s = 0; s < n; s += len)
            for (int k = 0; k < half; k++) {
                cf w = tw[k * step];
                if (inverse) w.im = -w.im;
                cf u = tmp[s + k], t = tmp[s + k + half];
                cf vv = {t.re * w.re - t.im * w.im, t.re * w.im + t.im * w.re};
                tmp[s + k].re = u.re + vv.re;
                tmp[s + k].im = u.im + vv.im;
                tmp[s + k + half].re = u.re - vv.re;
                tmp[s + k + half].im = u.im - vv.im;
            }
    }
    float sc = 1.0f / n;
    for (int i = 0; i < n; i++) {
        cf o = tmp[i];
        if (inverse) { o.re *= sc; o.im *= sc; }
        a[(long)i * stride] = o;
    }
}

static cf *make_twiddles(int n) {
    cf *tw = malloc(sizeof(cf) * (n / 2 + 1));
    for (int k = 0; k <= n / 2; k++) {
        double ang = -2.0 * M_PI * k / n;
        tw[k].re = (float)cos(ang);
        tw[k].im = (float)sin(ang);
    }
    return tw;
}

void ref_fft2d(cf *v, int w, int h, int inverse) {
    cf *tww = make_twiddles(w), *twh = make_twiddles(h);
    /* rows, then columns (FloatFFT_2D order) */
#pragma omp parallel
    {
        cf *tmp = malloc(sizeof(cf) * (w > h ? w : h));
#pragma omp for schedule(static)
        for (int y = 0; y < h; y++) fft1d(v + (long)y * w, w, 1, inverse, tww, tmp);
#pragma omp for schedule(static)
        for (int x = 0; x < w; x++) fft1d(v + x, h, w, inverse, twh, tmp);
        free(tmp);
    }
    free(tww); free(twh);
}

/* ------------------------------------------------------------------ */
/* Fast-shift reconstruction: tests/TestAccelSpeed.java:216-252 (input) */
/* and :563-693 (reconstruction), which is also the dataflow of         */
/* livemode/ReconstructionRunner.java:456-542.                          */
/* ------------------------------------------------------------------ */
typedef struct {
    int n;            /* raw width = height */
    int ndir, nphase, nband;
    double px[8], py[8];           /* band-1 shift per direction (reference coords) */
    double phases[8][16];          /* absolute phases (getPhases) per direction */
    double modul[8][4];            /* raw modulations per direction and band */
    /* OTF */
    int otf_n;
    double otf_cpm, otf_cutoff;    /* table spacing, cutoff (cycles/micron) */
    double vec_cpm;                /* image cycles/micron = 1/(n*pixel) */
    double att_str, att_fwhm;
    int use_att;
    double wiener, apo_bend, apo_cut;
    int fade_px;                   /* 10 in the reference; 0 = no fade */
    int fast_shift;                /* 1: SimUtils fast mode, 0: slow (2N FFTs) */
} ref_recon_params;

/* SimParam.Dir.getModulations clamp (SimParam.java:598-607) */
static double clamp_mod(double m) {
    if (m < 0.3) m = 0.3;
    if (m > 1.1) m = 1.1;
    return m;
}

/* Build the per-channel tables of setupFilters / TestAccelSpeed:
 * OTF with attenuation (N^2, complex), reciprocal Wiener denominator (2N^2),
 * apodization (2N^2). */
void ref_build_tables(const ref_recon_params *p, const cf *vals, cf *otf_vec,
                      float *wiener, float *apo) {
    int n = p->n, on = p->otf_n;
    cf *vatt = malloc(sizeof(cf) * on);
    float *oatt = malloc(sizeof(float) * on);
    ref_otf_set_attenuation(vals, on, p->otf_cpm, p->att_str, p->att_fwhm, vatt, oatt);
    ref_otf_to_vector(otf_vec, n, n, p->use_att ? vatt : vals, on, p->otf_cpm,
                      p->otf_cutoff, p->vec_cpm, 0, 0, 1);
    ref_wiener_denominator(wiener, 2 * n, 2 * n, vals, oatt, on, p->otf_cpm, p->otf_cutoff,
                           p->vec_cpm, p->ndir, p->nband, p->px, p->py, p->use_att, p->wiener);
    ref_write_apo(apo, 2 * n, 2 * n, p->vec_cpm, p->otf_cutoff, p->apo_bend, p->apo_cut);
    free(vatt); free(oatt);
}

/* SimUtils.pasteAndFourierShift  (SimUtils.java:64-91) */
static void paste_and_shift(cf *in, cf *out, int n, double kx, double ky, int fast) {
    if (!fast) {
        ref_paste_freq(out, 2 * n, 2 * n, in, n, n, 0, 0);
        ref_fft2d(out, 2 * n, 2 * n, 1);
        ref_fourier_shift(out, 2 * n, kx, -ky);
        ref_fft2d(out, 2 * n, 2 * n, 0);
    } else {
        double xf = kx - floor(kx);
        double yf = -ky - floor(-ky);
        ref_fft2d(in, n, n, 1);
        ref_fourier_shift(in, n, xf, yf);
        ref_fft2d(in, n, n, 0);
        int xi = (int)floor(kx);
        int yi = (int)floor(-ky);
        ref_paste_freq(out, 2 * n, 2 * n, in, n, n, xi, yi);
    }
}

/* Full reconstruction of one raw stack (ndir*nphase frames of n*n uint16,
 * angle-major as in ReconstructionRunner.java:456-468) to a 2n x 2n float
 * image (real part of the final inverse FFT, TestAccelSpeed.java:697-703).
 * Tables from ref_build_tables. Returns 0 on success. */
int ref_recon_fast(const ref_recon_params *p, const uint16_t *raw, const cf *otf_vec,
                   const float *wiener, const float *apo, float *out) {
    const int n = p->n, D = p->ndir, P = p->nphase, B = p->nband, C = 2 * B - 1;
    const long nn = (long)n * n, mm = 4 * nn;
    if (C > 8 || P > 16 || D > 8) return -1;
    float *fade = malloc(sizeof(float) * nn);
    for (long i = 0; i < nn; i++) fade[i] = 1.f;
    if (p->fade_px > 0) ref_fade_border(fade, n, n, p->fade_px);

    cf *in = malloc(sizeof(cf) * nn * P);
    cf *sep = malloc(sizeof(cf) * nn * C);
    cf *shifted = malloc(sizeof(cf) * mm * C);
    cf *full = calloc(mm, sizeof(cf));
    const cf *inp[16];
    cf *sepp[8];
    for (int i = 0; i < P; i++) inp[i] = in + nn * i;
    for (int c = 0; c < C; c++) sepp[c] = sep + nn * c;

    for (int d = 0; d < D; d++) {
        /* input: setFrom16bitPixels, times(fade), fft2d(false) */
        for (int ph = 0; ph < P; ph++) {
            const uint16_t *r = raw + ((long)d * P + ph) * nn;
            cf *v = in + nn * ph;
            for (long i = 0; i < nn; i++) {
                v[i].re = (float)r[i] * fade[i];
                v[i].im = 0.f * fade[i];
            }
            ref_fft2d(v, n, n, 0);
        }
        /* separateBands(inFFT, sep, getPhases, nrBand, getModulations) */
        double mods[4], S[2 * 8 * 16];
        float Sf[2 * 8 * 16];
        for (int b = 0; b < B; b++) mods[b] = clamp_mod(p->modul[d][b]);
        if (ref_separation_matrix(p->phases[d], P, B, mods, S)) return -2;
        for (int i = 0; i < 2 * C * P; i++) Sf[i] = (float)S[i];
        ref_separate_bands(inp, sepp, nn, Sf, P, C);
        /* separate[i].timesConj(otfV[(i+1)/2]) -- single-band OTF */
        for (int c = 0; c < C; c++) {
            cf *s = sepp[c];
#pragma omp parallel for schedule(static)
            for (long i = 0; i < nn; i++) {
                float xr = otf_vec[i].re, xi = -otf_vec[i].im;
                float yr = s[i].re, yi = s[i].im;
                s[i].re = (xr * yr) - (xi * yi);
                s[i].im = (xi * yr) + (xr * yi);
            }
        }
        /* shifted[0].pasteFreq(separate[0]) */
        ref_paste_freq(shifted, 2 * n, 2 * n, sepp[0], n, n, 0, 0);
        for (int b = 1; b < B; b++) {
            int pos = b * 2, neg = b * 2 - 1;
            paste_and_shift(sepp[pos], shifted + mm * pos, n, p->px[d] * b, p->py[d] * b,
                            p->fast_shift);
            paste_and_shift(sepp[neg], shifted + mm * neg, n, -p->px[d] * b, -p->py[d] * b,
                            p->fast_shift);
        }
        for (int c = 0; c < C; c++) {
            const cf *s = shifted + mm * c;
#pragma omp parallel for schedule(static)
            for (long i = 0; i < mm; i++) {
                full[i].re += s[i].re;
                full[i].im += s[i].im;
            }
        }
    }
    /* fullResult.times(wienerDenom); fullResult.timesConj(apoVector) */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < mm; i++) {
        full[i].re *= wiener[i];
        full[i].im *= wiener[i];
        float xr = apo[i], xi = -0.0f;
        float yr = full[i].re, yi = full[i].im;
        full[i].re = (xr * yr) - (xi * yi);
        full[i].im = (xi * yr) + (xr * yi);
    }
    ref_fft2d(full, 2 * n, 2 * n, 1);
    for (long i = 0; i < mm; i++) out[i] = full[i].re;
    free(fade); free(in); free(sep); free(shifted); free(full);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Correlation.locatePeak  (sim_algorithm/Correlation.java:144-167)    */
/* res: {xPos, yPos, max, phase}                                       */
/* ------------------------------------------------------------------ */
void ref_locate_peak(const cf *vec, int w, int h, double kmin, double *res) {
    double xpos = -0, ypos = -1, max = -1, phase = 0;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            double rad = sqrt((double)(((x < w / 2) ? (x * x) : ((x - w) * (x - w))) +
                                       ((y < h / 2) ? (y * y) : ((y - h) * (y - h)))));
            cf v = vec[x + y * w];
            float a = (float)sqrt((double)v.re * v.re + (double)v.im * v.im);
            if ((rad > kmin) && (a > max)) {
                max = a;
                xpos = x;
                ypos = y;
                phase = (float)atan2(v.im, v.re);
            }
        }
    if (xpos > w / 2) xpos = xpos - w;
    if (ypos > h / 2) ypos = h - ypos;
    else ypos *= -1;
    res[0] = xpos; res[1] = ypos; res[2] = max; res[3] = phase;
}

/* AbstractVectorCplx.norm2 (AbstractVectorCplx.java:241-247): float square,
 * double accumulate */
double ref_norm2(const cf *v, long n) {
    double r = 0;
    for (long i = 0; i < n; i++) {
        r += v[i].re * v[i].re;
        r += v[i].im * v[i].im;
    }
    return r;
}

/* AbstractVectorCplx.sumElements (AbstractVectorCplx.java:353-362) */
void ref_sum_elements(const cf *v, long n, double *res) {
    double re = 0, im = 0;
    for (long i = 0; i < n; i++) { re += v[i].re; im += v[i].im; }
    res[0] = re; res[1] = im;
}

/* ------------------------------------------------------------------ */
/* Parameter estimation (sim_algorithm/Correlation.java)               */
/* ------------------------------------------------------------------ */

/* Cplx.Float.abs (linalg/Cplx.java:159-161) */
static float cf_abs(cf v) { return (float)sqrt((double)v.re * v.re + (double)v.im * v.im); }

/* Cplx.Float.div = mult(a, reciproc(b)) (linalg/Cplx.java:298-319) */
static cf cf_div(cf a, cf b) {
    float s = b.re * b.re + b.im * b.im;
    float rr = b.re / s, ri = -b.im / s;
    cf r = {a.re * rr - a.im * ri, a.re * ri + a.im * rr};
    return r;
}

/* Correlation.commonRegion (Correlation.java:227-288): the weights are
 * writeOtfVector (no attenuation) of band bn0 at 0 and (kx, ky) and of band
 * bn1 at 0 and (-kx, -ky); the pixel loop runs in the reference's order,
 * y outer, x inner, including the zeros written into band1 at the shifted
 * position of each DC-ring pixel. */
void ref_common_region(cf *b0, cf *b1, int w, int h, const cf *tab0, const cf *tab1, int ns,
                       double cpm, double cutoff, double vec_cpm, double kx, double ky,
                       double dist, double wlim, int divide) {
    long n = (long)w * h;
    cf *g0 = malloc(sizeof(cf) * n), *g1 = malloc(sizeof(cf) * n);
    cf *t0 = malloc(sizeof(cf) * n), *t1 = malloc(sizeof(cf) * n);
    ref_otf_to_vector(g0, w, h, tab0, ns, cpm, cutoff, vec_cpm, 0, 0, 1);
    ref_otf_to_vector(g1, w, h, tab1, ns, cpm, cutoff, vec_cpm, 0, 0, 1);
    ref_otf_to_vector(t0, w, h, tab0, ns, cpm, cutoff, vec_cpm, kx, ky, 1);
    ref_otf_to_vector(t1, w, h, tab1, ns, cpm, cutoff, vec_cpm, -kx, -ky, 1);
    const cf zero = {0.f, 0.f};
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            long p = x + (long)y * w;
            double rad = sqrt((double)(((x < w / 2) ? (x * x) : ((x - w) * (x - w))) +
                                       ((y < h / 2) ? (y * y) : ((y - h) * (y - h)))));
            double ratio = rad / sqrt(kx * kx + ky * ky);
            if (cf_abs(g0[p]) < wlim || cf_abs(t0[p]) < wlim) b0[p] = zero;
            else if (divide) b0[p] = cf_div(b0[p], g0[p]);
            if (cf_abs(g1[p]) < wlim || cf_abs(t1[p]) < wlim) b1[p] = zero;
            else if (divide) b1[p] = cf_div(b1[p], g1[p]);
            if (ratio < dist || ratio > (1 - dist)) {
                b0[p] = zero;
                b1[(x - (int)kx + w) % w + (long)((y + (int)ky + h) % h) * w] = zero;
            }
        }
    free(g0); free(g1); free(t0); free(t1);
}

/* sum over pixels of (b1 shifted by (kx, ky)) * conj(b0): Vec2d.Cplx
 * fourierShift (BasicVectors.java:309-339), AbstractVectorCplx.timesConj
 * (AbstractVectorCplx.java:281-304) and sumElements (:353-362), on copies,
 * one candidate at a time in pixel order. */
static void shifted_corr(const cf *b0, const cf *b1, int N, double kx, double ky, double *res) {
    double re = 0, im = 0;
    for (int y = 0; y < N; y++)
        for (int x = 0; x < N; x++) {
            long p = x + (long)y * N;
            float pha = (float)(2 * M_PI * (kx * x + ky * y) / N);
            float co = (float)ref_fcos(pha), si = (float)ref_fsin(pha);
            float sr = (b1[p].re * co) - (b1[p].im * si);
            float sm = (b1[p].im * co) + (b1[p].re * si);
            float xr = b0[p].re, xi = -b0[p].im;
            float qr = (xr * sr) - (xi * sm);
            float qi = (xi * sr) + (xr * sm);
            re += qr;
            im += qi;
        }
    res[0] = re;
    res[1] = im;
}

/* one candidate of shifted_corr: sum(shift(b1, kx, ky) * conj(b0)) */
void ref_shifted_corr(const cf *b0, const cf *b1, int N, double kx, double ky, double *res) {
    shifted_corr(b0, b1, N, kx, ky, res);
}

/* copies -> commonRegion(0.15, wlim, divide) -> inverse FFT of both */
static void corr_pair(const cf *band0, const cf *band1, int N, const cf *tab0, const cf *tab1,
                      int ns, double cpm, double cutoff, double vec_cpm, double kx, double ky,
                      double wlim, cf *b0, cf *b1) {
    long n = (long)N * N;
    memcpy(b0, band0, sizeof(cf) * n);
    memcpy(b1, band1, sizeof(cf) * n);
    ref_common_region(b0, b1, N, N, tab0, tab1, ns, cpm, cutoff, vec_cpm, kx, ky, 0.15, wlim, 1);
    ref_fft2d(b0, N, N, 1);
    ref_fft2d(b1, N, N, 1);
}

/* Correlation.fitPeak (Correlation.java:43-136): res = {kx, ky, phase, mag};
 * cntrl: NULL or 30 x 10 floats (row-major, width 30). */
void ref_fit_peak(const cf *band0, const cf *band1, int N, const cf *tab0, const cf *tab1, int ns,
                  double cpm, double cutoff, double vec_cpm, double kx, double ky, double wlim,
                  double search, float *cntrl, double *res) {
    long n = (long)N * N;
    cf *b0 = malloc(sizeof(cf) * n), *b1 = malloc(sizeof(cf) * n);
    double res_phase = 0, res_mag = 0;
    for (int iter = 0; iter < 3; iter++) {
        corr_pair(band0, band1, N, tab0, tab1, ns, cpm, cutoff, vec_cpm, kx, ky, wlim, b0, b1);
        double scal = 1. / ref_norm2(b0, n);
        double cre[100], cim[100];
        const double tkx = kx, tky = ky, ts = search;
#pragma omp parallel for schedule(dynamic)
        for (int p = 0; p < 100; p++) {
            int xi = p % 10, yi = p / 10;
            double s[2];
            shifted_corr(b0, b1, N, tkx + ((xi - 4.5) / 4.5) * ts, -(tky + ((yi - 4.5) / 4.5) * ts), s);
            /* Cplx.mult(sum, Cplx.Double(scal)) */
            cre[p] = s[0] * scal - s[1] * 0.0;
            cim[p] = s[0] * 0.0 + s[1] * scal;
        }
        double max = 0, min = 1.7976931348623157e308, nkx = 0, nky = 0;
        for (int yi = 0; yi < 10; yi++)
            for (int xi = 0; xi < 10; xi++) {
                double m = hypot(cre[yi * 10 + xi], cim[yi * 10 + xi]);
                if (m > max) {
                    max = m;
                    nkx = tkx + ((xi - 4.5) / 4.5) * ts;
                    nky = tky + ((yi - 4.5) / 4.5) * ts;
                    res_phase = atan2(cim[yi * 10 + xi], cre[yi * 10 + xi]);
                    res_mag = m;
                }
                if (m < min) min = m;
            }
        if (cntrl)
            for (int yi = 0; yi < 10; yi++)
                for (int xi = 0; xi < 10; xi++)
                    cntrl[xi + iter * 10 + yi * 30] =
                        (float)((hypot(cre[yi * 10 + xi], cim[yi * 10 + xi]) - min) / (max - min));
        kx = nkx;
        ky = nky;
        search /= 5;
    }
    res[0] = kx; res[1] = ky; res[2] = res_phase; res[3] = res_mag;
    free(b0); free(b1);
}

/* Correlation.getPeak (Correlation.java:183-209): res = {re, im} */
void ref_get_peak(const cf *band0, const cf *band1, int N, const cf *tab0, const cf *tab1, int ns,
                  double cpm, double cutoff, double vec_cpm, double kx, double ky, double wlim,
                  double *res) {
    long n = (long)N * N;
    cf *b0 = malloc(sizeof(cf) * n), *b1 = malloc(sizeof(cf) * n);
    corr_pair(band0, band1, N, tab0, tab1, ns, cpm, cutoff, vec_cpm, kx, ky, wlim, b0, b1);
    double s[2];
    shifted_corr(b0, b1, N, kx, -ky, s);
    double scal = 1. / ref_norm2(b0, n);
    res[0] = s[0] * scal - s[1] * 0.0;
    res[1] = s[0] * 0.0 + s[1] * scal;
    free(b0); free(b1);
}
