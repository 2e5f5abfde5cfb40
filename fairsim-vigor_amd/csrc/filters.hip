// OTF tables and per-pixel filters (OtfProvider, WienerFilter, SimUtils
// fadeBorderCos), band separation and peak location, on the device.
#include "fsim_internal.h"
#include "fft_core.h"
#include "filters.h"
#include "simmath.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace fsim {

// ---------------------------------------------------------------------------
// host: OTF estimate tables (OtfProvider.java:74-127, 238-240, 276-287)
// ---------------------------------------------------------------------------
static double val_ideal_otf(double dist) {
    if ((dist < 0) || (dist > 1)) return 0;
    return (2 / M_PI) * (std::acos(dist) - dist * std::sqrt(1 - dist * dist));
}

static float val_attenuation(double dist, double str, double fwhm) {
    return (float)(1 - str * std::exp(-(std::pow(dist, 2)) / (2 * std::pow(fwhm / 2.355, 2))));
}

// ---------------------------------------------------------------------------
// host: separation matrix (BandSeparation.java:40-114).  The inverse follows
// MatrixComplex's LU with its pivot rule (MatrixComplex.java:56-104,
// 212-264) so that the float matrix handed to the kernels matches the
// reference's toFlt() values.
// ---------------------------------------------------------------------------
struct C64 {
    double re, im;
};
static C64 c_add(C64 a, C64 b) { return {a.re + b.re, a.im + b.im}; }
static C64 c_sub(C64 a, C64 b) { return {a.re - b.re, a.im - b.im}; }
static C64 c_mul(C64 a, C64 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
static C64 c_div(C64 a, C64 b) {
    double s = b.re * b.re + b.im * b.im;
    return c_mul(a, C64{b.re / s, -b.im / s});
}

using CMat = std::vector<C64>;  // row-major

static CMat mat_mul(const CMat &a, int am, int an, const CMat &b, int bn) {
    CMat c((size_t)am * bn, C64{0, 0});
    for (int i = 0; i < am; i++)
        for (int k = 0; k < an; k++)
            for (int j = 0; j < bn; j++) c[i * bn + j] = c_add(c[i * bn + j], c_mul(a[i * an + k], b[k * bn + j]));
    return c;
}

static CMat mat_inverse(const CMat &a, int n) {
    // permutation: rows ordered by magnitudes of the unpermuted columns
    CMat P((size_t)n * n, C64{0, 0});
    for (int i = 0; i < n; i++) P[i * n + i].re = 1;
    for (int i = 0; i < n; i++) {
        int mj = i;
        for (int j = i; j < n; j++)
            if (std::hypot(a[j * n + i].re, a[j * n + i].im) > std::hypot(a[mj * n + i].re, a[mj * n + i].im))
                mj = j;
        if (mj != i)
            for (int k = 0; k < n; k++) std::swap(P[i * n + k], P[mj * n + k]);
    }
    CMat Ap = mat_mul(P, n, n, a, n);
    CMat L((size_t)n * n, C64{0, 0}), U((size_t)n * n, C64{0, 0});
    for (int i = 0; i < n; i++) L[i * n + i].re = 1;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            if (j <= i) {
                C64 s{0, 0};
                for (int k = 0; k < j; k++) s = c_add(s, c_mul(L[j * n + k], U[k * n + i]));
                U[j * n + i] = c_sub(Ap[j * n + i], s);
            }
            if (j >= i) {
                C64 s{0, 0};
                for (int k = 0; k < i; k++) s = c_add(s, c_mul(L[j * n + k], U[k * n + i]));
                L[j * n + i] = c_div(c_sub(Ap[j * n + i], s), U[i * n + i]);
            }
        }
    CMat y((size_t)n * n, C64{0, 0}), x((size_t)n * n, C64{0, 0});
    for (int k = 0; k < n; k++)
        for (int i = 0; i < n; i++) {
            C64 r = P[i * n + k];
            for (int j = 0; j < i; j++) r = c_sub(r, c_mul(L[i * n + j], y[j * n + k]));
            y[i * n + k] = c_div(r, L[i * n + i]);
        }
    for (int k = 0; k < n; k++)
        for (int i = n - 1; i >= 0; i--) {
            C64 r = y[i * n + k];
            for (int j = i + 1; j < n; j++) r = c_sub(r, c_mul(U[i * n + j], x[j * n + k]));
            x[i * n + k] = c_div(r, U[i * n + i]);
        }
    return x;
}

int separation_matrix_host(const double *phases, int nphase, int nband, const double *fac_in, double *out) {
    const int C = 2 * nband - 1;
    if (nband < 2) return fail(FSIM_ERR_ARG, "separation matrix: #bands < 2, not useful");
    if (nphase < C) return fail(FSIM_ERR_ARG, "separation matrix: not enough phases for #bands");
    std::vector<double> fac(nband);
    for (int i = 0; i < nband; i++)
        fac[i] = fac_in ? (i == 0 ? fac_in[0] : fac_in[i] * 0.5) : (i == 0 ? 1.0 : 0.5);
    CMat M((size_t)nphase * C, C64{0, 0});
    for (int i = 0; i < nphase; i++) {
        M[i * C] = C64{fac[0], 0};
        for (int j = 0; j < nband - 1; j++) {
            const double ph = phases[i] * (j + 1);
            M[i * C + 2 * j + 1] = C64{fac[j + 1] * std::cos(ph), fac[j + 1] * std::sin(ph)};
            M[i * C + 2 * j + 2] = C64{fac[j + 1] * std::cos(-ph), fac[j + 1] * std::sin(-ph)};
        }
    }
    CMat inv;
    if (nphase == C) {
        inv = mat_inverse(M, C);
    } else {
        CMat adj((size_t)C * nphase);
        for (int i = 0; i < nphase; i++)
            for (int j = 0; j < C; j++) adj[j * nphase + i] = C64{M[i * C + j].re, -M[i * C + j].im};
        CMat axa = mat_mul(adj, C, nphase, M, C);
        CMat ai = mat_inverse(axa, C);
        inv = mat_mul(ai, C, C, adj, nphase);
    }
    for (int i = 0; i < C * nphase; i++) {
        out[2 * i] = inv[i].re;
        out[2 * i + 1] = inv[i].im;
    }
    return FSIM_OK;
}

// fadeBorderCos factors (SimUtils.java:162-185): fac[i] = sin^2 term of the
// border pixel i (top/left side) or of the mirrored pixel (bottom/right), 1
// in the interior.  Computed in double on the host.
void fade_factors_host(int n, int px, double *lo, double *hi) {
    const double fac = 1. / px * M_PI / 2.;
    for (int i = 0; i < n; i++) {
        lo[i] = (px > 0 && i < px) ? std::pow(std::sin(i * fac), 2) : 1.0;
        hi[i] = (px > 0 && i >= n - px) ? std::pow(std::sin((n - i - 1) * fac), 2) : 1.0;
    }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
#define GRID_STRIDE(i, n) \
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (size_t)gridDim.x * blockDim.x)

static inline dim3 grid_for(size_t n) {
    size_t b = (n + 1023) / 1024;
    if (b < 1) b = 1;
    if (b > 65535 * 8) b = 65535 * 8;
    return dim3((unsigned)b);
}

__global__ void k_otf_to_vector(float2 *__restrict__ v, int w, int h, const float2 *__restrict__ tab, int ns,
                                double cpm, double cutoff, double kx, double ky, double vec_cpm, int write) {
    GRID_STRIDE(i, (size_t)w * h) {
        const int x = (int)(i % w), y = (int)(i / w);
        const float2 o = otf_pixel(tab, ns, cpm, cutoff, x, y, w, h, kx, ky, vec_cpm);
        if (write) {
            v[i] = o;
        } else {
            const double cycl = cycl_at(x, y, w, h, kx, ky, vec_cpm);
            v[i] = (cycl > cutoff) ? make_float2(0.f, 0.f) : mul_conj_java(v[i], o);
        }
    }
}

__global__ void k_mask_otf(float2 *__restrict__ v, int w, int h, double cutoff, double kx, double ky,
                           double vec_cpm) {
    GRID_STRIDE(i, (size_t)w * h) {
        const int x = (int)(i % w), y = (int)(i / w);
        if (cycl_at(x, y, w, h, kx, ky, vec_cpm) > cutoff) v[i] = make_float2(0.f, 0.f);
    }
}

__device__ __forceinline__ double ideal_otf_d(double dist) {
#pragma clang fp contract(off)
    if ((dist < 0) || (dist > 1)) return 0;
    return (2 / M_PI) * (acos(dist) - dist * sqrt(1 - dist * dist));
}

// writeApoVector (OtfProvider.java:437-492); is_cplx selects the vector type
__global__ void k_apo(void *__restrict__ v, int is_cplx, int w, int h, double vec_cpm, double cutoff, double bend,
                      double apo_cut) {
#pragma clang fp contract(off)
    GRID_STRIDE(i, (size_t)w * h) {
        const int x = (int)(i % w), y = (int)(i / w);
        const double cycl = cycl_at(x, y, w, h, 0, 0, vec_cpm);
        const double frac = cycl / (cutoff * apo_cut);
        const float val = (float)pow(ideal_otf_d(frac), bend);
        if (is_cplx)
            reinterpret_cast<float2 *>(v)[i] = make_float2(val, 0.f);
        else
            reinterpret_cast<float *>(v)[i] = val;
    }
}

// writeAttenuationVector (OtfProvider.java:503-529)
__global__ void k_attenuation(void *__restrict__ v, int is_cplx, int w, int h, double vec_cpm, double str,
                              double fwhm, double kx, double ky) {
#pragma clang fp contract(off)
    GRID_STRIDE(i, (size_t)w * h) {
        const int x = (int)(i % w), y = (int)(i / w);
        const float val = attenuation_pixel(x, y, w, h, vec_cpm, str, fwhm, kx, ky);
        if (is_cplx)
            reinterpret_cast<float2 *>(v)[i] = make_float2(val, 0.f);
        else
            reinterpret_cast<float *>(v)[i] = val;
    }
}

// WienerFilter.updateCache + getDenominator (WienerFilter.java:46-122)
__global__ void k_wiener(float *__restrict__ out, int w, int h, const float2 *__restrict__ vals,
                         const float *__restrict__ only_att, int ns, int otf_nb, double cpm, double cutoff,
                         double vec_cpm, ShiftSet sh, int use_att, float w2) {
#pragma clang fp contract(off)
    GRID_STRIDE(i, (size_t)w * h) {
        const float acc =
            wiener_sum((int)(i % w), (int)(i / w), w, h, vals, only_att, ns, otf_nb, cpm, cutoff, vec_cpm, sh, use_att);
        out[i] = 1 / (acc + w2);
    }
}

// SimUtils.fadeBorderCos applied in the reference's order: top, bottom,
// left, right; each `float *= double` rounds once.
__global__ void k_fade(float *__restrict__ v, int w, int h, const double *__restrict__ ylo,
                       const double *__restrict__ yhi, const double *__restrict__ xlo,
                       const double *__restrict__ xhi) {
    GRID_STRIDE(i, (size_t)w * h) {
        const int x = (int)(i % w), y = (int)(i / w);
        float d = v[i];
        d = (float)((double)d * ylo[y]);
        d = (float)((double)d * yhi[y]);
        d = (float)((double)d * xlo[x]);
        d = (float)((double)d * xhi[x]);
        v[i] = d;
    }
}

// BandSeparation.separateBands (BandSeparation.java:157-163): out[c] = 0,
// then out[c] += S[c][p] * in[p] for p in order, each axpy in Java float
// order (AbstractVectorCplx.java:185-195).  One pass reads each input once.
struct SepArgs {
    const float2 *in[8];
    float2 *out[8];
    float2 S[8 * 8];  // [c][p]
    int nphase, ncomp;
};
__global__ void k_separate(SepArgs a, size_t n) {
#pragma clang fp contract(off)
    GRID_STRIDE(i, n) {
        float2 x[8];
        for (int p = 0; p < a.nphase; p++) x[p] = a.in[p][i];
        for (int c = 0; c < a.ncomp; c++) {
            float2 acc = make_float2(0.f, 0.f);
            for (int p = 0; p < a.nphase; p++) acc = sep_axpy(acc, a.S[c * 8 + p], x[p]);
            a.out[c][i] = acc;
        }
    }
}

// Correlation.locatePeak (Correlation.java:144-167): maximum of
// (float)sqrt((double)re*re + (double)im*im) over pixels with rad > kMin, first
// maximum in y-major order.  Key = abs bits << 32 | ~index; max key wins.
__global__ void k_locate_peak(const float2 *__restrict__ v, int w, int h, double kmin,
                              unsigned long long *__restrict__ best) {
    __shared__ unsigned long long sh[256];
    unsigned long long k = 0;
    GRID_STRIDE(i, (size_t)w * h) {
        const int x = (int)(i % w), y = (int)(i / w);
        const unsigned long long key = peak_key(v[i], x, y, w, h, kmin, (unsigned)i);
        if (key > k) k = key;
    }
    sh[threadIdx.x] = k;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s && sh[threadIdx.x + s] > sh[threadIdx.x]) sh[threadIdx.x] = sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && sh[0]) atomicMax(best, sh[0]);
}

int wiener_device(float *out, int w, int h, const Otf &o, double vec_cpm, const ShiftSet &sh, int use_att,
                  double wparam, hipStream_t s) {
    hipLaunchKernelGGL(k_wiener, grid_for((size_t)w * h), dim3(256), 0, s, out, w, h, o.vals, o.only_att, o.ns,
                       o.nband, o.cpm, o.cutoff, vec_cpm, sh, use_att, (float)(wparam * wparam));
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

int otf_table_device(float2 *out, int w, int h, const float2 *tab, int ns, double cpm, double cutoff,
                     double vec_cpm, hipStream_t s) {
    hipLaunchKernelGGL(k_otf_to_vector, grid_for((size_t)w * h), dim3(256), 0, s, out, w, h, tab, ns, cpm, cutoff,
                       0.0, 0.0, vec_cpm, 1);
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

int write_apo_device(float *out, int w, int h, double vec_cpm, double cutoff, double bend, double apo_cut,
                     hipStream_t s) {
    hipLaunchKernelGGL(k_apo, grid_for((size_t)w * h), dim3(256), 0, s, (void *)out, 0, w, h, vec_cpm, cutoff, bend,
                       apo_cut);
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

}  // namespace fsim

using namespace fsim;

static Vec *V(fsim_vec h) { return reinterpret_cast<Vec *>(h); }

extern "C" {

int fsim_otf_from_estimate(double na, double lambda, double a, int ns, double att_str, double att_fwhm,
                           float *vals, float *vals_att, float *only_att, double *cutoff, double *cpm) {
    if ((a < 0) || (a > 1) || (na < 0.3) || (na > 2.2) || (lambda < 300) || (lambda > 1500))
        return fail(FSIM_ERR_ARG, "otf_from_estimate: unphysical input parameters");
    if (ns < 2 || !vals || !vals_att || !only_att || !cutoff || !cpm)
        return fail(FSIM_ERR_ARG, "otf_from_estimate: bad arguments");
    const double cut = 1000 / (lambda / na / 2);
    const double c = cut / ns;
    for (int i = 0; i < ns; i++) {
        const double v = i / (double)ns;
        const float r = (float)(val_ideal_otf(v) * (std::pow(a, v)));
        vals[2 * i] = r;
        vals[2 * i + 1] = 0;
    }
    for (int v = 0; v < ns; v++) {
        const double dist = v * c;
        only_att[v] = val_attenuation(dist, att_str, att_fwhm);
        vals_att[2 * v] = vals[2 * v] * only_att[v];
        vals_att[2 * v + 1] = vals[2 * v + 1] * only_att[v];
    }
    *cutoff = cut;
    *cpm = c;
    return FSIM_OK;
}

int fsim_otf_attenuation(int nband, int ns, double cpm, double att_str, double att_fwhm, const float *vals,
                         float *vals_att, float *only_att) {
    if (nband < 1 || ns < 2 || !(cpm > 0) || !vals || !vals_att || !only_att)
        return fail(FSIM_ERR_ARG, "otf_attenuation: bad arguments");
    for (int b = 0; b < nband; b++)
        for (int v = 0; v < ns; v++) {
            const size_t i = (size_t)b * ns + v;
            const double dist = v * cpm;
            only_att[i] = val_attenuation(dist, att_str, att_fwhm);
            vals_att[2 * i] = vals[2 * i] * only_att[i];
            vals_att[2 * i + 1] = vals[2 * i + 1] * only_att[i];
        }
    return FSIM_OK;
}

int fsim_otf_create(int nband, int ns, double cutoff, double cpm, const float *vals, const float *vals_att,
                    const float *only_att, fsim_otf *out) {
    int rc = ensure_init();
    if (rc) return rc;
    if (nband < 1 || ns < 2 || !vals || !vals_att || !only_att || !out || cpm <= 0)
        return fail(FSIM_ERR_ARG, "otf_create: bad arguments");
    for (size_t i = 0; i < (size_t)nband * ns; i++)
        if (!std::isfinite(vals[2 * i]) || !std::isfinite(vals[2 * i + 1]) || !std::isfinite(vals_att[2 * i]) ||
            !std::isfinite(vals_att[2 * i + 1]) || !std::isfinite(only_att[i]))
            return fail(FSIM_ERR_ARG, "otf_create: table values must be finite");
    fsim_otf_s *h = new fsim_otf_s();
    Otf &o = h->o;
    o.nband = nband;
    o.ns = ns;
    o.cutoff = cutoff;
    o.cpm = cpm;
    const size_t n = (size_t)nband * ns;
    if (hipMalloc(&o.vals, n * sizeof(float2)) != hipSuccess ||
        hipMalloc(&o.vals_att, n * sizeof(float2)) != hipSuccess ||
        hipMalloc(&o.only_att, n * sizeof(float)) != hipSuccess) {
        (void)hipFree(o.vals);
        (void)hipFree(o.vals_att);
        (void)hipFree(o.only_att);
        delete h;
        return fail(FSIM_ERR_OOM, "otf_create: device allocation failed");
    }
    FSIM_HIP(hipMemcpy(o.vals, vals, n * sizeof(float2), hipMemcpyHostToDevice));
    FSIM_HIP(hipMemcpy(o.vals_att, vals_att, n * sizeof(float2), hipMemcpyHostToDevice));
    FSIM_HIP(hipMemcpy(o.only_att, only_att, n * sizeof(float), hipMemcpyHostToDevice));
    *out = h;
    return FSIM_OK;
}

int fsim_otf_destroy(fsim_otf h) {
    if (!h) return FSIM_OK;
    FSIM_HIP(hipStreamSynchronize(stream()));
    (void)hipFree(h->o.vals);
    (void)hipFree(h->o.vals_att);
    (void)hipFree(h->o.only_att);
    delete h;
    return FSIM_OK;
}

int fsim_otf_to_vector(fsim_vec vh, fsim_otf oh, int band, double kx, double ky, double vec_cpm, int use_att,
                       int write) {
    int rc = ensure_init();
    if (rc) return rc;
    Vec *v = V(vh);
    if (!v || !oh || !v->is_cplx) return fail(FSIM_ERR_ARG, "otf_to_vector: needs a complex vector and an OTF");
    if (vec_cpm <= 0) return fail(FSIM_ERR_STATE, "otf_to_vector: vector pixel size not initialized");
    const Otf &o = oh->o;
    if (o.nband == 1) band = 0;  // single-band OTF: getOtfVal ignores the band
    if (band < 0 || band >= o.nband) return fail(FSIM_ERR_ARG, "otf_to_vector: band idx too high or <0");
    const float2 *tab = (use_att ? o.vals_att : o.vals) + (size_t)band * o.ns;
    hipLaunchKernelGGL(k_otf_to_vector, grid_for((size_t)v->w * v->h), dim3(256), 0, stream(), (float2 *)v->d,
                       v->w, v->h, tab, o.ns, o.cpm, o.cutoff, kx, ky, vec_cpm, write);
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

int fsim_mask_otf(fsim_vec vh, fsim_otf oh, double kx, double ky, double vec_cpm) {
    int rc = ensure_init();
    if (rc) return rc;
    Vec *v = V(vh);
    if (!v || !oh || !v->is_cplx) return fail(FSIM_ERR_ARG, "mask_otf: needs a complex vector and an OTF");
    hipLaunchKernelGGL(k_mask_otf, grid_for((size_t)v->w * v->h), dim3(256), 0, stream(), (float2 *)v->d, v->w,
                       v->h, oh->o.cutoff, kx, ky, vec_cpm);
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

int fsim_write_apo(fsim_vec vh, fsim_otf oh, double vec_cpm, double bend, double cutoff_fac) {
    int rc = ensure_init();
    if (rc) return rc;
    Vec *v = V(vh);
    if (!v || !oh) return fail(FSIM_ERR_ARG, "write_apo: null argument");
    if (vec_cpm <= 0) return fail(FSIM_ERR_STATE, "write_apo: vector pixel size not initialized");
    hipLaunchKernelGGL(k_apo, grid_for((size_t)v->w * v->h), dim3(256), 0, stream(), v->d, v->is_cplx, v->w, v->h,
                       vec_cpm, oh->o.cutoff, bend, cutoff_fac);
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

int fsim_write_attenuation(fsim_vec vh, double vec_cpm, double str, double fwhm, double kx, double ky) {
    int rc = ensure_init();
    if (rc) return rc;
    Vec *v = V(vh);
    if (!v) return fail(FSIM_ERR_ARG, "write_attenuation: null vector");
    if (vec_cpm <= 0) return fail(FSIM_ERR_STATE, "write_attenuation: vector pixel size not initialized");
    hipLaunchKernelGGL(k_attenuation, grid_for((size_t)v->w * v->h), dim3(256), 0, stream(), v->d, v->is_cplx, v->w,
                       v->h, vec_cpm, str, fwhm, kx, ky);
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

int fsim_wiener_denominator(fsim_vec vh, fsim_otf oh, double vec_cpm, int ndir, int nband, const double *px,
                            const double *py, int use_att, double wparam) {
    int rc = ensure_init();
    if (rc) return rc;
    Vec *v = V(vh);
    if (!v || !oh || v->is_cplx) return fail(FSIM_ERR_ARG, "wiener_denominator: needs a real vector and an OTF");
    if (ndir < 1 || ndir > kMaxDir || nband < 1 || nband > kMaxBand || !px || !py)
        return fail(FSIM_ERR_ARG, "wiener_denominator: bad direction/band count");
    if (oh->o.nband != 1 && oh->o.nband < nband) return fail(FSIM_ERR_ARG, "wiener_denominator: OTF has too few bands");
    ShiftSet sh{};
    sh.ndir = ndir;
    sh.nband = nband;
    for (int d = 0; d < ndir; d++) {
        sh.px[d] = px[d];
        sh.py[d] = py[d];
    }
    return wiener_device((float *)v->d, v->w, v->h, oh->o, vec_cpm, sh, use_att, wparam, stream());
}

int fsim_fade_border(fsim_vec vh, int px) {
    int rc = ensure_init();
    if (rc) return rc;
    Vec *v = V(vh);
    if (!v || v->is_cplx) return fail(FSIM_ERR_ARG, "fade_border: needs a real vector");
    if (px < 0 || 2 * px > v->w || 2 * px > v->h) return fail(FSIM_ERR_ARG, "fade_border: border too wide");
    std::vector<double> f(2 * (size_t)(v->w + v->h));
    double *ylo = f.data(), *yhi = ylo + v->h, *xlo = yhi + v->h, *xhi = xlo + v->w;
    fade_factors_host(v->h, px, ylo, yhi);
    fade_factors_host(v->w, px, xlo, xhi);
    std::lock_guard<std::mutex> lk(scratch_mutex());
    void *scr = nullptr;
    rc = scratch_d(f.size() * sizeof(double), &scr);
    if (rc) return rc;
    double *d = reinterpret_cast<double *>(scr);
    FSIM_HIP(hipMemcpyAsync(d, f.data(), f.size() * sizeof(double), hipMemcpyHostToDevice, stream()));
    hipLaunchKernelGGL(k_fade, grid_for((size_t)v->w * v->h), dim3(256), 0, stream(), (float *)v->d, v->w, v->h, d,
                       d + v->h, d + 2 * v->h, d + 2 * v->h + v->w);
    FSIM_HIP(hipGetLastError());
    FSIM_HIP(hipStreamSynchronize(stream()));
    return FSIM_OK;
}

int fsim_separation_matrix(const double *phases, int nphase, int nband, const double *fac, double *out) {
    if (!phases || !out) return fail(FSIM_ERR_ARG, "separation_matrix: null argument");
    return separation_matrix_host(phases, nphase, nband, fac, out);
}

int fsim_separate_bands(const fsim_vec *in, int nphase, const fsim_vec *out, int ncomp, const float *S) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!in || !out || !S || nphase < 1 || nphase > 8 || ncomp < 1 || ncomp > 8)
        return fail(FSIM_ERR_ARG, "separate_bands: bad arguments");
    SepArgs a{};
    a.nphase = nphase;
    a.ncomp = ncomp;
    Vec *v0 = V(in[0]);
    if (!v0 || !v0->is_cplx) return fail(FSIM_ERR_ARG, "separate_bands: inputs must be complex");
    for (int p = 0; p < nphase; p++) {
        Vec *v = V(in[p]);
        if (!v || !v->is_cplx || v->w != v0->w || v->h != v0->h)
            return fail(FSIM_ERR_ARG, "separate_bands: input size mismatch");
        a.in[p] = (const float2 *)v->d;
    }
    for (int c = 0; c < ncomp; c++) {
        Vec *v = V(out[c]);
        if (!v || !v->is_cplx || v->w != v0->w || v->h != v0->h)
            return fail(FSIM_ERR_ARG, "separate_bands: output size mismatch");
        for (int p = 0; p < nphase; p++)
            if (v->d == a.in[p]) return fail(FSIM_ERR_ARG, "separate_bands: output aliases an input");
        a.out[c] = (float2 *)v->d;
        for (int p = 0; p < nphase; p++)
            a.S[c * 8 + p] = make_float2(S[2 * (c * nphase + p)], S[2 * (c * nphase + p) + 1]);
    }
    const size_t n = (size_t)v0->w * v0->h;
    hipLaunchKernelGGL(k_separate, grid_for(n), dim3(256), 0, stream(), a, n);
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

int fsim_locate_peak(fsim_vec vh, double kmin, double *res) {
    int rc = ensure_init();
    if (rc) return rc;
    Vec *v = V(vh);
    if (!v || !v->is_cplx || !res) return fail(FSIM_ERR_ARG, "locate_peak: needs a complex vector");
    const int w = v->w, h = v->h;
    unsigned long long best = 0;
    {
        std::lock_guard<std::mutex> lk(scratch_mutex());
        void *scr = nullptr;
        rc = scratch_d(sizeof(unsigned long long), &scr);
        if (rc) return rc;
        unsigned long long *d = reinterpret_cast<unsigned long long *>(scr);
        FSIM_HIP(hipMemsetAsync(d, 0, sizeof(*d), stream()));
        size_t n = (size_t)w * h;
        unsigned nb = (unsigned)((n + 2047) / 2048);
        if (nb > 4096) nb = 4096;
        hipLaunchKernelGGL(k_locate_peak, dim3(nb ? nb : 1), dim3(256), 0, stream(), (const float2 *)v->d, w, h, kmin,
                           d);
        FSIM_HIP(hipGetLastError());
        FSIM_HIP(hipMemcpyAsync(&best, d, sizeof(best), hipMemcpyDeviceToHost, stream()));
        FSIM_HIP(hipStreamSynchronize(stream()));
    }
    double xpos = -0.0, ypos = -1, max = -1, phase = 0;
    if (best) {
        const unsigned idx = 0xFFFFFFFFu - (unsigned)(best & 0xFFFFFFFFull);
        const unsigned bits = (unsigned)(best >> 32);
        float a;
        std::memcpy(&a, &bits, sizeof(a));
        float2 c;
        FSIM_HIP(hipMemcpy(&c, (const float2 *)v->d + idx, sizeof(c), hipMemcpyDeviceToHost));
        max = a;
        xpos = idx % w;
        ypos = idx / w;
        phase = (float)std::atan2(c.y, c.x);
    }
    if (xpos > w / 2) xpos = xpos - w;
    if (ypos > h / 2)
        ypos = h - ypos;
    else
        ypos *= -1;
    res[0] = xpos;
    res[1] = ypos;
    res[2] = max;
    res[3] = phase;
    return FSIM_OK;
}

}  // extern "C"
