// SIM parameter estimation on the device: the per-pixel loops of
// sim_algorithm/Correlation.java.
//
//  fsim_common_region  Correlation.commonRegion (Correlation.java:227-288)
//  fsim_fit_peak       Correlation.fitPeak      (Correlation.java:43-136)
//  fsim_get_peak       Correlation.getPeak      (Correlation.java:183-209)
//
// commonRegion is one gather kernel over both bands.  fitPeak's 10 x 10
// candidate shifts per iteration (a copy, a Fourier shift, a conjugate
// multiply and a double-precision sum per candidate in the reference) are
// evaluated together: each block keeps its pixels of both bands in registers
// and computes every candidate's shifted conjugate product per pixel, with
// the reference's float-rounded shift phase and float products, summing in
// double; per-block partial sums are reduced in a fixed order.  getPeak is
// the same kernel with one candidate.
#include "fsim_internal.h"
#include "filters.h"
#include "simmath.h"

#include <cmath>
#include <mutex>
#include <vector>

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

struct RegionArgs {
    const float2 *tab0, *tab1;   // radial OTF tables of bands bn0 and bn1 (no attenuation)
    int ns;
    double cpm, cutoff, vec_cpm;
    double kx, ky, dist, wlim;
    int ikx, iky;                // (int)kx, (int)ky: truncated toward zero
    int divide;
};

// Correlation.commonRegion.  band0(x, y): zero where either OTF weight
// (centred at 0 and at (kx, ky)) is below the limit, else divided by the
// weight; zero in the ring ratio < dist or ratio > 1 - dist around DC.
// band1(x, y): the same with the weights centred at 0 and (-kx, -ky); the
// ring pixels (xq, yq) also zero band1 at ((xq - ikx) mod w, (yq + iky) mod h),
// so band1(x, y) is zero when its source pixel ((x + ikx) mod w,
// (y - iky) mod h) lies in the ring.  The reference's loop order decides only
// the sign of such zeros.
__global__ void k_common_region(float2 *__restrict__ b0, float2 *__restrict__ b1, int w, int h, RegionArgs a) {
    const double kmax = sqrt(a.kx * a.kx + a.ky * a.ky);
    const size_t n = (size_t)w * h;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % w), y = (int)(i / w);
        const float2 w0 = otf_pixel(a.tab0, a.ns, a.cpm, a.cutoff, x, y, w, h, 0.0, 0.0, a.vec_cpm);
        const float2 wt0 = otf_pixel(a.tab0, a.ns, a.cpm, a.cutoff, x, y, w, h, a.kx, a.ky, a.vec_cpm);
        const float2 w1 = otf_pixel(a.tab1, a.ns, a.cpm, a.cutoff, x, y, w, h, 0.0, 0.0, a.vec_cpm);
        const float2 wt1 = otf_pixel(a.tab1, a.ns, a.cpm, a.cutoff, x, y, w, h, -a.kx, -a.ky, a.vec_cpm);
        float2 v0 = b0[i], v1 = b1[i];
        if (abs_java(w0) < a.wlim || abs_java(wt0) < a.wlim)
            v0 = make_float2(0.f, 0.f);
        else if (a.divide)
            v0 = div_java(v0, w0);
        if (abs_java(w1) < a.wlim || abs_java(wt1) < a.wlim)
            v1 = make_float2(0.f, 0.f);
        else if (a.divide)
            v1 = div_java(v1, w1);
        const double r0 = dc_rad(x, y, w, h) / kmax;
        if (r0 < a.dist || r0 > 1 - a.dist) v0 = make_float2(0.f, 0.f);
        const int xq = ((x + a.ikx) % w + w) % w, yq = ((y - a.iky) % h + h) % h;
        const double rq = dc_rad(xq, yq, w, h) / kmax;
        if (rq < a.dist || rq > 1 - a.dist) v1 = make_float2(0.f, 0.f);
        b0[i] = v0;
        b1[i] = v1;
    }
}

// Candidate shifts: b1 shifted by (xpos, -ypos) (Vec2d.Cplx.fourierShift,
// BasicVectors.java:309-339), times conj(b0) (AbstractVectorCplx.timesConj),
// summed over all pixels in double (sumElements).  One block per chunk of
// kPix pixels and candidate group; partial[c * nchunk + chunk].
constexpr int kMaxCand = 100;
constexpr int kCorrThreads = 256, kCorrPpt = 4, kCorrPix = kCorrThreads * kCorrPpt;
constexpr int kCandPerBlock = 25;
struct Cands {
    double2 p[kMaxCand];   // (xpos, ypos)
    int n;
};

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

__global__ void __launch_bounds__(kCorrThreads) k_corr_cands(const float2 *__restrict__ b0,
                                                             const float2 *__restrict__ b1, int N, Cands cands,
                                                             int nchunk, double2 *__restrict__ partial) {
    __shared__ double2 sh[kCorrThreads / 64];
    const size_t n = (size_t)N * N;
    const int chunk = blockIdx.x;
    float2 p0[kCorrPpt], p1[kCorrPpt];
    int px[kCorrPpt], py[kCorrPpt];
    bool ok[kCorrPpt];
#pragma unroll
    for (int j = 0; j < kCorrPpt; j++) {
        const size_t i = (size_t)chunk * kCorrPix + (size_t)j * kCorrThreads + threadIdx.x;
        ok[j] = i < n;
        const size_t ic = ok[j] ? i : 0;
        p0[j] = b0[ic];
        p1[j] = b1[ic];
        px[j] = (int)(ic % N);
        py[j] = (int)(ic / N);
    }
    const int c0 = blockIdx.y * kCandPerBlock;
    for (int c = c0; c < c0 + kCandPerBlock && c < cands.n; c++) {
        const double xpos = cands.p[c].x, ypos = cands.p[c].y;
        double2 acc = make_double2(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < kCorrPpt; j++) {
            float co, si;
            shift_cos_sin(xpos, -ypos, px[j], py[j], N, co, si);
            const float2 q = times_conj_java(rot_java(p1[j], co, si), p0[j]);
            if (ok[j]) {
                acc.x += (double)q.x;
                acc.y += (double)q.y;
            }
        }
        const double2 s = block_sum(acc, sh);
        if (threadIdx.x == 0) partial[(size_t)c * nchunk + chunk] = s;
    }
}

// One thread per candidate sums its partials in chunk order.
__global__ void k_corr_final(const double2 *__restrict__ partial, int nchunk, int ncand,
                             double2 *__restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncand) return;
    double2 s = make_double2(0.0, 0.0);
    for (int k = 0; k < nchunk; k++) {
        s.x += partial[(size_t)c * nchunk + k].x;
        s.y += partial[(size_t)c * nchunk + k].y;
    }
    out[c] = s;
}

// Workspace for the estimation calls: two band copies, partial sums,
// candidate results.  One user at a time.
struct EstWork {
    std::mutex mu;
    float2 *t0 = nullptr, *t1 = nullptr;
    double2 *part = nullptr, *res = nullptr;
    size_t nn = 0;
};
static EstWork g_est;

static int est_reserve(size_t nn) {
    if (nn <= g_est.nn) return FSIM_OK;
    (void)hipFree(g_est.t0);
    (void)hipFree(g_est.t1);
    (void)hipFree(g_est.part);
    (void)hipFree(g_est.res);
    g_est.t0 = g_est.t1 = nullptr;
    g_est.part = g_est.res = nullptr;
    g_est.nn = 0;
    const size_t nchunk = (nn + kCorrPix - 1) / kCorrPix;
    if (hipMalloc(&g_est.t0, nn * sizeof(float2)) != hipSuccess ||
        hipMalloc(&g_est.t1, nn * sizeof(float2)) != hipSuccess ||
        hipMalloc(&g_est.part, nchunk * kMaxCand * sizeof(double2)) != hipSuccess ||
        hipMalloc(&g_est.res, kMaxCand * sizeof(double2)) != hipSuccess)
        return fail(FSIM_ERR_OOM, "estimation workspace allocation failed");
    g_est.nn = nn;
    return FSIM_OK;
}

static int region_args(const Otf &o, int bn0, int bn1, double kx, double ky, double dist, double wlim,
                       double vec_cpm, int divide, RegionArgs &a) {
    if (vec_cpm <= 0) return fail(FSIM_ERR_STATE, "common_region: vector pixel size not initialized");
    if (o.nband == 1) bn0 = bn1 = 0;  // single-band OTF: getOtfVal ignores the band
    if (bn0 < 0 || bn0 >= o.nband || bn1 < 0 || bn1 >= o.nband)
        return fail(FSIM_ERR_ARG, "common_region: band idx too high or <0");
    a.tab0 = o.vals + (size_t)bn0 * o.ns;
    a.tab1 = o.vals + (size_t)bn1 * o.ns;
    a.ns = o.ns;
    a.cpm = o.cpm;
    a.cutoff = o.cutoff;
    a.vec_cpm = vec_cpm;
    a.kx = kx;
    a.ky = ky;
    a.dist = dist;
    a.wlim = wlim;
    a.ikx = (int)kx;
    a.iky = (int)ky;
    a.divide = divide;
    return FSIM_OK;
}

static int check_pair(const Vec *b0, const Vec *b1, const char *what) {
    if (!b0 || !b1 || !b0->is_cplx || !b1->is_cplx)
        return fail(FSIM_ERR_ARG, std::string(what) + ": needs two complex vectors");
    if (b0->w != b1->w || b0->h != b1->h) return fail(FSIM_ERR_ARG, std::string(what) + ": vector size mismatch");
    return FSIM_OK;
}

static int launch_region(float2 *b0, float2 *b1, int w, int h, const RegionArgs &a, hipStream_t s) {
    if (std::fabs(a.kx) >= w || std::fabs(a.ky) >= h)
        return fail(FSIM_ERR_ARG, "common_region: shift larger than the vector");
    const size_t n = (size_t)w * h;
    unsigned blocks = cdiv(n, 256);
    if (blocks > 65535 * 4) blocks = 65535 * 4;
    hipLaunchKernelGGL(k_common_region, dim3(blocks), dim3(256), 0, s, b0, b1, w, h, a);
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

// copies of both bands -> commonRegion -> inverse FFT of both -> 1/|b0|^2
static int prepare_pair(const Vec *v0, const Vec *v1, const RegionArgs &a, double &scal) {
    hipStream_t s = stream();
    const int w = v0->w, h = v0->h;
    const size_t nn = (size_t)w * h;
    FSIM_HIP(hipMemcpyAsync(g_est.t0, v0->d, nn * sizeof(float2), hipMemcpyDeviceToDevice, s));
    FSIM_HIP(hipMemcpyAsync(g_est.t1, v1->d, nn * sizeof(float2), hipMemcpyDeviceToDevice, s));
    int rc = launch_region(g_est.t0, g_est.t1, w, h, a, s);
    if (!rc) rc = fft2d_device(g_est.t0, w, h, 1, s);
    if (!rc) rc = fft2d_device(g_est.t1, w, h, 1, s);
    if (rc) return rc;
    double2 nrm;
    rc = reduce_device(reinterpret_cast<const float *>(g_est.t0), nn * 2, 0, &nrm, s);
    if (rc) return rc;
    scal = 1. / nrm.x;
    return FSIM_OK;
}

// sum over pixels of shift(b1) * conj(b0) for each candidate (unscaled)
static int corr_cands(int N, const Cands &c, std::vector<double2> &out) {
    hipStream_t s = stream();
    const size_t nn = (size_t)N * N;
    const int nchunk = (int)((nn + kCorrPix - 1) / kCorrPix);
    const int ngrp = (c.n + kCandPerBlock - 1) / kCandPerBlock;
    hipLaunchKernelGGL(k_corr_cands, dim3(nchunk, ngrp), dim3(kCorrThreads), 0, s, g_est.t0, g_est.t1, N, c,
                       nchunk, g_est.part);
    hipLaunchKernelGGL(k_corr_final, dim3(1), dim3(128), 0, s, g_est.part, nchunk, c.n, g_est.res);
    FSIM_HIP(hipGetLastError());
    out.resize(c.n);
    FSIM_HIP(hipMemcpyAsync(out.data(), g_est.res, c.n * sizeof(double2), hipMemcpyDeviceToHost, s));
    FSIM_HIP(hipStreamSynchronize(s));
    return FSIM_OK;
}

}  // namespace fsim

using namespace fsim;

static Vec *VV(fsim_vec h) { return reinterpret_cast<Vec *>(h); }

extern "C" int fsim_common_region(fsim_vec band0, fsim_vec band1, int bn0, int bn1, fsim_otf otf, double kx,
                                  double ky, double dist, double weight_limit, int divide_by_otf,
                                  double vec_cpm) {
    int rc = ensure_init();
    if (rc) return rc;
    Vec *v0 = VV(band0), *v1 = VV(band1);
    rc = check_pair(v0, v1, "common_region");
    if (rc) return rc;
    if (!otf) return fail(FSIM_ERR_ARG, "common_region: null OTF");
    RegionArgs a;
    rc = region_args(otf->o, bn0, bn1, kx, ky, dist, weight_limit, vec_cpm, divide_by_otf, a);
    if (rc) return rc;
    rc = launch_region((float2 *)v0->d, (float2 *)v1->d, v0->w, v0->h, a, stream());
    if (rc) return rc;
    FSIM_HIP(hipStreamSynchronize(stream()));
    return FSIM_OK;
}

extern "C" int fsim_fit_peak(fsim_vec band0, fsim_vec band1, int bn0, int bn1, fsim_otf otf, double kx, double ky,
                             double weight_limit, double search, double vec_cpm, float *cntrl, double *res) {
    int rc = ensure_init();
    if (rc) return rc;
    Vec *v0 = VV(band0), *v1 = VV(band1);
    rc = check_pair(v0, v1, "fit_peak");
    if (rc) return rc;
    if (!otf || !res) return fail(FSIM_ERR_ARG, "fit_peak: null argument");
    if (v0->w != v0->h) return fail(FSIM_ERR_ARG, "fit_peak: vectors must be square (fourierShift)");
    const int N = v0->w;
    std::lock_guard<std::mutex> lk(g_est.mu);
    rc = est_reserve((size_t)N * N);
    if (rc) return rc;
    double resPhase = 0, resMag = 0;
    for (int iter = 0; iter < 3; iter++) {
        RegionArgs a;
        rc = region_args(otf->o, bn0, bn1, kx, ky, 0.15, weight_limit, vec_cpm, 1, a);
        if (rc) return rc;
        double scal;
        rc = prepare_pair(v0, v1, a, scal);
        if (rc) return rc;
        // 10 x 10 points +- search around the current guess
        Cands c;
        c.n = 100;
        for (int p = 0; p < 100; p++) {
            const int xi = p % 10, yi = p / 10;
            c.p[p] = make_double2(kx + ((xi - 4.5) / 4.5) * search, ky + ((yi - 4.5) / 4.5) * search);
        }
        std::vector<double2> sum;
        rc = corr_cands(N, c, sum);
        if (rc) return rc;
        // corr = sum * (1 / |b0|^2), as Cplx.mult with a real Cplx.Double
        double corr_re[100], corr_im[100], mag[100];
        for (int p = 0; p < 100; p++) {
            corr_re[p] = sum[p].x * scal - sum[p].y * 0.0;
            corr_im[p] = sum[p].x * 0.0 + sum[p].y * scal;
            mag[p] = std::hypot(corr_re[p], corr_im[p]);
        }
        double max = 0, min = 1.7976931348623157e308, newKx = 0, newKy = 0;
        for (int yi = 0; yi < 10; yi++)
            for (int xi = 0; xi < 10; xi++) {
                const int p = yi * 10 + xi;
                if (mag[p] > max) {
                    max = mag[p];
                    newKx = kx + ((xi - 4.5) / 4.5) * search;
                    newKy = ky + ((yi - 4.5) / 4.5) * search;
                    resPhase = std::atan2(corr_im[p], corr_re[p]);
                    resMag = mag[p];
                }
                if (mag[p] < min) min = mag[p];
            }
        if (cntrl)
            for (int yi = 0; yi < 10; yi++)
                for (int xi = 0; xi < 10; xi++)
                    cntrl[yi * 30 + xi + iter * 10] = (float)((mag[yi * 10 + xi] - min) / (max - min));
        kx = newKx;
        ky = newKy;
        search /= 5;
    }
    res[0] = kx;
    res[1] = ky;
    res[2] = resPhase;
    res[3] = resMag;
    return FSIM_OK;
}

extern "C" int fsim_get_peak(fsim_vec band0, fsim_vec band1, int bn0, int bn1, fsim_otf otf, double kx, double ky,
                             double weight_limit, double vec_cpm, double *res) {
    int rc = ensure_init();
    if (rc) return rc;
    Vec *v0 = VV(band0), *v1 = VV(band1);
    rc = check_pair(v0, v1, "get_peak");
    if (rc) return rc;
    if (!otf || !res) return fail(FSIM_ERR_ARG, "get_peak: null argument");
    if (v0->w != v0->h) return fail(FSIM_ERR_ARG, "get_peak: vectors must be square (fourierShift)");
    const int N = v0->w;
    std::lock_guard<std::mutex> lk(g_est.mu);
    rc = est_reserve((size_t)N * N);
    if (rc) return rc;
    RegionArgs a;
    rc = region_args(otf->o, bn0, bn1, kx, ky, 0.15, weight_limit, vec_cpm, 1, a);
    if (rc) return rc;
    double scal;
    rc = prepare_pair(v0, v1, a, scal);
    if (rc) return rc;
    Cands c;
    c.n = 1;
    c.p[0] = make_double2(kx, ky);
    std::vector<double2> sum;
    rc = corr_cands(N, c, sum);
    if (rc) return rc;
    res[0] = sum[0].x * scal - sum[0].y * 0.0;
    res[1] = sum[0].x * 0.0 + sum[0].y * scal;
    return FSIM_OK;
}
