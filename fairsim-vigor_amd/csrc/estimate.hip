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
#include "estimate_dev.h"

#include <cmath>
#include <mutex>
#include <vector>

namespace fsim {

// Correlation.commonRegion over both bands in place.
__global__ void k_common_region(float2 *__restrict__ b0, float2 *__restrict__ b1, int w, int h, RegionArgs a) {
    const double kmax = k_len(a.kx, a.ky);
    const size_t n = (size_t)w * h;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % w), y = (int)(i / w);
        float2 v0 = b0[i], v1 = b1[i];
        region_pixel(a, x, y, w, h, kmax, v0, v1);
        b0[i] = v0;
        b1[i] = v1;
    }
}

// Candidate list of one launch; partial[c * nchunk + chunk].
struct Cands {
    double2 p[kMaxCand];   // (xpos, ypos)
    int n;
};

__global__ void __launch_bounds__(kCorrThreads) k_corr_cands(const float2 *__restrict__ b0,
                                                             const float2 *__restrict__ b1, int N, Cands cands,
                                                             int nchunk, double2 *__restrict__ partial) {
    __shared__ double2 sh[kCorrThreads / 64];
    const int chunk = blockIdx.x;
    CorrPix px;
    corr_load(px, b0, b1, N, chunk);
    const int c0 = blockIdx.y * kCandPerBlock;
    for (int c = c0; c < c0 + kCandPerBlock && c < cands.n; c++) {
        const double2 s = block_sum(corr_candidate(px, cands.p[c].x, cands.p[c].y, N), sh);
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
            c.p[p] = make_double2(grid_pos(kx, xi, search), grid_pos(ky, yi, search));
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
                    newKx = grid_pos(kx, xi, search);
                    newKy = grid_pos(ky, yi, search);
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
