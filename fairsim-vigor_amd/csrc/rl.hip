// Richardson-Lucy deconvolution on the device: RLDeconvolution.deconvolve
// (sim_algorithm/RLDeconvolution.java:43-103), the filter the reference uses
// for the "RL on input / output / both" filter styles.
//
// Per iteration the reference runs four 2-D FFTs and five element-wise
// vector operations; here the element-wise work between the transforms is
// fused into three kernels (same float operations, same order):
//   next = IFFT(FFT(next) * otf);  next = reciproc(next) * img
//   next = IFFT(FFT(next) * otf);  next = next * deconv; deconv = next
// The reference's per-iteration change estimate (errEst, a norm2 that only
// feeds a trace message) is not computed.
#include "fsim_internal.h"

#include <mutex>

namespace fsim {

#define RL_GRID_STRIDE(i, n) \
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (size_t)gridDim.x * blockDim.x)

// y = y * x as AbstractVectorCplx.times(Vec.Cplx) (AbstractVectorCplx.java:
// 281-304, Cplx.multReal / multImag)
__device__ __forceinline__ float2 times_java(float2 y, float2 x) {
#pragma clang fp contract(off)
    return make_float2((x.x * y.x) - (x.y * y.y), (x.y * y.x) + (x.x * y.y));
}

__global__ void k_rl_otf(float2 *__restrict__ v, const float2 *__restrict__ otf, size_t n) {
    RL_GRID_STRIDE(i, n) v[i] = times_java(v[i], otf[i]);
}

// reciproc() (AbstractVectorCplx.java:208-216), then times(img)
__global__ void k_rl_ratio(float2 *__restrict__ v, const float2 *__restrict__ img, size_t n) {
#pragma clang fp contract(off)
    RL_GRID_STRIDE(i, n) {
        const float2 a = v[i];
        const float s = a.x * a.x + a.y * a.y;
        v[i] = times_java(make_float2(a.x / s, -a.y / s), img[i]);
    }
}

// times(deconvImg), then deconvImg.copy(nextImg)
__global__ void k_rl_update(float2 *__restrict__ v, float2 *__restrict__ u, size_t n) {
    RL_GRID_STRIDE(i, n) {
        const float2 r = times_java(v[i], u[i]);
        v[i] = r;
        u[i] = r;
    }
}

static dim3 rl_grid(size_t n) {
    size_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    return dim3((unsigned)(b ? b : 1));
}

// work vectors (next, deconv), grown on demand; one call at a time
struct RlWork {
    std::mutex mu;
    float2 *buf = nullptr;
    size_t cap = 0;
};
static RlWork g_rl;

}  // namespace fsim

using namespace fsim;

extern "C" int fsim_rl_deconvolve(fsim_vec imgh, fsim_vec otfh, int steps, int input_in_freq) {
    int rc = ensure_init();
    if (rc) return rc;
    Vec *img = reinterpret_cast<Vec *>(imgh), *otf = reinterpret_cast<Vec *>(otfh);
    if (!img || !otf || !img->is_cplx || !otf->is_cplx || steps < 0)
        return fail(FSIM_ERR_ARG, "rl_deconvolve: needs two complex vectors and steps >= 0");
    if (img->w != otf->w || img->h != otf->h) return fail(FSIM_ERR_ARG, "rl_deconvolve: vector size mismatch");
    if (!fft_size_ok(img->w) || !fft_size_ok(img->h))
        return fail(FSIM_ERR_ARG, "rl_deconvolve: width and height must be powers of two in [2, 4096]");
    const int w = img->w, h = img->h;
    const size_t n = (size_t)w * h;
    hipStream_t s = stream();
    std::lock_guard<std::mutex> lk(g_rl.mu);
    if (2 * n > g_rl.cap) {
        FSIM_HIP(hipStreamSynchronize(s));
        (void)hipFree(g_rl.buf);
        g_rl.buf = nullptr;
        g_rl.cap = 0;
        if (hipMalloc(&g_rl.buf, 2 * n * sizeof(float2)) != hipSuccess)
            return fail(FSIM_ERR_OOM, "rl_deconvolve: work vector allocation failed");
        g_rl.cap = 2 * n;
    }
    float2 *d = reinterpret_cast<float2 *>(img->d);
    const float2 *o = reinterpret_cast<const float2 *>(otf->d);
    float2 *next = g_rl.buf, *deconv = g_rl.buf + n;
    if (input_in_freq) {
        rc = fft2d_device(d, w, h, 1, s);
        if (rc) return rc;
    }
    FSIM_HIP(hipMemcpyAsync(deconv, d, n * sizeof(float2), hipMemcpyDeviceToDevice, s));
    FSIM_HIP(hipMemcpyAsync(next, d, n * sizeof(float2), hipMemcpyDeviceToDevice, s));
    const dim3 g = rl_grid(n);
    for (int it = 0; it < steps; it++) {
        // 1: u_j * psf
        rc = fft2d_device(next, w, h, 0, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_rl_otf, g, dim3(256), 0, s, next, o, n);
        rc = fft2d_device(next, w, h, 1, s);
        if (rc) return rc;
        // 2: d / [1]
        hipLaunchKernelGGL(k_rl_ratio, g, dim3(256), 0, s, next, d, n);
        // 3: [2] * psf'
        rc = fft2d_device(next, w, h, 0, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_rl_otf, g, dim3(256), 0, s, next, o, n);
        rc = fft2d_device(next, w, h, 1, s);
        if (rc) return rc;
        // 4: u_j x [3] -> u_{j+1}
        hipLaunchKernelGGL(k_rl_update, g, dim3(256), 0, s, next, deconv, n);
        FSIM_HIP(hipGetLastError());
    }
    FSIM_HIP(hipMemcpyAsync(d, deconv, n * sizeof(float2), hipMemcpyDeviceToDevice, s));
    if (input_in_freq) {
        rc = fft2d_device(d, w, h, 0, s);
        if (rc) return rc;
    }
    FSIM_HIP(hipStreamSynchronize(s));
    return FSIM_OK;
}

namespace fsim {

// SimUtils.clipAndScale (SimUtils.java:244-264): max of the elements above 0
// (float bit patterns of positive values order as unsigned integers), then
// dat *= 255.f / max, then negative values to 0.
__global__ void k_pos_max(const float *__restrict__ v, size_t n, unsigned *__restrict__ mx) {
    unsigned m = 0;
    RL_GRID_STRIDE(i, n) {
        const float x = v[i];
        if (x > 0.f) {
            const unsigned b = __float_as_uint(x);
            if (b > m) m = b;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned t = __shfl_down(m, o, 64);
        if (t > m) m = t;
    }
    if (threadIdx.x % 64 == 0 && m) atomicMax(mx, m);
}

__global__ void k_scale_clip(float *__restrict__ v, size_t n, const unsigned *__restrict__ mx, int scale, int clip) {
    const float f = scale ? 255.f / __uint_as_float(*mx) : 1.f;
    RL_GRID_STRIDE(i, n) {
        float x = v[i];
        if (scale) x *= f;
        if (clip && x < 0) x = 0;
        v[i] = x;
    }
}

}  // namespace fsim

extern "C" int fsim_clip_scale(fsim_vec vh, int clip, int scale) {
    int rc = ensure_init();
    if (rc) return rc;
    Vec *v = reinterpret_cast<Vec *>(vh);
    if (!v || v->is_cplx) return fail(FSIM_ERR_ARG, "clip_scale: needs a real vector");
    const size_t n = (size_t)v->w * v->h;
    hipStream_t s = stream();
    std::lock_guard<std::mutex> lk(scratch_mutex());
    void *scr = nullptr;
    rc = scratch_d(sizeof(unsigned), &scr);
    if (rc) return rc;
    unsigned *mx = reinterpret_cast<unsigned *>(scr);
    FSIM_HIP(hipMemsetAsync(mx, 0, sizeof(unsigned), s));
    if (scale) hipLaunchKernelGGL(k_pos_max, rl_grid(n), dim3(256), 0, s, (const float *)v->d, n, mx);
    hipLaunchKernelGGL(k_scale_clip, rl_grid(n), dim3(256), 0, s, (float *)v->d, n, mx, scale, clip);
    FSIM_HIP(hipGetLastError());
    FSIM_HIP(hipStreamSynchronize(s));
    return FSIM_OK;
}
