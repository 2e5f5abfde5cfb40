// Richardson-Lucy deconvolution on the device: RLDeconvolution.deconvolve
// (sim_algorithm/RLDeconvolution.java:43-103), the filter the reference uses
// for the "RL on input / output / both" filter styles.
//
// Per iteration the reference runs four 2-D FFTs and five element-wise
// vector operations:
//   next = IFFT(FFT(next) * otf);  next = reciproc(next) * img
//   next = IFFT(FFT(next) * otf);  next = next * deconv; deconv = next
// Here an iteration is four passes over the image, each transform pair
// meeting its element-wise work in registers (the 2-D inverse transform
// runs columns first, then rows):
//   columns: FFT, * otf, IFFT            rows: IFFT, reciproc * img, FFT
//   columns: FFT, * otf, IFFT            rows: IFFT, * deconv (stored), FFT
// so next stays row-transformed between iterations (one row FFT before the
// first, none after the last): 104 bytes per pixel and iteration against
// 232 for the transforms and element-wise passes one at a time.  The float
// operations of the element-wise steps are the reference's.  The
// reference's per-iteration change estimate (errEst, a norm2 that only
// feeds a trace message) is not computed.
#include "fsim_internal.h"
#include "fft_core.h"
#include "fft_launch.h"

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

// reciproc() (AbstractVectorCplx.java:208-216), then times(img), one element
__device__ __forceinline__ float2 rl_ratio(float2 a, float2 im) {
#pragma clang fp contract(off)
    const float s = a.x * a.x + a.y * a.y;
    return times_java(make_float2(a.x / s, -a.y / s), im);
}

// Transform shapes of the RL steps.  One image is a small grid for the
// device (a 1024^2 image is 1M pixels), so the steps use 8 elements per
// thread up to L = 1024 (2048 for rows) and narrow column tiles, which
// spreads an image over more wavefronts than the batched transforms do.
// Register twiddles (each thread's, read once from the global table) for
// the transforms from 512 points on with 8 elements per thread, where the
// column step has no LDS table: 18.76k against 18.41k iterations/s with
// sincospi twiddles on a 1024^2 image (three alternating runs).
template <int L, int E>
struct RlTwReg {
    static constexpr bool value = L >= 512 && E <= 8;
};

template <int L>
struct RlShape {
    static constexpr int re = L <= 512 ? FftE<L>::value : (L <= 2048 ? 8 : 16);   // rows
    static constexpr int rt = L / re;
    static constexpr int rw = rt >= 256 ? 1 : 256 / rt;
    static constexpr int ce = L <= 512 ? FftE<L>::value : (L <= 1024 ? 8 : 16);   // columns
    static constexpr int ct = L / ce;
    static constexpr int cw0 = 512 / ct < 1 ? 1 : 512 / ct;
    static constexpr int cw = cw0 < ColW<L>::value ? cw0 : ColW<L>::value;
};

// Column step of an iteration on an h x w row-major image (L = h): forward
// column FFT, * otf, inverse column FFT scaled by 1/h.  Tiles of CW columns,
// neighbouring tiles on one XCD (they share cache lines).
template <int L>
__global__ void __launch_bounds__(RlShape<L>::ct * RlShape<L>::cw)
    k_rl_cols(float2 *__restrict__ data, const float2 *__restrict__ otf, int ncols, const float2 *__restrict__ tw) {
    constexpr int E = RlShape<L>::ce, T = L / E, CW = RlShape<L>::cw;
    constexpr int LB = FftLds<L, E>::len * CW * 8;   // bytes of the transform region
    __shared__ float2 lds[FftLds<L, E>::len * CW];
    // the twiddle table where it costs no block per CU
    using TW = TwLds<L, LB + L * 8 <= 163840 && 163840 / (LB + L * 8) == 163840 / LB>;
    __shared__ float2 twl[TW::len];
    const int ntiles = (ncols + CW - 1) / CW;
    const int tile = xcd_linear(ntiles);
    if (tile >= ntiles) return;   // the whole block: before any barrier
    const float2 *tws = tw_table<TW::tl, T * CW>(twl, tw);
    const int c = threadIdx.x % CW, t = threadIdx.x / CW;
    const int col = tile * CW + c;
    const bool active = col < ncols;
    float2 v[E];
    if (active) {
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = data[(size_t)(t + T * m) * ncols + col];
    }
    constexpr bool RG = RlTwReg<L, E>::value && TW::tl == 0;
    TwReg<L, E> twr;
    if constexpr (RG) {
        tw_reg_load<L, E, L>(twr, t, glb(tw));
        fft_regs_r<L, E, false, false>(v, lds + c, CW, t, twr, active);
    } else {
        fft_regs<L, E, false, false, TW::tl>(v, lds + c, CW, t, tws, active);
    }
    if (active) {
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = times_java(v[m], otf[(size_t)(t + T * m) * ncols + col]);
    }
    if constexpr (RG)
        fft_regs_r<L, E, true, false>(v, lds + c, CW, t, twr, active);
    else
        fft_regs<L, E, true, false, TW::tl>(v, lds + c, CW, t, tws, active);
    if (active) {
        const float sc = 1.0f / L;
#pragma unroll
        for (int m = 0; m < E; m++) data[(size_t)(t + T * m) * ncols + col] = cscale(v[m], sc);
    }
}

// Row step of an iteration (L = w): inverse row FFT scaled by 1/w, then
// MODE 0: reciproc * img, forward row FFT (next);
// MODE 1: * deconv, stored to deconv, forward row FFT (next);
// MODE 2: * deconv, stored to deconv only (the last iteration).
template <int L, int MODE>
__global__ void __launch_bounds__(RlShape<L>::rt * RlShape<L>::rw)
    k_rl_rows(float2 *__restrict__ next, float2 *__restrict__ deconv, const float2 *__restrict__ img, int nrows,
              const float2 *__restrict__ tw) {
    constexpr int E = RlShape<L>::re, T = L / E, RW = RlShape<L>::rw;
    constexpr int LP = FftLds<L, E>::len;
    constexpr bool WAVE = T <= 64;
    __shared__ float2 lds[RW * LP];
    // row transforms: small blocks, twiddles from sincospi (no table fill)
    using TW = TwLds<L, false>;
    const int t = threadIdx.x % T, rr = threadIdx.x / T;
    const int row = blockIdx.x * RW + rr;
    const bool active = row < nrows;
    const size_t base = (size_t)row * L;
    float2 v[E];
    if (active) {
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = next[base + t + T * m];
    }
    constexpr bool RG = RlTwReg<L, E>::value;
    TwReg<L, E> twr;
    if constexpr (RG) {
        tw_reg_load<L, E, L>(twr, t, glb(tw));
        fft_regs_r<L, E, true, WAVE>(v, lds + rr * LP, 1, t, twr, active);
    } else {
        fft_regs<L, E, true, WAVE, TW::tl>(v, lds + rr * LP, 1, t, tw, active);
    }
    if (active) {
        const float sc = 1.0f / L;
#pragma unroll
        for (int m = 0; m < E; m++) {
            const size_t i = base + t + T * m;
            const float2 a = cscale(v[m], sc);
            if constexpr (MODE == 0) {
                v[m] = rl_ratio(a, img[i]);
            } else {
                v[m] = times_java(a, deconv[i]);
                deconv[i] = v[m];
            }
        }
    }
    if constexpr (MODE != 2) {
        if constexpr (RG)
            fft_regs_r<L, E, false, WAVE>(v, lds + rr * LP, 1, t, twr, active);
        else
            fft_regs<L, E, false, WAVE, TW::tl>(v, lds + rr * LP, 1, t, tw, active);
        if (active) {
#pragma unroll
            for (int m = 0; m < E; m++) next[base + t + T * m] = v[m];
        }
    }
}

template <int L>
static int rl_cols_t(float2 *d, const float2 *otf, int ncols, hipStream_t s) {
    using S = RlShape<L>;
    const float2 *tw = twiddles(L);
    if (!tw) return fail(FSIM_ERR_OOM, "twiddle table allocation failed");
    hipLaunchKernelGGL(k_rl_cols<L>, dim3(xcd_grid(cdiv(ncols, S::cw))), dim3(S::ct * S::cw), 0, s, d, otf, ncols,
                       tw);
    return FSIM_OK;
}

template <int L>
static int rl_rows_t(float2 *next, float2 *deconv, const float2 *img, int nrows, int mode, hipStream_t s) {
    using S = RlShape<L>;
    const float2 *tw = twiddles(L);
    if (!tw) return fail(FSIM_ERR_OOM, "twiddle table allocation failed");
    const dim3 g(cdiv(nrows, S::rw)), b(S::rt * S::rw);
    if (mode == 0)
        hipLaunchKernelGGL((k_rl_rows<L, 0>), g, b, 0, s, next, deconv, img, nrows, tw);
    else if (mode == 1)
        hipLaunchKernelGGL((k_rl_rows<L, 1>), g, b, 0, s, next, deconv, img, nrows, tw);
    else
        hipLaunchKernelGGL((k_rl_rows<L, 2>), g, b, 0, s, next, deconv, img, nrows, tw);
    return FSIM_OK;
}

#define FSIM_RL_SWITCH(L, FN, ...)                                                          \
    switch (L) {                                                                            \
        case 2: return FN<2>(__VA_ARGS__);                                                  \
        case 4: return FN<4>(__VA_ARGS__);                                                  \
        case 8: return FN<8>(__VA_ARGS__);                                                  \
        case 16: return FN<16>(__VA_ARGS__);                                                \
        case 32: return FN<32>(__VA_ARGS__);                                                \
        case 64: return FN<64>(__VA_ARGS__);                                                \
        case 128: return FN<128>(__VA_ARGS__);                                              \
        case 256: return FN<256>(__VA_ARGS__);                                              \
        case 512: return FN<512>(__VA_ARGS__);                                              \
        case 1024: return FN<1024>(__VA_ARGS__);                                            \
        case 2048: return FN<2048>(__VA_ARGS__);                                            \
        case 4096: return FN<4096>(__VA_ARGS__);                                            \
        default: return fail(FSIM_ERR_ARG, "rl_deconvolve: size must be a power of two in [2, 4096]"); \
    }

// column step on an h x w image
static int rl_cols(float2 *d, const float2 *otf, int w, int h, hipStream_t s) {
    FSIM_RL_SWITCH(h, rl_cols_t, d, otf, w, s)
}
// row step (mode as k_rl_rows) on an h x w image
static int rl_rows(float2 *next, float2 *deconv, const float2 *img, int w, int h, int mode, hipStream_t s) {
    FSIM_RL_SWITCH(w, rl_rows_t, next, deconv, img, h, mode, s)
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
    if (steps > 0) {
        FSIM_HIP(hipMemcpyAsync(next, d, n * sizeof(float2), hipMemcpyDeviceToDevice, s));
        // next row-transformed; each iteration leaves it so for the next
        rc = fft_rows_batch_device(next, w, h, 1, 0, s);
        if (rc) return rc;
    }
    for (int it = 0; it < steps; it++) {
        // 1: u_j * psf (2-D FFT: the row part is in place; inverse: columns,
        // then rows); 2: d / [1], and the row FFT of step 3
        if ((rc = rl_cols(next, o, w, h, s)) || (rc = rl_rows(next, deconv, d, w, h, 0, s))) return rc;
        // 3: [2] * psf'; 4: u_j x [3] -> u_{j+1}, and the next iteration's row FFT
        if ((rc = rl_cols(next, o, w, h, s)) || (rc = rl_rows(next, deconv, d, w, h, it + 1 < steps ? 1 : 2, s)))
            return rc;
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
