// Stand-alone 2-D complex FFT of a device vector: one row pass and one column
// pass, each a single kernel that loads, transforms in registers/LDS and
// stores in place.  Used by the vector API (Vec2d.Cplx.fft2d); the fused
// reconstruction kernels call the same fft_regs core with their own
// prologues and epilogues.
#include "fsim_internal.h"
#include "fft_core.h"
#include "fft_launch.h"

namespace fsim {

// Rows of length L, RW rows per block of 256 threads (T = L/E per row).
template <int L, bool INV>
__global__ void __launch_bounds__(256) k_fft_rows(float2 *__restrict__ data, int nrows, int pitch,
                                                  const float2 *__restrict__ tw, float scale) {
    constexpr int E = FftE<L>::value, T = L / E;
    constexpr int RW = T >= 256 ? 1 : 256 / T;
    constexpr int LP = FftLds<L, E>::len;
    __shared__ float2 lds[RW * LP];
    // row transforms: small blocks, twiddles from sincospi (no table fill)
    using TW = TwLds<L, false>;
    const float2 *tws = tw;
    const int t = threadIdx.x % T, rr = threadIdx.x / T;
    const int row = blockIdx.x * RW + rr;
    const bool active = row < nrows;
    float2 v[E];
    float2 *p = data + (size_t)row * pitch;
    if (active) {
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = p[t + T * m];
    }
    fft_regs<L, E, INV, RowWave<L>::value, TW::tl>(v, lds + rr * LP, 1, t, tws, active);
    if (active) {
#pragma unroll
        for (int m = 0; m < E; m++) p[t + T * m] = INV ? cscale(v[m], scale) : v[m];
    }
}


template <int L, bool INV>
__global__ void __launch_bounds__(1024) k_fft_cols(float2 *__restrict__ data, int ncols, int pitch,
                                                   size_t img_stride, const float2 *__restrict__ tw,
                                                   float scale) {
    // blockIdx.y: image of a batch of equally shaped images img_stride apart
    data += (size_t)blockIdx.y * img_stride;
    constexpr int E = FftE<L>::value, T = L / E;
    constexpr int CW = ColW<L>::value;
    __shared__ float2 lds[FftLds<L, E>::len * CW];
    // the twiddle table where the LDS holds it beside the tile
    using TW = TwLds<L, (FftLds<L, E>::len * CW + L) * 8 <= 163840>;
    __shared__ float2 twl[TW::len];
    const float2 *tws = tw_table<TW::tl, T * CW>(twl, tw);
    const int c = threadIdx.x % CW, t = threadIdx.x / CW;
    const int col = blockIdx.x * CW + c;
    const bool active = col < ncols;
    float2 v[E];
    if (active) {
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = data[(size_t)(t + T * m) * pitch + col];
    }
    fft_regs<L, E, INV, false, TW::tl>(v, lds + c, CW, t, tws, active);
    if (active) {
#pragma unroll
        for (int m = 0; m < E; m++)
            data[(size_t)(t + T * m) * pitch + col] = INV ? cscale(v[m], scale) : v[m];
    }
}

template <int L>
static int launch_rows_t(float2 *d, int nrows, int pitch, int inverse, hipStream_t s) {
    constexpr int E = FftE<L>::value, T = L / E;
    constexpr int RW = T >= 256 ? 1 : 256 / T;
    const float2 *tw = twiddles(L);
    if (!tw) return fail(FSIM_ERR_OOM, "twiddle table allocation failed");
    dim3 grid(cdiv(nrows, RW)), block(T * RW);
    if (inverse)
        hipLaunchKernelGGL((k_fft_rows<L, true>), grid, block, 0, s, d, nrows, pitch, tw, 1.0f / L);
    else
        hipLaunchKernelGGL((k_fft_rows<L, false>), grid, block, 0, s, d, nrows, pitch, tw, 1.0f);
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

template <int L>
static int launch_cols_t(float2 *d, int ncols, int pitch, int nimg, size_t img_stride, int inverse,
                         hipStream_t s) {
    constexpr int E = FftE<L>::value, T = L / E;
    constexpr int CW = ColW<L>::value;
    const float2 *tw = twiddles(L);
    if (!tw) return fail(FSIM_ERR_OOM, "twiddle table allocation failed");
    dim3 grid(cdiv(ncols, CW), nimg), block(T * CW);
    if (inverse)
        hipLaunchKernelGGL((k_fft_cols<L, true>), grid, block, 0, s, d, ncols, pitch, img_stride, tw,
                           1.0f / L);
    else
        hipLaunchKernelGGL((k_fft_cols<L, false>), grid, block, 0, s, d, ncols, pitch, img_stride, tw, 1.0f);
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

bool fft_size_ok(int L) { return is_pow2(L) && L >= 2 && L <= 4096; }

#define FSIM_FFT_SWITCH(L, FN, ...)                      \
    switch (L) {                                         \
        case 2: return FN<2>(__VA_ARGS__);               \
        case 4: return FN<4>(__VA_ARGS__);               \
        case 8: return FN<8>(__VA_ARGS__);               \
        case 16: return FN<16>(__VA_ARGS__);             \
        case 32: return FN<32>(__VA_ARGS__);             \
        case 64: return FN<64>(__VA_ARGS__);             \
        case 128: return FN<128>(__VA_ARGS__);           \
        case 256: return FN<256>(__VA_ARGS__);           \
        case 512: return FN<512>(__VA_ARGS__);           \
        case 1024: return FN<1024>(__VA_ARGS__);         \
        case 2048: return FN<2048>(__VA_ARGS__);         \
        case 4096: return FN<4096>(__VA_ARGS__);         \
        default: return fail(FSIM_ERR_ARG, "FFT size must be a power of two in [2, 4096]"); \
    }

static int launch_rows(int L, float2 *d, int nrows, int pitch, int inverse, hipStream_t s) {
    FSIM_FFT_SWITCH(L, launch_rows_t, d, nrows, pitch, inverse, s)
}
static int launch_cols(int L, float2 *d, int ncols, int pitch, int nimg, size_t img_stride, int inverse,
                       hipStream_t s) {
    FSIM_FFT_SWITCH(L, launch_cols_t, d, ncols, pitch, nimg, img_stride, inverse, s)
}

int fft2d_batch_device(float2 *data, int w, int h, int nimg, int inverse, hipStream_t s) {
    if (!fft_size_ok(w) || !fft_size_ok(h))
        return fail(FSIM_ERR_ARG, "fft2d: width and height must be powers of two in [2, 4096]");
    if (nimg < 1 || nimg > 65535) return fail(FSIM_ERR_ARG, "fft2d: batch of 1 to 65535 images");
    int rc = launch_rows(w, data, h * nimg, w, inverse, s);
    if (rc) return rc;
    return launch_cols(h, data, w, w, nimg, (size_t)w * h, inverse, s);
}

int fft_cols_batch_device(float2 *data, int w, int h, int nimg, int inverse, hipStream_t s) {
    if (!fft_size_ok(w) || !fft_size_ok(h))
        return fail(FSIM_ERR_ARG, "fft2d: width and height must be powers of two in [2, 4096]");
    if (nimg < 1 || nimg > 65535) return fail(FSIM_ERR_ARG, "fft2d: batch of 1 to 65535 images");
    return launch_cols(h, data, w, w, nimg, (size_t)w * h, inverse, s);
}

int fft_rows_batch_device(float2 *data, int w, int h, int nimg, int inverse, hipStream_t s) {
    if (!fft_size_ok(w) || !fft_size_ok(h))
        return fail(FSIM_ERR_ARG, "fft2d: width and height must be powers of two in [2, 4096]");
    if (nimg < 1 || nimg > 65535) return fail(FSIM_ERR_ARG, "fft2d: batch of 1 to 65535 images");
    return launch_rows(w, data, h * nimg, w, inverse, s);
}

int fft2d_device(float2 *data, int w, int h, int inverse, hipStream_t s) {
    return fft2d_batch_device(data, w, h, 1, inverse, s);
}

}  // namespace fsim

using namespace fsim;

extern "C" int fsim_vec_fft2d(fsim_vec vh, int inverse) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!vh) return fail(FSIM_ERR_ARG, "fft2d: null vector");
    Vec *v = reinterpret_cast<Vec *>(vh);
    if (!v->is_cplx) return fail(FSIM_ERR_ARG, "fft2d: vector is not complex");
    return fft2d_device(reinterpret_cast<float2 *>(v->d), v->w, v->h, inverse, stream());
}
