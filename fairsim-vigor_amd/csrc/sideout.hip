// Live-mode side outputs on the device: the widefield image of every stack
// (livemode/ReconstructionRunner.java:452-481) and the registration of
// widefield or reconstructed images through per-pixel translation tables
// (registration/Registration.java).
//
// Both are one pass over HBM.  The widefield pass reads the D*P raw frames
// of a stack (2 bytes per pixel each) and writes one float image; it
// reproduces the reference's float arithmetic: each frame is faded by the
// dampBorder vector (fadeBorderCos of a vector of ones), the frames are added
// in (direction, phase) order and the sum is scaled by 1.f / (D*P).  The
// registration pass is a bilinear gather with the reference's float
// operation order.
#include "fsim_internal.h"
#include "recon.h"

#include <cmath>
#include <vector>

namespace fsim {

// Widefield images of `count` stacks: grid (blocks of PX-pixel groups,
// stacks); n is a multiple of PX, so a group stays in one row.  A group's
// pixels of one frame are one load of PX / 2 32-bit words (two pixels each,
// the lower half first).
// raw_output in 1..D*P: that raw frame instead of the average
// (ReconstructionRunner.java:473-479, setFrom16bitPixels of it).
template <int PX>
struct alignas(2 * PX) WfWords {
    uint32_t w[PX / 2];
};
template <int PX>
__device__ __forceinline__ float wf_pixel(const WfWords<PX> &v, int j) {
    return (float)((v.w[j / 2] >> (16 * (j & 1))) & 0xffffu);
}
template <int PX>
__global__ void k_widefield(const uint16_t *__restrict__ raw, float *__restrict__ out, int n, int nf,
                            const double *__restrict__ lo, const double *__restrict__ hi, float scale,
                            int raw_output) {
#pragma clang fp contract(off)
    const size_t nn = (size_t)n * n;
    const size_t nq = nn / PX;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const WfWords<PX> *r = reinterpret_cast<const WfWords<PX> *>(raw + (size_t)blockIdx.y * nf * nn);
    float4 *o = reinterpret_cast<float4 *>(out + (size_t)blockIdx.y * nn) + q * (PX / 4);
    float a[PX];
    if (raw_output > 0) {
        const WfWords<PX> v = r[(size_t)(raw_output - 1) * nq + q];
#pragma unroll
        for (int j = 0; j < PX; j++) a[j] = wf_pixel(v, j);
    } else {
        const int x0 = (int)((PX * q) % n), y = (int)((PX * q) / n);
        // dampBorder: fadeBorderCos on a vector of ones, top, bottom, left,
        // right (SimUtils.java:162-185; ReconstructionRunner.java:775-780)
        float d[PX];
#pragma unroll
        for (int j = 0; j < PX; j++) {
            float e = 1.f;
            e = (float)((double)e * lo[y]);
            e = (float)((double)e * hi[y]);
            e = (float)((double)e * lo[x0 + j]);
            e = (float)((double)e * hi[x0 + j]);
            d[j] = e;
            a[j] = 0.f;
        }
        // frames in groups of G: a group's loads in flight together, the
        // sums in frame order
        constexpr int G = 9;
        for (int f0 = 0; f0 < nf; f0 += G) {
            WfWords<PX> v[G];
#pragma unroll
            for (int g = 0; g < G; g++)
                if (f0 + g < nf) v[g] = r[(size_t)(f0 + g) * nq + q];
#pragma unroll
            for (int g = 0; g < G; g++)
                if (f0 + g < nf) {
#pragma unroll
                    for (int j = 0; j < PX; j++) a[j] = a[j] + wf_pixel(v[g], j) * d[j];
                }
        }
#pragma unroll
        for (int j = 0; j < PX; j++) a[j] = a[j] * scale;
    }
#pragma unroll
    for (int k = 0; k < PX / 4; k++) o[k] = make_float4(a[4 * k], a[4 * k + 1], a[4 * k + 2], a[4 * k + 3]);
}

// Registration.transPixel (Registration.java:239-268) at target (x, y) of a
// w x h image in tables of width tw (the image centred in them at offx,
// offy), in two parts: the source position, its four neighbours and their
// weights (reg_tap, from the tables alone), then the weighted sum over one
// image (reg_gather).  0 where the source falls outside [0, w-1) x [0, h-1)
// (the reference's fresh result vector).
struct RegTap {
    int tl, tr, bl, br;          // neighbour indices; tl < 0: outside
    float tlf, trf, blf, brf;    // their weights
};
__device__ __forceinline__ RegTap reg_tap(const float *__restrict__ xt, const float *__restrict__ yt, int tw, int w,
                                          int h, int offx, int offy, int x, int y) {
#pragma clang fp contract(off)
    const size_t t = (size_t)(y + offy) * tw + (x + offx);
    const float sx = xt[t] - offx, sy = yt[t] - offy;
    RegTap r;
    r.tl = -1;
    r.tr = r.bl = r.br = 0;
    r.tlf = r.trf = r.blf = r.brf = 0.f;
    if (sx >= 0 && sy >= 0 && sx < w - 1 && sy < h - 1) {
        const int left = (int)sx, top = (int)sy, right = (int)sx + 1, bottom = (int)sy + 1;
        const float xf = sx - left, yf = sy - top;
        r.tlf = (1 - xf) * (1 - yf);
        r.trf = xf * (1 - yf);
        r.blf = (1 - xf) * yf;
        r.brf = xf * yf;
        r.tl = top * w + left;
        r.tr = top * w + right;
        r.bl = bottom * w + left;
        r.br = bottom * w + right;
    }
    return r;
}
__device__ __forceinline__ float reg_gather(const float *__restrict__ s, const RegTap &r) {
#pragma clang fp contract(off)
    if (r.tl < 0) return 0.f;
    const float tlv = s[r.tl] * r.tlf;
    const float trv = s[r.tr] * r.trf;
    const float blv = s[r.bl] * r.blf;
    const float brv = s[r.br] * r.brf;
    return tlv + trv + blv + brv;
}

// Target pixels of `count` images: grid (blocks of threads, groups of
// kRegImg images).  Each thread takes kRegPix pixels a block width apart
// (coalesced table reads and result writes, several gathers in flight) and
// applies each pixel's tap to every image of its group, so the tables are
// read once per group instead of once per image.  Measured on 64 images of
// 1024^2 (fraction of 8 TB/s on the images read and written and the tables
// once): 4 pixels x 4 images 0.47, 4 x 16 0.56, 4 x 32 0.58, 2 x 64 and
// 1 x 64 0.59 -- the tables' re-reads, not the gathers, set the rate.
constexpr int kRegPix = 2, kRegImg = 64;
__global__ void __launch_bounds__(256) k_register(float *__restrict__ dst, const float *__restrict__ src,
                                                  const float *__restrict__ xt, const float *__restrict__ yt, int tw,
                                                  int w, int h, int offx, int offy, int count) {
    const size_t nn = (size_t)w * h;
    const size_t base = (size_t)blockIdx.x * 256 * kRegPix + threadIdx.x;
    RegTap tap[kRegPix];
#pragma unroll
    for (int j = 0; j < kRegPix; j++) {
        const size_t i = base + (size_t)j * 256;
        tap[j] = reg_tap(xt, yt, tw, w, h, offx, offy, i < nn ? (int)(i % w) : 0, i < nn ? (int)(i / w) : 0);
    }
    const int g0 = blockIdx.y * kRegImg;
#pragma unroll
    for (int g = 0; g < kRegImg; g++) {
        if (g0 + g >= count) break;
        const float *s = src + (size_t)(g0 + g) * nn;
        float *d = dst + (size_t)(g0 + g) * nn;
#pragma unroll
        for (int j = 0; j < kRegPix; j++) {
            const size_t i = base + (size_t)j * 256;
            if (i < nn) d[i] = reg_gather(s, tap[j]);
        }
    }
}
}  // namespace fsim

using namespace fsim;

struct fsim_registration_s {
    int w = 0, h = 0;            // reconstruction-size tables
    float *xr = nullptr, *yr = nullptr;   // w x h
    float *xw = nullptr, *yw = nullptr;   // w/2 x h/2 (widefield)
};

extern "C" {

int fsim_widefield(fsim_channel ch, const uint16_t *raw, float *out, int count, int raw_output, void *strm) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!ch || !raw || !out || count < 1 || count > 65535) return fail(FSIM_ERR_ARG, "widefield: bad arguments");
    const ReconConst &K = channel_host_const(ch);
    const int nf = K.D * K.P;
    if (raw_output < 0 || raw_output > nf)
        return fail(FSIM_ERR_ARG, "widefield: raw_output must be 0 (average) or a frame number 1.." +
                                      std::to_string(nf));
    hipStream_t s = strm ? (hipStream_t)strm : stream();
    const size_t nn = (size_t)K.n * K.n;
    // 1.f / (nrDirs * nrPhases), in float as the reference divides
    const float scale = 1.f / nf;
    if (((uintptr_t)raw & 7) || ((uintptr_t)out & 15))
        return fail(FSIM_ERR_ARG, "widefield: raw must be 8-byte and out 16-byte aligned");
    // 4 pixels per thread (8-byte loads): 8 pixels (16-byte loads) measured
    // 0.63 against 0.68 of 8 TB/s on 64 stacks of 512^2
    hipLaunchKernelGGL(k_widefield<4>, dim3(cdiv(nn / 4, 256), count), dim3(256), 0, s, raw, out, K.n, nf,
                       K.fade_lo, K.fade_hi, scale, raw_output);
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

int fsim_registration_create(const float *xtrans, const float *ytrans, int width, int height,
                             fsim_registration *out) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!xtrans || !ytrans || !out || width < 2 || height < 2)
        return fail(FSIM_ERR_ARG, "registration_create: bad arguments");
    fsim_registration_s *r = new fsim_registration_s();
    r->w = width;
    r->h = height;
    const int ww = width / 2, wh = height / 2;
    // widefield tables: the reconstruction tables at even (col, row), halved
    // (Registration.writeRowToTransVec, Registration.java:371-385)
    std::vector<float> xw((size_t)ww * wh), yw((size_t)ww * wh);
    for (int y = 0; y < wh; y++)
        for (int x = 0; x < ww; x++) {
            xw[(size_t)y * ww + x] = xtrans[(size_t)(2 * y) * width + 2 * x] / 2;
            yw[(size_t)y * ww + x] = ytrans[(size_t)(2 * y) * width + 2 * x] / 2;
        }
    const size_t nb = (size_t)width * height * sizeof(float), wb = xw.size() * sizeof(float);
    if (hipMalloc(&r->xr, nb) != hipSuccess || hipMalloc(&r->yr, nb) != hipSuccess ||
        hipMalloc(&r->xw, wb) != hipSuccess || hipMalloc(&r->yw, wb) != hipSuccess) {
        (void)hipFree(r->xr);
        (void)hipFree(r->yr);
        (void)hipFree(r->xw);
        (void)hipFree(r->yw);
        delete r;
        return fail(FSIM_ERR_OOM, "registration_create: device allocation failed");
    }
    FSIM_HIP(hipMemcpy(r->xr, xtrans, nb, hipMemcpyHostToDevice));
    FSIM_HIP(hipMemcpy(r->yr, ytrans, nb, hipMemcpyHostToDevice));
    FSIM_HIP(hipMemcpy(r->xw, xw.data(), wb, hipMemcpyHostToDevice));
    FSIM_HIP(hipMemcpy(r->yw, yw.data(), wb, hipMemcpyHostToDevice));
    *out = r;
    return FSIM_OK;
}

int fsim_registration_destroy(fsim_registration r) {
    if (!r) return FSIM_OK;
    FSIM_HIP(hipStreamSynchronize(stream()));
    (void)hipFree(r->xr);
    (void)hipFree(r->yr);
    (void)hipFree(r->xw);
    (void)hipFree(r->yw);
    delete r;
    return FSIM_OK;
}

int fsim_register_images(fsim_registration r, int widefield, const float *src, float *dst, int w, int h, int count,
                         void *strm) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!r || !src || !dst || w < 1 || h < 1 || count < 1 || count > 65535)
        return fail(FSIM_ERR_ARG, "register_images: bad arguments");
    if (src == dst) return fail(FSIM_ERR_ARG, "register_images: the result needs its own buffer");
    const int tw = widefield ? r->w / 2 : r->w, th = widefield ? r->h / 2 : r->h;
    if (w > tw || h > th)
        return fail(FSIM_ERR_ARG, "register_images: Image registration failed, because of differences in dimensions");
    hipStream_t s = strm ? (hipStream_t)strm : stream();
    const size_t nn = (size_t)w * h;
    hipLaunchKernelGGL(k_register, dim3(cdiv(nn, 256 * kRegPix), cdiv(count, kRegImg)), dim3(256), 0, s, dst, src,
                       widefield ? r->xw : r->xr, widefield ? r->yw : r->yr, tw, w, h, (tw - w) / 2, (th - h) / 2,
                       count);
    FSIM_HIP(hipGetLastError());
    return FSIM_OK;
}

}  // extern "C"
