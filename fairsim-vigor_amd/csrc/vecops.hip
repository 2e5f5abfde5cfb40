// Vector API element-wise operations, reductions, Fourier shift and
// pasteFreq.  Each entry point is one stream-ordered launch over the vector,
// mirroring one AccelVector native.
#include "fsim_internal.h"
#include "fft_core.h"
#include "simmath.h"

#include <mutex>

namespace fsim {

static inline dim3 grid_for(size_t n, int per_thread = 4) {
    size_t b = (n + 256 * per_thread - 1) / (256 * per_thread);
    if (b < 1) b = 1;
    if (b > 65535 * 8) b = 65535 * 8;
    return dim3((unsigned)b);
}

#define GRID_STRIDE(i, n) \
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (size_t)gridDim.x * blockDim.x)

__global__ void k_copy_c_from_r(float2 *__restrict__ d, const float *__restrict__ s, size_t n) {
    GRID_STRIDE(i, n) d[i] = make_float2(s[i], 0.f);
}
__global__ void k_copy_r_from_c(float *__restrict__ d, const float2 *__restrict__ s, size_t n) {
    GRID_STRIDE(i, n) d[i] = s[i].x;
}
__global__ void k_add_f(float *__restrict__ d, const float *__restrict__ s, size_t n) {
    GRID_STRIDE(i, n) d[i] += s[i];
}
__global__ void k_add_const_c(float2 *__restrict__ d, float re, float im, size_t n) {
    GRID_STRIDE(i, n) d[i] = make_float2(d[i].x + re, d[i].y + im);
}
__global__ void k_add_const_r(float *__restrict__ d, float re, size_t n) {
    GRID_STRIDE(i, n) d[i] += re;
}
// AbstractVectorCplx.axpy(Cplx.Float) (AbstractVectorCplx.java:185-195)
__global__ void k_axpy_c(float2 *__restrict__ d, float ar, float ai, const float2 *__restrict__ x, size_t n) {
#pragma clang fp contract(off)
    GRID_STRIDE(i, n) {
        float2 xv = x[i], dv = d[i];
        dv.x += (ar * xv.x) - (ai * xv.y);
        dv.y += (ai * xv.x) + (ar * xv.y);
        d[i] = dv;
    }
}
__global__ void k_axpy_r(float *__restrict__ d, float a, const float *__restrict__ x, size_t n) {
#pragma clang fp contract(off)
    GRID_STRIDE(i, n) d[i] += a * x[i];
}
// AbstractVectorCplx.times(Cplx, conj) (AbstractVectorCplx.java:281-295)
__global__ void k_times_cc(float2 *__restrict__ d, const float2 *__restrict__ x, int conj, size_t n) {
#pragma clang fp contract(off)
    GRID_STRIDE(i, n) {
        float2 y = d[i], xv = x[i];
        float xr = xv.x, xi = conj ? -xv.y : xv.y;
        d[i] = make_float2((xr * y.x) - (xi * y.y), (xi * y.x) + (xr * y.y));
    }
}
__global__ void k_times_cr(float2 *__restrict__ d, const float *__restrict__ x, size_t n) {
    GRID_STRIDE(i, n) {
        float2 y = d[i];
        float xv = x[i];
        d[i] = make_float2(y.x * xv, y.y * xv);
    }
}
__global__ void k_times_rr(float *__restrict__ d, const float *__restrict__ x, size_t n) {
    GRID_STRIDE(i, n) d[i] *= x[i];
}
// AbstractVectorCplx.scal(Cplx.Float) (AbstractVectorCplx.java:226-235)
__global__ void k_scal_c(float2 *__restrict__ d, float re, float im, size_t n) {
#pragma clang fp contract(off)
    GRID_STRIDE(i, n) {
        float2 y = d[i];
        d[i] = make_float2((re * y.x) - (im * y.y), (im * y.x) + (re * y.y));
    }
}
__global__ void k_scal_r(float *__restrict__ d, float a, size_t n) {
    GRID_STRIDE(i, n) d[i] *= a;
}
__global__ void k_set_u16(float2 *__restrict__ d, const uint16_t *__restrict__ s, size_t n) {
    GRID_STRIDE(i, n) d[i] = make_float2((float)s[i], 0.f);
}
__global__ void k_set_u16_r(float *__restrict__ d, const uint16_t *__restrict__ s, size_t n) {
    GRID_STRIDE(i, n) d[i] = (float)s[i];
}

// Double-precision reductions: per-block partials, then one block sums them
// in a fixed order.  mode 0: sum of float squares (norm2), 1: sum re/im.
constexpr int kRedBlock = 256;
__global__ void __launch_bounds__(kRedBlock) k_reduce(const float *__restrict__ d, size_t nfloat, int mode,
                                                      double2 *__restrict__ part) {
    __shared__ double2 sh[kRedBlock];
    double a = 0, b = 0;
    if (mode == 0) {
        GRID_STRIDE(i, nfloat) {
            float v = d[i];
            float sq = v * v;  // float square as in AbstractVectorCplx.norm2
            a += (double)sq;
        }
    } else {
        size_t nc = nfloat / 2;
        const float2 *c = reinterpret_cast<const float2 *>(d);
        GRID_STRIDE(i, nc) {
            float2 v = c[i];
            a += (double)v.x;
            b += (double)v.y;
        }
    }
    sh[threadIdx.x] = make_double2(a, b);
    __syncthreads();
    for (int s = kRedBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sh[threadIdx.x].x += sh[threadIdx.x + s].x;
            sh[threadIdx.x].y += sh[threadIdx.x + s].y;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(kRedBlock) k_reduce_final(double2 *__restrict__ part, int np) {
    __shared__ double2 sh[kRedBlock];
    double a = 0, b = 0;
    for (int i = threadIdx.x; i < np; i += kRedBlock) {
        a += part[i].x;
        b += part[i].y;
    }
    sh[threadIdx.x] = make_double2(a, b);
    __syncthreads();
    for (int s = kRedBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sh[threadIdx.x].x += sh[threadIdx.x + s].x;
            sh[threadIdx.x].y += sh[threadIdx.x + s].y;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[np] = sh[0];
}

int reduce_device(const float *d, size_t nfloat, int mode, double2 *out, hipStream_t s) {
    std::lock_guard<std::mutex> lk(scratch_mutex());
    int nb = (int)((nfloat + kRedBlock * 8 - 1) / (kRedBlock * 8));
    if (nb < 1) nb = 1;
    if (nb > 2048) nb = 2048;
    void *scr = nullptr;
    int rc = scratch_d(sizeof(double2) * (nb + 1), &scr);
    if (rc) return rc;
    double2 *part = reinterpret_cast<double2 *>(scr);
    hipLaunchKernelGGL(k_reduce, dim3(nb), dim3(kRedBlock), 0, s, d, nfloat, mode, part);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kRedBlock), 0, s, part, nb);
    FSIM_HIP(hipGetLastError());
    FSIM_HIP(hipMemcpyAsync(out, part + nb, sizeof(double2), hipMemcpyDeviceToHost, s));
    FSIM_HIP(hipStreamSynchronize(s));
    return FSIM_OK;
}

// Vec2d.Cplx.fourierShift (BasicVectors.java:309-339): phase rounded to
// float, fast sin/cos polynomial in double.
__global__ void k_fourier_shift(float2 *__restrict__ d, int N, double kx, double ky) {
    GRID_STRIDE(i, (size_t)N * N) {
        int x = (int)(i % N), y = (int)(i / N);
        float co, si;
        shift_cos_sin(kx, ky, x, y, N, co, si);
        d[i] = rot_java(d[i], co, si);
    }
}

// Vec2d.Cplx.pasteFreq (BasicVectors.java:259-285), written as a gather over
// the output so that zeroing and pasting are one pass.
__global__ void k_paste_freq(float2 *__restrict__ out, int wo, int ho, const float2 *__restrict__ in,
                             int wi, int hi, int xoff, int yoff) {
    GRID_STRIDE(i, (size_t)wo * ho) {
        int xo = (int)(i % wo), yo = (int)(i / wo);
        int xs = paste_src(xo, wo, wi, xoff);
        int ys = paste_src(yo, ho, hi, yoff);
        out[i] = (xs >= 0 && ys >= 0) ? in[(size_t)ys * wi + xs] : make_float2(0.f, 0.f);
    }
}

}  // namespace fsim

using namespace fsim;

static Vec *V(fsim_vec h) { return reinterpret_cast<Vec *>(h); }
static size_t nel(const Vec *v) { return (size_t)v->w * v->h; }

#define CHECK_INIT()             \
    do {                         \
        int rc_ = ensure_init(); \
        if (rc_) return rc_;     \
    } while (0)
#define LAUNCH_CHECK() FSIM_HIP(hipGetLastError())

static int same_size(Vec *a, Vec *b, const char *op) {
    if (!a || !b) return fail(FSIM_ERR_ARG, std::string(op) + ": null vector");
    if (a->w != b->w || a->h != b->h)
        return fail(FSIM_ERR_ARG, std::string(op) + ": vector size mismatch");
    return FSIM_OK;
}

extern "C" {

int fsim_vec_zero(fsim_vec vh) {
    CHECK_INIT();
    Vec *v = V(vh);
    if (!v) return fail(FSIM_ERR_ARG, "zero: null vector");
    FSIM_HIP(hipMemsetAsync(v->d, 0, v->bytes, stream()));
    return FSIM_OK;
}

int fsim_vec_copy(fsim_vec dh, fsim_vec sh) {
    CHECK_INIT();
    Vec *d = V(dh), *s = V(sh);
    int rc = same_size(d, s, "copy");
    if (rc) return rc;
    size_t n = nel(d);
    if (d->is_cplx == s->is_cplx) {
        FSIM_HIP(hipMemcpyAsync(d->d, s->d, d->bytes, hipMemcpyDeviceToDevice, stream()));
    } else if (d->is_cplx) {
        hipLaunchKernelGGL(k_copy_c_from_r, grid_for(n), dim3(256), 0, stream(), (float2 *)d->d,
                           (const float *)s->d, n);
    } else {
        hipLaunchKernelGGL(k_copy_r_from_c, grid_for(n), dim3(256), 0, stream(), (float *)d->d,
                           (const float2 *)s->d, n);
    }
    LAUNCH_CHECK();
    return FSIM_OK;
}

int fsim_vec_add(fsim_vec dh, fsim_vec sh) {
    CHECK_INIT();
    Vec *d = V(dh), *s = V(sh);
    int rc = same_size(d, s, "add");
    if (rc) return rc;
    if (d->is_cplx != s->is_cplx) return fail(FSIM_ERR_ARG, "add: mixed real/complex");
    size_t nf = nel(d) * (d->is_cplx ? 2 : 1);
    hipLaunchKernelGGL(k_add_f, grid_for(nf), dim3(256), 0, stream(), (float *)d->d, (const float *)s->d, nf);
    LAUNCH_CHECK();
    return FSIM_OK;
}

int fsim_vec_add_const(fsim_vec dh, float re, float im) {
    CHECK_INIT();
    Vec *d = V(dh);
    if (!d) return fail(FSIM_ERR_ARG, "add_const: null vector");
    size_t n = nel(d);
    if (d->is_cplx)
        hipLaunchKernelGGL(k_add_const_c, grid_for(n), dim3(256), 0, stream(), (float2 *)d->d, re, im, n);
    else
        hipLaunchKernelGGL(k_add_const_r, grid_for(n), dim3(256), 0, stream(), (float *)d->d, re, n);
    LAUNCH_CHECK();
    return FSIM_OK;
}

int fsim_vec_axpy(fsim_vec dh, float are, float aim, fsim_vec xh) {
    CHECK_INIT();
    Vec *d = V(dh), *x = V(xh);
    int rc = same_size(d, x, "axpy");
    if (rc) return rc;
    if (d->is_cplx != x->is_cplx) return fail(FSIM_ERR_ARG, "axpy: mixed real/complex");
    size_t n = nel(d);
    if (d->is_cplx)
        hipLaunchKernelGGL(k_axpy_c, grid_for(n), dim3(256), 0, stream(), (float2 *)d->d, are, aim,
                           (const float2 *)x->d, n);
    else
        hipLaunchKernelGGL(k_axpy_r, grid_for(n), dim3(256), 0, stream(), (float *)d->d, are,
                           (const float *)x->d, n);
    LAUNCH_CHECK();
    return FSIM_OK;
}

int fsim_vec_times(fsim_vec dh, fsim_vec xh, int conj) {
    CHECK_INIT();
    Vec *d = V(dh), *x = V(xh);
    int rc = same_size(d, x, "times");
    if (rc) return rc;
    size_t n = nel(d);
    if (d->is_cplx && x->is_cplx)
        hipLaunchKernelGGL(k_times_cc, grid_for(n), dim3(256), 0, stream(), (float2 *)d->d,
                           (const float2 *)x->d, conj, n);
    else if (d->is_cplx)
        hipLaunchKernelGGL(k_times_cr, grid_for(n), dim3(256), 0, stream(), (float2 *)d->d,
                           (const float *)x->d, n);
    else if (!x->is_cplx)
        hipLaunchKernelGGL(k_times_rr, grid_for(n), dim3(256), 0, stream(), (float *)d->d,
                           (const float *)x->d, n);
    else
        return fail(FSIM_ERR_ARG, "times: real vector times complex vector");
    LAUNCH_CHECK();
    return FSIM_OK;
}

int fsim_vec_scal(fsim_vec dh, float re, float im) {
    CHECK_INIT();
    Vec *d = V(dh);
    if (!d) return fail(FSIM_ERR_ARG, "scal: null vector");
    size_t n = nel(d);
    if (d->is_cplx)
        hipLaunchKernelGGL(k_scal_c, grid_for(n), dim3(256), 0, stream(), (float2 *)d->d, re, im, n);
    else
        hipLaunchKernelGGL(k_scal_r, grid_for(n), dim3(256), 0, stream(), (float *)d->d, re, n);
    LAUNCH_CHECK();
    return FSIM_OK;
}

int fsim_vec_norm2(fsim_vec vh, double *out) {
    CHECK_INIT();
    Vec *v = V(vh);
    if (!v || !out) return fail(FSIM_ERR_ARG, "norm2: null argument");
    double2 r;
    int rc = reduce_device((const float *)v->d, nel(v) * (v->is_cplx ? 2 : 1), 0, &r, stream());
    if (rc) return rc;
    *out = r.x;
    return FSIM_OK;
}

int fsim_vec_sum(fsim_vec vh, double *re, double *im) {
    CHECK_INIT();
    Vec *v = V(vh);
    if (!v || !re || !im) return fail(FSIM_ERR_ARG, "sum: null argument");
    double2 r;
    if (v->is_cplx) {
        int rc = reduce_device((const float *)v->d, nel(v) * 2, 1, &r, stream());
        if (rc) return rc;
    } else {
        // real vector: reuse the complex path on (re, 0) pairs is not possible
        // in place; sum real values via mode 1 over pairs and add halves.
        int rc = reduce_device((const float *)v->d, nel(v) & ~(size_t)1, 1, &r, stream());
        if (rc) return rc;
        r.x += r.y;
        r.y = 0;
        if (nel(v) & 1) {
            float last;
            FSIM_HIP(hipMemcpy(&last, (const float *)v->d + nel(v) - 1, sizeof(float), hipMemcpyDeviceToHost));
            r.x += last;
        }
    }
    *re = r.x;
    *im = r.y;
    return FSIM_OK;
}

int fsim_vec_set_u16(fsim_vec vh, const uint16_t *host) {
    CHECK_INIT();
    Vec *v = V(vh);
    if (!v || !host) return fail(FSIM_ERR_ARG, "set_u16: null argument");
    size_t n = nel(v);
    std::lock_guard<std::mutex> lk(scratch_mutex());
    void *scr = nullptr;
    int rc = scratch_d(n * sizeof(uint16_t), &scr);
    if (rc) return rc;
    uint16_t *d = reinterpret_cast<uint16_t *>(scr);
    FSIM_HIP(hipMemcpyAsync(d, host, n * sizeof(uint16_t), hipMemcpyHostToDevice, stream()));
    if (v->is_cplx)
        hipLaunchKernelGGL(k_set_u16, grid_for(n), dim3(256), 0, stream(), (float2 *)v->d, d, n);
    else
        hipLaunchKernelGGL(k_set_u16_r, grid_for(n), dim3(256), 0, stream(), (float *)v->d, d, n);
    LAUNCH_CHECK();
    FSIM_HIP(hipStreamSynchronize(stream()));
    return FSIM_OK;
}

int fsim_vec_set_u16_device(fsim_vec vh, const uint16_t *dev) {
    CHECK_INIT();
    Vec *v = V(vh);
    if (!v || !dev) return fail(FSIM_ERR_ARG, "set_u16_device: null argument");
    size_t n = nel(v);
    if (v->is_cplx)
        hipLaunchKernelGGL(k_set_u16, grid_for(n), dim3(256), 0, stream(), (float2 *)v->d, dev, n);
    else
        hipLaunchKernelGGL(k_set_u16_r, grid_for(n), dim3(256), 0, stream(), (float *)v->d, dev, n);
    LAUNCH_CHECK();
    return FSIM_OK;
}

int fsim_vec_fourier_shift(fsim_vec vh, double kx, double ky) {
    CHECK_INIT();
    Vec *v = V(vh);
    if (!v || !v->is_cplx) return fail(FSIM_ERR_ARG, "fourier_shift: needs a complex vector");
    if (v->w != v->h) return fail(FSIM_ERR_ARG, "fourier_shift: vector not square");
    size_t n = nel(v);
    hipLaunchKernelGGL(k_fourier_shift, grid_for(n, 2), dim3(256), 0, stream(), (float2 *)v->d, v->w, kx, ky);
    LAUNCH_CHECK();
    return FSIM_OK;
}

int fsim_vec_paste_freq(fsim_vec oh, fsim_vec ih, int xoff, int yoff) {
    CHECK_INIT();
    Vec *o = V(oh), *in = V(ih);
    if (!o || !in || !o->is_cplx || !in->is_cplx) return fail(FSIM_ERR_ARG, "paste_freq: needs complex vectors");
    // the reference only pastes into the doubled grid (SimUtils.java:33-34, 84-86)
    if (o->w < 2 * in->w || o->h < 2 * in->h)
        return fail(FSIM_ERR_ARG, "paste_freq: output must be at least twice the input size");
    // (xo + xOffset + wo) % wo must stay non-negative as in the Java loop
    if (xoff < -o->w || yoff < -o->h) return fail(FSIM_ERR_ARG, "paste_freq: offset out of range");
    size_t n = nel(o);
    hipLaunchKernelGGL(k_paste_freq, grid_for(n), dim3(256), 0, stream(), (float2 *)o->d, o->w, o->h,
                       (const float2 *)in->d, in->w, in->h, xoff, yoff);
    LAUNCH_CHECK();
    return FSIM_OK;
}

int fsim_vec_get(fsim_vec vh, int x, int y, float *re, float *im) {
    CHECK_INIT();
    Vec *v = V(vh);
    if (!v || x < 0 || y < 0 || x >= v->w || y >= v->h) return fail(FSIM_ERR_ARG, "get: index out of range");
    size_t i = (size_t)y * v->w + x;
    FSIM_HIP(hipStreamSynchronize(stream()));
    if (v->is_cplx) {
        float2 c;
        FSIM_HIP(hipMemcpy(&c, (float2 *)v->d + i, sizeof(c), hipMemcpyDeviceToHost));
        if (re) *re = c.x;
        if (im) *im = c.y;
    } else {
        float r;
        FSIM_HIP(hipMemcpy(&r, (float *)v->d + i, sizeof(r), hipMemcpyDeviceToHost));
        if (re) *re = r;
        if (im) *im = 0;
    }
    return FSIM_OK;
}

int fsim_vec_set(fsim_vec vh, int x, int y, float re, float im) {
    CHECK_INIT();
    Vec *v = V(vh);
    if (!v || x < 0 || y < 0 || x >= v->w || y >= v->h) return fail(FSIM_ERR_ARG, "set: index out of range");
    size_t i = (size_t)y * v->w + x;
    FSIM_HIP(hipStreamSynchronize(stream()));
    if (v->is_cplx) {
        float2 c = make_float2(re, im);
        FSIM_HIP(hipMemcpy((float2 *)v->d + i, &c, sizeof(c), hipMemcpyHostToDevice));
    } else {
        FSIM_HIP(hipMemcpy((float *)v->d + i, &re, sizeof(re), hipMemcpyHostToDevice));
    }
    return FSIM_OK;
}

}  // extern "C"
