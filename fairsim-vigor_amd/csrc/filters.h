// Device-resident OTF tables and helpers shared by filters.hip and recon.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "simmath.h"

namespace fsim {

constexpr int kMaxDir = 5;
constexpr int kMaxBand = 3;
constexpr int kMaxPhase = 7;
constexpr int kMaxComp = 2 * kMaxBand - 1;

// OtfProvider's per-band 1-D tables (OtfProvider.java:39-45) on the device.
struct Otf {
    int nband = 0, ns = 0;
    double cutoff = 0, cpm = 0;
    float2 *vals = nullptr, *vals_att = nullptr;
    float *only_att = nullptr;
};

// Band-1 shift per direction (SimParam.Dir pX, pY), passed by value.
struct ShiftSet {
    int ndir, nband;
    double px[kMaxDir], py[kMaxDir];
};

// WienerFilter.updateCache (WienerFilter.java:46-122) at pixel (x, y): over
// directions d and bands b, |OTF_b|^2 (times the attenuation when on) at
// the distances of (xh, yh) to -+ (px_d, py_d) * b, accumulated in float as
// (acc + v1) + v2, d outer, b inner.  Band 0 counts twice (px(0) = 0).
__device__ __forceinline__ float wiener_sum(int x, int y, int w, int h, const float2 *__restrict__ vals,
                                            const float *__restrict__ only_att, int ns, int otf_nb, double cpm,
                                            double cutoff, double vec_cpm, const ShiftSet &sh, int use_att) {
#pragma clang fp contract(off)
    const double xh = xh_of(x, w), yh = yh_of(y, h);
    float acc = 0.f;
    for (int d = 0; d < sh.ndir; d++)
        for (int b = 0; b < sh.nband; b++) {
            const double bx = sh.px[d] * b, by = sh.py[d] * b;
            const double r1 = sqrt((xh - bx) * (xh - bx) + (yh - by) * (yh - by)) * vec_cpm;
            const double r2 = sqrt((xh + bx) * (xh + bx) + (yh + by) * (yh + by)) * vec_cpm;
            // a single-band OTF serves every band (OtfProvider.java:192-193)
            const size_t tb = (size_t)(otf_nb > 1 ? b : 0) * ns;
            const float2 o1 = otf_val(vals + tb, ns, cpm, cutoff, r1);
            const float2 o2 = otf_val(vals + tb, ns, cpm, cutoff, r2);
            float v1 = o1.x * o1.x + o1.y * o1.y;
            float v2 = o2.x * o2.x + o2.y * o2.y;
            if (use_att) {
                v1 *= att_val(only_att + tb, cpm, cutoff, r1);
                v2 *= att_val(only_att + tb, cpm, cutoff, r2);
            }
            acc = acc + v1 + v2;
        }
    return acc;
}

int wiener_device(float *out, int w, int h, const Otf &o, double vec_cpm, const ShiftSet &sh, int use_att,
                  double wparam, hipStream_t s);
void fade_factors_host(int n, int px, double *lo, double *hi);
int separation_matrix_host(const double *phases, int nphase, int nband, const double *fac, double *out);
int reduce_device(const float *d, size_t nfloat, int mode, double2 *out, hipStream_t s);
// otfToVector(write) of one band table into a w x h complex buffer
int otf_table_device(float2 *out, int w, int h, const float2 *tab, int ns, double cpm, double cutoff,
                     double vec_cpm, hipStream_t s);
// writeApoVector into a real buffer (OtfProvider.java:469-492)
int write_apo_device(float *out, int w, int h, double vec_cpm, double cutoff, double bend, double apo_cut,
                     hipStream_t s);

}  // namespace fsim

// handle behind fsim_otf
struct fsim_otf_s {
    fsim::Otf o;
};
