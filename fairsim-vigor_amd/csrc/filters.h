// Device-resident OTF tables and helpers shared by filters.hip and recon.hip.
#pragma once
#include <hip/hip_runtime.h>

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
