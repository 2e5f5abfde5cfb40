// Internal declarations shared by the libfsim translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <mutex>
#include <string>

#include "../../include/fsim.h"

namespace fsim {

// ---- error reporting: status code + thread-local message ----------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
#define FSIM_HIP(call)                                                        \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess)                                                 \
            return ::fsim::fail(FSIM_ERR_HIP, std::string("HIP: ") +          \
                                hipGetErrorString(e_) + " at " + __FILE__ +  \
                                ":" + std::to_string(__LINE__));              \
    } while (0)

// ---- device vectors (the Vec2d.Real / Vec2d.Cplx device side) -----------
struct Vec {
    int is_cplx;
    int w, h;
    void *d;       // float2* or float*
    size_t bytes;
};

hipStream_t stream();            // the library's ordered stream
int ensure_init();
const float2 *twiddles(int L);   // exp(-2 pi i m / L), m < L, device resident
int scratch_d(size_t bytes, void **p);   // growing device scratch (reductions, staging)
// counted device allocations (fsim_device_bytes; fsim_debug_fail_alloc)
int dev_alloc(void **p, size_t bytes);
void dev_free(void *p);
template <class T>
inline int dev_alloc_n(T **p, size_t n) {
    void *v = nullptr;
    const int rc = dev_alloc(&v, n * sizeof(T));
    *p = static_cast<T *>(v);
    return rc;
}
// device address of the fixed work-buffer array on the bound device
// (kNoAllocScratchBytes bytes), nullptr before fsim_init
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;
void *noalloc_scratch(std::string *why);   // why: the reason when it returns nullptr
// a twiddle table already created by twiddles(L), or nullptr (never allocates)
const float2 *twiddles_cached(int L);
std::mutex &scratch_mutex();            // held by every user of scratch_d for its whole use

// ---- FFT entry (fft.hip) -------------------------------------------------
bool fft_size_ok(int L);
int fft2d_device(float2 *data, int w, int h, int inverse, hipStream_t s);
// nimg images of w x h, w*h elements apart, transformed in the same launches
int fft2d_batch_device(float2 *data, int w, int h, int nimg, int inverse, hipStream_t s);
// the column pass alone (the second half of fft2d_batch_device, whose row
// pass runs first): for callers that fuse their own row pass
int fft_cols_batch_device(float2 *data, int w, int h, int nimg, int inverse, hipStream_t s);
// the row pass alone (the first half of fft2d_batch_device)
int fft_rows_batch_device(float2 *data, int w, int h, int nimg, int inverse, hipStream_t s);

// small helpers
inline bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace fsim
