// Library state: device selection, the ordered stream, error strings,
// twiddle tables, device scratch and the vector lifecycle.
#include "fsim_internal.h"

#include <cstdlib>

#include <atomic>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

namespace fsim {

// Work buffers of fsim_reconstruct_noalloc (the A/B intermediates of the
// fused pipeline), for callers that cannot allocate: one array per GPU,
// held until the process ends.  Host code never passes `scratch` itself
// (on the host it is a stand-in, not a device address): fsim_init looks up
// the device address once, with the device current, and everything uses
// that pointer.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static constexpr int kMaxDevices = 64;
static void *g_scratch_dev[kMaxDevices] = {};
static std::string g_scratch_err = "fsim_init has not been called";

static thread_local std::string t_err;
static std::mutex g_mu;
static bool g_init = false;
static int g_device = 0;
static hipStream_t g_stream = nullptr;
static std::map<int, float2 *> g_tw;
static void *g_scratch = nullptr;
static size_t g_scratch_bytes = 0;
std::atomic<int64_t> g_bytes{0};

void set_error(const std::string &msg) { t_err = msg; }
int fail(int code, const std::string &msg) {
    t_err = msg;
    return code;
}

int ensure_init() {
    if (!g_init) return fail(FSIM_ERR_STATE, "fsim_init has not been called");
    return FSIM_OK;
}

hipStream_t stream() { return g_stream; }

const float2 *twiddles(int L) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_tw.find(L);
    if (it != g_tw.end()) return it->second;
    std::vector<float2> h(L);
    for (int m = 0; m < L; m++) {
        double a = -2.0 * M_PI * (double)m / (double)L;
        h[m] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    float2 *d = nullptr;
    if (hipMalloc(&d, sizeof(float2) * L) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), sizeof(float2) * L, hipMemcpyHostToDevice) != hipSuccess)
        return nullptr;
    g_tw[L] = d;
    return d;
}

// Counted device allocations of the library's own state (channel tables,
// work buffers): their bytes are part of fsim_device_bytes, and a test hook
// (fsim_debug_fail_alloc) can make the k-th of them fail to exercise the
// error paths.
static std::map<void *, size_t> g_owned;
static int g_fail_after = -1;

int dev_alloc(void **p, size_t bytes) {
    *p = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (g_fail_after == 0) {
            g_fail_after = -1;
            return FSIM_ERR_OOM;
        }
        if (g_fail_after > 0) g_fail_after--;
    }
    if (hipMalloc(p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return FSIM_ERR_OOM;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_owned[*p] = bytes;
    g_bytes += (int64_t)bytes;
    return FSIM_OK;
}

void dev_free(void *p) {
    if (!p) return;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_owned.find(p);
        if (it != g_owned.end()) {
            g_bytes -= (int64_t)it->second;
            g_owned.erase(it);
        }
    }
    (void)hipFree(p);
}

void *noalloc_scratch(std::string *why) {
    std::lock_guard<std::mutex> lk(g_mu);
    void *p = g_init ? g_scratch_dev[g_device] : nullptr;
    if (!p && why) *why = g_scratch_err;
    return p;
}

const float2 *twiddles_cached(int L) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_tw.find(L);
    return it == g_tw.end() ? nullptr : it->second;
}

std::mutex &scratch_mutex() {
    static std::mutex m;
    return m;
}

int scratch_d(size_t bytes, void **p) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (bytes > g_scratch_bytes) {
        if (g_scratch) {
            FSIM_HIP(hipStreamSynchronize(g_stream));
            FSIM_HIP(hipFree(g_scratch));
            g_scratch = nullptr;
        }
        size_t nb = bytes < (1 << 20) ? (1 << 20) : bytes;
        if (hipMalloc(&g_scratch, nb) != hipSuccess) {
            g_scratch_bytes = 0;
            return fail(FSIM_ERR_OOM, "scratch allocation failed");
        }
        g_scratch_bytes = nb;
    }
    *p = g_scratch;
    return FSIM_OK;
}

}  // namespace fsim

using namespace fsim;

extern "C" {

const char *fsim_last_error(void) { return t_err.c_str(); }

int fsim_init(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_init) {
        if (device != g_device) return fail(FSIM_ERR_STATE, "fsim_init: already bound to another device");
        return FSIM_OK;
    }
    int n = 0;
    FSIM_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(FSIM_ERR_ARG, "fsim_init: no such device");
    if (device >= kMaxDevices) return fail(FSIM_ERR_ARG, "fsim_init: device index out of range");
    FSIM_HIP(hipSetDevice(device));
    if (!g_scratch_dev[device]) {
        // a failed lookup leaves the pointer unset: fsim_reconstruct_noalloc
        // then refuses with this message, every other entry point works
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            g_scratch_err = std::string("no device address for the work-buffer array: ") + hipGetErrorString(e);
        } else {
            g_scratch_dev[device] = ptr;
        }
    }
    FSIM_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    g_device = device;
    g_init = true;
    // a saved autotuning cache named by FSIM_AUTOTUNE_CACHE is merged in; a
    // missing or malformed file leaves the defaults (and the message in
    // fsim_last_error) without failing the initialisation
    if (const char *path = std::getenv("FSIM_AUTOTUNE_CACHE")) (void)fsim_autotune_load(path);
    return FSIM_OK;
}

int fsim_sync(void) {
    int rc = ensure_init();
    if (rc) return rc;
    FSIM_HIP(hipStreamSynchronize(g_stream));
    return FSIM_OK;
}

int64_t fsim_device_bytes(void) { return g_bytes.load(); }

int fsim_debug_fail_alloc(int nth) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_fail_after = nth < 0 ? -1 : nth;
    return FSIM_OK;
}

int fsim_vec_create(int is_cplx, int w, int h, fsim_vec *out) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!out || w <= 0 || h <= 0) return fail(FSIM_ERR_ARG, "fsim_vec_create: bad size");
    Vec *v = new Vec();
    v->is_cplx = is_cplx ? 1 : 0;
    v->w = w;
    v->h = h;
    v->bytes = (size_t)w * h * (is_cplx ? sizeof(float2) : sizeof(float));
    if (hipMalloc(&v->d, v->bytes) != hipSuccess) {
        delete v;
        (void)hipGetLastError();
        return fail(FSIM_ERR_OOM, "fsim_vec_create: device allocation failed");
    }
    if (hipMemsetAsync(v->d, 0, v->bytes, g_stream) != hipSuccess) {
        (void)hipFree(v->d);
        delete v;
        return fail(FSIM_ERR_HIP, "fsim_vec_create: memset failed");
    }
    g_bytes += (int64_t)v->bytes;
    *out = reinterpret_cast<fsim_vec>(v);
    return FSIM_OK;
}

int fsim_vec_destroy(fsim_vec vh) {
    if (!vh) return FSIM_OK;
    Vec *v = reinterpret_cast<Vec *>(vh);
    FSIM_HIP(hipStreamSynchronize(g_stream));
    FSIM_HIP(hipFree(v->d));
    g_bytes -= (int64_t)v->bytes;
    delete v;
    return FSIM_OK;
}

int fsim_vec_shape(fsim_vec vh, int *is_cplx, int *w, int *h) {
    if (!vh) return fail(FSIM_ERR_ARG, "null vector");
    Vec *v = reinterpret_cast<Vec *>(vh);
    if (is_cplx) *is_cplx = v->is_cplx;
    if (w) *w = v->w;
    if (h) *h = v->h;
    return FSIM_OK;
}

int fsim_vec_reshape(fsim_vec vh, int w, int h) {
    if (!vh) return fail(FSIM_ERR_ARG, "null vector");
    Vec *v = reinterpret_cast<Vec *>(vh);
    if (w <= 0 || h <= 0 || (size_t)w * h != (size_t)v->w * v->h)
        return fail(FSIM_ERR_ARG, "reshape: element count mismatch");
    v->w = w;
    v->h = h;
    return FSIM_OK;
}

void *fsim_vec_device_ptr(fsim_vec vh) { return vh ? reinterpret_cast<Vec *>(vh)->d : nullptr; }

int fsim_vec_upload(fsim_vec vh, const float *host) {
    if (!vh || !host) return fail(FSIM_ERR_ARG, "fsim_vec_upload: null argument");
    Vec *v = reinterpret_cast<Vec *>(vh);
    FSIM_HIP(hipMemcpyAsync(v->d, host, v->bytes, hipMemcpyHostToDevice, g_stream));
    FSIM_HIP(hipStreamSynchronize(g_stream));
    return FSIM_OK;
}

int fsim_vec_download(fsim_vec vh, float *host) {
    if (!vh || !host) return fail(FSIM_ERR_ARG, "fsim_vec_download: null argument");
    Vec *v = reinterpret_cast<Vec *>(vh);
    FSIM_HIP(hipMemcpyAsync(host, v->d, v->bytes, hipMemcpyDeviceToHost, g_stream));
    FSIM_HIP(hipStreamSynchronize(g_stream));
    return FSIM_OK;
}

int fsim_alloc_device(int64_t bytes, void **out) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!out || bytes <= 0) return fail(FSIM_ERR_ARG, "alloc_device: bad arguments");
    if (hipMalloc(out, (size_t)bytes) != hipSuccess) {
        (void)hipGetLastError();
        *out = nullptr;
        return fail(FSIM_ERR_OOM, "alloc_device: device allocation failed");
    }
    g_bytes += bytes;
    return FSIM_OK;
}

int fsim_free_device(void *p) {
    if (!p) return FSIM_OK;
    FSIM_HIP(hipStreamSynchronize(g_stream));
    size_t sz = 0;
    if (hipMemPtrGetInfo(p, &sz) == hipSuccess) g_bytes -= (int64_t)sz;
    FSIM_HIP(hipFree(p));
    return FSIM_OK;
}

int fsim_alloc_host(int64_t bytes, void **out) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!out || bytes <= 0) return fail(FSIM_ERR_ARG, "alloc_host: bad arguments");
    if (hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        *out = nullptr;
        return fail(FSIM_ERR_OOM, "alloc_host: pinned allocation failed");
    }
    return FSIM_OK;
}

int fsim_free_host(void *p) {
    if (!p) return FSIM_OK;
    FSIM_HIP(hipHostFree(p));
    return FSIM_OK;
}

int fsim_memcpy(void *dst, const void *src, int64_t bytes, int kind) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!dst || !src || bytes < 0) return fail(FSIM_ERR_ARG, "memcpy: bad arguments");
    hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                                : kind == 2 ? hipMemcpyDeviceToHost
                                            : kind == 3 ? hipMemcpyDeviceToDevice : hipMemcpyDefault;
    if (kind < 1 || kind > 3) return fail(FSIM_ERR_ARG, "memcpy: kind must be 1, 2 or 3");
    FSIM_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, k, g_stream));
    FSIM_HIP(hipStreamSynchronize(g_stream));
    return FSIM_OK;
}

}  // extern "C"
