// pam_noalloc.hip -- the misfit without any device allocation: the same two
// reduction passes as pam_misfit (k_residual, k_sum_parts), with the per-block
// partials and the sum in a fixed device array of this file instead of
// library-allocated memory, for callers that cannot allocate (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>

#include "../../include/pam.h"
#include "pam_internal.h"

using namespace pam;

namespace {

// the kMisfitParts float64 partials of the first pass (8 KiB per GPU, held
// until the process ends); the second pass writes the sum straight into the
// caller's output.  Every call on a GPU shares that GPU's copy.
constexpr int kScratchDoubles = kMisfitParts;
static __device__ __align__(256) double scratch[kScratchDoubles];

// the device address of `scratch` on each GPU (each holds its own copy)
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_ptr[kMaxDevices] = {nullptr};

}  // namespace

// the calling thread's last failed context-free call (defined in pam_api.cpp)
extern thread_local std::string g_free_err;

static int na_fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
static int na_fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_free_err = buf;
    return code;
}

// The address of the current device's `scratch`, looked up once per device
// with that device current (pam_create calls this for its device, so a
// context's creation already resolves it).  The host-side symbol `scratch` is
// never passed to a kernel or a copy: only this address is.
int pam::pam_noalloc_scratch(double **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return na_fail(PAM_E_HIP, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return na_fail(PAM_E_UNSUPPORTED, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return na_fail(PAM_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_ptr[dev];
    return PAM_OK;
}

extern "C" int pam_misfit_noalloc(int nshots, int nt_data, int nrec, const float *dobs, const float *dcal,
                                  float *residual, double *misfit, void *hip_stream) {
    if (!dobs || !dcal || !misfit) return na_fail(PAM_E_ARG, "pam_misfit_noalloc: NULL argument");
    if (nshots < 1 || nt_data < 1 || nrec < 1)
        return na_fail(PAM_E_ARG, "pam_misfit_noalloc: nshots, nt_data and nrec must be >= 1");
    const long long n = (long long)nshots * nt_data * nrec;
    if (n > PAM_MISFIT_NOALLOC_MAX_SAMPLES)
        return na_fail(PAM_E_ARG, "pam_misfit_noalloc: %lld samples exceed the maximum of %lld", n,
                       (long long)PAM_MISFIT_NOALLOC_MAX_SAMPLES);
    if (misfit_parts(n) > kScratchDoubles)
        return na_fail(PAM_E_ARG, "pam_misfit_noalloc: %d partial sums exceed the scratch of %d", misfit_parts(n),
                       kScratchDoubles);
    double *sp = nullptr;
    const int r = pam_noalloc_scratch(&sp);
    if (r) return r;
    // every partial the second pass reads is written by the first: nothing to clear
    const int e = launch_residual(dobs, dcal, residual, n, sp, misfit, (hipStream_t)hip_stream);
    if (e) return na_fail(PAM_E_HIP, "pam_misfit_noalloc: launch failed: %s", hipGetErrorString((hipError_t)e));
    return PAM_OK;
}
