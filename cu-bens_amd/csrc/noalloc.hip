// cb_dots_noalloc: the fused dot products of cb_dots for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops).
//
// The partial sums live in `scratch`, a static device array of
// DOTS_PARTS + DOTS_MAX doubles (33 KB) that every GPU the library runs on
// holds until the process ends.  In host code the array's name is a host-side
// stand-in, not a device address, so the entry point looks the address up once
// per device with hipGetSymbolAddress (with that device current) and only ever
// passes that address to the kernels.  Every call on a GPU shares that GPU's
// array: calls that overlap on one GPU race.  The kernels write every partial
// sum they read, so the array needs no clearing between calls.
#include <mutex>
#include <string>

#include "dev.h"

static __device__ __align__(256) double scratch[cb::DOTS_PARTS + cb::DOTS_MAX];

namespace cb {
int set_last_error(int code, const std::string& msg);
}

namespace {

constexpr int MAX_DEVICES = 64;
double* g_scratch[MAX_DEVICES];
hipError_t g_status[MAX_DEVICES];
std::once_flag g_once[MAX_DEVICES];

// address of this device's copy of scratch (dev must be current)
double* scratch_on(int dev, hipError_t* status) {
    std::call_once(g_once[dev], [dev]() {
        void* p = nullptr;
        g_status[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch[dev] = g_status[dev] == hipSuccess ? static_cast<double*>(p) : nullptr;
    });
    *status = g_status[dev];
    return g_scratch[dev];
}

}  // namespace

extern "C" int cb_dots_noalloc(int n, const double* const* a, const double* const* b, long len, double* out,
                               void* stream) {
    if (n < 1 || n > cb::DOTS_MAX)
        return cb::set_last_error(-1, "cb_dots_noalloc: 1 to 8 products per call (its static scratch "
                                      "holds the partial sums of 8)");
    if (!a || !b || !out || len < 0) return cb::set_last_error(-1, "null argument or negative length");
    for (int k = 0; k < n; ++k)
        if (len > 0 && (!a[k] || !b[k])) return cb::set_last_error(-1, "null vector");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0)
        return cb::set_last_error(-5, "no HIP device available");
    if (dev >= MAX_DEVICES) return cb::set_last_error(-5, "device index beyond the 64 this entry point serves");
    hipError_t st = hipSuccess;
    double* part = scratch_on(dev, &st);
    if (!part)
        return cb::set_last_error(-2, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(st));
    try {
        cb::launch_dots(static_cast<hipStream_t>(stream), n, a, b, len, part, out);
    } catch (const cb::DeviceError& e) {
        return cb::set_last_error(-2, e.msg);
    }
    return 0;
}
