// igaos_ls_dots_f64_noalloc: the line-search dots of igaos_ls_dots_f64
// with the library's own scratch instead of a caller workspace, for
// callers that cannot allocate device memory (code under HIP graph
// capture, serving loops that forbid hipMalloc).
//
// The scratch is one file-scope device array per GPU, held until the
// process ends.  Host code never passes the array's name to a kernel: in
// host code that name is a stand-in, not a device address.  The device
// address is looked up once per device with hipGetSymbolAddress, with that
// device current, and kept per device.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "exact.hpp"
#include "igaos_pdhg.h"

namespace igaos {
int set_last_error(const std::string& what);
}

namespace {

// 3 partials per workgroup of the reference's dot layout, whose grid is
// capped at 4096 workgroups: 96 KiB covers every n, m.
constexpr long SCRATCH_DOUBLES = 3L * igaos::dev::DOTS_MAX_BLOCKS;
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double* g_ptr[MAX_DEVICES] = {};

int fail(const std::string& what) { return igaos::set_last_error(what); }

// Device address of this device's scratch (looked up on first use).
int scratch_ptr(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEVICES) return fail("device index beyond 63");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_ptr[dev];
    return 0;
}

}  // namespace

extern "C" {

int igaos_ls_dots_f64_noalloc(int n, int m, const double* dx, const double* dy,
                              const double* adx, double* out, void* stream) {
    if (n < 0 || m < 0) return fail("negative size");
    if ((n > 0 && !dx) || (m > 0 && (!dy || !adx)) || !out) return fail("null device array");
    const long need = 3L * igaos::dev::dots_blocks(n, m);
    if (need > SCRATCH_DOUBLES)
        return fail("partials of " + std::to_string(need) + " doubles exceed the " +
                    std::to_string(SCRATCH_DOUBLES) + "-double scratch");
    double* part = nullptr;
    if (scratch_ptr(&part) != 0) return -1;
    // every workgroup of the first stage writes all three of its partials,
    // so nothing in the scratch needs clearing between calls
    igaos::dev::ls_dots(n, m, dx, dy, adx, part, out, static_cast<hipStream_t>(stream));
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail(hipGetErrorString(e));
}

}  // extern "C"
