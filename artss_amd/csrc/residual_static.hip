// Pressure residual norm ||rhs - lap p||_2 with no device allocation.
//
// The same kernels as VCycleMG::residual (k_residual_partials, then the
// fixed-order k_reduce_sum), but the per-workgroup partial sums and the
// total go to `scratch`, a file-scope device array of fixed size, instead of
// the domain's buffers.  Each GPU the library runs on holds its own copy of
// the array (1 MiB) until the process ends.  Its device address is looked up
// once per device, with that device current, when the first domain on the
// device is created; host code uses only that address, never the array's
// name.  Calls on one GPU share its array: calls that overlap on one GPU
// (two streams, two threads) race.  Every slot a call reads is written
// earlier in the same call, so nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <stdexcept>
#include <string>

#include "domain.h"
#include "kernels.h"
#include "residual_static.h"

namespace artss {

static __device__ __align__(256) double scratch[(1 << 17) + 32];

namespace {

constexpr int kMaxDevices = 64;
// the total sits on its own 256-byte line after the partial sums
constexpr int kTotalSlot = kStaticMaxPartials;
static_assert(sizeof(scratch) == sizeof(double) * (kStaticMaxPartials + 32), "scratch size");

double *g_scratch[kMaxDevices] = {};
std::once_flag g_once[kMaxDevices];

}  // namespace

void static_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device id out of range for static scratch");
    std::call_once(g_once[device], [device] {
        int current = 0;
        ARTSS_HIP(hipGetDevice(&current));
        ARTSS_HIP(hipSetDevice(device));
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        ARTSS_HIP(hipSetDevice(current));
        ARTSS_HIP(e);
        g_scratch[device] = static_cast<double *>(ptr);
    });
    if (!g_scratch[device]) throw hip_error("static scratch of device " + std::to_string(device) + " has no address");
}

double residual_static(Domain &d, const double *p, const double *rhs) {
    const int dev = d.device();
    if (dev < 0 || dev >= kMaxDevices || !g_scratch[dev]) throw hip_error("static scratch not set up on this device");
    const LevelDesc &L = d.level(0);
    const int n = k::partial_count(L);
    if (n > kStaticMaxPartials)
        throw std::invalid_argument("residual_static: " + std::to_string(n) + " workgroups exceed the " +
                                    std::to_string(kStaticMaxPartials) + " partial sums of the static scratch");
    double *base = g_scratch[dev];
    k::residual_partials(d.stream(), L, p, rhs, base);
    k::reduce_sum(d.stream(), base, n, base + kTotalSlot);
    double host = 0.0;
    ARTSS_HIP(hipMemcpyAsync(&host, base + kTotalSlot, sizeof(double), hipMemcpyDeviceToHost, d.stream()));
    ARTSS_HIP(hipStreamSynchronize(d.stream()));
    return std::sqrt(host);
}

}  // namespace artss
