// mcb_reduce_fields_noalloc (include/mcb.h): the FP64 diagnostics of
// mcb_reduce over caller-owned device fields, for callers that cannot
// allocate device memory (code under hipGraph capture, serving loops that
// forbid allocation). It runs mcb_reduce's kernels (k_reduce_partial,
// k_reduce_final); their per-block partial sums live in `scratch`, a
// file-scope device array, instead of a buffer the library allocates.
//
// `scratch` names a host-side stand-in in host code: its device address is
// taken once per device with hipGetSymbolAddress, with that device current,
// and only that address is passed to kernels. Every call on a device shares
// that device's array: calls that overlap on one device (two streams, or two
// host threads) race on it and must be ordered by the caller. Each call
// overwrites every partial it reads, so nothing has to be cleared between
// calls.
#include <mutex>

#include "fused.hpp"

namespace mcb {
namespace {

// Two doubles (sum of phi, kinetic energy) per reduction block: 16 KiB.
constexpr int kScratchDoubles = 2 * kReduceBlocks;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (resolved on first use).
hipError_t scratch_on_current_device(double** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace
}  // namespace mcb

using namespace mcb;

extern "C" mcb_status mcb_reduce_fields_noalloc(const float* phi, const float* rho, const float* ux, const float* uy,
                                                const float* uz, int NX, int NY, int NZ, double* out, void* stream) {
    if (!phi || !rho || !ux || !uy || !uz || !out) return MCB_ERR_INVALID_ARG;
    if (NX < 1 || NY < 1 || NZ < 1) return MCB_ERR_INVALID_ARG;
    // The same lattice limit as every context: below 2^30 cells.
    const long long n = (long long)NX * NY * NZ;
    if (n >= (1LL << 30)) return MCB_ERR_INVALID_ARG;
    double* partial = nullptr;
    hipError_t e = scratch_on_current_device(&partial);
    if (e != hipSuccess) return hip_status(e);
    return hip_status(launch_reduce(phi, rho, ux, uy, uz, n, partial, out, static_cast<hipStream_t>(stream)));
}
