// GMRES that allocates nothing: bem_gmres_ex with its workspace taken from a
// static device array instead of a caller's buffer.  For callers that cannot
// allocate device memory (serving loops, harnesses that forbid hipMalloc).
//
// The array below exists once per GPU (each device loads its own copy of the
// code object) and is held until the process ends.  Host code never uses the
// array's name as an address: on the first call on a device, with that device
// current, hipGetSymbolAddress gives the device's own address of its copy,
// which is kept per device and passed to the kernels as their workspace.
// Every call on a GPU shares that GPU's copy, so calls that overlap on one
// GPU (other streams, other threads) race; the GMRES kernels write every
// workspace entry before reading it, so nothing has to start at zero.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/bem_hip.h"
#include "bem_internal.hpp"

static __device__ __align__(256) unsigned char scratch[BEM_GMRES_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// the current device's address of its copy of `scratch`, looked up once per
// device
int scratch_address(void** out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return bem::report_failure(BEM_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return bem::report_failure(BEM_ERR_BAD_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return bem::report_failure(BEM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        if (!p) return bem::report_failure(BEM_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return BEM_OK;
}

}  // namespace

extern "C" int64_t bem_gmres_noalloc_capacity(void) { return (int64_t)BEM_GMRES_NOALLOC_SCRATCH_BYTES; }

extern "C" int bem_gmres_noalloc(const void* A, int64_t n, int64_t lda, int trans, const void* b, void* x, float tol,
                                 int restart, int maxiter, bem_solve_info* info, void* stream)
{
    if (!A || !b || !x) return bem::report_failure(BEM_ERR_BAD_ARG, "null pointer");
    if (n <= 0 || restart <= 0 || restart > (1 << 20))
        return bem::report_failure(BEM_ERR_BAD_ARG, "bad GMRES argument");
    const int64_t need = bem_gmres_work_size(n, restart);
    if (need > (int64_t)BEM_GMRES_NOALLOC_SCRATCH_BYTES)
        return bem::report_failure(BEM_ERR_BAD_ARG,
                                   "GMRES workspace of n = " + std::to_string(n) + ", restart = " +
                                       std::to_string(restart) + " needs " + std::to_string(need) +
                                       " bytes, beyond the static scratch of " +
                                       std::to_string((int64_t)BEM_GMRES_NOALLOC_SCRATCH_BYTES) + " bytes");
    void* work = nullptr;
    if (int rc = scratch_address(&work)) return rc;
    return bem_gmres_timed(A, n, lda, trans, b, x, tol, restart, maxiter, work, info, nullptr, nullptr, stream);
}
