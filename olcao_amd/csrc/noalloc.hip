// The drop-in complex generalised eigensolver without device allocation:
// olcao_solve_zhegv_noalloc, for callers that may not allocate device memory
// (serving loops, harnesses that forbid hipMalloc).  Instead of per-call
// device buffers it stages A, B, the eigenvalues and the status in a
// file-scope device array of fixed size, one per GPU, and solves with the
// one-workgroup LDS solver (ksolve.hip), which needs no other memory and no
// library handle.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "../../include/olcao_gpu.h"
#include "olcao_internal.h"

namespace olcao {

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;   // 64 MiB per GPU
constexpr int kMaxDevices = 64;

// Device staging of the no-allocation solver.  Host code never passes this
// name to a kernel or a copy: it is a host-side stand-in, not a device
// address; scratch_ptr() gives the device address.
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

double *g_scratch[kMaxDevices] = {};
std::once_flag g_scratchOnce[kMaxDevices];
hipError_t g_scratchErr[kMaxDevices] = {};

// The calling thread's current device's copy of `scratch`, looked up once
// per device with that device current.
hipError_t scratch_ptr(double **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::call_once(g_scratchOnce[dev], [dev] {
        void *p = nullptr;
        g_scratchErr[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch[dev] = g_scratchErr[dev] == hipSuccess ? static_cast<double *>(p) : nullptr;
    });
    *out = g_scratch[dev];
    return g_scratchErr[dev];
}

olcao_status fail(olcao_status st, const std::string &msg)
{
    set_noctx_error(msg);
    return st;
}

}  // namespace

}  // namespace olcao

using namespace olcao;

extern "C" {

// Largest batch of n x n problems the scratch array holds (0 when the LDS
// solver does not take n).
static long long noalloc_capacity(int n)
{
    if (n <= 0 || !small_hegv_fits(n, true)) return 0;
    const size_t per = (4 * (size_t)n * n + n + 1) * sizeof(double);   // A, B, w, status
    return (long long)(kScratchBytes / per);
}

olcao_status olcao_solve_zhegv_noalloc(int n, int nb, double *A, double *B, double *w)
{
    if (!A || !B || !w || n <= 0 || nb <= 0) return fail(OLCAO_ERR_INPUT, "null buffer or non-positive size");
    if (!small_hegv_fits(n, true))
        return fail(OLCAO_ERR_UNSUPPORTED, "the allocation-free solver takes n <= 70 (the LDS solver); n = " +
                                               std::to_string(n));
    const long long cap = noalloc_capacity(n);
    if (nb > cap)
        return fail(OLCAO_ERR_INPUT, "batch of " + std::to_string(nb) + " exceeds the " +
                                         std::to_string(kScratchBytes >> 20) + " MiB scratch (at most " +
                                         std::to_string(cap) + " matrices of n = " + std::to_string(n) + ")");
    double *base = nullptr;
    hipError_t e = scratch_ptr(&base);
    if (e != hipSuccess) return fail(OLCAO_ERR_HIP, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
    const size_t mat = 2 * (size_t)n * n * nb;   // doubles of one batch of complex matrices
    double *dA = base, *dB = dA + mat, *dw = dB + mat;
    int *dinfo = reinterpret_cast<int *>(dw + (size_t)n * nb);
    hipStream_t s = nullptr;   // the null stream: no stream object is created
    if ((e = hipMemcpyAsync(dA, A, mat * sizeof(double), hipMemcpyHostToDevice, s)) == hipSuccess &&
        (e = hipMemcpyAsync(dB, B, mat * sizeof(double), hipMemcpyHostToDevice, s)) == hipSuccess) {
        // every matrix's status starts at 0 (the last call's stays otherwise)
        launch_zero_words(dinfo, nb, s);
        launch_small_hegv(true, n, nb, dA, dB, dw, dinfo, s);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(A, dA, mat * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(w, dw, (size_t)n * nb * sizeof(double), hipMemcpyDeviceToHost, s);
    int bad = 0, badAt = -1;
    for (int k0 = 0; e == hipSuccess && k0 < nb; k0 += 256) {
        int part[256];
        const int m = nb - k0 < 256 ? nb - k0 : 256;
        e = hipMemcpyAsync(part, dinfo + k0, m * sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        for (int k = 0; e == hipSuccess && k < m && !bad; ++k)
            if (part[k] != 0) {
                bad = part[k];
                badAt = k0 + k;
            }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(OLCAO_ERR_HIP, std::string("HIP: ") + hipGetErrorString(e));
    if (bad > n) return fail(OLCAO_ERR_NOT_SPD, "B not positive definite in batch " + std::to_string(badAt));
    if (bad != 0) return fail(OLCAO_ERR_SOLVER, "no convergence in batch " + std::to_string(badAt));
    set_noctx_error(std::string());
    return OLCAO_OK;
}

}  // extern "C"
