// Substitution with a device-resident factor that allocates nothing.
//
// hpl_gesl uploads the factor into its context and keeps the diagonal-block
// inverses in context buffers allocated on demand.  Callers that cannot
// allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc) pass the factor and the right-hand side in
// their own device memory instead; the inverses and the state of the
// single-launch solve (slice sums, flags -- which each launch clears itself)
// then live in one fixed device array, `scratch`, of kScratchBytes per GPU.
// The kernels are hpl_gesl's, in the same order, so the result is hpl_gesl's
// bit for bit.
//
// The array's host-side name is not a device address: the entry point looks
// up its address once per device (hipGetSymbolAddress, with that device
// current) and passes only that address to the kernels.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "../../include/hpl_hip.h"
#include "kernels.hpp"

namespace {

constexpr size_t kScratchBytes = size_t(40) << 20;  // 40 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// The address of this device's copy of `scratch` (looked up on the first call
// on the device, with the device current).
hpl_status scratch_address(int dev, float** out) {
    if (dev < 0 || dev >= kMaxDevices) return HPL_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return HPL_ERR_HIP;
        }
        g_scratch[dev] = p;
    }
    *out = static_cast<float*>(g_scratch[dev]);
    return HPL_SUCCESS;
}

}  // namespace

extern "C" hpl_status hpl_gesl_noalloc(const float* lu, uint64_t n, uint64_t lda, float* x,
                                       void* stream) {
    if (!lu || !x || n == 0 || lda < n || (lda & 3) ||
        (reinterpret_cast<uintptr_t>(lu) & 15) || (reinterpret_cast<uintptr_t>(x) & 15))
        return HPL_ERR_INVALID_ARG;
    // hpl_gesl's block: the largest power of two <= 256 dividing n
    uint32_t B = 256;
    while (B >= 32 && n % B) B /= 2;
    if (B < 32) return HPL_ERR_UNSUPPORTED;
    const uint64_t N = n / B;
    const int64_t inv_stride = 3 * (B / 32) * 1024;        // floats per block
    const uint64_t inv_floats = N * (uint64_t)inv_stride;  // n * 96 floats
    const uint64_t binv_floats = 2 * n * (uint64_t)B;      // B x B inverses of L and U
    const uint64_t ws_bytes = gesl_ws_bytes((int64_t)n);  // the single-launch solve's state
    if ((inv_floats + binv_floats) * sizeof(float) + ws_bytes > kScratchBytes)
        return HPL_ERR_BUDGET;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return HPL_ERR_NO_DEVICE;
    float* base = nullptr;
    hpl_status st = scratch_address(dev, &base);
    if (st) return st;
    float* inv = base;
    float* binv = base + inv_floats;  // inv_floats is a multiple of 1024: 4 KiB aligned
    // binv_floats = 2 n B, a multiple of 64: the workspace stays 256-byte aligned
    const GeslWs ws = gesl_ws_carve(binv + binv_floats, (int64_t)n);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (launch_prep_all(lu, (int64_t)lda, (int)B, (int)N, inv, inv_stride, s) != hipSuccess ||
        launch_block_inverse(lu, (int64_t)lda, (int64_t)n, (int)B, inv, inv_stride, binv, s) !=
            hipSuccess ||
        launch_gesl(lu, (int64_t)lda, (int64_t)n, (int)B, inv, inv_stride, x, s, binv, &ws) !=
            hipSuccess)
        return HPL_ERR_HIP;
    return HPL_SUCCESS;
}
