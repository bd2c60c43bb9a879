// FP64 LINPACK behind the C ABI: the reference built with DATA_TYPE=double
// (its "_DP" build; gefa_ref_nopvt / gesl_ref_nopvt with HOST_DATA_TYPE
// double, LINPACK/src/host/linpack_data.cpp:166-229).  Kernels in f64.hip.
//
// Schedule: blocked right-looking LU with block 64 and the FP32 path's two
// schedules: pipelined (default) -- the panels of one group of `depth` steps
// and their row/column-block updates on s0 while s1 applies the previous
// group's steps (K = depth * 64) to the trailing matrix beyond, in column
// slices -- or grouped (HPL_SCHEDULE=grouped: depth-1 look-ahead with a bulk
// update every `depth` steps).  K = 64 alone would leave the FP64 update at 8
// flop per byte of C traffic, under the device's ~10.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "../../include/hpl_hip.h"
#include "ctx.hpp"
#include "kernels.hpp"

namespace {

using hplx::fail;

// Update depth (steps per bulk update, K = depth * 64): the context's
// depth_f64.
int f64_depth(const hpl_ctx* c) { return c->depth_f64; }

#define FCHK(ctx, expr)                                                                   \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail((ctx), e_ == hipErrorOutOfMemory ? HPL_ERR_OUT_OF_MEMORY : HPL_ERR_HIP, \
                        "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                        \
    } while (0)

template <typename T>
hpl_status grow(hpl_ctx* c, T*& p, size_t& cap, size_t bytes) {
    void* v = p;
    hpl_status st = hplx::ensure_mem(c, &v, &cap, bytes);
    p = static_cast<T*>(v);
    return st;
}

hpl_status alloc_f64(hpl_ctx* c, uint64_t n) {
    if (n == 0 || n % (uint64_t)f64_block())
        return fail(c, HPL_ERR_INVALID_ARG, "FP64 path: n=%llu must be a positive multiple of %d",
                    (unsigned long long)n, f64_block());
    const uint64_t lda = hplx::device_lda(n);
    hpl_status st = grow(c, c->dA64, c->capA64, n * lda * sizeof(double));
    if (!st) st = grow(c, c->dA064, c->capA064, n * lda * sizeof(double));
    if (!st && n * sizeof(double) > c->capV64) {
        size_t cap = 0;
        st = grow(c, c->dB64s, cap, n * sizeof(double));
        cap = 0;
        if (!st) st = grow(c, c->dX64s, cap, n * sizeof(double));
        cap = 0;
        if (!st) st = grow(c, c->dOut64, cap, 4 * sizeof(double));
        if (!st) c->capV64 = n * sizeof(double);
    }
    if (!st) st = grow(c, c->dWork64, c->capWork64, residual_f64_work((int64_t)n) * sizeof(double));
    if (!st) st = hplx::ensure_mem(c, reinterpret_cast<void**>(&c->dStatus), &c->capStatus, 64);
    if (st) return st;
    c->n64 = n;
    c->lda64 = lda;
    c->have_orig64 = false;
    c->factored64 = false;
    return HPL_SUCCESS;
}

// Slot of step k's transposed left strip (64 x n): one per step in the
// pipelined schedule (any window of steps is contiguous), else two groups of
// `depth` reused in turn.
int64_t lt_slot64(const hpl_ctx* c, int64_t k) {
    if (c->f64_linear) return k;
    const int D = f64_depth(c);
    return ((k / D) % 2) * D + (k % D);
}

// C += L T over rows [r0, r0+M) and columns [c0, c0+N), steps k0 .. k0+K/64-1
hpl_status update64(hpl_ctx* c, hipStream_t st, int64_t k0, int64_t r0, int64_t c0, int64_t M,
                    int64_t N, int K) {
    if (M <= 0 || N <= 0) return HPL_SUCCESS;
    const int64_t n = (int64_t)c->n64, lda = (int64_t)c->lda64, B = f64_block();
    const double* Lt = c->dLt64 + lt_slot64(c, k0) * B * n + r0;
    const double* T = c->dA64 + k0 * B * lda + c0;
    FCHK(c, launch_update_f64(c->dA64 + r0 * lda + c0, lda, Lt, n, T, lda, (int)M, (int)N, K, st));
    return HPL_SUCCESS;
}

hpl_status panel64(hpl_ctx* c, hipStream_t st, int64_t k) {
    const int64_t n = (int64_t)c->n64, lda = (int64_t)c->lda64, B = f64_block();
    const int64_t off = k * B;
    FCHK(c, launch_lu_f64(c->dA64, lda, off, c->dStatus, st));
    const int64_t m = n - off - B;
    if (m > 0)
        FCHK(c, launch_strips_f64(c->dA64, lda, off, m,
                                  c->dLt64 + lt_slot64(c, k) * B * n + off + B, n, st));
    return HPL_SUCCESS;
}

// Depth-1 look-ahead as the FP32 schedule: once the next panel's row and
// column blocks are up to date (event evA), that panel is factorised on s1
// while s0 runs the bulk update; s0 waits for it (evP) before the next step.
// The panel writes its own Lt slot, which the bulk update of this step does
// not read, and touches only its row and column blocks.
hpl_status schedule64(hpl_ctx* c) {
    const int64_t n = (int64_t)c->n64, B = f64_block(), N = n / B;
    hpl_status r = panel64(c, c->s0, 0);
    if (r) return r;
    for (int64_t k = 0; k + 1 < N; ++k) {
        const int64_t k1 = (k + 1) * B, m = n - k1;
        const int64_t j = k % f64_depth(c);
        const int K = (int)((j + 1) * B);
        const int64_t k0 = k - j;
        if (m <= B) {
            r = update64(c, c->s0, k0, k1, k1, m, m, K);
            if (!r) r = panel64(c, c->s0, k + 1);
            if (r) return r;
            continue;
        }
        r = update64(c, c->s0, k0, k1, k1, B, m, K);  // row block of the next panel
        if (!r) r = update64(c, c->s0, k0, k1 + B, k1, m - B, B, K);  // column block
        if (r) return r;
        FCHK(c, hipEventRecord(c->evA, c->s0));
        FCHK(c, hipStreamWaitEvent(c->s1, c->evA, 0));
        r = panel64(c, c->s1, k + 1);
        if (r) return r;
        FCHK(c, hipEventRecord(c->evP, c->s1));
        if (j == f64_depth(c) - 1) {
            r = update64(c, c->s0, k0, k1 + B, k1 + B, m - B, m - B, K);
            if (r) return r;
        }
        FCHK(c, hipStreamWaitEvent(c->s0, c->evP, 0));
    }
    return HPL_SUCCESS;
}

// Pipelined schedule, as the FP32 one (gefa_pipelined in api.hip): phase g
// factorises the blocks of group g on s0, each after the update of its row
// and column blocks by the steps of groups g-1 and g before it, while s1
// applies group g-1's steps to rows and columns >= (g+1) depth 64 in column
// slices; the streams join at the end of the phase.
hpl_status schedule64_pipelined(hpl_ctx* c) {
    const int64_t n = (int64_t)c->n64, B = f64_block(), N = n / B;
    // tail groups of at most 3 steps
    const std::vector<int64_t> bnd = hplx::group_bounds(c, N, B, f64_depth(c), 3);
    const int64_t G = (int64_t)bnd.size() - 1;
    hpl_status r;
    for (int64_t g = 0; g < G; ++g) {
        const int64_t s = bnd[g + 1] * B;
        if (g >= 1 && s < n) {
            const int64_t steps = bnd[g] - bnd[g - 1];
            const int64_t slices = c->bulk_slices > 0 ? c->bulk_slices : steps;
            const int64_t w = n - s, tiles = w / 128;
            const int64_t S = std::max<int64_t>(1, std::min(slices, tiles));
            for (int64_t i = 0; i < S; ++i) {
                const int64_t a = s + tiles * i / S * 128;
                const int64_t b = i + 1 == S ? n : s + tiles * (i + 1) / S * 128;
                r = update64(c, c->s1, bnd[g - 1], s, a, w, b - a, (int)(steps * B));
                if (r) return r;
            }
        }
        FCHK(c, hipEventRecord(c->evP, c->s1));
        const int64_t k0 = g >= 1 ? bnd[g - 1] : 0;
        for (int64_t k = bnd[g]; k < bnd[g + 1]; ++k) {
            const int64_t kb = k * B, m = n - kb;
            if (k > k0) {
                const int K = (int)((k - k0) * B);
                r = update64(c, c->s0, k0, kb, kb, B, m, K);  // row block k
                if (!r) r = update64(c, c->s0, k0, kb + B, kb, m - B, B, K);  // column block k
                if (r) return r;
            }
            r = panel64(c, c->s0, k);
            if (r) return r;
        }
        FCHK(c, hipEventRecord(c->evA, c->s0));
        FCHK(c, hipStreamWaitEvent(c->s1, c->evA, 0));
        FCHK(c, hipStreamWaitEvent(c->s0, c->evP, 0));
    }
    return HPL_SUCCESS;
}

double seconds(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms * 1e-3;
}

}  // namespace

extern "C" {

hpl_status hpl_upload_f64(hpl_ctx* c, const double* a, uint64_t n, uint64_t lda, const double* b,
                          hpl_timing* t) {
    if (!c) return HPL_ERR_INVALID_ARG;
    if (!a || lda < n) return fail(c, HPL_ERR_INVALID_ARG, "bad upload_f64 arguments");
    FCHK(c, hipSetDevice(c->device));
    hpl_status st = alloc_f64(c, n);
    if (st) return st;
    FCHK(c, hipEventRecord(c->e0, c->s0));
    FCHK(c, hipMemcpy2DAsync(c->dA64, c->lda64 * sizeof(double), a, lda * sizeof(double),
                             n * sizeof(double), n, hipMemcpyHostToDevice, c->s0));
    if (b)
        FCHK(c, hipMemcpyAsync(c->dB64s, b, n * sizeof(double), hipMemcpyHostToDevice, c->s0));
    else
        FCHK(c, launch_fill_f64(c->dB64s, (int64_t)n, 0.0, c->s0));
    FCHK(c, hipMemcpyAsync(c->dA064, c->dA64, n * c->lda64 * sizeof(double),
                           hipMemcpyDeviceToDevice, c->s0));
    FCHK(c, hipEventRecord(c->e1, c->s0));
    FCHK(c, hipStreamSynchronize(c->s0));
    c->have_orig64 = true;
    if (t) t->h2d_s = seconds(c->e0, c->e1);
    return HPL_SUCCESS;
}

hpl_status hpl_generate_f64(hpl_ctx* c, uint64_t n, int mode, uint64_t seed, double d) {
    if (!c) return HPL_ERR_INVALID_ARG;
    // the FP32 generator's system, widened: M exactly as hpl_generate makes
    // it, b = M x_true in FP64 from those values
    hpl_status st = hpl_generate(c, n, mode, seed, d);
    if (st) return st;
    st = alloc_f64(c, n);
    if (st) return st;
    FCHK(c, launch_f32_to_f64(c->dA, (int64_t)c->lda, c->dA64, (int64_t)c->lda64, (int64_t)n,
                              c->s0));
    FCHK(c, hipMemcpyAsync(c->dB64s, c->db64, n * sizeof(double), hipMemcpyDeviceToDevice, c->s0));
    FCHK(c, hipMemcpyAsync(c->dA064, c->dA64, n * c->lda64 * sizeof(double),
                           hipMemcpyDeviceToDevice, c->s0));
    FCHK(c, hipStreamSynchronize(c->s0));
    c->have_orig64 = true;
    return HPL_SUCCESS;
}

hpl_status hpl_restore_f64(hpl_ctx* c) {
    if (!c) return HPL_ERR_INVALID_ARG;
    if (!c->have_orig64) return fail(c, HPL_ERR_STATE, "no FP64 system resident");
    FCHK(c, hipMemcpyAsync(c->dA64, c->dA064, c->n64 * c->lda64 * sizeof(double),
                           hipMemcpyDeviceToDevice, c->s0));
    FCHK(c, hipStreamSynchronize(c->s0));
    c->factored64 = false;
    return HPL_SUCCESS;
}

hpl_status hpl_gefa_dev_f64(hpl_ctx* c, hpl_timing* t) {
    if (!c) return HPL_ERR_INVALID_ARG;
    if (!c->dA64 || c->n64 == 0) return fail(c, HPL_ERR_STATE, "no FP64 system resident");
    FCHK(c, hipSetDevice(c->device));
    FCHK(c, hipMemsetAsync(c->dStatus, 0, sizeof(int), c->s0));
    c->factored64 = false;
    // strip slots: one per step (pipelined) or 2 * depth
    c->f64_linear = c->pipelined && c->lookahead;
    const uint64_t n = c->n64;
    const size_t slots = c->f64_linear ? n / f64_block() : 2 * (size_t)f64_depth(c);
    hpl_status st = grow(c, c->dLt64, c->capLt64, slots * f64_block() * n * sizeof(double));
    if (st) return st;
    FCHK(c, hipEventRecord(c->e0, c->s0));
    st = c->f64_linear ? schedule64_pipelined(c) : schedule64(c);
    if (st) return st;
    FCHK(c, hipEventRecord(c->e1, c->s0));
    FCHK(c, hipStreamSynchronize(c->s0));
    if (t) t->gefa_s = seconds(c->e0, c->e1);
    int h = 0;
    FCHK(c, hipMemcpy(&h, c->dStatus, sizeof(int), hipMemcpyDeviceToHost));
    if (h) return fail(c, HPL_ERR_ZERO_PIVOT, "zero or non-finite pivot met during FP64 GEFA");
    c->factored64 = true;
    return HPL_SUCCESS;
}

hpl_status hpl_gesl_dev_f64(hpl_ctx* c, hpl_timing* t) {
    if (!c) return HPL_ERR_INVALID_ARG;
    if (!c->factored64) return fail(c, HPL_ERR_STATE, "hpl_gesl_dev_f64 before a successful GEFA");
    FCHK(c, hipSetDevice(c->device));
    FCHK(c, hipEventRecord(c->e0, c->s0));
    FCHK(c, hipMemcpyAsync(c->dX64s, c->dB64s, c->n64 * sizeof(double), hipMemcpyDeviceToDevice,
                           c->s0));
    FCHK(c, launch_gesl_f64(c->dA64, (int64_t)c->lda64, (int64_t)c->n64, c->dX64s, c->s0));
    FCHK(c, hipEventRecord(c->e1, c->s0));
    FCHK(c, hipStreamSynchronize(c->s0));
    if (t) t->gesl_s = seconds(c->e0, c->e1);
    return HPL_SUCCESS;
}

hpl_status hpl_download_f64(hpl_ctx* c, double* a, uint64_t lda, double* x, hpl_timing* t) {
    if (!c) return HPL_ERR_INVALID_ARG;
    if (!c->dA64) return fail(c, HPL_ERR_STATE, "no FP64 system resident");
    if (a && lda < c->n64) return fail(c, HPL_ERR_INVALID_ARG, "lda < n");
    FCHK(c, hipSetDevice(c->device));
    FCHK(c, hipEventRecord(c->e0, c->s0));
    if (a)
        FCHK(c, hipMemcpy2DAsync(a, lda * sizeof(double), c->dA64, c->lda64 * sizeof(double),
                                 c->n64 * sizeof(double), c->n64, hipMemcpyDeviceToHost, c->s0));
    if (x)
        FCHK(c, hipMemcpyAsync(x, c->dX64s, c->n64 * sizeof(double), hipMemcpyDeviceToHost, c->s0));
    FCHK(c, hipEventRecord(c->e1, c->s0));
    FCHK(c, hipStreamSynchronize(c->s0));
    if (t) t->d2h_s = seconds(c->e0, c->e1);
    return HPL_SUCCESS;
}

hpl_status hpl_residual_f64(hpl_ctx* c, double out[4]) {
    if (!c || !out) return HPL_ERR_INVALID_ARG;
    if (!c->have_orig64) return fail(c, HPL_ERR_STATE, "no FP64 system resident");
    FCHK(c, hipSetDevice(c->device));
    FCHK(c, launch_residual_f64(c->dA064, (int64_t)c->lda64, (int64_t)c->n64, c->dX64s, c->dB64s,
                                c->dWork64, c->dOut64, c->s0));
    FCHK(c, hipMemcpyAsync(out, c->dOut64, 4 * sizeof(double), hipMemcpyDeviceToHost, c->s0));
    FCHK(c, hipStreamSynchronize(c->s0));
    return HPL_SUCCESS;
}

hpl_status hpl_gefa_f64(hpl_ctx* c, double* a, uint64_t n, uint64_t lda, hpl_timing* t) {
    if (!c) return HPL_ERR_INVALID_ARG;
    hpl_status st = hpl_upload_f64(c, a, n, lda, nullptr, t);
    if (!st) st = hpl_gefa_dev_f64(c, t);
    if (!st) st = hpl_download_f64(c, a, lda, nullptr, t);
    return st;
}

hpl_status hpl_gesl_f64(hpl_ctx* c, const double* lu, double* b, uint64_t n, uint64_t lda,
                        hpl_timing* t) {
    if (!c) return HPL_ERR_INVALID_ARG;
    if (!lu || !b) return fail(c, HPL_ERR_INVALID_ARG, "bad gesl_f64 arguments");
    hpl_status st = hpl_upload_f64(c, lu, n, lda, b, t);
    if (st) return st;
    c->factored64 = true;
    st = hpl_gesl_dev_f64(c, t);
    if (st) return st;
    FCHK(c, hipMemcpy(b, c->dX64s, n * sizeof(double), hipMemcpyDeviceToHost));
    return HPL_SUCCESS;
}

}  // extern "C"
