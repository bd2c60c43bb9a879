// C ABI of libhpl_hip.so: context, resident system and the GEFA/GESL schedule.
//
// This file replaces the OpenCL execution backend
// (LINPACK/src/host/execution_types/execution_pcie.hpp:53-728).  Where the
// reference launches one single-work-item kernel per block and round-trips
// every panel through the host, here one step of the blocked right-looking
// factorisation is four stream-ordered launches on device-resident data:
//   lu (diagonal block) -> prep (inverses of 32x32 diagonal sub-systems)
//   -> trsm (whole top and left strips) -> gemm (whole trailing matrix).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hpl_hip.h"
#include "ctx.hpp"
#include "kernels.hpp"

namespace hplx {

hpl_status fail(hpl_ctx* c, hpl_status st, const char* fmt, ...) {
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return st;
}

}  // namespace hplx

namespace {

using hplx::fail;

constexpr int kMaxLeafLU = 128;

#define HIPCHK(ctx, expr)                                                                 \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail((ctx), e_ == hipErrorOutOfMemory ? HPL_ERR_OUT_OF_MEMORY : HPL_ERR_HIP, \
                        "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                        \
    } while (0)

template <typename T>
hpl_status ensure(hpl_ctx* c, T*& p, size_t& cap, size_t bytes) {
    if (bytes <= cap && p) return HPL_SUCCESS;
    if (c->external) {
        if (bytes == 0) return HPL_SUCCESS;
        return fail(c, HPL_ERR_BUDGET, "workspace of %zu bytes exceeds the reserved %zu", bytes,
                    cap);
    }
    if (p) {
        (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    if (bytes == 0) return HPL_SUCCESS;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), bytes);
    if (e != hipSuccess) {
        p = nullptr;
        (void)hipGetLastError();
        return fail(c, HPL_ERR_OUT_OF_MEMORY, "hipMalloc of %zu bytes failed: %s", bytes,
                    hipGetErrorString(e));
    }
    cap = bytes;
    return HPL_SUCCESS;
}

bool valid_block(uint64_t n, uint32_t B) {
    if (B < 32 || B > 512 || (B & (B - 1))) return false;
    return n > 0 && n % B == 0;
}

using hplx::device_lda;

hpl_status alloc_system(hpl_ctx* c, uint64_t n) {
    const uint64_t lda = device_lda(n);
    hpl_status st = ensure(c, c->dA, c->capA, n * lda * sizeof(float));
    if (st) return st;
    if (c->keep_orig) {
        st = ensure(c, c->dA0, c->capA0, n * lda * sizeof(float));
        if (st) return st;
    }
    const size_t vb = n * sizeof(double);
    if (vb > c->capVec) {
        for (void* p : {(void*)c->db, (void*)c->dx, (void*)c->db64, (void*)c->dxt})
            if (p) (void)hipFree(p);
        c->db = c->dx = nullptr;
        c->db64 = c->dxt = nullptr;
        c->capVec = 0;
        HIPCHK(c, hipMalloc(&c->db, n * sizeof(float)));
        HIPCHK(c, hipMalloc(&c->dx, n * sizeof(float)));
        HIPCHK(c, hipMalloc(&c->db64, vb));
        HIPCHK(c, hipMalloc(&c->dxt, vb));
        c->capVec = vb;
    }
    c->n = n;
    c->lda = lda;
    c->factored = false;
    c->pivoted = false;
    c->have_orig = false;
    c->B = 0;
    return HPL_SUCCESS;
}

size_t lp_elem_bytes(int prec) { return prec == HPL_PREC_FP8 ? 1 : 2; }

hpl_status alloc_factor_ws(hpl_ctx* c, uint32_t B) {
    const hplx::WsPlan w = hplx::plan_factor_ws(c->n, B, c->prec, c->depth, c->lt_linear);
    hpl_status st = ensure(c, c->dLt, c->capLt, w.lt);
    if (st) return st;
    st = ensure(c, c->dDT, c->capDT, w.dt);
    if (st) return st;
    st = ensure(c, c->dInv, c->capInv, w.inv);
    if (st) return st;
    st = ensure(c, c->dInner, c->capInner, w.inner);
    if (st) return st;
    st = ensure(c, c->dTinv, c->capTinv, c->strip_inverse ? w.tinv : 0);
    if (st) return st;
    st = ensure(c, c->dLpT, c->capLpT, w.lp);
    if (st) return st;
    st = ensure(c, c->dLpL, c->capLpL, w.lp);
    if (st) return st;
    st = ensure(c, c->dScT, c->capScT, w.sc);
    if (st) return st;
    st = ensure(c, c->dScL, c->capScL, w.sc);
    if (st) return st;
    return ensure(c, c->dStatus, c->capStatus, w.status);
}

// Slot of step k's transposed left strip: groups of `depth` consecutive steps
// fill one half of the buffer in order, so a group's strips are contiguous
// rows (k-major) and can feed one GEMM with K = depth * B.  The linear layout
// (one slot per step) makes any window of steps contiguous.
int64_t lt_slot(const hpl_ctx* c, int64_t k) {
    if (c->lt_linear) return k;
    return ((k / c->depth) % 2) * c->depth + (k % c->depth);
}

// Trailing update C += L T over rows [r0, r0+M) and columns [c0, c0+N) of A,
// applying the pending steps k0 .. k0 + K/B - 1 (their operands are stacked
// contiguously by the slot layout).  FP32 operands: Lt slots (k-major) and the
// top strips in A; 16-bit operands: the k-contiguous copies.
// An update region: rows [r0, r0+M), columns [c0, c0+N) of the trailing matrix.
struct UpdRegion {
    int64_t r0, c0;
    int M, N;
};

// Trailing update of one region, or of two in a single launch (b.M > 0).
hpl_status launch_update2(hpl_ctx* c, hipStream_t st, int64_t k0, UpdRegion a, UpdRegion b, int K,
                          uint32_t B, int max_groups = 0);

hpl_status launch_update(hpl_ctx* c, hipStream_t st, int64_t k0, int64_t r0, int64_t c0, int M,
                         int N, int K, uint32_t B, int max_groups = 0) {
    return launch_update2(c, st, k0, UpdRegion{r0, c0, M, N}, UpdRegion{0, 0, 0, 0}, K, B,
                          max_groups);
}

hpl_status launch_update2(hpl_ctx* c, hipStream_t st, int64_t k0, UpdRegion a, UpdRegion b, int K,
                          uint32_t B, int max_groups) {
    if (a.M <= 0 || a.N <= 0) {
        a = b;
        b = UpdRegion{0, 0, 0, 0};
    }
    if (a.M <= 0 || a.N <= 0) return HPL_SUCCESS;
    if (b.M <= 0 || b.N <= 0) b = UpdRegion{0, 0, 0, 0};
    const int64_t n = (int64_t)c->n, lda = (int64_t)c->lda;
    const int64_t r0 = a.r0, c0 = a.c0;
    const int M = a.M, N = a.N;
    float* C = c->dA + r0 * lda + c0;
    float* C2 = c->dA + b.r0 * lda + b.c0;
    hipEvent_t ev_beg = nullptr, ev_end = nullptr;
    if (c->prof) {
        if (c->ev_used == c->ev.size()) {
            hipEvent_t x, y;
            HIPCHK(c, hipEventCreate(&x));
            HIPCHK(c, hipEventCreate(&y));
            c->ev.emplace_back(x, y);
        }
        ev_beg = c->ev[c->ev_used].first;
        ev_end = c->ev[c->ev_used].second;
        c->ev_used++;
        HIPCHK(c, hipEventRecord(ev_beg, st));
    }
    const int64_t slot = lt_slot(c, k0);
    if (c->prec == 0) {
        const float* Lt = c->dLt + slot * (int64_t)B * n;
        const float* T = c->dA + k0 * B * lda;
        HIPCHK(c, launch_gemm_update_f32_2(C, lda, Lt + r0, n, T + c0, lda, M, N, C2, Lt + b.r0,
                                           T + b.c0, b.M, b.N, K, st));
    } else {
        const size_t esz = lp_elem_bytes(c->prec);
        const char* Lb = static_cast<const char*>(c->dLpL) + slot * n * B * esz;
        const char* Tb = static_cast<const char*>(c->dLpT) + slot * n * B * esz;
        const int64_t ns = n / 32;
        const float* sL = c->dScL ? c->dScL + slot * ns : nullptr;
        const float* sT = c->dScT ? c->dScT + slot * ns : nullptr;
        LpRegion2 r2{C2, Lb + b.r0 * B * esz, Tb + b.c0 * B * esz,
                     sL ? sL + b.r0 / 32 : nullptr, sT ? sT + b.c0 / 32 : nullptr, b.M, b.N};
        if (c->prec == HPL_PREC_FP8) {
            HIPCHK(c, launch_gemm_update_fp8(C, lda, Lb + r0 * B * esz, B, n * B,
                                             Tb + c0 * B * esz, B, n * B, sL + r0 / 32,
                                             sT + c0 / 32, ns, M, N, K, (int)B, st, &r2));
        } else {
            HIPCHK(c, launch_gemm_update_lp(c->prec, C, lda, Lb + r0 * B * esz, B, n * B,
                                            Tb + c0 * B * esz, B, n * B, M, N, K, (int)B, st,
                                            &r2, max_groups));
        }
    }
    if (c->prof) {
        HIPCHK(c, hipEventRecord(ev_end, st));
        c->prof_flops += 2.0 * ((double)M * N + (double)b.M * b.N) * K;
        const double esz = c->prec == 0 ? 4.0 : (double)lp_elem_bytes(c->prec);
        c->prof_bytes += 8.0 * ((double)M * N + (double)b.M * b.N) +
                         esz * K * ((double)M + N + (b.M > 0 ? (double)b.M + b.N : 0.0));
    }
    return HPL_SUCCESS;
}

// Factorise the B x B diagonal block at (off, off) in place.  Blocks above the
// leaf size are split 2 x 2: LU(D11), solve D12/D21, D22 -= D21 D12, LU(D22).
hpl_status factor_diag(hpl_ctx* c, hipStream_t st, int64_t off, uint32_t B, float* scratch) {
    if (B <= (uint32_t)kMaxLeafLU) {
        HIPCHK(c, launch_lu_diag(c->dA, c->lda, off, B, c->dStatus, st));
        return HPL_SUCCESS;
    }
    const uint32_t h = B / 2;
    float* Lt = scratch;
    float* DT = Lt + (size_t)h * h;
    float* inv = DT + (size_t)h * h;
    float* next = inv + 3 * (h / 32) * 1024;
    hpl_status r = factor_diag(c, st, off, h, next);
    if (r) return r;
    HIPCHK(c, launch_prep(c->dA, c->lda, off, h, DT, inv, st));
    HIPCHK(c, launch_trsm(c->dA, c->lda, off, h, h, DT, inv, Lt, h, st));
    const int64_t lda = (int64_t)c->lda;
    HIPCHK(c, launch_gemm_update_f32(c->dA + (off + h) * lda + off + h, lda, Lt, h,
                                     c->dA + off * lda + off + h, lda, h, h, h, st));
    return factor_diag(c, st, off + h, h, next);
}

// Step k's buffer of diagonal-block inverses for the strip solves of the
// BF16 factorisation, or null where the strips take the recurrence.  (The
// split product is exact to about 2^-17, far below BF16's rounding of the
// update operands (2^-9) but not far enough below FP16's (2^-12): the FP16
// factorisation keeps the FP32 recurrence.)
float* step_tinv(hpl_ctx* c, int64_t k, uint32_t B) {
    if (!c->dTinv || !c->strip_inverse || B > 256 || c->prec != HPL_PREC_BF16) return nullptr;
    return c->dTinv + (k & 1) * tinv_floats((int)B);
}

// Panel of step k, part 1: factor the diagonal block and prepare its
// inverses (latency-bound, a few workgroups).
hpl_status panel_diag(hpl_ctx* c, hipStream_t st, int64_t k, uint32_t B) {
    const int64_t lda = (int64_t)c->lda;
    const int64_t off = k * B;
    hpl_status r = factor_diag(c, st, off, B, c->dInner);
    if (r) return r;
    float* inv = c->dInv + k * (int64_t)(3 * (B / 32) * 1024);
    float* DT = c->dDT + (k & 1) * (int64_t)B * B;
    HIPCHK(c, launch_prep(c->dA, lda, off, B, DT, inv, st));
    // BF16: the inverses for the strip solve are the substitution's own, so
    // they go to their slot of dBInv when the factorisation fills it
    if (float* t = step_tinv(c, k, B))
        HIPCHK(c, launch_tinv_prepare(c->dA, lda, off, B, inv, t,
                                      c->binv_in_gefa ? c->dBInv + k * 2 * (int64_t)B * B : nullptr,
                                      st));
    return HPL_SUCCESS;
}

// Panel of step k, part 2: solve the top and left strips (the left one also,
// transposed, into its Lt slot with absolute row indices) -- one wave per 32
// vectors, as wide as the trailing matrix.
hpl_status panel_strips(hpl_ctx* c, hipStream_t st, int64_t k, uint32_t B) {
    const int64_t n = (int64_t)c->n, lda = (int64_t)c->lda;
    const int64_t off = k * B;
    float* inv = c->dInv + k * (int64_t)(3 * (B / 32) * 1024);
    float* DT = c->dDT + (k & 1) * (int64_t)B * B;
    const int64_t m = n - off - B;
    if (m > 0) {
        const int64_t slot = lt_slot(c, k);
        // the FP32 transposed copy feeds only the FP32 update
        float* Lt = c->prec == 0 ? c->dLt + slot * (int64_t)B * n + off + B : nullptr;
        void* lpT = nullptr;
        void* lpL = nullptr;
        float* scT = nullptr;
        float* scL = nullptr;
        if (c->prec != 0) {
            const size_t esz = lp_elem_bytes(c->prec);
            lpT = static_cast<char*>(c->dLpT) + (slot * n * B + (off + B) * B) * esz;
            lpL = static_cast<char*>(c->dLpL) + (slot * n * B + (off + B) * B) * esz;
        }
        if (c->prec == HPL_PREC_FP8) {
            scT = c->dScT + slot * (n / 32) + (off + B) / 32;
            scL = c->dScL + slot * (n / 32) + (off + B) / 32;
        }
        if (const float* t = step_tinv(c, k, B))
            HIPCHK(c, launch_trsm_inv(c->dA, lda, off, m, B, t, st, c->prec, lpT, lpL, B));
        else
            HIPCHK(c, launch_trsm(c->dA, lda, off, m, B, DT, inv, Lt, n, st, c->prec, lpT, lpL,
                                  B, scT, scL));
    }
    return HPL_SUCCESS;
}

hpl_status panel(hpl_ctx* c, hipStream_t st, int64_t k, uint32_t B) {
    hpl_status r = panel_diag(c, st, k, B);
    return r ? r : panel_strips(c, st, k, B);
}

// Grouped schedule (HPL_SCHEDULE=grouped; also the one used without a linear
// strip buffer): blocked right-looking LU with depth-1 look-ahead (the pattern of the
// reference's XRT backend, execution_xrt_pcie.hpp:328-415) and a delayed bulk
// update.  At step k the row and column blocks of the next panel receive all
// pending updates; the panel of step k+1 (diagonal block and strip solves)
// then runs on stream s1 while stream s0 updates the rest of the trailing
// matrix -- but only every `depth` steps, applying `depth` steps' updates in
// one GEMM with K = depth * B.  Every element still receives each update
// exactly once, in step order.  (Running the strip solves on s0 after the
// bulk update instead, with the whole device, measured 1% slower at
// n = 32768: they then sit on the critical path.)
hpl_status gefa_schedule(hpl_ctx* c, uint32_t B) {
    const int64_t n = (int64_t)c->n;
    const int64_t N = n / B;
    const int D = c->lookahead ? c->depth : 1;
    hpl_status r = panel(c, c->s0, 0, B);
    if (r) return r;
    for (int64_t k = 0; k + 1 < N; ++k) {
        const int64_t k1 = (k + 1) * B;  // first trailing row/column
        const int64_t m = n - k1;
        const int64_t j = k % D;         // updates k-j .. k are pending
        const int K = (int)((j + 1) * B);
        const int64_t k0 = k - j;
        if (!c->lookahead || m <= (int64_t)B) {
            r = launch_update(c, c->s0, k0, k1, k1, (int)m, (int)m, K, B);
            if (r) return r;
            r = panel(c, c->s0, k + 1, B);
            if (r) return r;
            continue;
        }
        // next panel's region: row block k+1 and column block k+1, one launch
        r = launch_update2(c, c->s0, k0, UpdRegion{k1, k1, (int)B, (int)m},
                           UpdRegion{k1 + B, k1, (int)(m - B), (int)B}, K, B);
        if (r) return r;
        HIPCHK(c, hipEventRecord(c->evA, c->s0));
        HIPCHK(c, hipStreamWaitEvent(c->s1, c->evA, 0));
        r = panel(c, c->s1, k + 1, B);
        if (r) return r;
        HIPCHK(c, hipEventRecord(c->evP, c->s1));
        // the rest of the trailing matrix, every D steps; pending updates left
        // at the end are applied by the final step's full update (m == B)
        if (j == D - 1) {
            r = launch_update(c, c->s0, k0, k1 + B, k1 + B, (int)(m - B), (int)(m - B), K, B);
            if (r) return r;
        }
        HIPCHK(c, hipStreamWaitEvent(c->s0, c->evP, 0));
    }
    return HPL_SUCCESS;
}

// Pipelined schedule.  Steps form groups of D = depth; group g holds the
// blocks [gD, gD + D).  Phase g runs at once:
//   s1: the bulk update of group g-1's steps (K = D B) on the trailing
//       matrix beyond group g's blocks, rows and columns >= (g+1) D B;
//   for each block k of group g, which still needs the steps of group g-1
//   and those of group g before k:
//   s0: the diagonal block's update by those steps, its factorisation and
//       the strip solves -- the critical path, on one stream;
//   s2: the update of the row block (columns kB+B..n) and column block (rows
//       kB+B..n) by the same steps, beside the factorisation (the strip
//       solves wait for it);
//   s3: the look-ahead: block k+1's diagonal, row and column blocks updated
//       by all its pending steps but step k, so that at step k+1 only one
//       step (K = B) is left on s0 and s2 (the group's first block takes
//       all of its pending steps there, K = (k - max(0, (g-1)D)) B over the
//       contiguous strip slots).
// The streams write disjoint parts of A (or wait for each other where they
// do not: the look-ahead before the next step's updates of the same block)
// and read only finished strips; at the phase's end s1 and s0 wait for each
// other (s1 for the panels whose strips the next bulk update reads, s0 for
// the bulk update that covers the next group's blocks).  Every element still
// receives each step's update exactly once, in step order.
//
// A bulk launch fills every workgroup slot while it has tiles left to start,
// and the chain's kernels (whose waves do not fit beside two update
// workgroups on a CU) wait until then.  The bulk update is therefore issued
// as `bulk_slices` launches over column slices: while one slice drains, the
// chain's waiting kernel takes the freed slots, so each slice boundary lets
// the chain advance.  With as many slices as steps per group, n = 32768,
// B = 256, D = 3 measured 207 ms against 225 ms for the grouped schedule and
// 221 ms for this one with single bulk launches (D = 4: 207 ms, D = 2:
// 210 ms; 2 slices at D = 3: 215 ms).  The 16-bit update takes at most 6
// slices per group (below).
//
// Groups need not be equal: hplx::group_bounds gives their boundaries, so the
// last part of the factorisation can use shorter groups (tail_m, tail_depth)
// and, for the 16-bit update, the first part deeper ones (head_m,
// head_depth).
// FP32 defaults to D = 4 with single steps in the tail (n = 32768: GEFA
// 203.6-204.9 ms against 204.2-207.9 for D = 3; D = 4 with tail groups of 2:
// 204.3-205.1; D = 5 and 6 with single steps: 203.8-205.6, 204.7-205.9).
// The substitution's inverses of diagonal blocks [k0, k1) (launch_block_inverse
// on that range of the factor and of the 32 x 32 inverses).
hipError_t binv_range(hpl_ctx* c, hipStream_t st, int64_t k0, int64_t k1, uint32_t B) {
    const int64_t lda = (int64_t)c->lda, is = 3 * (B / 32) * 1024;
    return launch_block_inverse(c->dA + k0 * B * lda + k0 * B, lda, (k1 - k0) * B, (int)B,
                                c->dInv + k0 * is, is, c->dBInv + k0 * 2 * (int64_t)B * B, st);
}

hpl_status gefa_pipelined(hpl_ctx* c, uint32_t B) {
    hpl_status r = hplx::init_pipeline_streams(c);
    if (r) return r;
    const int64_t n = (int64_t)c->n;
    const int64_t N = n / B;
    const std::vector<int64_t> bnd =
        hplx::group_bounds(c, N, B, c->depth, c->prec == HPL_PREC_FP32 ? 1 : 16,
                           c->prec == HPL_PREC_BF16 || c->prec == HPL_PREC_FP16);
    const int64_t G = (int64_t)bnd.size() - 1;
    for (int64_t g = 0; g < G; ++g) {
        const int64_t s = bnd[g + 1] * (int64_t)B;
        const bool wide = (c->prec == HPL_PREC_BF16 || c->prec == HPL_PREC_FP16) &&
                          (n - s) % 256 == 0;
        if (g >= 1 && s < n && wide && c->chain_cus > 0 && c->num_cu > c->chain_cus) {
            // one launch on at most num_cu - chain_cus persistent workgroups:
            // the chain's launches always find chain_cus CUs free
            const int64_t steps = bnd[g] - bnd[g - 1];
            r = launch_update(c, c->s1, bnd[g - 1], s, s, (int)(n - s), (int)(n - s),
                              (int)(steps * B), B, c->num_cu - c->chain_cus);
            if (r) return r;
        } else if (g >= 1 && s < n) {
            // group g-1's steps in column slices of whole tiles (256 wide for
            // the 16-bit update's 256 x 256 tiles, else 128): one slice per
            // step, at most 6 for the 16-bit update (n = 65536, 16 steps per
            // group, on one device: GEFA 251.8 ms with 16 slices, 248.6 with
            // 10, 247.3-248.3 with 8, 245.6 with 6, 246.2 with 4)
            const int64_t steps = bnd[g] - bnd[g - 1];
            const int64_t slices = c->bulk_slices > 0 ? c->bulk_slices
                                   : wide             ? std::min<int64_t>(steps, 6)
                                                      : steps;
            const int64_t tw = wide ? 256 : 128;
            const int64_t w = n - s, tiles = w / tw;
            const int64_t S = std::max<int64_t>(1, std::min(slices, tiles));
            for (int64_t i = 0; i < S; ++i) {
                const int64_t a = s + tiles * i / S * tw;
                const int64_t b = i + 1 == S ? n : s + tiles * (i + 1) / S * tw;
                r = launch_update(c, c->s1, bnd[g - 1], s, a, (int)w, (int)(b - a),
                                  (int)(steps * B), B);
                if (r) return r;
            }
        }
        // the diagonal-block inverses of the substitution for group g-1's
        // blocks (final; their 32 x 32 inverses come from panel_diag), off
        // the critical path behind the bulk update; BF16 computes them per
        // step for its strip solves
        if (g >= 1 && c->binv_in_gefa && !step_tinv(c, bnd[g - 1], B))
            HIPCHK(c, binv_range(c, c->s1, bnd[g - 1], bnd[g], B));
        HIPCHK(c, hipEventRecord(c->evP, c->s1));
        const int64_t k0 = g >= 1 ? bnd[g - 1] : 0;
        for (int64_t k = bnd[g]; k < bnd[g + 1]; ++k) {
            const int64_t kb = k * (int64_t)B, m = n - kb;
            if (k == k0) {
                r = panel(c, c->s0, k, B);
                if (r) return r;
                continue;
            }
            // Block k still needs steps k0 .. k-1.  For all but a group's
            // first block, steps k0 .. k-2 were applied one step ahead (on
            // s3, below), so only step k-1 is left on the critical path: the
            // strip solves of step k-1, the diagonal block's update and
            // factorisation, the strip solves of step k -- all on s0, so the
            // path crosses no stream -- while s2 updates the rest of the row
            // and column blocks by the same steps and the strip solves wait
            // for it (evD).  (Each element still receives the steps in order,
            // and an update split over launches gives the bits of one launch:
            // the accumulators start from C.)  With the diagonal block on s2
            // instead and the row and column blocks on s0, GEFA at n = 65536
            // measured 245.5 / 246.6 ms against 244.5 / 243.3 (A/B on one
            // device).
            const bool ahead = k > bnd[g];
            const int64_t from = ahead ? k - 1 : k0;
            const int K = (int)((k - from) * B);
            HIPCHK(c, hipEventRecord(c->evS, c->s0));
            if (m > (int64_t)B) {
                HIPCHK(c, hipStreamWaitEvent(c->s2, c->evS, 0));
                if (ahead) HIPCHK(c, hipStreamWaitEvent(c->s2, c->evL, 0));
                r = launch_update2(c, c->s2, from, UpdRegion{kb, kb + B, (int)B, (int)(m - B)},
                                   UpdRegion{kb + B, kb, (int)(m - B), (int)B}, K, B);
                if (r) return r;
                HIPCHK(c, hipEventRecord(c->evD, c->s2));
            }
            if (ahead) HIPCHK(c, hipStreamWaitEvent(c->s0, c->evL, 0));
            r = launch_update(c, c->s0, from, kb, kb, (int)B, (int)B, K, B);
            if (!r) r = panel_diag(c, c->s0, k, B);
            if (r) return r;
            // steps k0 .. k-1 of the group's next block, on s3 beside this
            // step's factorisation (they need only the strips up to k-1)
            if (k + 1 < bnd[g + 1]) {
                const int64_t kb1 = kb + B, m1 = m - B;
                const int K1 = (int)((k - k0) * B);
                HIPCHK(c, hipStreamWaitEvent(c->s3, c->evS, 0));
                r = launch_update(c, c->s3, k0, kb1, kb1, (int)B, (int)B, K1, B);
                if (!r && m1 > (int64_t)B)
                    r = launch_update2(c, c->s3, k0,
                                       UpdRegion{kb1, kb1 + B, (int)B, (int)(m1 - B)},
                                       UpdRegion{kb1 + B, kb1, (int)(m1 - B), (int)B}, K1, B);
                if (r) return r;
                HIPCHK(c, hipEventRecord(c->evL, c->s3));
            }
            if (m > (int64_t)B) HIPCHK(c, hipStreamWaitEvent(c->s0, c->evD, 0));
            r = panel_strips(c, c->s0, k, B);
            if (r) return r;
        }
        HIPCHK(c, hipEventRecord(c->evA, c->s0));
        HIPCHK(c, hipStreamWaitEvent(c->s1, c->evA, 0));
        HIPCHK(c, hipStreamWaitEvent(c->s0, c->evP, 0));
    }
    if (G >= 1 && c->binv_in_gefa && !step_tinv(c, bnd[G - 1], B))
        HIPCHK(c, binv_range(c, c->s0, bnd[G - 1], bnd[G], B));
    return HPL_SUCCESS;
}

hpl_status check_status(hpl_ctx* c) {
    int h = 0;
    HIPCHK(c, hipMemcpyAsync(&h, c->dStatus, sizeof(int), hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    if (h & 2)
        return fail(c, HPL_ERR_HIP,
                    "pivot panel: a grid barrier timed out (workgroups not co-resident)");
    if (h & 4) return fail(c, HPL_ERR_HIP, "pivot panel: inconsistent pivot key");
    if (h) return fail(c, HPL_ERR_ZERO_PIVOT, "zero or non-finite pivot met during GEFA");
    return HPL_SUCCESS;
}

double elapsed_s(hpl_ctx* c) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->e0, c->e1);
    return ms * 1e-3;
}

}  // namespace

namespace hplx {

std::vector<int64_t> group_bounds(const hpl_ctx* c, int64_t N, int64_t B, int64_t depth,
                                  int64_t tail_cap, bool deep_head) {
    const int64_t n = N * B;
    const int64_t tail_m = c->tail_m < 0 ? n / 2 : c->tail_m;
    const int64_t tail_d = c->tail_depth > 0
                               ? c->tail_depth
                               : std::max<int64_t>(1, std::min<int64_t>(depth / 2, tail_cap));
    const int64_t head_m = c->head_m >= 0 ? c->head_m : deep_head ? n * 5 / 8 : 0;
    const int64_t head_d = c->head_depth > 0 ? c->head_depth : deep_head ? depth + depth / 4 : 0;
    std::vector<int64_t> b{0};
    while (b.back() < N) {
        const int64_t left = n - b.back() * B;
        const bool tail = tail_m > 0 && left <= tail_m;
        const bool head = head_m > 0 && head_d > 0 && left > head_m;
        const int64_t d = std::max<int64_t>(1, tail ? tail_d : head ? head_d : depth);
        b.push_back(std::min(N, b.back() + d));
    }
    return b;
}

uint64_t device_lda(uint64_t n) {
    uint64_t lda = (n + 63) / 64 * 64;
    // optional padding (experiments on the row stride), kept a multiple of 4
    // floats: the kernels move 16-byte vectors along rows
    if (const char* e = getenv("HPL_LDA_PAD")) lda += ((uint64_t)std::max(0LL, atoll(e)) + 3) / 4 * 4;
    return lda;
}

WsPlan plan_factor_ws(uint64_t n, uint32_t B, int prec, int depth, bool linear) {
    WsPlan w{};
    const uint64_t N = n / B;
    const size_t slots = linear ? (size_t)N : 2 * (size_t)depth;
    w.lt = prec == HPL_PREC_FP32 ? slots * B * n * sizeof(float) : 0;
    w.dt = 2 * (size_t)B * B * sizeof(float);
    w.inv = N * (size_t)(3 * (B / 32) * 1024) * sizeof(float);
    w.inner = 0;
    if (B > (uint32_t)kMaxLeafLU) {
        // recursion scratch per level h = B/2, B/4, .., 128: Lt and DT (h x h
        // each) and the 3 h/32 inverses of 32 x 32
        size_t need = 0;
        for (uint32_t h = B / 2; h >= (uint32_t)kMaxLeafLU; h /= 2)
            need += 2 * (size_t)h * h + 3 * (h / 32) * 1024;
        w.inner = need * sizeof(float);
    }
    w.lp = prec == HPL_PREC_FP32 ? 0 : slots * n * B * lp_elem_bytes(prec);
    w.sc = prec == HPL_PREC_FP8 ? slots * (n / 32) * sizeof(float) : 0;
    w.status = 64;
    w.binv = B <= 256 ? 2 * (size_t)n * B * sizeof(float) : 0;
    w.tinv = prec == HPL_PREC_BF16 && B <= 256
                 ? 2 * (size_t)tinv_floats((int)B) * sizeof(float)
                 : 0;
    return w;
}

hpl_status ensure_mem(hpl_ctx* c, void** p, size_t* cap, size_t bytes) {
    return ensure(c, *p, *cap, bytes);
}

hpl_status alloc_factor_workspace(hpl_ctx* c, uint32_t B) { return alloc_factor_ws(c, B); }

hpl_status ensure_block_inverses(hpl_ctx* c) {
    if (c->binv_ready) return HPL_SUCCESS;
    const uint32_t B = c->B;
    if (B > 256) return HPL_SUCCESS;  // sequential diagonal solves
    // a pivoted factor keeps the 32 x 32 diagonal solves: the diagonal blocks
    // of U of a matrix that is not diagonally dominant can be poorly
    // conditioned, and applying a whole B x B inverse loses up to a digit
    // there (B = 256, uniform matrices: 1.8e-3 against 4e-4 with 32 x 32)
    if (c->pivoted) return HPL_SUCCESS;
    const hplx::WsPlan w = plan_factor_ws(c->n, B, c->prec, 1);
    hpl_status st = ensure(c, c->dBInv, c->capBInv, w.binv);
    if (st) return st;
    HIPCHK(c, launch_block_inverse(c->dA, (int64_t)c->lda, (int64_t)c->n, (int)B, c->dInv,
                                   3 * (B / 32) * 1024, c->dBInv, c->s0));
    c->binv_ready = true;
    return HPL_SUCCESS;
}

const GeslWs* gesl_workspace(hpl_ctx* c, GeslWs* ws) {
    // own mode only: the serving engine runs several contexts' solves at
    // once, and their single launches could not all be co-resident
    if (c->external || !c->binv_ready || gesl_persist_groups((int64_t)c->n, (int)c->B) <= 0)
        return nullptr;
    if (ensure(c, c->dGeslWs, c->capGeslWs, gesl_ws_bytes((int64_t)c->n)) != HPL_SUCCESS) {
        c->err.clear();
        return nullptr;  // no room: the per-step launches solve
    }
    *ws = gesl_ws_carve(c->dGeslWs, (int64_t)c->n);
    return ws;
}

hpl_status gesl_check(hpl_ctx* c, const GeslWs* ws) {
    if (!ws) return HPL_SUCCESS;
    int h = 0;
    HIPCHK(c, hipMemcpyAsync(&h, ws->status, sizeof(int), hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    if (h & 2)
        return fail(c, HPL_ERR_HIP,
                    "substitution: a bounded wait gave up (workgroups not co-resident)");
    return HPL_SUCCESS;
}

IrPlan plan_ir(uint64_t n, int nrhs) {
    IrPlan p{};
    p.work = (size_t)residual_multi_work((int64_t)n, nrhs) * sizeof(double);
    p.r = (size_t)n * nrhs * sizeof(double);
    p.xs = (size_t)n * nrhs * sizeof(float);
    p.out = 4 * (size_t)nrhs * sizeof(double);
    return p;
}

hpl_status init_streams(hpl_ctx* c) {
    if (hipStreamCreateWithFlags(&c->s0, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->s1, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->e0) != hipSuccess || hipEventCreate(&c->e1) != hipSuccess ||
        hipEventCreateWithFlags(&c->evA, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->evP, hipEventDisableTiming) != hipSuccess)
        return fail(c, HPL_ERR_HIP, "stream or event creation failed");
    return HPL_SUCCESS;
}

// The pipelined schedule's two further streams and their events, made on its
// first use: contexts that never run it (the serving engine's slots) keep
// two streams each, so several of them do not crowd the device's few hardware
// queues.
hpl_status init_pipeline_streams(hpl_ctx* c) {
    if (c->s2) return HPL_SUCCESS;
    if (hipStreamCreateWithFlags(&c->s2, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->s3, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->evS, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->evD, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->evL, hipEventDisableTiming) != hipSuccess)
        return fail(c, HPL_ERR_HIP, "stream or event creation failed");
    return HPL_SUCCESS;
}

hpl_status ir_solve(hpl_ctx* c, int nrhs, const double* dB, double* dX, int max_iter, double tol,
                    hpl_ir_result* res) {
    if (nrhs < 1 || nrhs > 16) return fail(c, HPL_ERR_INVALID_ARG, "nrhs %d not in [1,16]", nrhs);
    const int64_t n = (int64_t)c->n;
    const IrPlan p = plan_ir(c->n, nrhs);
    hpl_status st = ensure(c, c->dWork, c->capWork, p.work);
    if (st) return st;
    st = ensure(c, c->dIrR, c->capIrR, p.r);
    if (st) return st;
    st = ensure(c, c->dIrXs, c->capIrXs, p.xs);
    if (st) return st;
    st = ensure(c, c->dIrOut, c->capIrOut, p.out);
    if (st) return st;
    st = ensure_block_inverses(c);
    if (st) return st;
    GeslWs wsbuf;
    const GeslWs* ws = nrhs == 1 ? gesl_workspace(c, &wsbuf) : nullptr;
    const int B = (int)c->B;
    const int64_t inv_stride = 3 * (B / 32) * 1024;
    double sc_in[16], sc_out[16], h[64];
    int iters[16];
    for (int j = 0; j < nrhs; ++j) {
        sc_in[j] = sc_out[j] = 1.0;
        iters[j] = 0;
    }
    // x0 = F^-1 b in FP32, promoted (a pivoted factor solves for P b)
    const int* perm = c->pivoted ? c->dPerm : nullptr;
    HIPCHK(c, launch_d2f_multi(dB, n, c->dIrXs, n, n, nrhs, sc_in, c->s0, perm));
    HIPCHK(c, launch_gesl_multi(c->dA, (int64_t)c->lda, n, B, c->dInv, inv_stride, c->dIrXs, n,
                                nrhs, c->s0, c->binv_ready ? c->dBInv : nullptr, ws));
    HIPCHK(c, launch_axpy_multi(dX, n, c->dIrXs, n, n, nrhs, sc_out, 1, c->s0));
    for (int it = 0;; ++it) {
        HIPCHK(c, launch_residual_multi(c->dA0, (int64_t)c->lda, n, dX, n, dB, n, nrhs, c->dWork,
                                        c->dIrR, n, c->dIrOut, c->s0));
        HIPCHK(c, hipMemcpyAsync(h, c->dIrOut, 4 * nrhs * sizeof(double), hipMemcpyDeviceToHost,
                                 c->s0));
        HIPCHK(c, hipStreamSynchronize(c->s0));
        bool any = false;
        for (int j = 0; j < nrhs; ++j) {
            const double* hj = h + 4 * j;
            const double berr = hj[0] / (hj[1] * hj[2] + hj[3]);
            res[j].iterations = iters[j];
            res[j].converged = berr <= tol ? 1 : 0;
            res[j].backward_error = berr;
            res[j].residual_inf = hj[0];
            res[j].norm_a = hj[1];
            res[j].norm_x = hj[2];
            res[j].norm_b = hj[3];
            res[j].gefa_s = res[j].total_s = 0.0;
            // a column is corrected while it misses tol, has corrections left
            // and a finite, non-zero residual
            const bool active = !(berr <= tol) && it < max_iter && hj[0] > 0.0 &&
                                hj[0] < HUGE_VAL;
            sc_in[j] = active ? 1.0 / hj[0] : 0.0;
            sc_out[j] = active ? hj[0] : 0.0;
            iters[j] += active ? 1 : 0;
            any = any || active;
        }
        if (!any) break;
        // correction: F d = r / ||r|| in FP32, x += ||r|| d in FP64 (inactive
        // columns get scale 0 and stay unchanged)
        HIPCHK(c, launch_d2f_multi(c->dIrR, n, c->dIrXs, n, n, nrhs, sc_in, c->s0, perm));
        HIPCHK(c, launch_gesl_multi(c->dA, (int64_t)c->lda, n, B, c->dInv, inv_stride, c->dIrXs,
                                    n, nrhs, c->s0, c->binv_ready ? c->dBInv : nullptr, ws));
        HIPCHK(c, launch_axpy_multi(dX, n, c->dIrXs, n, n, nrhs, sc_out, 0, c->s0));
    }
    st = gesl_check(c, ws);
    if (st) return st;
    for (int j = 0; j < nrhs; ++j)
        if (!res[j].converged)
            return fail(c, HPL_ERR_NOT_CONVERGED,
                        "refinement of column %d stopped after %d iterations at backward error "
                        "%.3e (tol %.3e)",
                        j, res[j].iterations, res[j].backward_error, tol);
    return HPL_SUCCESS;
}

}  // namespace hplx

extern "C" {

const char* hpl_version(void) { return "hpl_hip 1.0 (gfx950, FP32 MFMA)"; }

const char* hpl_status_string(hpl_status s) {
    switch (s) {
        case HPL_SUCCESS: return "success";
        case HPL_ERR_INVALID_ARG: return "invalid argument";
        case HPL_ERR_HIP: return "HIP runtime error";
        case HPL_ERR_OUT_OF_MEMORY: return "out of device memory";
        case HPL_ERR_ZERO_PIVOT: return "zero or non-finite pivot";
        case HPL_ERR_BUDGET: return "HBM budget exceeded";
        case HPL_ERR_STATE: return "call out of order";
        case HPL_ERR_NOT_CONVERGED: return "refinement did not converge";
        case HPL_ERR_UNSUPPORTED: return "unsupported";
        case HPL_ERR_NO_DEVICE: return "no HIP device";
    }
    return "unknown status";
}

hpl_status hpl_create(int device, hpl_ctx** out) {
    if (!out) return HPL_ERR_INVALID_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return HPL_ERR_NO_DEVICE;
    if (device >= count) return HPL_ERR_NO_DEVICE;
    if (device < 0) device = 0 % count;
    if (hipSetDevice(device) != hipSuccess) return HPL_ERR_NO_DEVICE;
    hpl_ctx* c = new hpl_ctx();
    c->device = device;
    if (const char* e = getenv("HPL_LOOKAHEAD")) c->lookahead = atoi(e) != 0;
    if (const char* e = getenv("HPL_SCHEDULE")) c->pipelined = strcmp(e, "grouped") != 0;
    if (const char* e = getenv("HPL_BULK_SLICES")) c->bulk_slices = std::max(0, atoi(e));
    if (const char* e = getenv("HPL_STRIP_SOLVE")) c->strip_inverse = strcmp(e, "recurrence") != 0;
    if (const char* e = getenv("HPL_CHAIN_CUS")) c->chain_cus = std::max(0, atoi(e));
    if (hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, device) !=
        hipSuccess)
        c->num_cu = 0;
    if (const char* e = getenv("HPL_TAIL_M")) c->tail_m = std::max(0LL, atoll(e));
    if (const char* e = getenv("HPL_HEAD_M")) c->head_m = std::max(0LL, atoll(e));
    if (const char* e = getenv("HPL_HEAD_DEPTH")) c->head_depth = std::max(0, std::min(32, atoi(e)));
    if (const char* e = getenv("HPL_TAIL_DEPTH")) c->tail_depth = std::max(0, std::min(32, atoi(e)));
    if (const char* e = getenv("HPL_PANEL_SPIN_LIMIT")) c->pv_spin_limit = atol(e);
    if (const char* e = getenv("HPL_F64_DEPTH")) c->depth_f64 = std::max(1, std::min(16, atoi(e)));
    if (const char* e = getenv("HPL_UPDATE_DEPTH"))
        c->depth_fp32 = c->depth_lp = std::max(1, std::min(32, atoi(e)));
    if (hplx::init_streams(c) != HPL_SUCCESS) {
        hpl_destroy(c);
        return HPL_ERR_HIP;
    }
    *out = c;
    return HPL_SUCCESS;
}

void hpl_destroy(hpl_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->s0) (void)hipStreamSynchronize(c->s0);
    if (c->s1) (void)hipStreamSynchronize(c->s1);
    if (c->s2) (void)hipStreamSynchronize(c->s2);
    if (c->s3) (void)hipStreamSynchronize(c->s3);
    if (!c->external)
        for (void* p : {(void*)c->dA, (void*)c->dA0, (void*)c->db, (void*)c->dx, (void*)c->db64,
                        (void*)c->dxt, (void*)c->dLt, (void*)c->dDT, (void*)c->dInv,
                        (void*)c->dInner, (void*)c->dTinv, (void*)c->dStatus, (void*)c->dWork, (void*)c->dOut4,
                        c->dLpT, c->dLpL, (void*)c->dScT, (void*)c->dScL, (void*)c->dX64,
                        (void*)c->dIrR, (void*)c->dIrXs, (void*)c->dIrOut, (void*)c->dRhsB,
                        (void*)c->dRhsX, (void*)c->dGmV, (void*)c->dGmW, (void*)c->dGmC,
                        (void*)c->dGmWork, (void*)c->dBInv, c->dPv, (void*)c->dA64,
                        (void*)c->dA064, (void*)c->dB64s, (void*)c->dX64s, (void*)c->dLt64,
                        (void*)c->dWork64, (void*)c->dOut64, c->dGeslWs})
            if (p) (void)hipFree(p);
    for (auto& e : c->ev) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (c->e0) (void)hipEventDestroy(c->e0);
    if (c->e1) (void)hipEventDestroy(c->e1);
    if (c->evA) (void)hipEventDestroy(c->evA);
    if (c->evP) (void)hipEventDestroy(c->evP);
    if (c->evS) (void)hipEventDestroy(c->evS);
    if (c->evD) (void)hipEventDestroy(c->evD);
    if (c->evL) (void)hipEventDestroy(c->evL);
    if (c->s3) (void)hipStreamDestroy(c->s3);
    if (c->s2) (void)hipStreamDestroy(c->s2);
    if (c->s1) (void)hipStreamDestroy(c->s1);
    if (c->s0) (void)hipStreamDestroy(c->s0);
    delete c;
}

const char* hpl_last_error(const hpl_ctx* c) { return c ? c->err.c_str() : "null context"; }

hpl_status hpl_device_name(hpl_ctx* c, char* buf, size_t len) {
    if (!c || !buf || !len) return HPL_ERR_INVALID_ARG;
    hipDeviceProp_t p;
    HIPCHK(c, hipGetDeviceProperties(&p, c->device));
    snprintf(buf, len, "%s (%s)", p.name, p.gcnArchName);
    return HPL_SUCCESS;
}

hpl_status hpl_keep_original(hpl_ctx* c, int keep) {
    if (!c) return HPL_ERR_INVALID_ARG;
    c->keep_orig = keep != 0;
    return HPL_SUCCESS;
}

hpl_status hpl_upload(hpl_ctx* c, const float* a, uint64_t n, uint64_t lda, const float* b,
                      hpl_timing* t) {
    if (!c || !a || n == 0 || lda < n) return fail(c, HPL_ERR_INVALID_ARG, "bad upload arguments");
    HIPCHK(c, hipSetDevice(c->device));
    hpl_status st = alloc_system(c, n);
    if (st) return st;
    HIPCHK(c, hipEventRecord(c->e0, c->s0));
    HIPCHK(c, hipMemcpy2DAsync(c->dA, c->lda * sizeof(float), a, lda * sizeof(float),
                               n * sizeof(float), n, hipMemcpyHostToDevice, c->s0));
    if (b)
        HIPCHK(c, hipMemcpyAsync(c->db, b, n * sizeof(float), hipMemcpyHostToDevice, c->s0));
    else
        HIPCHK(c, hipMemsetAsync(c->db, 0, n * sizeof(float), c->s0));
    if (c->keep_orig) {
        HIPCHK(c, hipMemcpyAsync(c->dA0, c->dA, n * c->lda * sizeof(float),
                                 hipMemcpyDeviceToDevice, c->s0));
        c->have_orig = true;
    }
    HIPCHK(c, hipEventRecord(c->e1, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    if (t) t->h2d_s = elapsed_s(c);
    return HPL_SUCCESS;
}

hpl_status hpl_generate(hpl_ctx* c, uint64_t n, int mode, uint64_t seed, double d) {
    if (!c || n == 0 || mode < 0 || mode > 2)
        return fail(c, HPL_ERR_INVALID_ARG, "bad generate arguments (n=%llu mode=%d)",
                    (unsigned long long)n, mode);
    HIPCHK(c, hipSetDevice(c->device));
    hpl_status st = alloc_system(c, n);
    if (st) return st;
    HIPCHK(c, launch_generate(c->dA, (int64_t)c->lda, (int64_t)n, mode, seed, d, c->db, c->db64,
                              c->dxt, c->s0));
    if (c->keep_orig) {
        HIPCHK(c, hipMemcpyAsync(c->dA0, c->dA, n * c->lda * sizeof(float),
                                 hipMemcpyDeviceToDevice, c->s0));
        c->have_orig = true;
    }
    HIPCHK(c, hipStreamSynchronize(c->s0));
    return HPL_SUCCESS;
}

hpl_status hpl_restore(hpl_ctx* c) {
    if (!c) return HPL_ERR_INVALID_ARG;
    if (!c->have_orig) return fail(c, HPL_ERR_STATE, "no original copy kept");
    HIPCHK(c, hipMemcpyAsync(c->dA, c->dA0, c->n * c->lda * sizeof(float),
                             hipMemcpyDeviceToDevice, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    c->factored = false;
    c->pivoted = false;
    return HPL_SUCCESS;
}

hpl_status hpl_gefa_dev(hpl_ctx* c, uint32_t block, hpl_prec prec, hpl_timing* t) {
    if (!c) return HPL_ERR_INVALID_ARG;
    if (!c->dA || c->n == 0) return fail(c, HPL_ERR_STATE, "no resident matrix");
    if (!valid_block(c->n, block))
        return fail(c, HPL_ERR_INVALID_ARG,
                    "block size %u invalid for n=%llu (power of two in [32,512] dividing n)",
                    block, (unsigned long long)c->n);
    if (prec != HPL_PREC_FP32 && prec != HPL_PREC_BF16 && prec != HPL_PREC_FP16 &&
        prec != HPL_PREC_FP8)
        return fail(c, HPL_ERR_UNSUPPORTED, "precision %d not available", (int)prec);
    if (prec != HPL_PREC_FP32 && block % 64)
        return fail(c, HPL_ERR_INVALID_ARG,
                    "low-precision trailing updates need a block multiple of 64");
    if (c->pivoting && (prec != HPL_PREC_FP32 || block > 256))
        return fail(c, HPL_ERR_UNSUPPORTED,
                    "partial pivoting runs the FP32 trailing update with blocks up to 256");
    HIPCHK(c, hipSetDevice(c->device));
    c->prec = (int)prec;
    c->depth = prec == HPL_PREC_FP32 ? c->depth_fp32 : c->depth_lp;
    c->lt_linear = c->pipelined && c->lookahead && !c->external && !c->pivoting;
    hpl_status st = alloc_factor_ws(c, block);
    if (st) return st;
    HIPCHK(c, hipMemsetAsync(c->dStatus, 0, sizeof(int), c->s0));
    c->factored = false;
    c->pivoted = false;
    c->binv_ready = false;
    // the pipelined schedule also fills the substitution's diagonal-block
    // inverses (off its critical path), so the first solve need not
    c->binv_in_gefa = c->lt_linear && block <= 256;
    if (c->binv_in_gefa) {
        const hplx::WsPlan w = hplx::plan_factor_ws(c->n, block, c->prec, 1);
        if (ensure(c, c->dBInv, c->capBInv, w.binv) != HPL_SUCCESS) {
            c->err.clear();
            c->binv_in_gefa = false;  // no room: the first solve computes them
        }
    }
    HIPCHK(c, hipEventRecord(c->e0, c->s0));
    st = c->pivoting ? hplx::gefa_pivot(c, block)
         : c->lt_linear ? gefa_pipelined(c, block)
                        : gefa_schedule(c, block);
    if (st) return st;
    HIPCHK(c, hipEventRecord(c->e1, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    if (t) t->gefa_s = elapsed_s(c);
    st = check_status(c);
    if (st) return st;
    c->B = block;
    c->factored = true;
    c->pivoted = c->pivoting;
    c->binv_ready = c->binv_in_gefa;
    return HPL_SUCCESS;
}

hpl_status hpl_gesl_dev(hpl_ctx* c, hpl_timing* t) {
    if (!c) return HPL_ERR_INVALID_ARG;
    if (!c->factored) return fail(c, HPL_ERR_STATE, "hpl_gesl_dev before a successful GEFA");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipEventRecord(c->e0, c->s0));
    hpl_status st = hplx::ensure_block_inverses(c);
    if (st) return st;
    if (c->pivoted)
        HIPCHK(c, launch_gather_f32(c->dx, c->db, c->dPerm, (int64_t)c->n, c->s0));
    else
        HIPCHK(c, hipMemcpyAsync(c->dx, c->db, c->n * sizeof(float), hipMemcpyDeviceToDevice,
                                 c->s0));
    GeslWs wsbuf;
    const GeslWs* ws = hplx::gesl_workspace(c, &wsbuf);
    HIPCHK(c, launch_gesl(c->dA, (int64_t)c->lda, (int64_t)c->n, (int)c->B, c->dInv,
                          3 * (c->B / 32) * 1024, c->dx, c->s0, c->binv_ready ? c->dBInv : nullptr,
                          ws));
    HIPCHK(c, hipEventRecord(c->e1, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    if (t) t->gesl_s = elapsed_s(c);
    return hplx::gesl_check(c, ws);
}

hpl_status hpl_download(hpl_ctx* c, float* a, uint64_t lda, float* x, hpl_timing* t) {
    if (!c) return HPL_ERR_INVALID_ARG;
    if (!c->dA) return fail(c, HPL_ERR_STATE, "no resident system");
    if (a && lda < c->n) return fail(c, HPL_ERR_INVALID_ARG, "lda < n");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipEventRecord(c->e0, c->s0));
    if (a)
        HIPCHK(c, hipMemcpy2DAsync(a, lda * sizeof(float), c->dA, c->lda * sizeof(float),
                                   c->n * sizeof(float), c->n, hipMemcpyDeviceToHost, c->s0));
    if (x) HIPCHK(c, hipMemcpyAsync(x, c->dx, c->n * sizeof(float), hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipEventRecord(c->e1, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    if (t) t->d2h_s = elapsed_s(c);
    if (a && c->factored && c->pivoted) return hplx::pivot_factor_to_host(c, a, lda);
    return HPL_SUCCESS;
}

hpl_status hpl_download_vectors(hpl_ctx* c, float* b, double* b64, double* xtrue) {
    if (!c) return HPL_ERR_INVALID_ARG;
    if (!c->dA) return fail(c, HPL_ERR_STATE, "no resident system");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t n = c->n;
    if (b) HIPCHK(c, hipMemcpyAsync(b, c->db, n * sizeof(float), hipMemcpyDeviceToHost, c->s0));
    if (b64)
        HIPCHK(c, hipMemcpyAsync(b64, c->db64, n * sizeof(double), hipMemcpyDeviceToHost, c->s0));
    if (xtrue)
        HIPCHK(c, hipMemcpyAsync(xtrue, c->dxt, n * sizeof(double), hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    return HPL_SUCCESS;
}

hpl_status hpl_gefa(hpl_ctx* c, float* a, uint64_t n, uint64_t lda, uint32_t block,
                    hpl_timing* t) {
    if (!c) return HPL_ERR_INVALID_ARG;
    if (!a || lda < n || !valid_block(n, block))
        return fail(c, HPL_ERR_INVALID_ARG, "bad gefa arguments (n=%llu lda=%llu block=%u)",
                    (unsigned long long)n, (unsigned long long)lda, block);
    hpl_status st = hpl_upload(c, a, n, lda, nullptr, t);
    if (st) return st;
    // gefa_ref_nopvt's call: no pivoting whatever the context's setting
    const bool saved = c->pivoting;
    c->pivoting = false;
    st = hpl_gefa_dev(c, block, HPL_PREC_FP32, t);
    c->pivoting = saved;
    if (st) return st;
    return hpl_download(c, a, lda, nullptr, t);
}

hpl_status hpl_gesl(hpl_ctx* c, const float* lu, float* b, uint64_t n, uint64_t lda,
                    hpl_timing* t) {
    if (!c) return HPL_ERR_INVALID_ARG;
    if (!lu || !b || lda < n || n == 0) return fail(c, HPL_ERR_INVALID_ARG, "bad gesl arguments");
    // block size for the solve: the largest power of two <= 256 dividing n
    uint32_t B = 256;
    while (B >= 32 && n % B) B /= 2;
    if (B < 32) return fail(c, HPL_ERR_UNSUPPORTED, "n=%llu is not a multiple of 32",
                            (unsigned long long)n);
    hpl_status st = hpl_upload(c, lu, n, lda, b, t);
    if (st) return st;
    st = alloc_factor_ws(c, B);
    if (st) return st;
    HIPCHK(c, launch_prep_all(c->dA, (int64_t)c->lda, (int)B, (int)(n / B), c->dInv,
                              3 * (B / 32) * 1024, c->s0));
    c->B = B;
    c->factored = true;
    c->pivoted = false;
    c->binv_ready = false;
    st = hpl_gesl_dev(c, t);
    if (st) return st;
    HIPCHK(c, hipMemcpy(b, c->dx, n * sizeof(float), hipMemcpyDeviceToHost));
    return HPL_SUCCESS;
}

hpl_status hpl_residual(hpl_ctx* c, double out[4]) {
    if (!c || !out) return HPL_ERR_INVALID_ARG;
    if (!c->have_orig) return fail(c, HPL_ERR_STATE, "hpl_residual needs hpl_keep_original(1)");
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = (int64_t)c->n;
    hpl_status st = ensure(c, c->dWork, c->capWork,
                           2 * (size_t)residual_chunks(n) * n * sizeof(double));
    if (st) return st;
    if (!c->dOut4) {
        if (c->external) return fail(c, HPL_ERR_STATE, "no residual buffer in external mode");
        HIPCHK(c, hipMalloc(&c->dOut4, 4 * sizeof(double)));
    }
    HIPCHK(c, launch_residual_f32x(c->dA0, (int64_t)c->lda, n, c->dx, c->db, c->dWork, nullptr,
                                   c->dOut4, c->s0));
    HIPCHK(c, hipMemcpyAsync(out, c->dOut4, 4 * sizeof(double), hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    return HPL_SUCCESS;
}

hpl_status hpl_solve_ir(hpl_ctx* c, uint32_t block, hpl_prec prec, int max_iter, double tol,
                        hpl_ir_result* out) {
    if (!c || !out) return HPL_ERR_INVALID_ARG;
    if (!c->have_orig) return fail(c, HPL_ERR_STATE, "hpl_solve_ir needs hpl_keep_original(1)");
    if (max_iter < 0) return fail(c, HPL_ERR_INVALID_ARG, "max_iter < 0");
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = (int64_t)c->n;
    hpl_status st = ensure(c, c->dX64, c->capIR, n * sizeof(double));
    if (st) return st;
    // fresh matrix from the kept original (not timed)
    HIPCHK(c, hipMemcpyAsync(c->dA, c->dA0, n * c->lda * sizeof(float), hipMemcpyDeviceToDevice,
                             c->s0));
    hipEvent_t t0, t1;
    HIPCHK(c, hipEventCreate(&t0));
    if (hipEventCreate(&t1) != hipSuccess) {
        (void)hipEventDestroy(t0);
        return fail(c, HPL_ERR_HIP, "hipEventCreate failed");
    }
    hpl_timing tg{};
    st = hipEventRecord(t0, c->s0) == hipSuccess ? HPL_SUCCESS
                                                 : fail(c, HPL_ERR_HIP, "hipEventRecord failed");
    if (!st) st = hpl_gefa_dev(c, block, prec, &tg);
    if (!st) {
        st = hplx::ir_solve(c, 1, c->db64, c->dX64, max_iter, tol, out);
        if (st == HPL_SUCCESS || st == HPL_ERR_NOT_CONVERGED) {
            float ms = 0.f;
            if (hipEventRecord(t1, c->s0) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
                hipEventElapsedTime(&ms, t0, t1) != hipSuccess)
                st = fail(c, HPL_ERR_HIP, "refinement timing failed");
            out->gefa_s = tg.gefa_s;
            out->total_s = ms * 1e-3;
        }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return st;
}

namespace {

// z = F^-1 v for an FP64 vector v (FP32 substitution with the resident
// factor, result promoted); xs is the FP32 scratch of length n.
hpl_status precond_apply(hpl_ctx* c, const double* v, double* z) {
    const int64_t n = (int64_t)c->n;
    const int B = (int)c->B;
    const double one = 1.0;
    hpl_status st = hplx::ensure_block_inverses(c);
    if (st) return st;
    HIPCHK(c, launch_d2f_multi(v, n, c->dIrXs, n, n, 1, &one, c->s0,
                               c->pivoted ? c->dPerm : nullptr));
    GeslWs wsbuf;
    const GeslWs* ws = hplx::gesl_workspace(c, &wsbuf);
    HIPCHK(c, launch_gesl_multi(c->dA, (int64_t)c->lda, n, B, c->dInv, 3 * (B / 32) * 1024,
                                c->dIrXs, n, 1, c->s0, c->binv_ready ? c->dBInv : nullptr, ws));
    st = hplx::gesl_check(c, ws);
    if (st) return st;
    HIPCHK(c, launch_axpy_multi(z, n, c->dIrXs, n, n, 1, &one, 1, c->s0));
    return HPL_SUCCESS;
}

// Givens rotation (c, s) with c a + s b = r, -s a + c b = 0, following the
// reference's rotmat (LINPACK/src/host/gmres.c:29-43).
void givens(double a, double b, double* cs, double* sn) {
    if (b == 0.0) {
        *cs = 1.0;
        *sn = 0.0;
    } else if (std::fabs(b) > std::fabs(a)) {
        const double t = a / b;
        *sn = 1.0 / std::sqrt(1.0 + t * t);
        *cs = t * *sn;
    } else {
        const double t = b / a;
        *cs = 1.0 / std::sqrt(1.0 + t * t);
        *sn = t * *cs;
    }
}

}  // namespace

/* GMRES-based refinement: each correction solves M d = r with restarted,
 * left-preconditioned GMRES in FP64 (the reference's gmres_ref,
 * LINPACK/src/host/gmres.c:53-232, with the low-precision factor as the
 * preconditioner and classical Gram-Schmidt applied twice), then x += d. */
hpl_status hpl_solve_gmres(hpl_ctx* c, uint32_t block, hpl_prec prec, int restart, int max_iter,
                           double tol, hpl_ir_result* out) {
    if (!c || !out) return HPL_ERR_INVALID_ARG;
    if (!c->have_orig) return fail(c, HPL_ERR_STATE, "hpl_solve_gmres needs hpl_keep_original(1)");
    if (max_iter < 0 || restart < 1 || restart > 64)
        return fail(c, HPL_ERR_INVALID_ARG, "restart must be in [1,64] and max_iter >= 0");
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = (int64_t)c->n;
    const int m = restart;
    hpl_status st = ensure(c, c->dX64, c->capIR, n * sizeof(double));
    if (!st) st = ensure(c, c->dGmV, c->capGmV, (size_t)(m + 1) * n * sizeof(double));
    if (!st) st = ensure(c, c->dGmW, c->capGmW, (size_t)n * sizeof(double));
    if (!st) st = ensure(c, c->dGmC, c->capGmC, (size_t)(2 * m + 2) * sizeof(double));
    if (!st)
        st = ensure(c, c->dGmWork, c->capGmWork,
                    (size_t)gmres_work_doubles(n, m + 1) * sizeof(double));
    if (!st) {
        const hplx::IrPlan p = hplx::plan_ir(c->n, 1);
        st = ensure(c, c->dWork, c->capWork, p.work);
        if (!st) st = ensure(c, c->dIrR, c->capIrR, p.r);
        if (!st) st = ensure(c, c->dIrXs, c->capIrXs, p.xs);
        if (!st) st = ensure(c, c->dIrOut, c->capIrOut, p.out);
    }
    if (st) return st;
    HIPCHK(c, hipMemcpyAsync(c->dA, c->dA0, n * c->lda * sizeof(float), hipMemcpyDeviceToDevice,
                             c->s0));
    hipEvent_t t0, t1;
    HIPCHK(c, hipEventCreate(&t0));
    if (hipEventCreate(&t1) != hipSuccess) {
        (void)hipEventDestroy(t0);
        return fail(c, HPL_ERR_HIP, "hipEventCreate failed");
    }
    struct EventGuard {
        hipEvent_t a, b;
        ~EventGuard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } guard{t0, t1};
    HIPCHK(c, hipEventRecord(t0, c->s0));
    hpl_timing tg{};
    st = hpl_gefa_dev(c, block, prec, &tg);
    if (st) return st;
    double* V = c->dGmV;
    double* W = c->dGmW;
    double* dh = c->dGmC;  // coefficients: [0, m+1) pass 1 / solution, [m+1, 2m+2) pass 2
    double* work = c->dGmWork;
    // x0 = F^-1 b
    st = precond_apply(c, c->db64, c->dX64);
    if (st) return st;
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), h1(m + 1), h2(m + 1),
        y(m);
    double hres[4];
    int inner_total = 0, outer = 0;
    double berr = 0.0;
    for (;; ++outer) {
        HIPCHK(c, launch_residual_multi(c->dA0, (int64_t)c->lda, n, c->dX64, n, c->db64, n, 1,
                                        c->dWork, c->dIrR, n, c->dIrOut, c->s0));
        HIPCHK(c, hipMemcpyAsync(hres, c->dIrOut, sizeof hres, hipMemcpyDeviceToHost, c->s0));
        HIPCHK(c, hipStreamSynchronize(c->s0));
        berr = hres[0] / (hres[1] * hres[2] + hres[3]);
        if (berr <= tol || outer >= max_iter || !(hres[0] > 0.0) || !(hres[0] < HUGE_VAL)) break;
        // GMRES on F^-1 M d = F^-1 r, d0 = 0: V0 = z / beta
        st = precond_apply(c, c->dIrR, W);
        if (st) return st;
        HIPCHK(c, launch_multi_dot(W, n, 1, W, n, dh, work, c->s0));
        double beta2 = 0.0;
        HIPCHK(c, hipMemcpyAsync(&beta2, dh, sizeof(double), hipMemcpyDeviceToHost, c->s0));
        HIPCHK(c, hipStreamSynchronize(c->s0));
        const double beta = std::sqrt(beta2);
        if (!(beta > 0.0) || !(beta < HUGE_VAL)) break;
        HIPCHK(c, launch_scale_copy(V, W, 1.0 / beta, n, c->s0));
        std::fill(H.begin(), H.end(), 0.0);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int k = 0;
        for (int i = 0; i < m; ++i) {
            // W = F^-1 M V_i
            HIPCHK(c, launch_matvec_f64(c->dA0, (int64_t)c->lda, n, V + (int64_t)i * n, W, work,
                                        c->s0));
            st = precond_apply(c, W, W);
            if (st) return st;
            // classical Gram-Schmidt, twice: h = V^T W, W -= V h
            HIPCHK(c, launch_multi_dot(V, n, i + 1, W, n, dh, work, c->s0));
            HIPCHK(c, launch_multi_axpy(W, 1.0, -1.0, V, n, i + 1, dh, n, c->s0));
            HIPCHK(c, launch_multi_dot(V, n, i + 1, W, n, dh + m + 1, work, c->s0));
            HIPCHK(c, launch_multi_axpy(W, 1.0, -1.0, V, n, i + 1, dh + m + 1, n, c->s0));
            double nw2 = 0.0;
            HIPCHK(c, hipMemcpyAsync(h1.data(), dh, (i + 1) * sizeof(double),
                                     hipMemcpyDeviceToHost, c->s0));
            HIPCHK(c, hipMemcpyAsync(h2.data(), dh + m + 1, (i + 1) * sizeof(double),
                                     hipMemcpyDeviceToHost, c->s0));
            HIPCHK(c, launch_multi_dot(W, n, 1, W, n, dh + 2 * m + 1, work, c->s0));
            HIPCHK(c, hipMemcpyAsync(&nw2, dh + 2 * m + 1, sizeof(double), hipMemcpyDeviceToHost,
                                     c->s0));
            HIPCHK(c, hipStreamSynchronize(c->s0));
            const double hn = std::sqrt(nw2);
            double* col = &H[(size_t)i * (m + 1)];  // column i of H, rows 0..i+1
            for (int q = 0; q <= i; ++q) col[q] = h1[q] + h2[q];
            col[i + 1] = hn;
            if (hn > 0.0 && i + 1 < m + 1)
                HIPCHK(c, launch_scale_copy(V + (int64_t)(i + 1) * n, W, 1.0 / hn, n, c->s0));
            // previous rotations, then the new one (reference: gmres.c:145-162)
            for (int q = 0; q < i; ++q) {
                const double tq = cs[q] * col[q] + sn[q] * col[q + 1];
                col[q + 1] = -sn[q] * col[q] + cs[q] * col[q + 1];
                col[q] = tq;
            }
            givens(col[i], col[i + 1], &cs[i], &sn[i]);
            const double gi = cs[i] * g[i];
            g[i + 1] = -sn[i] * g[i];
            g[i] = gi;
            col[i] = cs[i] * col[i] + sn[i] * col[i + 1];
            col[i + 1] = 0.0;
            k = i + 1;
            // the preconditioned system is solved to the FP32 accuracy of the
            // preconditioner; the outer loop measures the true residual
            if (std::fabs(g[i + 1]) <= 1e-10 * beta || !(hn > 0.0)) break;
        }
        // y = H(0:k, 0:k)^-1 g, x += V y
        for (int q = k - 1; q >= 0; --q) {
            double sum = g[q];
            for (int p2 = q + 1; p2 < k; ++p2) sum -= H[(size_t)p2 * (m + 1) + q] * y[p2];
            y[q] = sum / H[(size_t)q * (m + 1) + q];
        }
        HIPCHK(c, hipMemcpyAsync(dh, y.data(), k * sizeof(double), hipMemcpyHostToDevice, c->s0));
        HIPCHK(c, launch_multi_axpy(c->dX64, 1.0, 1.0, V, n, k, dh, n, c->s0));
        inner_total += k;
    }
    float ms = 0.f;
    HIPCHK(c, hipEventRecord(t1, c->s0));
    HIPCHK(c, hipEventSynchronize(t1));
    HIPCHK(c, hipEventElapsedTime(&ms, t0, t1));
    out->iterations = inner_total;
    out->converged = berr <= tol ? 1 : 0;
    out->backward_error = berr;
    out->residual_inf = hres[0];
    out->norm_a = hres[1];
    out->norm_x = hres[2];
    out->norm_b = hres[3];
    out->gefa_s = tg.gefa_s;
    out->total_s = ms * 1e-3;
    c->factored = true;
    if (!out->converged)
        return fail(c, HPL_ERR_NOT_CONVERGED,
                    "GMRES refinement stopped after %d outer / %d inner iterations at backward "
                    "error %.3e (tol %.3e)",
                    outer, inner_total, berr, tol);
    return HPL_SUCCESS;
}

/* Solve nrhs further right-hand sides with the resident factor and refine
 * each in FP64 against the kept original. */
hpl_status hpl_solve_rhs64(hpl_ctx* c, int nrhs, const double* b, uint64_t ldb, double* x,
                           uint64_t ldx, int max_iter, double tol, hpl_ir_result* res) {
    if (!c) return HPL_ERR_INVALID_ARG;
    if (!b || !x || !res || nrhs < 1 || nrhs > 16 || max_iter < 0)
        return fail(c, HPL_ERR_INVALID_ARG, "bad hpl_solve_rhs64 arguments (nrhs=%d)", nrhs);
    if (!c->factored || !c->have_orig)
        return fail(c, HPL_ERR_STATE, "hpl_solve_rhs64 needs a factor of a kept original");
    const uint64_t n = c->n;
    if (ldb < n || ldx < n) return fail(c, HPL_ERR_INVALID_ARG, "ldb or ldx < n");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t bytes = n * (size_t)nrhs * sizeof(double);
    hpl_status st = ensure(c, c->dRhsB, c->capRhsB, bytes);
    if (st) return st;
    st = ensure(c, c->dRhsX, c->capRhsX, bytes);
    if (st) return st;
    HIPCHK(c, hipMemcpy2DAsync(c->dRhsB, n * sizeof(double), b, ldb * sizeof(double),
                               n * sizeof(double), (size_t)nrhs, hipMemcpyHostToDevice, c->s0));
    st = hplx::ir_solve(c, nrhs, c->dRhsB, c->dRhsX, max_iter, tol, res);
    if (st && st != HPL_ERR_NOT_CONVERGED) return st;
    HIPCHK(c, hipMemcpy2DAsync(x, ldx * sizeof(double), c->dRhsX, n * sizeof(double),
                               n * sizeof(double), (size_t)nrhs, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    return st;
}

hpl_status hpl_download_solution64(hpl_ctx* c, double* x) {
    if (!c || !x) return HPL_ERR_INVALID_ARG;
    if (!c->dX64) return fail(c, HPL_ERR_STATE, "no refined solution");
    HIPCHK(c, hipMemcpy(x, c->dX64, c->n * sizeof(double), hipMemcpyDeviceToHost));
    return HPL_SUCCESS;
}

hpl_status hpl_upload_rhs64(hpl_ctx* c, const double* b) {
    if (!c || !b) return HPL_ERR_INVALID_ARG;
    if (!c->dA) return fail(c, HPL_ERR_STATE, "no resident system");
    HIPCHK(c, hipMemcpy(c->db64, b, c->n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, launch_d2f_scaled(c->db64, c->db, (int64_t)c->n, 1.0, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    return HPL_SUCCESS;
}

hpl_status hpl_set_profiling(hpl_ctx* c, int enable) {
    if (!c) return HPL_ERR_INVALID_ARG;
    c->prof = enable != 0;
    c->ev_used = 0;
    c->prof_flops = 0.0;
    c->prof_bytes = 0.0;
    return HPL_SUCCESS;
}

hpl_status hpl_kernel_bytes(hpl_ctx* c, double* bytes) {
    if (!c || !bytes) return HPL_ERR_INVALID_ARG;
    *bytes = c->prof_bytes;
    return HPL_SUCCESS;
}

hpl_status hpl_kernel_stats(hpl_ctx* c, uint64_t* launches, double* seconds, double* flops) {
    if (!c) return HPL_ERR_INVALID_ARG;
    HIPCHK(c, hipStreamSynchronize(c->s0));
    double s = 0.0;
    for (size_t i = 0; i < c->ev_used; ++i) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev[i].first, c->ev[i].second));
        s += ms * 1e-3;
    }
    if (launches) *launches = c->ev_used;
    if (seconds) *seconds = s;
    if (flops) *flops = c->prof_flops;
    return HPL_SUCCESS;
}

hpl_status hpl_kernel_busy(hpl_ctx* c, double* seconds) {
    if (!c || !seconds) return HPL_ERR_INVALID_ARG;
    HIPCHK(c, hipStreamSynchronize(c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s1));
    std::vector<std::pair<double, double>> iv;
    iv.reserve(c->ev_used);
    for (size_t i = 0; i < c->ev_used; ++i) {
        float a = 0.f, b = 0.f;
        HIPCHK(c, hipEventElapsedTime(&a, c->ev[0].first, c->ev[i].first));
        HIPCHK(c, hipEventElapsedTime(&b, c->ev[0].first, c->ev[i].second));
        iv.emplace_back(a * 1e-3, b * 1e-3);
    }
    std::sort(iv.begin(), iv.end());
    double busy = 0.0, lo = 0.0, hi = -1.0;
    for (const auto& x : iv) {
        if (x.first > hi) {
            if (hi > lo) busy += hi - lo;
            lo = x.first;
            hi = x.second;
        } else {
            hi = std::max(hi, x.second);
        }
    }
    if (hi > lo) busy += hi - lo;
    *seconds = busy;
    return HPL_SUCCESS;
}

}  // extern "C"
