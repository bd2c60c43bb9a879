// Partial pivoting behind the C ABI: the factorisation schedule on the
// transposed matrix (kernels in pivot.hip), the host-buffer entry points that
// replace gefa_ref and gesl_ref (LINPACK/src/host/linpack_data.cpp:72-156),
// and the conversions between the device factor and gefa_ref's layout.
//
// Device factor.  P M = L U with L carrying every row exchange (LAPACK's
// convention), stored like the unpivoted factor (negated multipliers below
// the diagonal, U above, -1/u_kk on it), so the substitution, the block
// inverses and the refinement of the unpivoted path apply unchanged to P b;
// perm gives P b as b[perm[i]].  gefa_ref's layout differs only by data
// movement: its column k has the exchanges of steps j > k undone, and its
// diagonal holds u_kk (kept in dDiagU).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/hpl_hip.h"
#include "ctx.hpp"
#include "kernels.hpp"

namespace {

using hplx::fail;

#define PCHK(ctx, expr)                                                                   \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail((ctx), e_ == hipErrorOutOfMemory ? HPL_ERR_OUT_OF_MEMORY : HPL_ERR_HIP, \
                        "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                        \
    } while (0)

// Byte offsets inside the context's pivot block.
struct PvLayout {
    size_t piv, perm, src, diag, xw, list, y, d, lt2, total;
};

size_t al256(size_t x) { return (x + 255) / 256 * 256; }

PvLayout pv_layout(uint64_t n, uint32_t B) {
    PvLayout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t r = o;
        o += al256(bytes);
        return r;
    };
    L.piv = take(n * sizeof(int32_t));
    L.perm = take(n * sizeof(int32_t));
    L.src = take(n * sizeof(int32_t));
    L.diag = take(n * sizeof(float));
    L.xw = take(panel_exchange_words((int64_t)n, (int)B) * sizeof(unsigned long long));
    L.list = take((4 * (size_t)B + 1) * sizeof(int));
    L.y = take(2 * (size_t)B * n * sizeof(float));
    L.d = take((size_t)B * B * sizeof(float));
    L.lt2 = take((size_t)B * n * sizeof(float));
    L.total = o;
    return L;
}

// Allocate the pivot block for order c->n at block B and point dPiv, dPerm
// and dDiagU into it.
hpl_status bind_pivot_block(hpl_ctx* c, uint32_t B, PvLayout* out) {
    const PvLayout L = pv_layout(c->n, B);
    hpl_status st = hplx::ensure_mem(c, &c->dPv, &c->capPv, L.total);
    if (st) return st;
    char* base = static_cast<char*>(c->dPv);
    c->dPiv = reinterpret_cast<int32_t*>(base + L.piv);
    c->dPerm = reinterpret_cast<int32_t*>(base + L.perm);
    c->dDiagU = reinterpret_cast<float*>(base + L.diag);
    *out = L;
    return HPL_SUCCESS;
}

// Run f(k) for k in [0, n) on up to 16 host threads, rows interleaved so the
// triangular workloads balance.
template <typename F>
void parallel_rows(uint64_t n, F f) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 512) nt = 1;
    std::vector<std::thread> th;
    for (unsigned w = 0; w < nt; ++w)
        th.emplace_back([=, &f]() {
            for (uint64_t k = w; k < n; k += nt) f(k);
        });
    for (auto& x : th) x.join();
}

}  // namespace

namespace hplx {

hpl_status gefa_pivot(hpl_ctx* c, uint32_t B) {
    const int64_t n = (int64_t)c->n, lda = (int64_t)c->lda;
    if (B < 32 || B > 256) return fail(c, HPL_ERR_UNSUPPORTED, "pivoting needs a block in [32, 256]");
    if (c->pv_B != B) {
        PCHK(c, panel_pivot_prepare((int)B, &c->pv_max_groups));
        c->pv_B = B;
    }
    const int R = panel_pivot_rows((int)B);
    const int64_t groups = (n + R - 1) / R;
    if (groups > c->pv_max_groups)
        return fail(c, HPL_ERR_UNSUPPORTED,
                    "pivot panel of %lld rows needs %lld co-resident workgroups of %d rows, the "
                    "device holds %d (use a smaller block)",
                    (long long)n, (long long)groups, R, c->pv_max_groups);
    PvLayout L{};
    hpl_status st = bind_pivot_block(c, B, &L);
    if (st) return st;
    char* base = static_cast<char*>(c->dPv);
    int* src = reinterpret_cast<int*>(base + L.src);
    // the panel exchange: candidate rows, row k, key slots (groups + 1 per
    // parity, the extra one for the error-path test's phantom workgroup)
    auto* pub = reinterpret_cast<unsigned long long*>(base + L.xw);
    auto* krow = pub + 2 * groups * (B + 1);
    auto* kslot = krow + 2 * (B + 1);
    int* list = reinterpret_cast<int*>(base + L.list);
    int* count = list + 4 * B;
    float* Y = reinterpret_cast<float*>(base + L.y);
    float* D = reinterpret_cast<float*>(base + L.d);
    hipStream_t s = c->s0;
    const int64_t N = n / B;
    const int64_t inv_stride = 3 * (B / 32) * 1024;
    PCHK(c, launch_zero_words(pub, panel_exchange_words(n, (int)B), s));
    PCHK(c, launch_transpose_square(c->dA, lda, n, s));
    PCHK(c, launch_iota(c->dPerm, n, s));
    // Look-ahead: step k updates the next panel's B columns on s0 first; the
    // rest of its trailing update runs on s1 while panel k+1 factorises on s0
    // (on the CUs the update leaves).  Panel k+1's row moves touch every
    // column, so they wait for that update.  The panels' transposed
    // multipliers alternate between two buffers: panel k+1 writes its own
    // while step k's update still reads panel k's.
    float* lts[2] = {c->dLt, reinterpret_cast<float*>(base + L.lt2)};
    bool rest_pending = false;
    for (int64_t k = 0; k < N; ++k) {
        const int64_t k0 = k * B, m1 = n - k0 - B;
        float* lt = lts[k & 1];
        PCHK(c, launch_panel_pivot(c->dA, lda, n, k0, (int)B, pub, krow, kslot, (int)groups + 1,
                                   lt, n, c->dDiagU, c->dPiv, src, c->dStatus, c->pv_max_groups,
                                   c->num_cu, s, c->pv_spin_limit));
        if (rest_pending) {
            PCHK(c, hipStreamWaitEvent(s, c->evP, 0));
            rest_pending = false;
        }
        PCHK(c, launch_row_moves(c->dA, lda, n, k0, (int)B, c->dPiv, src, list, count, Y, n,
                                 c->dPerm, c->dStatus, s));
        if (m1 > 0) {
            float* inv = c->dInv + k * inv_stride;
            PCHK(c, launch_unit_lower(c->dA, lda, k0, (int)B, D, s));
            PCHK(c, launch_prep(D, B, 0, (int)B, c->dDT, inv, s));
            PCHK(c, launch_trsm_top(c->dA, lda, k0, m1, (int)B, c->dDT, inv, s));
            float* C = c->dA + (k0 + B) * lda + k0 + B;
            const float* T = c->dA + k0 * lda + k0 + B;
            const int64_t nb = std::min<int64_t>(m1, B);
            PCHK(c, launch_gemm_update_f32(C, lda, lt + k0 + B, n, T, lda, (int)m1, (int)nb,
                                           (int)B, s));
            if (m1 > nb) {
                PCHK(c, hipEventRecord(c->evA, s));
                PCHK(c, hipStreamWaitEvent(c->s1, c->evA, 0));
                PCHK(c, launch_gemm_update_f32(C + nb, lda, lt + k0 + B, n, T + nb, lda, (int)m1,
                                               (int)(m1 - nb), (int)B, c->s1));
                PCHK(c, hipEventRecord(c->evP, c->s1));
                rest_pending = true;
            }
        }
    }
    if (rest_pending) PCHK(c, hipStreamWaitEvent(s, c->evP, 0));
    PCHK(c, launch_transpose_square(c->dA, lda, n, s));
    // the substitution's 32 x 32 inverses, from the final factor
    PCHK(c, launch_prep_all(c->dA, lda, (int)B, (int)N, c->dInv, inv_stride, s));
    return HPL_SUCCESS;
}

hpl_status pivot_factor_to_host(hpl_ctx* c, float* a, uint64_t lda) {
    const uint64_t n = c->n;
    std::vector<int32_t> ipvt(n);
    std::vector<float> u(n);
    PCHK(c, hipMemcpy(ipvt.data(), c->dPiv, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    PCHK(c, hipMemcpy(u.data(), c->dDiagU, n * sizeof(float), hipMemcpyDeviceToHost));
    // stored row k is M's column k: u_kk on the diagonal, and the exchanges
    // of the later steps j > k (positions j and ipvt[j] > j) undone in
    // reverse order
    parallel_rows(n, [&](uint64_t k) {
        float* row = a + k * lda;
        row[k] = u[k];
        for (uint64_t j = n - 1; j > k; --j) {
            const uint64_t p = (uint64_t)ipvt[j];
            if (p != j) std::swap(row[j], row[p]);
        }
    });
    return HPL_SUCCESS;
}

}  // namespace hplx

extern "C" {

hpl_status hpl_set_pivoting(hpl_ctx* c, int enable) {
    if (!c) return HPL_ERR_INVALID_ARG;
    c->pivoting = enable != 0;
    return HPL_SUCCESS;
}

hpl_status hpl_get_pivots(hpl_ctx* c, int32_t* ipvt) {
    if (!c || !ipvt) return HPL_ERR_INVALID_ARG;
    if (!c->factored || !c->pivoted) return fail(c, HPL_ERR_STATE, "no pivoted factor resident");
    PCHK(c, hipSetDevice(c->device));
    PCHK(c, hipMemcpy(ipvt, c->dPiv, c->n * sizeof(int32_t), hipMemcpyDeviceToHost));
    return HPL_SUCCESS;
}

hpl_status hpl_gefa_pvt(hpl_ctx* c, float* a, uint64_t n, uint64_t lda, uint32_t block,
                        int32_t* ipvt, hpl_timing* t) {
    if (!c) return HPL_ERR_INVALID_ARG;
    if (!a || !ipvt || lda < n || n == 0)
        return fail(c, HPL_ERR_INVALID_ARG, "bad gefa_pvt arguments (n=%llu lda=%llu)",
                    (unsigned long long)n, (unsigned long long)lda);
    hpl_status st = hpl_upload(c, a, n, lda, nullptr, t);
    if (st) return st;
    const bool saved = c->pivoting;
    c->pivoting = true;
    st = hpl_gefa_dev(c, block, HPL_PREC_FP32, t);
    c->pivoting = saved;
    if (st) return st;
    st = hpl_download(c, a, lda, nullptr, t);
    if (st) return st;
    return hpl_get_pivots(c, ipvt);
}

hpl_status hpl_gesl_pvt(hpl_ctx* c, const float* lu, float* b, const int32_t* ipvt, uint64_t n,
                        uint64_t lda, hpl_timing* t) {
    if (!c) return HPL_ERR_INVALID_ARG;
    if (!lu || !b || !ipvt || lda < n || n == 0)
        return fail(c, HPL_ERR_INVALID_ARG, "bad gesl_pvt arguments");
    for (uint64_t j = 0; j < n; ++j)
        if (ipvt[j] < (int32_t)j || (uint64_t)ipvt[j] >= n)
            return fail(c, HPL_ERR_INVALID_ARG, "ipvt[%llu] = %d outside [%llu, n)",
                        (unsigned long long)j, ipvt[j], (unsigned long long)j);
    uint32_t B = 256;
    while (B >= 32 && n % B) B /= 2;
    if (B < 32)
        return fail(c, HPL_ERR_UNSUPPORTED, "n=%llu is not a multiple of 32",
                    (unsigned long long)n);
    // gefa_ref's layout -> the device factor: every later exchange applied to
    // each column's multipliers, -1/u_kk on the diagonal
    std::vector<float> f(n * n);
    std::vector<float> u(n);
    parallel_rows(n, [&](uint64_t k) {
        float* row = f.data() + k * n;
        std::memcpy(row, lu + k * lda, n * sizeof(float));
        for (uint64_t j = k + 1; j < n; ++j) {
            const uint64_t p = (uint64_t)ipvt[j];
            if (p != j) std::swap(row[j], row[p]);
        }
        u[k] = row[k];
        row[k] = (float)(-1.0 / (double)row[k]);
    });
    std::vector<int32_t> perm(n);
    for (uint64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
    for (uint64_t j = 0; j < n; ++j) std::swap(perm[j], perm[(uint64_t)ipvt[j]]);
    hpl_status st = hpl_upload(c, f.data(), n, n, b, t);
    if (st) return st;
    st = hplx::alloc_factor_workspace(c, B);
    if (st) return st;
    PvLayout L{};
    st = bind_pivot_block(c, B, &L);
    if (st) return st;
    PCHK(c, hipMemcpy(c->dPiv, ipvt, n * sizeof(int32_t), hipMemcpyHostToDevice));
    PCHK(c, hipMemcpy(c->dPerm, perm.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
    PCHK(c, hipMemcpy(c->dDiagU, u.data(), n * sizeof(float), hipMemcpyHostToDevice));
    PCHK(c, launch_prep_all(c->dA, (int64_t)c->lda, (int)B, (int)(n / B), c->dInv,
                            3 * (B / 32) * 1024, c->s0));
    c->B = B;
    c->factored = true;
    c->pivoted = true;
    c->binv_ready = false;
    st = hpl_gesl_dev(c, t);
    if (st) return st;
    PCHK(c, hipMemcpy(b, c->dx, n * sizeof(float), hipMemcpyDeviceToHost));
    return HPL_SUCCESS;
}

}  // extern "C"
