// C ABI for batched solves of many systems of uneven order (hpl_batch_*).
//
// Systems are grouped in size buckets by their padded order (<= 128, 256, 512,
// 1024).  A bucket is solved either by one persistent work-queue launch (one
// workgroup per system) or blocked: one launch per 32-wide step for the
// diagonal blocks, the strips and the trailing updates of all its systems,
// then the persistent kernel for the substitutions and refinement only.  The
// autotuner picks the mode and the workgroup size by timing; the buckets'
// launches run concurrently on forked streams and the whole solve -- restore
// of the matrices (skipped when every bucket is blocked: the first step then
// reads the originals), queue resets, the bucket launches -- is captured once
// into a HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/hpl_hip.h"
#include "kernels.hpp"

namespace {

constexpr int kMaxBatchOrder = 1024;  // LDS holds the two strips of the largest system
constexpr int kBuckets = 4;

struct Bucket {
    int hi = 0;         // largest padded order in the bucket
    int begin = 0;      // first entry in the order array
    int count = 0;
    int maxm = 0;       // largest trailing size (npad - 32)
    int threads = 256;  // workgroup size (autotuned)
    bool blocked = false;  // blocked factorisation (autotuned)
    // scratch of the blocked factorisation: diagonal blocks (count x 1056
    // floats), strip copies of two consecutive steps (count x maxm x 32
    // floats each)
    float* dDg = nullptr;
    void* dLg = nullptr;
    void* dTg = nullptr;
    void* dLg2 = nullptr;
    void* dTg2 = nullptr;
};

}  // namespace

struct hpl_batch {
    hpl_ctx* ctx = nullptr;
    int device = 0;
    int nsys = 0;
    std::vector<BatchSystem> sys;
    std::vector<int> order;
    Bucket buckets[kBuckets];
    int64_t a_total = 0, v_total = 0;
    BatchSystem* dSys = nullptr;
    int* dOrder = nullptr;
    int* dCounters = nullptr;
    float* dA = nullptr;
    float* dA0 = nullptr;
    double* db64 = nullptr;
    double* dx64 = nullptr;
    double* drs = nullptr;
    float* dxs = nullptr;
    BatchResult* dRes = nullptr;
    int* dBad = nullptr;  // zero-pivot flags of the blocked factorisation
    hipStream_t s0 = nullptr;
    hipStream_t sb[kBuckets] = {};
    hipEvent_t fork = nullptr, join[kBuckets] = {}, t0 = nullptr, t1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_prec = -1, graph_iter = -1;
    double graph_tol = -1.0;
    int cus = 256;
    std::string err;
};

namespace {

hpl_status bfail(hpl_batch* b, hpl_status st, const char* fmt, ...) {
    if (b) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        b->err = buf;
    }
    return st;
}

#define BCHK(b, expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return bfail((b), e_ == hipErrorOutOfMemory ? HPL_ERR_OUT_OF_MEMORY : HPL_ERR_HIP, \
                         "%s failed: %s", #expr, hipGetErrorString(e_));               \
    } while (0)

void drop_graph(hpl_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int grid_for(const hpl_batch* b, const Bucket& k, bool f16) {
    const size_t lds = batch_lds_bytes(k.maxm, f16);
    int per_cu = (int)std::max<size_t>(1, kBatchLdsLimit / std::max<size_t>(lds, 1));
    per_cu = std::min(per_cu, 2048 / k.threads);
    return std::max(1, std::min(k.count, b->cus * per_cu));
}

// the blocked scratch of bucket k, allocated on first use (never under capture)
hpl_status ensure_blocked(hpl_batch* b, Bucket& k) {
    if (k.dDg || !k.count) return HPL_SUCCESS;
    const size_t strip = (size_t)k.count * std::max(k.maxm, 32) * 32 * sizeof(float);
    if (!b->dBad) BCHK(b, hipMalloc(&b->dBad, b->nsys * sizeof(int)));
    BCHK(b, hipMalloc(&k.dDg, (size_t)k.count * 1056 * sizeof(float)));
    BCHK(b, hipMalloc(&k.dLg, strip));
    BCHK(b, hipMalloc(&k.dTg, strip));
    BCHK(b, hipMalloc(&k.dLg2, strip));
    BCHK(b, hipMalloc(&k.dTg2, strip));
    return HPL_SUCCESS;
}

// from_a0: the matrices were not restored into dA; a blocked bucket's first
// step reads the originals from dA0 instead
hpl_status enqueue_bucket(hpl_batch* b, int i, bool f16, int max_iter, double tol,
                          hipStream_t s, bool from_a0 = false) {
    const Bucket& k = b->buckets[i];
    if (!k.count) return HPL_SUCCESS;
    if (!k.blocked) {
        BCHK(b, launch_batch_solve(k.threads, f16, grid_for(b, k, f16), b->dSys,
                                   b->dOrder + k.begin, k.count, b->dCounters + i, b->dA, b->dA0,
                                   b->db64, b->dx64, b->dxs, b->drs, max_iter, tol, b->dRes,
                                   std::max(k.maxm, 32), s));
        return HPL_SUCCESS;
    }
    if (!k.dDg) return bfail(b, HPL_ERR_STATE, "blocked scratch of bucket %d missing", i);
    // pairs of steps over the systems still longer than the step: a prefix
    // of the bucket's largest-first order
    const int* order = b->order.data() + k.begin;
    const int top = b->sys[order[0]].npad;
    const size_t sstride = (size_t)std::max(k.maxm, 32) * 32;
    auto active = [&](int above) {
        int a = 0;
        while (a < k.count && b->sys[order[a]].npad > above) ++a;
        return a;
    };
    for (int k0 = 0; k0 < top; k0 += 64) {
        const int act_diag = active(k0), act = active(k0 + 32), act2 = active(k0 + 64);
        const int mmax = act ? top - k0 - 32 : 0;
        BCHK(b, launch_batch_step2(f16, b->dSys, b->dOrder + k.begin, k0, act_diag, act, act2,
                                   mmax, b->dA, from_a0 ? b->dA0 : nullptr, k.dDg, b->dBad, k.dLg,
                                   k.dTg, k.dLg2, k.dTg2, sstride, s));
    }
    // substitutions and refinement: LDS for npad doubles and npad floats
    Bucket solve = k;
    solve.maxm = std::max(32, (k.hi + 4) / 5);
    BCHK(b, launch_batch_solve(k.threads, f16, grid_for(b, solve, f16), b->dSys,
                               b->dOrder + k.begin, k.count, b->dCounters + i, b->dA, b->dA0,
                               b->db64, b->dx64, b->dxs, b->drs, max_iter, tol, b->dRes,
                               solve.maxm, s, b->dBad));
    return HPL_SUCCESS;
}

// restore (unless every bucket is blocked) + all buckets, forked over the
// bucket streams, on s0
hpl_status enqueue_solve(hpl_batch* b, bool f16, int max_iter, double tol) {
    // every element of a matrix is written by its first blocked step, which
    // then reads the original: with every bucket blocked no restore is needed
    bool all_blocked = true;
    for (const auto& k : b->buckets) all_blocked = all_blocked && (k.blocked || !k.count);
    if (!all_blocked)
        BCHK(b, hipMemcpyAsync(b->dA, b->dA0, b->a_total * sizeof(float),
                               hipMemcpyDeviceToDevice, b->s0));
    BCHK(b, hipEventRecord(b->fork, b->s0));
    for (int i = 0; i < kBuckets; ++i) {
        BCHK(b, hipStreamWaitEvent(b->sb[i], b->fork, 0));
        hpl_status st = enqueue_bucket(b, i, f16, max_iter, tol, b->sb[i], all_blocked);
        if (st) return st;
        BCHK(b, hipEventRecord(b->join[i], b->sb[i]));
        BCHK(b, hipStreamWaitEvent(b->s0, b->join[i], 0));
    }
    return HPL_SUCCESS;
}

}  // namespace

extern "C" {

hpl_status hpl_batch_create(hpl_ctx* ctx, const int32_t* n, int nsys, hpl_batch** out) {
    if (!ctx || !n || nsys <= 0 || !out) return HPL_ERR_INVALID_ARG;
    *out = nullptr;
    for (int i = 0; i < nsys; ++i)
        if (n[i] < 1 || n[i] > kMaxBatchOrder) return HPL_ERR_UNSUPPORTED;
    hpl_batch* b = new hpl_batch();
    b->ctx = ctx;
    if (hipGetDevice(&b->device) != hipSuccess) {
        delete b;
        return HPL_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->device) == hipSuccess) b->cus = prop.multiProcessorCount;
    if (batch_prepare() != hipSuccess) {
        (void)hipGetLastError();
        delete b;
        return HPL_ERR_HIP;
    }
    b->nsys = nsys;
    b->sys.resize(nsys);
    for (int i = 0; i < nsys; ++i) {
        BatchSystem& s = b->sys[i];
        s.n = n[i];
        s.npad = (n[i] + 31) / 32 * 32;
        s.a_off = b->a_total;
        s.v_off = b->v_total;
        b->a_total += (int64_t)s.npad * s.npad;
        b->v_total += s.npad;
    }
    // buckets by padded order, largest-first inside each
    const int his[kBuckets] = {128, 256, 512, 1024};
    std::vector<int> idx(nsys);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int x, int y) { return b->sys[x].n > b->sys[y].n; });
    for (int k = kBuckets - 1; k >= 0; --k) {
        Bucket& bk = b->buckets[k];
        bk.hi = his[k];
        bk.begin = (int)b->order.size();
        for (int i : idx) {
            const int np = b->sys[i].npad;
            const int lo = k ? his[k - 1] : 0;
            if (np > lo && np <= bk.hi) {
                b->order.push_back(i);
                bk.maxm = std::max(bk.maxm, np - 32);
            }
        }
        bk.count = (int)b->order.size() - bk.begin;
        bk.threads = bk.hi <= 128 ? 128 : 256;
    }
    auto fail_alloc = [&]() {
        hpl_batch_destroy(b);
        return HPL_ERR_OUT_OF_MEMORY;
    };
    if (hipMalloc(&b->dSys, nsys * sizeof(BatchSystem)) != hipSuccess ||
        hipMalloc(&b->dOrder, nsys * sizeof(int)) != hipSuccess ||
        hipMalloc(&b->dCounters, kBuckets * sizeof(int)) != hipSuccess ||
        hipMalloc(&b->dA, b->a_total * sizeof(float)) != hipSuccess ||
        hipMalloc(&b->dA0, b->a_total * sizeof(float)) != hipSuccess ||
        hipMalloc(&b->db64, b->v_total * sizeof(double)) != hipSuccess ||
        hipMalloc(&b->dx64, b->v_total * sizeof(double)) != hipSuccess ||
        hipMalloc(&b->drs, b->v_total * sizeof(double)) != hipSuccess ||
        hipMalloc(&b->dxs, b->v_total * sizeof(float)) != hipSuccess ||
        hipMalloc(&b->dRes, nsys * sizeof(BatchResult)) != hipSuccess)
        return fail_alloc();
    bool ok = hipStreamCreateWithFlags(&b->s0, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreate(&b->t0) == hipSuccess && hipEventCreate(&b->t1) == hipSuccess;
    for (int i = 0; i < kBuckets && ok; ++i)
        ok = hipStreamCreateWithFlags(&b->sb[i], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) == hipSuccess;
    if (!ok ||
        hipMemcpy(b->dSys, b->sys.data(), nsys * sizeof(BatchSystem), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(b->dOrder, b->order.data(), nsys * sizeof(int), hipMemcpyHostToDevice) !=
            hipSuccess) {
        hpl_batch_destroy(b);
        return HPL_ERR_HIP;
    }
    *out = b;
    return HPL_SUCCESS;
}

void hpl_batch_destroy(hpl_batch* b) {
    if (!b) return;
    if (b->s0) (void)hipStreamSynchronize(b->s0);
    drop_graph(b);
    for (void* p : {(void*)b->dSys, (void*)b->dOrder, (void*)b->dCounters, (void*)b->dA,
                    (void*)b->dA0, (void*)b->db64, (void*)b->dx64, (void*)b->drs, (void*)b->dxs,
                    (void*)b->dRes, (void*)b->dBad})
        if (p) (void)hipFree(p);
    for (auto& k : b->buckets)
        for (void* p : {(void*)k.dDg, k.dLg, k.dTg, k.dLg2, k.dTg2})
            if (p) (void)hipFree(p);
    for (int i = 0; i < kBuckets; ++i) {
        if (b->sb[i]) (void)hipStreamDestroy(b->sb[i]);
        if (b->join[i]) (void)hipEventDestroy(b->join[i]);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->t0) (void)hipEventDestroy(b->t0);
    if (b->t1) (void)hipEventDestroy(b->t1);
    if (b->s0) (void)hipStreamDestroy(b->s0);
    delete b;
}

const char* hpl_batch_last_error(const hpl_batch* b) { return b ? b->err.c_str() : "null batch"; }

hpl_status hpl_batch_generate(hpl_batch* b, uint64_t seed) {
    if (!b) return HPL_ERR_INVALID_ARG;
    BCHK(b, launch_batch_generate(b->dSys, b->nsys, seed, b->dA0, b->db64, b->s0));
    BCHK(b, hipStreamSynchronize(b->s0));
    return HPL_SUCCESS;
}

hpl_status hpl_batch_set_system(hpl_batch* b, int s, const float* a, uint64_t lda,
                                const double* rhs) {
    if (!b || s < 0 || s >= b->nsys || !a || !rhs) return HPL_ERR_INVALID_ARG;
    const BatchSystem& sy = b->sys[s];
    if (lda < (uint64_t)sy.n) return bfail(b, HPL_ERR_INVALID_ARG, "lda < n");
    std::vector<float> pad((size_t)sy.npad * sy.npad, 0.0f);
    for (int r = 0; r < sy.npad; ++r)
        for (int c = 0; c < sy.npad; ++c)
            pad[(size_t)r * sy.npad + c] =
                (r < sy.n && c < sy.n) ? a[r * lda + c] : (r == c ? 1.0f : 0.0f);
    std::vector<double> bv(sy.npad, 0.0);
    std::copy(rhs, rhs + sy.n, bv.begin());
    BCHK(b, hipMemcpy(b->dA0 + sy.a_off, pad.data(), pad.size() * sizeof(float),
                      hipMemcpyHostToDevice));
    BCHK(b, hipMemcpy(b->db64 + sy.v_off, bv.data(), bv.size() * sizeof(double),
                      hipMemcpyHostToDevice));
    return HPL_SUCCESS;
}

hpl_status hpl_batch_set_blocked(hpl_batch* b, int blocked) {
    if (!b) return HPL_ERR_INVALID_ARG;
    drop_graph(b);
    for (auto& k : b->buckets) {
        if (blocked) {
            hpl_status st = ensure_blocked(b, k);
            if (st) return st;
        }
        k.blocked = blocked != 0 && k.count > 0;
    }
    return HPL_SUCCESS;
}

hpl_status hpl_batch_autotune(hpl_batch* b, hpl_prec prec, int max_iter, double tol_per_n,
                              int* threads_out) {
    if (!b) return HPL_ERR_INVALID_ARG;
    if (prec != HPL_PREC_FP16 && prec != HPL_PREC_FP32)
        return bfail(b, HPL_ERR_UNSUPPORTED, "batched solves run FP16 or FP32 trailing updates");
    const bool f16 = prec == HPL_PREC_FP16;
    drop_graph(b);
    const int cands[3] = {128, 256, 512};
    for (int i = 0; i < kBuckets; ++i) {
        Bucket& k = b->buckets[i];
        if (!k.count) continue;
        hpl_status st0 = ensure_blocked(b, k);
        if (st0) return st0;
        const bool lds_ok = f16 || batch_lds_bytes(std::max(k.maxm, 32), false) <= kBatchLdsLimit;
        float best = 1e30f;
        int best_t = k.threads;
        bool best_blk = k.blocked;
        for (int cand = 0; cand < 6; ++cand) {
            const int t = cands[cand % 3];
            const bool blk = cand >= 3;
            if (!blk && !lds_ok) continue;
            k.threads = t;
            k.blocked = blk;
            float ms_sum = 0.f;
            for (int rep = 0; rep < 2; ++rep) {
                BCHK(b, hipMemcpyAsync(b->dA, b->dA0, b->a_total * sizeof(float),
                                       hipMemcpyDeviceToDevice, b->s0));
                BCHK(b, hipEventRecord(b->t0, b->s0));
                hpl_status st = enqueue_bucket(b, i, f16, max_iter, tol_per_n, b->s0);
                if (st) return st;
                BCHK(b, hipEventRecord(b->t1, b->s0));
                BCHK(b, hipEventSynchronize(b->t1));
                float ms = 0.f;
                BCHK(b, hipEventElapsedTime(&ms, b->t0, b->t1));
                if (rep) ms_sum += ms;  // first run warms up
            }
            if (ms_sum < best) {
                best = ms_sum;
                best_t = t;
                best_blk = blk;
            }
        }
        k.threads = best_t;
        k.blocked = best_blk;
    }
    // The buckets run concurrently: a bucket timed alone can prefer one
    // workgroup per system where, beside the others, the blocked mode (which
    // spreads each step over the device) finishes first.  Time the whole
    // solve with the per-bucket choice and with every bucket blocked (same
    // workgroup sizes), and keep the faster.
    auto time_all = [&](float* out_ms) -> hpl_status {
        float sum = 0.f;
        for (int rep = 0; rep < 3; ++rep) {
            BCHK(b, hipEventRecord(b->t0, b->s0));
            hpl_status st = enqueue_solve(b, f16, max_iter, tol_per_n);
            if (st) return st;
            BCHK(b, hipEventRecord(b->t1, b->s0));
            BCHK(b, hipEventSynchronize(b->t1));
            float ms = 0.f;
            BCHK(b, hipEventElapsedTime(&ms, b->t0, b->t1));
            if (rep) sum += ms;  // first run warms up
        }
        *out_ms = sum;
        return HPL_SUCCESS;
    };
    bool chosen[kBuckets];
    for (int i = 0; i < kBuckets; ++i) chosen[i] = b->buckets[i].blocked;
    float t_chosen = 0.f, t_blocked = 0.f;
    hpl_status st1 = time_all(&t_chosen);
    if (st1) return st1;
    for (auto& k : b->buckets) k.blocked = k.count > 0;
    st1 = time_all(&t_blocked);
    if (st1) return st1;
    if (t_chosen <= t_blocked)
        for (int i = 0; i < kBuckets; ++i) b->buckets[i].blocked = chosen[i];
    if (threads_out)
        for (int i = 0; i < kBuckets; ++i) {
            const Bucket& k = b->buckets[i];
            threads_out[i] = k.count ? (k.blocked ? -k.threads : k.threads) : 0;
        }
    return HPL_SUCCESS;
}

hpl_status hpl_batch_solve(hpl_batch* b, hpl_prec prec, int max_iter, double tol_per_n,
                           hpl_batch_stats* stats) {
    if (!b || max_iter < 0) return HPL_ERR_INVALID_ARG;
    if (prec != HPL_PREC_FP16 && prec != HPL_PREC_FP32)
        return bfail(b, HPL_ERR_UNSUPPORTED, "batched solves run FP16 or FP32 trailing updates");
    const bool f16 = prec == HPL_PREC_FP16;
    if (!f16)
        for (auto& k : b->buckets)
            if (k.count && !k.blocked &&
                batch_lds_bytes(std::max(k.maxm, 32), false) > kBatchLdsLimit)
                return bfail(b, HPL_ERR_UNSUPPORTED,
                             "FP32 batched updates need n <= 480 (or the blocked factorisation)");
    if (!b->exec || b->graph_prec != (int)prec || b->graph_iter != max_iter ||
        b->graph_tol != tol_per_n) {
        drop_graph(b);
        BCHK(b, hipStreamBeginCapture(b->s0, hipStreamCaptureModeGlobal));
        hpl_status st = enqueue_solve(b, f16, max_iter, tol_per_n);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->s0, &g);
        if (st) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        BCHK(b, e);
        b->graph = g;
        BCHK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->graph_prec = (int)prec;
        b->graph_iter = max_iter;
        b->graph_tol = tol_per_n;
    }
    BCHK(b, hipEventRecord(b->t0, b->s0));
    BCHK(b, hipGraphLaunch(b->exec, b->s0));
    BCHK(b, hipEventRecord(b->t1, b->s0));
    BCHK(b, hipEventSynchronize(b->t1));
    float ms = 0.f;
    BCHK(b, hipEventElapsedTime(&ms, b->t0, b->t1));
    if (stats) {
        std::vector<BatchResult> r(b->nsys);
        BCHK(b, hipMemcpy(r.data(), b->dRes, b->nsys * sizeof(BatchResult), hipMemcpyDeviceToHost));
        double flops = 0.0, worst = 0.0;
        int maxit = 0, failed = 0, pivots = 0;
        for (int i = 0; i < b->nsys; ++i) {
            const double n = b->sys[i].n;
            flops += 2.0 / 3.0 * n * n * n + 2.0 * n * n;
            worst = std::max(worst, r[i].backward_error / n);
            maxit = std::max(maxit, r[i].iterations);
            if (!(r[i].backward_error <= tol_per_n * n)) failed++;
            if (r[i].status) pivots++;
        }
        stats->seconds = ms * 1e-3;
        stats->flops = flops;
        stats->gflops = flops / (ms * 1e-3) / 1e9;
        stats->max_iterations = maxit;
        stats->max_backward_error_per_n = worst;
        stats->failed = failed;
        stats->zero_pivots = pivots;
    }
    return HPL_SUCCESS;
}

hpl_status hpl_batch_get(hpl_batch* b, int s, double* x, int* iterations, double* berr) {
    if (!b || s < 0 || s >= b->nsys) return HPL_ERR_INVALID_ARG;
    const BatchSystem& sy = b->sys[s];
    if (x) BCHK(b, hipMemcpy(x, b->dx64 + sy.v_off, sy.n * sizeof(double), hipMemcpyDeviceToHost));
    BatchResult r;
    BCHK(b, hipMemcpy(&r, b->dRes + s, sizeof r, hipMemcpyDeviceToHost));
    if (iterations) *iterations = r.iterations;
    if (berr) *berr = r.backward_error;
    return HPL_SUCCESS;
}

hpl_status hpl_batch_get_system(hpl_batch* b, int s, float* a, double* rhs) {
    if (!b || s < 0 || s >= b->nsys) return HPL_ERR_INVALID_ARG;
    const BatchSystem& sy = b->sys[s];
    if (a)
        BCHK(b, hipMemcpy(a, b->dA0 + sy.a_off, (size_t)sy.npad * sy.npad * sizeof(float),
                          hipMemcpyDeviceToHost));
    if (rhs)
        BCHK(b, hipMemcpy(rhs, b->db64 + sy.v_off, sy.npad * sizeof(double), hipMemcpyDeviceToHost));
    return HPL_SUCCESS;
}

}  // extern "C"
