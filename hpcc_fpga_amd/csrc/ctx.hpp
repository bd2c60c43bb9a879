// The solver context behind the C ABI (internal to libhpl_hip.so).
//
// A context owns two HIP streams and the device buffers of one resident
// system: the matrix being factorised (dA), an untouched copy for residuals
// (dA0), vectors, the factorisation workspaces and the per-block inverses.
// In its own mode every buffer is allocated on demand (hipMalloc) and freed by
// hpl_destroy.  In external mode (the serving engine) the buffers are regions
// of a larger allocation set up by the owner: nothing is allocated or freed,
// and a workspace that would have to grow is an error.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/hpl_hip.h"

struct hpl_ctx {
    int device = 0;
    bool external = false;
    hipStream_t s0 = nullptr;   // trailing updates (and everything else)
    hipStream_t s1 = nullptr;   // look-ahead panel factorisation
    hipStream_t s2 = nullptr;   // row and column blocks' last update in the pipelined schedule
    hipStream_t s3 = nullptr;   // its look-ahead updates of the next block
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipEvent_t evA = nullptr, evP = nullptr;  // next-panel region updated / panel ready
    hipEvent_t evS = nullptr, evD = nullptr;  // strips solved / row and column blocks updated
    hipEvent_t evL = nullptr;                 // next block's look-ahead update done
    bool lookahead = true;
    // Pipelined schedule (with look-ahead): the bulk update of one group of
    // `depth` steps runs on s1 while the next group's panels run on s0; the
    // transposed left strips then stay in one linear buffer (slot = step).
    bool pipelined = true;
    bool lt_linear = false;
    // Column slices (separate launches) of each bulk update of the pipelined
    // schedule; 0: one per step of the group, at most 6 for the 16-bit update.
    int bulk_slices = 0;
    // Pipelined 16-bit schedule: CUs kept free of the bulk update for the
    // chain (0: bulk in column slices instead); num_cu = the device's CUs.
    int chain_cus = 0;
    int num_cu = 0;
    // Tail of the pipelined schedule: once the order still to factorise is at
    // most tail_m (-1: half the order; 0: no tail), groups hold tail_depth
    // steps (0: half the depth, at least 1).  The row and column-block updates
    // on the critical path then carry fewer steps where the bulk update no
    // longer hides the panels: n = 32768 FP32 204.6 ms against 207.3, n =
    // 65536 BF16 320 ms against 324.5.
    int64_t tail_m = -1;
    int tail_depth = 0;
    // Head of the pipelined schedule for the BF16/FP16 update: while the order
    // still to factorise exceeds head_m (-1: 5n/8; 0: no head), groups hold
    // head_depth steps (0: 5/4 of the depth).  There the chain of a group
    // finishes well before its bulk update, and deeper groups give fewer,
    // longer bulk updates: n = 65536, depth 16, GEFA 239.3 / 240.7 / 239.9 ms
    // without the head against 238.0 / 238.1 / 237.3 with 20 steps beyond
    // order 40000 (A/B on one device; 24 or 32 steps were slower).
    int64_t head_m = -1;
    int head_depth = 0;
    // Steps per bulk trailing update (K = depth * B).  FP32: 3 -- at
    // n = 32768, B = 256 the bulk launches run at 0.75 of the FP32 MFMA peak
    // instead of 0.69-0.70 at depth 1, and with the pipelined schedule depth
    // 3 (207 ms) is within noise of 4 (207 ms) and ahead of 2 (210 ms).  The
    // 16-bit and FP8 updates are bound by the FP32 C traffic (8 bytes per
    // element per pass), which `depth` divides: 12 (n = 65536, BF16: 327 ms
    // against 334 at 8 and 329 at 16; FP8: 299 against 314 at 8).
    int depth = 1;
    int depth_fp32 = 4, depth_lp = 16;
    std::string err;

    // resident system
    uint64_t n = 0, lda = 0;
    float* dA = nullptr;
    size_t capA = 0;
    float* dA0 = nullptr;  // untouched copy (keep_original)
    size_t capA0 = 0;
    bool keep_orig = false;
    bool have_orig = false;
    float* db = nullptr;
    float* dx = nullptr;
    double* db64 = nullptr;
    double* dxt = nullptr;
    size_t capVec = 0;

    // GEFA workspaces
    uint32_t B = 0;             // block size of the resident factor (0: none)
    float* dLt = nullptr;       // transposed left strips, 2 * depth (linear: n / B) slots of B x n
    size_t capLt = 0;
    float* dDT = nullptr;       // transposed diagonal block, 2 x (B x B), by step parity
    size_t capDT = 0;
    float* dInv = nullptr;      // per block: 3 * (B/32) inverses of 32 x 32
    size_t capInv = 0;
    float* dBInv = nullptr;     // full inverses of the diagonal blocks (substitution), N x 2 x B x B
    size_t capBInv = 0;
    void* dGeslWs = nullptr;    // workspace of the single-launch substitution
    size_t capGeslWs = 0;
    bool binv_ready = false;    // dBInv holds the inverses of the current factor
    bool binv_in_gefa = false;  // the running factorisation fills dBInv
    float* dInner = nullptr;    // recursion scratch for blocks > 128
    size_t capInner = 0;
    // BF16 factorisation: the current steps' diagonal-block inverses for
    // the strip solves (two buffers, step parity), and whether to use them
    // (HPL_STRIP_SOLVE=recurrence: the 32 x 32 block recurrence instead)
    float* dTinv = nullptr;
    size_t capTinv = 0;
    bool strip_inverse = true;
    void* dLpT = nullptr;       // low-precision k-contiguous top-strip copies, slots of n x B (as dLt)
    void* dLpL = nullptr;       // low-precision k-contiguous left-strip copies, same layout
    size_t capLpT = 0, capLpL = 0;
    float* dScT = nullptr;      // FP8 block scales of the copies, slots of n/32 (as dLt)
    float* dScL = nullptr;
    size_t capScT = 0, capScL = 0;
    int prec = 0;               // trailing-update precision of the current factorisation
    int* dStatus = nullptr;
    size_t capStatus = 0;
    double* dWork = nullptr;    // residual partial sums
    size_t capWork = 0;
    double* dX64 = nullptr;     // refined solution of hpl_solve_ir (FP64)
    size_t capIR = 0;
    double* dOut4 = nullptr;
    // refinement of nrhs columns (leading dimension n): residuals (FP64),
    // FP32 corrections, norms (4 per column)
    double* dIrR = nullptr;
    float* dIrXs = nullptr;
    double* dIrOut = nullptr;
    size_t capIrR = 0, capIrXs = 0, capIrOut = 0;
    // GMRES refinement: Krylov basis (restart + 1 vectors of n), work vector,
    // basis coefficients and reduction scratch
    double* dGmV = nullptr;
    double* dGmW = nullptr;
    double* dGmC = nullptr;
    double* dGmWork = nullptr;
    size_t capGmV = 0, capGmW = 0, capGmC = 0, capGmWork = 0;
    // right-hand sides and solutions of hpl_solve_rhs64
    double* dRhsB = nullptr;
    double* dRhsX = nullptr;
    size_t capRhsB = 0, capRhsX = 0;
    bool factored = false;

    // partial pivoting (api_pivot.hip): hpl_gefa_dev factorises P M = L U when
    // `pivoting` is set; the factor then holds L (every row exchange applied),
    // U and -1/u_kk like the unpivoted encoding, and `pivoted` is set.
    bool pivoting = false;
    bool pivoted = false;
    void* dPv = nullptr;        // one block: the arrays below and the panel scratch
    size_t capPv = 0;
    int32_t* dPiv = nullptr;    // ipvt in gefa_ref's meaning
    int32_t* dPerm = nullptr;   // (P M)(i, :) = M(perm[i], :)
    float* dDiagU = nullptr;    // u_kk
    int pv_max_groups = 0;      // co-resident panel workgroups at block pv_B
    uint32_t pv_B = 0;
    // polls of the panel's grid barrier before it gives up and fails the
    // factorisation (0: the kernel's default, about a second); a negative
    // value -k makes every barrier give up after k polls, which exercises
    // that error path with every kernel still exiting normally
    long pv_spin_limit = 0;

    // FP64 system (api_f64.hip): matrix, kept original, b, x, transposed
    // left strips (slots of 64 x n) and the residual partials
    double* dA64 = nullptr;
    double* dA064 = nullptr;
    double* dB64s = nullptr;
    double* dX64s = nullptr;
    double* dLt64 = nullptr;
    double* dWork64 = nullptr;
    double* dOut64 = nullptr;
    size_t capA64 = 0, capA064 = 0, capV64 = 0, capLt64 = 0, capWork64 = 0;
    uint64_t n64 = 0, lda64 = 0;
    // Steps per FP64 bulk update (K = depth_f64 * 64); f64_linear: one strip
    // slot per step (pipelined schedule) instead of 2 * depth_f64 reused.
    // At n = 32768, depth 8 with tail groups of 3 measured 472-477 ms
    // against 482-488 for depth 6 (tail 3), 478-479 for depth 8 (tail 4).
    int depth_f64 = 8;
    bool f64_linear = false;
    bool have_orig64 = false;
    bool factored64 = false;

    // roofline accounting of the trailing update
    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    size_t ev_used = 0;
    double prof_flops = 0.0;
    double prof_bytes = 0.0;  // algorithmic HBM bytes: C read and written once, operands once
};

struct GeslWs;  // kernels.hpp

namespace hplx {

hpl_status fail(hpl_ctx* c, hpl_status st, const char* fmt, ...);

// Streams and events of a context (default priority, non-blocking).
hpl_status init_streams(hpl_ctx* c);
hpl_status init_pipeline_streams(hpl_ctx* c);

// Device bytes of the factorisation workspaces of an n x n system at block B
// with the given trailing-update precision and update depth.
struct WsPlan {
    size_t lt, dt, inner, lp, sc, status;
    size_t inv;   // per-block 32 x 32 inverses (kept with the factor)
    size_t binv;  // full diagonal-block inverses for the substitution (kept with the factor)
    size_t tinv;  // 16-bit factorisations: two steps' inverses for the strip solves
};
// `linear`: one strip slot per step (the pipelined schedule) instead of
// 2 * depth reused slots.
WsPlan plan_factor_ws(uint64_t n, uint32_t B, int prec, int depth, bool linear = false);
// Device bytes of the refinement buffers for nrhs columns of order n.
struct IrPlan {
    size_t work, r, xs, out;
};
IrPlan plan_ir(uint64_t n, int nrhs);

// Classic refinement of nrhs columns with the resident factor: dB (FP64,
// column j at dB + j * n) -> dX.  res[j] receives column j's result (its
// gefa_s / total_s are left 0).  Returns HPL_ERR_NOT_CONVERGED when a column
// misses tol.
hpl_status ir_solve(hpl_ctx* c, int nrhs, const double* dB, double* dX, int max_iter, double tol,
                    hpl_ir_result* res);

uint64_t device_lda(uint64_t n);

// Group boundaries, in blocks, of the pipelined schedule over N blocks of B:
// b[0] = 0 < b[1] < ... = N, `depth` blocks per group until the order left
// is at most the tail threshold, then the tail depth (the context's tail_m and
// tail_depth, or their defaults: half the order, and half the depth but at
// most tail_cap).  With deep_head, groups before that hold the head depth
// while the order left exceeds the head threshold (head_m, head_depth).
std::vector<int64_t> group_bounds(const hpl_ctx* c, int64_t N, int64_t B, int64_t depth,
                                  int64_t tail_cap = 16, bool deep_head = false);

// The full diagonal-block inverses of the resident factor (computed once per
// factorisation, for blocks up to 256); dBInv stays null for larger blocks.
hpl_status ensure_block_inverses(hpl_ctx* c);

// Workspace of the single-launch substitution of one right-hand side with the
// resident factor (in *ws), or null where the per-step launches solve instead
// (external mode, no full inverses, or a block the single launch does not
// take).  gesl_check reads its status after the solve (synchronises s0).
const GeslWs* gesl_workspace(hpl_ctx* c, GeslWs* ws);
hpl_status gesl_check(hpl_ctx* c, const GeslWs* ws);

// Device buffer of at least `bytes` (grown with hipMalloc in the context's own
// mode, an error in external mode).
hpl_status ensure_mem(hpl_ctx* c, void** p, size_t* cap, size_t bytes);
// Factorisation workspaces of the resident system at block B.
hpl_status alloc_factor_workspace(hpl_ctx* c, uint32_t B);
// Partial-pivoting factorisation of the resident matrix (FP32, B <= 256) on
// stream s0, leaving the factor, dPiv, dPerm, dDiagU and the GESL inverses.
hpl_status gefa_pivot(hpl_ctx* c, uint32_t B);
// Rewrite a downloaded pivoted factor (host, row-major, lda) into gefa_ref's
// layout: u_kk on the diagonal, each column's multipliers in the row order of
// its own step.
hpl_status pivot_factor_to_host(hpl_ctx* c, float* a, uint64_t lda);

}  // namespace hplx
