// Launch wrappers of the gfx950 kernels (host-side declarations).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// diagonal-block LU (lu.hip); status[0] is raised to 1 on a zero/non-finite pivot
hipError_t launch_lu_diag(float* A, int64_t lda, int64_t off, int B, int* status, hipStream_t s);

// panel solves (trsm.hip)
hipError_t launch_prep(const float* A, int64_t lda, int64_t off, int B, float* DT, float* inv,
                       hipStream_t s);
hipError_t launch_prep_all(const float* A, int64_t lda, int B, int nblocks, float* inv,
                           int64_t inv_stride, hipStream_t s);
// lp_prec 0: FP32 only; 1/2: also write BF16/FP16 k-contiguous copies of the
// top strip (lpT[c][k]) and left strip (lpL[r][k]), leading dimension ldlp;
// 3: FP8 E4M3 copies with the scale of vectors 32 g .. 32 g + 31 at scT[g] /
// scL[g] (vectors counted from the strip start).
hipError_t launch_trsm(float* A, int64_t lda, int64_t off, int64_t m, int B, const float* DT,
                       const float* inv, float* Lt, int64_t ldt, hipStream_t s, int lp_prec = 0,
                       void* lpT = nullptr, void* lpL = nullptr, int64_t ldlp = 0,
                       float* scT = nullptr, float* scL = nullptr);

// Strip solves of the BF16 factorisation through the diagonal block's full
// inverses (B <= 256): launch_tinv_prepare turns the block at (off, off) --
// factorised, with its 32 x 32 inverses in inv (launch_prep) -- into both
// inverses (OutL then OutU as launch_block_inverse writes them: at `out`, the
// block's slot of the substitution's inverses, or, when out is null, at the
// start of `tinv`) and their MFMA operand copies in `tinv` (tinv_floats(B)
// floats, the copies after the first 2 B^2);
// launch_trsm_inv then solves both strips as one product each, y = G v, on
// BF16 MFMA with the operands split into high and low BF16 halves (three
// products: about 2^-17 relative), writing the FP32 strips and their
// k-contiguous 16-bit copies like launch_trsm (lp_prec 1: BF16; 2: FP16).
int64_t tinv_floats(int B);
hipError_t launch_tinv_prepare(const float* A, int64_t lda, int64_t off, int B, const float* inv,
                               float* tinv, float* out, hipStream_t s);
hipError_t launch_trsm_inv(float* A, int64_t lda, int64_t off, int64_t m, int B,
                           const float* tinv, hipStream_t s, int lp_prec, void* lpT, void* lpL,
                           int64_t ldlp);

// the top strip alone (rows off..off+B-1, columns off+B..off+B+m-1), FP32, B <= 256
hipError_t launch_trsm_top(float* A, int64_t lda, int64_t off, int64_t m, int B, const float* DT,
                           const float* inv, hipStream_t s);

// trailing update (gemm.hip)
hipError_t launch_gemm_update_f32(float* C, int64_t ldc, const float* Lt, int64_t ldl,
                                  const float* T, int64_t ldt, int M, int N, int K,
                                  hipStream_t s);
// the same over two regions in one launch (same leading dimensions and K)
hipError_t launch_gemm_update_f32_2(float* C, int64_t ldc, const float* Lt, int64_t ldl,
                                    const float* T, int64_t ldt, int M, int N, float* C2,
                                    const float* Lt2, const float* T2, int M2, int N2, int K,
                                    hipStream_t s);

// Second region of a low-precision trailing-update launch (same leading
// dimensions, slots and K as the first; FP8 scales offset like L and T).
struct LpRegion2 {
    float* C;
    const void* L;
    const void* T;
    const float* scL;
    const float* scT;
    int M, N;
};

// mixed-precision trailing update (gemm_lp.hip): prec 1 = BF16, 2 = FP16.
// max_groups > 0 caps the workgroups of the 256 x 256-tile kernel (one per CU,
// persistent over the tiles), leaving the other CUs to concurrent launches.
hipError_t launch_gemm_update_lp(int prec, float* C, int64_t ldc, const void* L, int64_t l_ld,
                                 int64_t l_slot, const void* Tt, int64_t t_ld, int64_t t_slot,
                                 int M, int N, int K, int Bblk, hipStream_t s,
                                 const LpRegion2* r2 = nullptr, int max_groups = 0);

// FP8 trailing update (gemm_lp.hip): C[r][c] += sum over the K / Bblk steps p
// of scL[p][r/32] * scT[p][c/32] * sum_k L[r][k] T[c][k]  (E4M3 codes, FP32
// accumulation); step p's scales at scL + p * sc_slot, scT + p * sc_slot
hipError_t launch_gemm_update_fp8(float* C, int64_t ldc, const void* L, int64_t l_ld,
                                  int64_t l_slot, const void* Tt, int64_t t_ld, int64_t t_slot,
                                  const float* scL, const float* scT, int64_t sc_slot, int M,
                                  int N, int K, int Bblk, hipStream_t s,
                                  const LpRegion2* r2 = nullptr);

// full inverses of the diagonal blocks of a factor for the substitution
// (trsm.hip): per block k, out + k * 2 * B * B holds OutL (row r = L^-1 e_r)
// then OutU (OutU[i][c] = -(U^-1)(c, i)); B <= 256
hipError_t launch_block_inverse(const float* A, int64_t lda, int64_t n, int B, const float* inv,
                                int64_t inv_stride, float* out, hipStream_t s);

// substitution (gesl.hip)
// Workspace of the single-launch substitution of one right-hand side: the
// forward and backward slice sums zf, zb and the forward solution y (n floats
// each), two flags per slice and a status word (bit 2: a bounded wait gave
// up).  Every launch clears the flags and the status itself.
struct GeslWs {
    float* pf;  // forward partial solutions, B per slice
    float* pb;  // backward partial solutions, B per slice
    unsigned* flags;
    int* status;
};
size_t gesl_ws_bytes(int64_t n);
GeslWs gesl_ws_carve(void* base, int64_t n);  // base: gesl_ws_bytes(n), 16-byte aligned
// workgroups of the single launch for order n at block B; 0: it does not
// apply (block above 256, or too many slices per workgroup)
int64_t gesl_persist_groups(int64_t n, int B);
// binv: the full diagonal-block inverses of launch_block_inverse (B <= 256),
// or null for the sequential diagonal solve with the 32 x 32 inverses.  With
// binv, one right-hand side and a workspace, the substitution is a single
// persistent launch; otherwise two launches per block step.
hipError_t launch_gesl(const float* A, int64_t lda, int64_t n, int B, const float* inv,
                       int64_t inv_stride, float* x, hipStream_t s, const float* binv = nullptr,
                       const GeslWs* ws = nullptr);
// nrhs right-hand sides, column r at X + r * ldx (ldx a multiple of 4, >= n);
// blocks up to 256 for more than 4 right-hand sides
hipError_t launch_gesl_multi(const float* A, int64_t lda, int64_t n, int B, const float* inv,
                             int64_t inv_stride, float* X, int64_t ldx, int nrhs, hipStream_t s,
                             const float* binv = nullptr, const GeslWs* ws = nullptr);

// partial pivoting (pivot.hip); X is the transposed matrix X[i][j] = M(i, j)
hipError_t launch_transpose_square(float* A, int64_t ld, int64_t n, hipStream_t s);
// rows per workgroup of the panel kernel at block B; workgroups of a panel
// of m rows
int panel_pivot_rows(int B);
int panel_pivot_groups(int64_t m, int B);
// set the panel kernel up and return how many of its workgroups are co-resident
hipError_t panel_pivot_prepare(int B, int* max_groups);
// factorise panel k0 (m = n - k0 rows of X, B columns) with partial pivoting;
// pub, krow, kslot: the exchange words of the panel's workgroups (G =
// panel_pivot_groups(m, B) <= max_groups; kslot has gslots > G slots per
// parity), carved from panel_exchange_words(n, B) words that
// launch_zero_words clears before each factorisation; writes the panel (-1/u
// on the diagonal), Lt[c][pos] for pos >= k0 + B, diagU/ipvt at k0.., and the
// row of origin src[pos] of positions k0..n-1.  status |= 1: zero or
// non-finite pivot; |= 2: an exchange timed out; |= 4: a pivot key outside
// the panel was read (a broken exchange; the row is left in place).
// spin_limit: polls of a wait before a workgroup gives up (0: about a second;
// < 0: give up after -spin_limit polls on every exchange, the error path).
// A grid of at most num_cu workgroups gets one CU per workgroup.
hipError_t launch_panel_pivot(float* X, int64_t ldx, int64_t n, int64_t k0, int B,
                              unsigned long long* pub, unsigned long long* krow,
                              unsigned long long* kslot, int gslots, float* Lt, int64_t ldl,
                              float* diagU, int* ipvt, int* src, int* status, int max_groups,
                              int num_cu, hipStream_t s, long spin_limit = 0);
// 64-bit words of the panel exchange for order n at block B: pub (2 G (B+1)),
// krow (2 (B+1)), kslot (2 (G+1) slots of 8), G = panel_pivot_groups(n, B)
size_t panel_exchange_words(int64_t n, int B);
hipError_t launch_zero_words(unsigned long long* p, size_t words, hipStream_t s);
// apply panel k0's row moves to the columns outside it and to perm; list:
// 4 B ints, count: one int, Y: 2 B x ldy floats.  Nothing moves when status
// carries a panel failure (bits 2, 4) or a listed position lies outside
// [k0, n) (then status |= 4).
hipError_t launch_row_moves(float* X, int64_t ldx, int64_t n, int64_t k0, int B, const int* ipvt,
                            const int* src, int* list, int* count, float* Y, int64_t ldy,
                            int* perm, int* status, hipStream_t s);
hipError_t launch_iota(int* p, int64_t n, hipStream_t s);
// unit lower triangle of X's diagonal block at k0 into D (B x B)
hipError_t launch_unit_lower(const float* X, int64_t ldx, int64_t k0, int B, float* D,
                             hipStream_t s);
// x[i] = b[perm[i]]
hipError_t launch_gather_f32(float* x, const float* b, const int* perm, int64_t n, hipStream_t s);

// FP64 path (f64.hip): block f64_block() = 64
int f64_block();
hipError_t launch_lu_f64(double* A, int64_t lda, int64_t off, int* status, hipStream_t s);
// top strip (rows off.., columns off+64..off+64+m-1) and left strip (rows
// off+64.., columns off..off+63); the left strip also into Lt[k][v] (ldl)
hipError_t launch_strips_f64(double* A, int64_t lda, int64_t off, int64_t m, double* Lt,
                             int64_t ldl, hipStream_t s);
// C[r][c] += sum_k Lt[k][r] T[k][c] (K a multiple of 16, M and N even)
hipError_t launch_update_f64(double* C, int64_t ldc, const double* Lt, int64_t ldl,
                             const double* T, int64_t ldt, int M, int N, int K, hipStream_t s);
// x := M^-1 x with the factor (gesl_ref_nopvt's order; n a multiple of 64)
hipError_t launch_gesl_f64(const double* A, int64_t lda, int64_t n, double* x, hipStream_t s);
int64_t residual_f64_work(int64_t n);
hipError_t launch_residual_f64(const double* A, int64_t lda, int64_t n, const double* x,
                               const double* b, double* work, double* out4, hipStream_t s);
hipError_t launch_f32_to_f64(const float* src, int64_t lds, double* dst, int64_t ldd, int64_t n,
                             hipStream_t s);

// generation and residuals (gen.hip)
hipError_t launch_generate(float* A, int64_t lda, int64_t n, int mode, uint64_t seed, double d,
                           float* b, double* b64, double* xt, hipStream_t s);
int64_t residual_chunks(int64_t n);
hipError_t launch_residual_f32x(const float* A, int64_t lda, int64_t n, const float* x,
                                const float* b, double* work, double* r, double* out4,
                                hipStream_t s);
hipError_t launch_residual_f64x(const float* A, int64_t lda, int64_t n, const double* x,
                                const double* b, double* work, double* r, double* out4,
                                bool with_norm, hipStream_t s);

// refinement vector helpers (gen.hip): x = (float)(r * scale); x = d*scale (+ x)
hipError_t launch_d2f_scaled(const double* r, float* x, int64_t n, double scale, hipStream_t s);
hipError_t launch_axpy_f2d(double* x, const float* d, int64_t n, double scale, int assign,
                           hipStream_t s);

// several right-hand sides (gen.hip), nrhs <= 16, column r at offset r * ld:
// R = B - M X in FP64 with out[4 r + 0..3] = (||r||, ||M||, ||x||, ||b||)_inf;
// work holds residual_multi_work(n, nrhs) doubles
int64_t residual_multi_work(int64_t n, int nrhs);
hipError_t launch_residual_multi(const float* A, int64_t lda, int64_t n, const double* X,
                                 int64_t ldx, const double* Bv, int64_t ldb, int nrhs,
                                 double* work, double* R, int64_t ldr, double* out,
                                 hipStream_t s);
// Xs[r][i] = (float)(R[r][perm[i]] * scale[r]) (perm null: identity);
// X[r] = scale[r] * D[r] (+ X[r])
hipError_t launch_d2f_multi(const double* R, int64_t ldr, float* Xs, int64_t ldxs, int64_t n,
                            int nrhs, const double* scale, hipStream_t s,
                            const int* perm = nullptr);
hipError_t launch_axpy_multi(double* X, int64_t ldx, const float* D, int64_t ldd, int64_t n,
                             int nrhs, const double* scale, int assign, hipStream_t s);

// identity extension of an n x n system to npad x npad; fill (kernels, no
// memset nodes) (gen.hip)
hipError_t launch_pad_identity(float* A, int64_t lda, int64_t n, int64_t npad, hipStream_t s);
hipError_t launch_fill_f64(double* p, int64_t count, double v, hipStream_t s);

// FP64 Krylov-basis kernels of the GMRES refinement (gmres.hip); work holds
// gmres_work_doubles(n, kmax) doubles for up to kmax basis vectors
int64_t gmres_work_doubles(int64_t n, int kmax);
// y = M x with M(i, j) = A[j][i] (FP32 matrix, FP64 vectors)
hipError_t launch_matvec_f64(const float* A, int64_t lda, int64_t n, const double* x, double* y,
                             double* work, hipStream_t s);
// out[k] = V[k] . w for k < kc (basis vector k at V + k * ldv)
hipError_t launch_multi_dot(const double* V, int64_t ldv, int kc, const double* w, int64_t n,
                            double* out, double* work, hipStream_t s);
// y = alpha y + beta sum_k c[k] V[k]  (c on the device)
hipError_t launch_multi_axpy(double* y, double alpha, double beta, const double* V, int64_t ldv,
                             int kc, const double* c, int64_t n, hipStream_t s);
hipError_t launch_scale_copy(double* dst, const double* src, double a, int64_t n, hipStream_t s);

// batched solves (batch.hip)
struct BatchSystem {
    int64_t a_off;  // matrix offset (elements) in the matrix pool
    int64_t v_off;  // vector offset (elements) in the vector pools
    int n;          // order
    int npad;       // order rounded up to 32 (identity extension)
};
struct BatchResult {
    int iterations;
    int status;  // 1: zero or non-finite pivot
    double backward_error;
};
constexpr size_t kBatchLdsLimit = 160 * 1024;  // LDS per workgroup on gfx950
hipError_t batch_prepare();  // allow kBatchLdsLimit bytes of dynamic LDS
size_t batch_lds_bytes(int maxm, bool f16);
hipError_t launch_batch_generate(const BatchSystem* sys, int nsys, uint64_t seed, float* A,
                                 double* b64, hipStream_t s);
// pre_bad non-null: the systems were factorised by launch_batch_step (their
// zero-pivot flags in pre_bad); only the substitution and refinement run, and
// maxm only sizes their LDS vectors (npad doubles and npad floats)
hipError_t launch_batch_solve(int threads, bool f16, int grid, const BatchSystem* sys,
                              const int* order, int nsys, int* counter, float* A, const float* A0,
                              const double* b64, double* x64, float* xs, double* rs, int max_iter,
                              double tol_per_n, BatchResult* out, int maxm, hipStream_t s,
                              const int* pre_bad = nullptr);
// Step k0 of the blocked factorisation over the first act_diag systems of
// order (all with npad > k0; the first act of them have npad > k0 + 32 and a
// trailing size of at most mmax): diagonal blocks (Dg: 32 x 33 floats per
// system; bad[s]: zero or non-finite pivot, set at k0 = 0, raised after),
// strips (Lg, Tg: sstride >= mmax * 32 elements per system, FP16 or FP32 as
// f16) and trailing updates.
hipError_t launch_batch_step(bool f16, const BatchSystem* sys, const int* order, int k0,
                             int act_diag, int act, int mmax, float* A, float* Dg, int* bad,
                             void* Lg, void* Tg, size_t sstride, hipStream_t s);
// Steps k0 and k0 + 32 together: act_diag systems with npad > k0, act with
// npad > k0 + 32 (trailing size <= mmax after step k0), act2 with npad > k0 +
// 64.  Step k0 updates only the next step's diagonal block and strips; the
// other tiles get both steps' updates in one pass (the trailing matrix is
// read and written once for two steps), bit-identical to two single steps.
// Step k0's strips in Lg / Tg, step k0 + 32's in Lg2 / Tg2.  A0 non-null: at
// k0 = 0 every read of the original elements comes from A0 (which then
// needs no copy into A first).
hipError_t launch_batch_step2(bool f16, const BatchSystem* sys, const int* order, int k0,
                              int act_diag, int act, int act2, int mmax, float* A,
                              const float* A0, float* Dg, int* bad, void* Lg, void* Tg, void* Lg2,
                              void* Tg2, size_t sstride, hipStream_t s);
