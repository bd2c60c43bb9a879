/*
 * hpl_hip.h -- C ABI of libhpl_hip.so, the MI355X (gfx950) execution backend
 * of the LINPACK benchmark: blocked LU factorisation without pivoting (GEFA)
 * followed by forward and back substitution (GESL).
 *
 * Matrix convention (identical to the reference): the n x n matrix is stored
 * row-major, a[row*lda + col], and is factorised as the column-major matrix
 * M = A^T, M(i,j) = a[j*lda + i].  The system solved is M x = b.  After GEFA
 * the buffer holds the LINPACK dgefa encoding without pivoting: the
 * strictly-lower part of M holds the negated multipliers -l_ik, the
 * strictly-upper part holds u_kj and the diagonal holds -1/u_kk
 * (reference: LINPACK/src/host/linpack_data.cpp:166-229,
 *  LINPACK/src/device/hpl_torus_PCIE.cl:62-71).
 *
 * Errors: every call returns hpl_status (HPL_SUCCESS == 0) and never throws;
 * hpl_last_error() gives the message of the last failure on a context.  This
 * replaces the reference's ASSERT_CL -> OpenClException convention
 * (shared/include/setup/fpga_setup.hpp:125, shared/setup/fpga_setup.cpp:112-122);
 * the C++ host maps a non-zero status to an exception.
 *
 * Threading: a context is not thread-safe; use one context per host thread.
 * The caller owns every host buffer; the context owns every device buffer and
 * frees them in hpl_destroy().
 */
#ifndef HPL_HIP_H_
#define HPL_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum hpl_status {
    HPL_SUCCESS = 0,
    HPL_ERR_INVALID_ARG = 1,    /* bad size, block size, pointer or mode      */
    HPL_ERR_HIP = 2,            /* a HIP runtime call failed                   */
    HPL_ERR_OUT_OF_MEMORY = 3,  /* device allocation failed                    */
    HPL_ERR_ZERO_PIVOT = 4,     /* zero or non-finite pivot (no pivoting)      */
    HPL_ERR_BUDGET = 5,         /* request exceeds the HBM admission budget    */
    HPL_ERR_STATE = 6,          /* call out of order (e.g. solve before factor)*/
    HPL_ERR_NOT_CONVERGED = 7,  /* iterative refinement missed its tolerance   */
    HPL_ERR_UNSUPPORTED = 8,    /* valid request this build does not implement */
    HPL_ERR_NO_DEVICE = 9       /* no HIP device, or device index out of range */
} hpl_status;

/* Input precision of the trailing-update GEMM.  Accumulation is always FP32;
 * the panel LU and the triangular solves always run in FP32. */
typedef enum hpl_prec {
    HPL_PREC_FP32 = 0,
    HPL_PREC_BF16 = 1,
    HPL_PREC_FP16 = 2,
    HPL_PREC_FP8 = 3 /* OCP E4M3 with one power-of-two FP32 scale per block of
                        32 rows (columns) x B of each panel operand */
} hpl_prec;

/* Wall times of the last call, measured with HIP events on the library's
 * stream (seconds).  Fields a call does not touch are left unchanged. */
typedef struct hpl_timing {
    double gefa_s;
    double gesl_s;
    double h2d_s;
    double d2h_s;
} hpl_timing;

typedef struct hpl_ctx hpl_ctx;

/* ---- context ------------------------------------------------------------ */

/* Select device `device` (>= 0) or, for device < 0, device (rank % count) as
 * the reference's round-robin selection does (shared/setup/fpga_setup.cpp:291-293),
 * with rank taken from 0.  Replaces fpga_setup::selectFPGADevice +
 * fpga_setup::fpgaSetup (shared/setup/fpga_setup.cpp:132-181,200-312). */
hpl_status hpl_create(int device, hpl_ctx** ctx);
void hpl_destroy(hpl_ctx* ctx);
const char* hpl_last_error(const hpl_ctx* ctx);
const char* hpl_status_string(hpl_status status);
/* Device name as reported by the runtime (ExecutionSettings::getDeviceName). */
hpl_status hpl_device_name(hpl_ctx* ctx, char* buf, size_t len);
/* Library build string (version and target). */
const char* hpl_version(void);

/* ---- host-buffer path: the drop-in for execution::pcie::calculate -------
 * (LINPACK/src/host/execution_types/execution_pcie.hpp:53-728).
 * hpl_gefa uploads a, factorises it on the device and writes the factor back
 * into a (in place, dgefa encoding), as calculate() leaves data.A
 * (execution_pcie.hpp:695-703).  block must be a power of two in [32, 512]
 * dividing n (blocks above 128 factor their diagonal block by a 2 x 2
 * recursion; the pivoted calls below take blocks up to 256, and the
 * substitution applies whole diagonal-block inverses only up to 256, above
 * it the 32 x 32 sequential solves).  hpl_gesl solves M x = b with a factor
 * produced by hpl_gefa (host buffer lu) and overwrites b with x, as gesl_ref_nopvt does
 * (LINPACK/src/host/linpack_data.cpp:199-229). */
hpl_status hpl_gefa(hpl_ctx* ctx, float* a, uint64_t n, uint64_t lda, uint32_t block,
                    hpl_timing* timing);
hpl_status hpl_gesl(hpl_ctx* ctx, const float* lu, float* b, uint64_t n, uint64_t lda,
                    hpl_timing* timing);

/* ---- partial pivoting (the reference's --uniform path) -------------------
 * LU with partial pivoting by rows of M as gefa_ref / gesl_ref
 * (LINPACK/src/host/linpack_data.cpp:72-156, chosen by --uniform,
 * linpack_benchmark.hpp:69,440-447): column k takes the first row of largest
 * |M(i,k)|, i >= k.  FP32 trailing update; block a power of two in [32, 256]
 * dividing n.
 * hpl_gefa_pvt replaces gefa_ref: it factorises a in place into gefa_ref's
 * layout (the multipliers of column k in the row order of step k, u_kk on the
 * diagonal) and writes ipvt (ipvt[k] = the row exchanged with row k at step
 * k, ipvt[n-1] = n-1).  hpl_gesl_pvt replaces gesl_ref: it solves M x = b with
 * such a factor and overwrites b with x.  A zero or non-finite pivot returns
 * HPL_ERR_ZERO_PIVOT (gefa_ref divides by it). */
hpl_status hpl_gefa_pvt(hpl_ctx* ctx, float* a, uint64_t n, uint64_t lda, uint32_t block,
                        int32_t* ipvt, hpl_timing* timing);
hpl_status hpl_gesl_pvt(hpl_ctx* ctx, const float* lu, float* b, const int32_t* ipvt, uint64_t n,
                        uint64_t lda, hpl_timing* timing);
/* With pivoting on (default off), hpl_gefa_dev factorises P M = L U (FP32
 * trailing update only), hpl_gesl_dev / hpl_solve_ir / hpl_solve_gmres /
 * hpl_solve_rhs64 solve with that factor, and hpl_download returns it in
 * gefa_ref's layout.  No separate reference counterpart: it selects the
 * gefa_ref path of the device-resident calls below. */
hpl_status hpl_set_pivoting(hpl_ctx* ctx, int enable);
/* ipvt of the resident pivoted factor (n entries, gefa_ref's meaning). */
hpl_status hpl_get_pivots(hpl_ctx* ctx, int32_t* ipvt);

/* ---- FP64: the reference built with DATA_TYPE=double ----------------------
 * The "_DP" build of the reference runs the same LINPACK in double
 * (HOST_DATA_TYPE double: gefa_ref_nopvt / gesl_ref_nopvt,
 * LINPACK/src/host/linpack_data.cpp:166-229; CI builds it, .gitlab-ci.yml:220-228).
 * Same storage and encoding as the FP32 calls; blocks of 64 (n must be a
 * multiple of 64), FP64 MFMA trailing update.  hpl_gefa_f64 / hpl_gesl_f64
 * are the host-buffer calls, as hpl_gefa / hpl_gesl; the device-resident
 * calls keep one FP64 system (with an untouched copy for restore and the
 * residual) beside the FP32 one. */
hpl_status hpl_gefa_f64(hpl_ctx* ctx, double* a, uint64_t n, uint64_t lda, hpl_timing* timing);
hpl_status hpl_gesl_f64(hpl_ctx* ctx, const double* lu, double* b, uint64_t n, uint64_t lda,
                        hpl_timing* timing);
hpl_status hpl_upload_f64(hpl_ctx* ctx, const double* a, uint64_t n, uint64_t lda,
                          const double* b, hpl_timing* timing);
/* The system of hpl_generate(n, mode, seed, d) widened to FP64 (b = M x_true
 * in FP64 from the same values); the FP32 resident system is that of
 * hpl_generate afterwards. */
hpl_status hpl_generate_f64(hpl_ctx* ctx, uint64_t n, int mode, uint64_t seed, double d);
hpl_status hpl_gefa_dev_f64(hpl_ctx* ctx, hpl_timing* timing);
hpl_status hpl_gesl_dev_f64(hpl_ctx* ctx, hpl_timing* timing);
hpl_status hpl_restore_f64(hpl_ctx* ctx);
hpl_status hpl_download_f64(hpl_ctx* ctx, double* a, uint64_t lda, double* x, hpl_timing* timing);
/* out = (||b - M x||inf, ||M||inf, ||x||inf, ||b||inf) against the untouched copy. */
hpl_status hpl_residual_f64(hpl_ctx* ctx, double out[4]);

/* ---- device-resident path -------------------------------------------------
 * The context holds one resident system (A, b, x).  Repetitions then time
 * only device work, as the reference times GEFA after the upload
 * (execution_pcie.hpp:123-127,162-163). */

/* Copy a host system to the device (b may be NULL: then b = 0). */
hpl_status hpl_upload(hpl_ctx* ctx, const float* a, uint64_t n, uint64_t lda, const float* b,
                      hpl_timing* timing);
/* Generate a system on the device with the counter-based generator:
 *  mode 0: M(i,j) = U[-0.5,0.5) off the diagonal, M(i,i) += n, b = M*1 (FP64
 *          sums rounded to FP32);
 *  mode 1: M(i,i) = d * sum_{j!=i} |M(i,j)|, x_true = U[-1,1), b = M x_true
 *          in FP64 (kept in FP64 for refinement, rounded copy in FP32);
 *  mode 2: M(i,j) = U[-0.5,0.5) everywhere (not diagonally dominant: the
 *          pivoting workload), b = M*1 as in mode 0.
 * Element (r,c) of the row-major storage is a pure function of (seed, r*n+c),
 * so the result is bit-identical to the CPU generator of the same name. */
hpl_status hpl_generate(hpl_ctx* ctx, uint64_t n, int mode, uint64_t seed, double d);
/* Keep (1) or drop (0) an untouched FP32 copy of A for residuals and
 * refinement.  Must be set before hpl_upload/hpl_generate. */
hpl_status hpl_keep_original(hpl_ctx* ctx, int keep);
/* Factorise the resident matrix in place. */
hpl_status hpl_gefa_dev(hpl_ctx* ctx, uint32_t block, hpl_prec prec, hpl_timing* timing);
/* Solve with the resident factor: x = M^-1 b (b is kept). */
hpl_status hpl_gesl_dev(hpl_ctx* ctx, hpl_timing* timing);
/* Restore the resident matrix from the kept original (device copy). */
hpl_status hpl_restore(hpl_ctx* ctx);
/* Copy back the resident matrix (a may be NULL) and/or solution x (may be
 * NULL).  lda is the host leading dimension. */
hpl_status hpl_download(hpl_ctx* ctx, float* a, uint64_t lda, float* x, hpl_timing* timing);
/* Copy back the resident right-hand side (b, FP32), its FP64 version (b64)
 * and the generator's exact solution (xtrue, FP64); any pointer may be NULL. */
hpl_status hpl_download_vectors(hpl_ctx* ctx, float* b, double* b64, double* xtrue);
/* FP64 norms on the device against the kept original:
 *  out[0] = ||b - M x||inf, out[1] = ||M||inf, out[2] = ||x||inf, out[3] = ||b||inf. */
hpl_status hpl_residual(hpl_ctx* ctx, double out[4]);

/* ---- mixed precision + iterative refinement --------------------------------
 * Classic (GMRES-free) refinement on the resident system, which must have
 * been created with hpl_keep_original(1) and carries an FP64 right-hand side
 * (hpl_generate, or hpl_upload_rhs64): the kept FP32 matrix is factorised with
 * the trailing-update GEMM in `prec` (FP32 accumulation; panel LU and
 * triangular solves in FP32), x0 = F^-1 b in FP32, then
 *   r = b - M x (FP64 accumulation on the device), d = F^-1 (r/||r||) in FP32,
 *   x += ||r|| d (FP64)
 * until the normwise backward error ||r||/(||M|| ||x|| + ||b||) <= tol or
 * max_iter corrections.  No reference counterpart: the reference's only
 * refinement is the test-only FGMRES (LINPACK/src/host/gmres.c:53-232), whose
 * HPL-AI residual (:181-198) this stopping test follows. */
typedef struct hpl_ir_result {
    int iterations;        /* corrections applied */
    int converged;         /* backward error <= tol */
    double backward_error; /* ||b - Mx||inf / (||M||inf ||x||inf + ||b||inf) */
    double residual_inf, norm_a, norm_x, norm_b;
    double gefa_s;         /* factorisation time */
    double total_s;        /* factor + solve + refinement, device time */
} hpl_ir_result;
hpl_status hpl_solve_ir(hpl_ctx* ctx, uint32_t block, hpl_prec prec, int max_iter, double tol,
                        hpl_ir_result* result);
/* As hpl_solve_ir, but each correction M d = r is solved with restarted,
 * left-preconditioned GMRES(restart) in FP64 (restart 1..64), the
 * low-precision factor applied as the preconditioner through the FP32
 * substitution, before x += d.  This follows the reference's test-only
 * FGMRES (LINPACK/src/host/gmres.c:53-232: Arnoldi, Givens rotations of
 * rotmat :29-43, triangular solve of the Hessenberg system) with classical
 * Gram-Schmidt applied twice; it converges where the classic iteration's
 * contraction is too weak (FP8 factors of poorly conditioned matrices).
 * result->iterations counts Arnoldi steps over all max_iter outer steps. */
hpl_status hpl_solve_gmres(hpl_ctx* ctx, uint32_t block, hpl_prec prec, int restart,
                           int max_iter, double tol, hpl_ir_result* result);
/* Solve nrhs (1..16) further right-hand sides with the resident factor (of
 * hpl_solve_ir, or of hpl_gefa_dev on a kept original) and refine each column
 * as above: b and x are n x nrhs FP64, column j at b + j*ldb / x + j*ldx;
 * res[j] receives column j's result (gefa_s and total_s 0).  No reference
 * counterpart: the reference solves one right-hand side per factorisation. */
hpl_status hpl_solve_rhs64(hpl_ctx* ctx, int nrhs, const double* b, uint64_t ldb, double* x,
                           uint64_t ldx, int max_iter, double tol, hpl_ir_result* res);
/* The refined FP64 solution of the last hpl_solve_ir. */
hpl_status hpl_download_solution64(hpl_ctx* ctx, double* x);
/* Replace the resident right-hand side by an FP64 vector (FP32 copy derived). */
hpl_status hpl_upload_rhs64(hpl_ctx* ctx, const double* b);

/* ---- batched solves of many small systems of uneven order ----------------
 * No reference counterpart (the reference solves one matrix per run); every
 * system is solved as above: LU without pivoting in the dgefa encoding, the
 * trailing update in FP16 (or FP32) with FP32 accumulation, substitution in
 * FP32, classic refinement with FP64 residuals until the normwise backward
 * error <= tol_per_n * n.  Orders 1..1024; one persistent work-queue launch per
 * size bucket, all captured into one HIP graph. */
typedef struct hpl_batch hpl_batch;
typedef struct hpl_batch_stats {
    double seconds;                  /* device time of the graph replay        */
    double flops;                    /* sum of 2/3 n^3 + 2 n^2                 */
    double gflops;                   /* flops / seconds / 1e9                  */
    int max_iterations;              /* refinement corrections, worst system   */
    double max_backward_error_per_n; /* max over systems of berr / n           */
    int failed;                      /* systems above the tolerance            */
    int zero_pivots;                 /* systems that met a zero pivot          */
} hpl_batch_stats;
hpl_status hpl_batch_create(hpl_ctx* ctx, const int32_t* n, int nsys, hpl_batch** batch);
void hpl_batch_destroy(hpl_batch* batch);
const char* hpl_batch_last_error(const hpl_batch* batch);
/* Counter-based generator: system s gets M(i,j) = U[-0.5,0.5) (key of seed
 * and s), M(i,i) += n, b = M * 1 in FP64. */
hpl_status hpl_batch_generate(hpl_batch* batch, uint64_t seed);
/* Host system s: row-major a (lda >= n, M = a^T), FP64 right-hand side. */
hpl_status hpl_batch_set_system(hpl_batch* batch, int s, const float* a, uint64_t lda,
                                const double* b);
/* Time each candidate per size bucket -- workgroup sizes 128, 256, 512, each
 * with the persistent factorisation (one workgroup per system) and with the
 * blocked one (one launch per 32-wide step over the bucket's systems) --
 * and keep the fastest; then time the whole solve with that choice and with
 * every bucket blocked, and keep the faster.  threads[4] (may be NULL)
 * receives the choice per bucket: the workgroup size, negated for the
 * blocked factorisation (0: empty bucket).  No counterpart in the reference
 * (one system per run). */
hpl_status hpl_batch_autotune(hpl_batch* batch, hpl_prec prec, int max_iter, double tol_per_n,
                              int* threads);
/* Factorise every bucket blocked (1) or one workgroup per system (0) until
 * the next autotune; both give bit-identical factors. */
hpl_status hpl_batch_set_blocked(hpl_batch* batch, int blocked);
hpl_status hpl_batch_solve(hpl_batch* batch, hpl_prec prec, int max_iter, double tol_per_n,
                           hpl_batch_stats* stats);
/* Refined solution (n doubles), corrections and backward error of system s. */
hpl_status hpl_batch_get(hpl_batch* batch, int s, double* x, int* iterations, double* berr);
/* The padded matrix (npad x npad, npad = n rounded up to 32) and right-hand
 * side (npad doubles) of system s as stored on the device. */
hpl_status hpl_batch_get_system(hpl_batch* batch, int s, float* a, double* b);

/* ---- serving: a request loop over one device -----------------------------
 * No reference counterpart (the reference runs one benchmark matrix per
 * process).  A server admits solves of varying order n and right-hand-side
 * count nrhs as they arrive.  `slots` worker threads each run one solve at a
 * time on their own HIP stream, so requests that arrive while others run are
 * admitted as soon as a slot and memory free up, and their device work
 * overlaps (continuous batching).  Admission is strictly in arrival order and
 * by memory: the server holds `hbm_budget` bytes of device memory, of which
 * the slot workspaces are reserved at creation and the rest holds the factors
 * (original + factor + inverses, about 8 n^2 bytes each).
 *
 * A new system is factorised with the trailing update in `prec` (FP32
 * accumulation, FP32 panel and solves), then every right-hand side is refined
 * in FP64 to a backward error <= tol_per_n * n.  A factor can be kept resident
 * (`keep`) and later requests solve new right-hand sides with it (`factor`)
 * without refactorising.  The block size and padding depend on n alone, so a
 * request's result does not depend on what else is running.
 *
 * Host buffers of a request (a, b, x) must stay valid until hpl_serve_wait
 * returns for its ticket; each ticket is waited for once. */
typedef struct hpl_server hpl_server;
typedef struct hpl_serve_config {
    int device;          /* HIP device                                       */
    int slots;           /* concurrent solves, 1..8                          */
    uint64_t hbm_budget; /* device bytes held by the server                  */
    hpl_prec prec;       /* trailing-update precision of new factorisations */
    int max_iter;        /* refinement corrections per right-hand side       */
    double tol_per_n;    /* backward-error target per unit of n              */
    uint64_t max_n;      /* largest order admitted                           */
    int max_nrhs;        /* largest right-hand-side count, 1..16             */
} hpl_serve_config;
typedef struct hpl_request {
    uint64_t n;       /* order                                                 */
    int nrhs;         /* right-hand sides                                      */
    const float* a;   /* row-major n x n (M = a^T), or NULL                    */
    uint64_t lda;
    uint64_t seed;    /* a == NULL, factor == 0: generate M on the device:
                         U[-0.5,0.5) off the diagonal, M(i,i) += n (the
                         counter-based generator of hpl_generate mode 0)       */
    uint64_t factor;  /* non-zero: solve with this resident factor            */
    int keep;         /* keep the new factor resident (handle in the response)*/
    const double* b;  /* n x nrhs, column j at b + j*ldb                       */
    uint64_t ldb;
    double* x;        /* n x nrhs solution, column j at x + j*ldx              */
    uint64_t ldx;
} hpl_request;
typedef struct hpl_response {
    hpl_status status;
    uint64_t factor;       /* resident handle (kept or reused), else 0        */
    int reused;            /* 1: solved with a resident factor                */
    int iterations;        /* most refinement corrections over the columns    */
    double backward_error; /* largest backward error over the columns         */
    double queue_s;        /* submit -> admitted                              */
    double service_s;      /* admitted -> done                                */
    double latency_s;      /* submit -> done                                  */
    uint32_t block;        /* block size used                                 */
    uint64_t npad;         /* order after identity padding to a block multiple*/
} hpl_response;
typedef struct hpl_serve_stats {
    uint64_t submitted, completed, failed;
    int queued, in_flight, max_in_flight;
    uint64_t resident_factors;
    uint64_t factor_bytes_in_use, peak_factor_bytes, factor_budget_bytes;
    uint64_t workspace_bytes; /* slot workspaces reserved at creation */
} hpl_serve_stats;
/* Defaults: device 0, 4 slots, 200 GB, FP8, 30 corrections, 1e-14, n <= 8192,
 * nrhs <= 16. */
void hpl_serve_config_default(hpl_serve_config* cfg);
hpl_status hpl_serve_create(const hpl_serve_config* cfg, hpl_server** server);
/* Finishes the admitted and queued requests, then frees everything. */
void hpl_serve_destroy(hpl_server* server);
const char* hpl_serve_last_error(const hpl_server* server);
hpl_status hpl_serve_submit(hpl_server* server, const hpl_request* req, uint64_t* ticket);
/* Block until the request is done; returns its status (also in resp). */
hpl_status hpl_serve_wait(hpl_server* server, uint64_t ticket, hpl_response* resp);
/* Give up a resident factor (freed once no queued request uses it). */
hpl_status hpl_serve_release(hpl_server* server, uint64_t factor);
hpl_status hpl_serve_stats_get(hpl_server* server, hpl_serve_stats* stats);

/* ---- kernel timing for roofline accounting -------------------------------
 * No reference counterpart (measurement only; the reference times whole
 * GEFA/GESL calls, linpack_benchmark.hpp:306-324).
 * When enabled, every trailing-update GEMM launch is bracketed by HIP events
 * on the stream it runs on.  hpl_kernel_stats returns the number of launches
 * since the last reset, their summed duration (s) and their algorithmic
 * flops (2*m_rows*m_cols*K per launch, K the steps applied times B). */
hpl_status hpl_set_profiling(hpl_ctx* ctx, int enable);
hpl_status hpl_kernel_stats(hpl_ctx* ctx, uint64_t* launches, double* seconds, double* flops);
/* Time (s) during which at least one of those launches ran: the union of
 * their intervals.  Launches on the two streams overlap, so this is below
 * the summed duration; flops / busy is the device's trailing-update rate. */
hpl_status hpl_kernel_busy(hpl_ctx* ctx, double* seconds);
/* Algorithmic HBM bytes of those launches: per launch the updated region of C
 * read and written once (8 bytes per element) and each operand strip read
 * once (K values per row, in the update's input precision). */
hpl_status hpl_kernel_bytes(hpl_ctx* ctx, double* bytes);

/* ---- substitution that allocates nothing ----------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc -- hpl_gesl's solve on the caller's device memory, with no
 * context: lu is a device pointer to a factor in hpl_gefa's encoding (n x n,
 * row-major, leading dimension lda, a multiple of 4, 16-byte aligned) and x a
 * device pointer to n floats (16-byte aligned) holding b on entry and x on
 * return.  stream is the hipStream_t to enqueue on (NULL: the null stream);
 * the call only enqueues work, so synchronise that stream before reading x.
 * The block is hpl_gesl's (the largest power of two <= 256 dividing n; n must
 * be a multiple of 32) and the result equals hpl_gesl's bit for bit.  The
 * diagonal-block inverses and the solve's state, about n * (416 + 8 * block)
 * bytes, live in one fixed 40 MiB device array per GPU (held until the
 * process ends): a larger system returns HPL_ERR_BUDGET before anything is
 * enqueued (n <= 16896 when 256 divides n).  Every call on a GPU shares that array and overwrites it, so
 * calls that overlap on one GPU (several streams or host threads) race. */
hpl_status hpl_gesl_noalloc(const float* lu, uint64_t n, uint64_t lda, float* x, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HPL_HIP_H_ */
