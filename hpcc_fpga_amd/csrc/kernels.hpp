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
hipError_t launch_trsm(float* A, int64_t lda, int64_t off, int64_t m, int B, const float* DT,
                       const float* inv, float* Lt, int64_t ldt, hipStream_t s);

// trailing update (gemm.hip)
hipError_t launch_gemm_update_f32(float* C, int64_t ldc, const float* Lt, int64_t ldl,
                                  const float* T, int64_t ldt, int M, int N, int K,
                                  hipStream_t s);

// substitution (gesl.hip)
hipError_t launch_gesl(const float* A, int64_t lda, int64_t n, int B, const float* inv,
                       int64_t inv_stride, float* x, hipStream_t s);

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
