// Forward and back substitution with the dgefa-encoded factor
// ("gesl", LINPACK/src/host/linpack_data.cpp:199-229; the reference runs it on
// the host, LINPACK/src/host/linpack_benchmark.hpp:78-201).
//
// With D_q the q-th diagonal block and A the row-major factor:
//   forward  (L y = b, L unit lower, -l stored):  x_e += sum_{p in q} A[p][e] y_p
//   backward (U x = y):                           x_e -= sum_{p in q} A[p][e] x_p
// Both read rows of the factor, so every update is a coalesced GEMV (float4
// loads, the p range split over thread groups and reduced in LDS).
//
// One launch per block step.  Workgroup 0 first applies the just-solved block
// to the NEXT diagonal block's entries and then solves that diagonal system;
// the other workgroups apply the same solved block to all further entries in
// parallel.  The diagonal solve works on 32-row sub-blocks: the contribution
// of solved rows is a dot product spread over the workgroup, and the 32 x 32
// triangular system is applied as its inverse (computed by the panel prep).
#include "common.hpp"
#include "kernels.hpp"

namespace {

constexpr int NT = 256;
constexpr int CHUNK = 256;  // entries per non-diagonal workgroup (64 float4 columns)

// red[0][e] (e < 4*ncols) = sum_{p < B} A[(pbase + p) * lda + ebase + e] * v[p] for the
// entries ebase .. ebase + 4*ncols - 1 below elim (ncols <= 64).
__device__ __forceinline__ void block_gemv(const float* __restrict__ A, int64_t lda, int64_t pbase,
                                           int64_t ebase, int ncols, int B, const float* v,
                                           float (*red)[4 * 64], int64_t elim) {
    const int t = threadIdx.x;
    const int groups = min(NT / ncols, NT / 16);  // p-groups (rows of red)
    const int col = t % ncols, g = t / ncols;
    f32x4 a = {0, 0, 0, 0};
    const int64_t e = ebase + 4 * col;
    if (g < groups && e < elim) {
        const float* src = A + pbase * lda + e;
#pragma unroll 8
        for (int p = g; p < B; p += groups) {
            const f32x4 m = *reinterpret_cast<const f32x4*>(src + (int64_t)p * lda);
            const float vp = v[p];
            a[0] = fmaf(m[0], vp, a[0]);
            a[1] = fmaf(m[1], vp, a[1]);
            a[2] = fmaf(m[2], vp, a[2]);
            a[3] = fmaf(m[3], vp, a[3]);
        }
    }
    if (g < groups) *reinterpret_cast<f32x4*>(&red[g][4 * col]) = a;
    __syncthreads();
    if (t < 4 * ncols) {
        float s = 0.0f;
        for (int gg = 0; gg < groups; ++gg) s += red[gg][t];
        red[0][t] = s;
    }
    __syncthreads();
}

// Solve the B x B diagonal system starting at Qoff in x (in place), forward
// (unit lower) or backward (upper).  256 threads.
template <bool FWD>
__device__ void diag_solve(const float* __restrict__ A, int64_t lda, int64_t Qoff, int B,
                           const float* __restrict__ invT, float* __restrict__ x, float* ybuf,
                           float (*part)[32]) {
    const int t = threadIdx.x;
    const int NQ = B / 32;
    for (int e = t; e < B; e += NT) ybuf[e] = x[Qoff + e];
    __syncthreads();
    for (int step = 0; step < NQ; ++step) {
        const int s = FWD ? step : NQ - 1 - step;
        const int i = t & 31, g = t >> 5;  // 8 groups split the dot product
        // rows already solved: forward p < 32 s ; backward p >= 32 (s+1)
        const int p_lo = FWD ? 0 : 32 * (s + 1);
        const int p_hi = FWD ? 32 * s : B;
        float a = 0.0f;
        const float* col = A + Qoff * lda + Qoff + 32 * s + i;
#pragma unroll 4
        for (int p = p_lo + g; p < p_hi; p += 8) a = fmaf(col[(int64_t)p * lda], ybuf[p], a);
        part[g][i] = a;
        __syncthreads();
        if (t < 32) {
            float z = 0.0f;
#pragma unroll
            for (int gg = 0; gg < 8; ++gg) z += part[gg][t];
            z = FWD ? ybuf[32 * s + t] + z : ybuf[32 * s + t] - z;
            part[0][t] = z;
        }
        __syncthreads();
        if (t < 32) {
            // y_s = X z with X^T stored: X[t][j] = invT[j][t]
            float y = 0.0f;
            const float* X = invT + s * 1024;
#pragma unroll 8
            for (int j = 0; j < 32; ++j) y = fmaf(X[j * 32 + t], part[0][j], y);
            ybuf[32 * s + t] = y;
        }
        __syncthreads();
    }
    for (int e = t; e < B; e += NT) x[Qoff + e] = ybuf[e];
}

// Forward step q (q = -1 .. N-2): apply solved block q, solve block q+1.
__global__ __launch_bounds__(NT) void gesl_fwd_kernel(const float* __restrict__ A, int64_t lda,
                                                      int64_t n, int B, int q,
                                                      const float* __restrict__ inv,
                                                      int64_t inv_stride, float* __restrict__ x) {
    __shared__ float ybuf[512];
    __shared__ float part[8][32];
    __shared__ float ys[512];
    __shared__ float red[NT / 16][4 * 64];
    const int t = threadIdx.x;
    const int NQ = B / 32;
    const int64_t qoff = (int64_t)q * B;
    if (q >= 0) {
        for (int e = t; e < B; e += NT) ys[e] = x[qoff + e];
        __syncthreads();
    }
    if (blockIdx.x == 0) {
        const int64_t Q = (int64_t)(q + 1) * B;
        if (q >= 0) {
            // entries of block q+1: x_e += sum_p A[qoff+p][e] y_p
            for (int e0 = 0; e0 < B; e0 += 256) {
                const int ncols = min(64, (B - e0) / 4);
                block_gemv(A, lda, qoff, Q + e0, ncols, B, ys, red, n);
                if (t < 4 * ncols) x[Q + e0 + t] += red[0][t];
                __syncthreads();
            }
        }
        diag_solve<true>(A, lda, Q, B, inv + (int64_t)(q + 1) * inv_stride + NQ * 1024, x, ybuf,
                         part);
        return;
    }
    // entries [(q+2)B, n) in chunks of CHUNK
    const int64_t e0 = (int64_t)(q + 2) * B + (int64_t)(blockIdx.x - 1) * CHUNK;
    block_gemv(A, lda, qoff, e0, 64, B, ys, red, n);
    if (t < CHUNK && e0 + t < n) x[e0 + t] += red[0][t];
}

// Backward step q (q = N .. 1): apply solved block q (if q < N), solve q-1.
__global__ __launch_bounds__(NT) void gesl_bwd_kernel(const float* __restrict__ A, int64_t lda,
                                                      int B, int q, int N,
                                                      const float* __restrict__ inv,
                                                      int64_t inv_stride, float* __restrict__ x) {
    __shared__ float ybuf[512];
    __shared__ float part[8][32];
    __shared__ float xs[512];
    __shared__ float red[NT / 16][4 * 64];
    const int t = threadIdx.x;
    const int NQ = B / 32;
    const int64_t qoff = (int64_t)q * B;
    if (q < N) {
        for (int e = t; e < B; e += NT) xs[e] = x[qoff + e];
        __syncthreads();
    }
    if (blockIdx.x == 0) {
        const int64_t Q = (int64_t)(q - 1) * B;
        if (q < N) {
            for (int e0 = 0; e0 < B; e0 += 256) {
                const int ncols = min(64, (B - e0) / 4);
                block_gemv(A, lda, qoff, Q + e0, ncols, B, xs, red, Q + B);
                if (t < 4 * ncols) x[Q + e0 + t] -= red[0][t];
                __syncthreads();
            }
        }
        diag_solve<false>(A, lda, Q, B, inv + (int64_t)(q - 1) * inv_stride + 2 * NQ * 1024, x,
                          ybuf, part);
        return;
    }
    // entries [0, (q-1)B)
    const int64_t e0 = (int64_t)(blockIdx.x - 1) * CHUNK;
    const int64_t lim = (int64_t)(q - 1) * B;
    block_gemv(A, lda, qoff, e0, 64, B, xs, red, lim);
    if (t < CHUNK && e0 + t < lim) x[e0 + t] -= red[0][t];
}

}  // namespace

hipError_t launch_gesl(const float* A, int64_t lda, int64_t n, int B, const float* inv,
                       int64_t inv_stride, float* x, hipStream_t s) {
    if (B > 512 || B % 32) return hipErrorInvalidValue;
    const int N = (int)(n / B);
    for (int q = -1; q <= N - 2; ++q) {
        const int64_t rest = n - (int64_t)(q + 2) * B;
        const int grid = 1 + (q >= 0 ? (int)((rest + CHUNK - 1) / CHUNK) : 0);
        gesl_fwd_kernel<<<grid, NT, 0, s>>>(A, lda, n, B, q, inv, inv_stride, x);
    }
    for (int q = N; q >= 1; --q) {
        const int64_t rest = (int64_t)(q - 1) * B;
        const int grid = 1 + (q < N ? (int)((rest + CHUNK - 1) / CHUNK) : 0);
        gesl_bwd_kernel<<<grid, NT, 0, s>>>(A, lda, B, q, N, inv, inv_stride, x);
    }
    return hipGetLastError();
}
