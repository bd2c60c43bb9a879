// Forward and back substitution with the dgefa-encoded factor
// ("gesl", LINPACK/src/host/linpack_data.cpp:199-229; the reference runs it on
// the host, LINPACK/src/host/linpack_benchmark.hpp:78-201), for one or several
// right-hand sides at once.
//
// With D_q the q-th diagonal block and A the row-major factor:
//   forward  (L y = b, L unit lower, -l stored):  x_e += sum_{p in q} A[p][e] y_p
//   backward (U x = y):                           x_e -= sum_{p in q} A[p][e] x_p
// Both read rows of the factor, so every update is a coalesced GEMV (float4
// loads); with several right-hand sides each factor element read feeds all of
// them, so the factor is streamed once per sweep whatever the count.
//
// One launch per block step.  Workgroup 0 first applies the just-solved block
// to the NEXT diagonal block's entries and then solves that diagonal system;
// the other workgroups apply the same solved block to all further entries in
// parallel.  The diagonal solve works on 32-row sub-blocks: the contribution
// of solved rows is a dot product spread over the workgroup, and the 32 x 32
// triangular system is applied as its inverse (computed by the panel prep);
// with the full inverses of the diagonal blocks (launch_block_inverse, B <=
// 256) the diagonal solve is instead one coalesced GEMV.
//
// Right-hand sides: column r of X at X + r * ldx.  NR is the compile-time
// capacity (1, 4, 16), nr <= NR the count in use.
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace {

constexpr int NT = 256;
constexpr int kDiagPass = 128;  // entries of the next diagonal block per gemv pass

// Entries per off-diagonal workgroup: 4*ncols entries, the B rows split over
// NT/ncols thread groups, so each thread streams B*ncols/NT rows with all of
// its loads in flight (the sweep is bound by load latency, not bandwidth).
template <int NR>
struct gesl_cfg {
    static constexpr int MAXB = NR >= 16 ? 256 : 512;  // largest block
    static constexpr int CH = NR >= 16 ? 256 : NR >= 4 ? 128 : 64;
};

// sum_{p < B} A[(pbase + p) * lda + e] * v[r][p] for the entries
// e = ebase .. ebase + 4*ncols - 1 below elim, for r < nr, added (SIGN=+1) or
// subtracted (SIGN=-1) to X[r][e].  The p range is split over NT / ncols
// thread groups and reduced through red (1024 floats).  Every thread of the
// workgroup must call it.
template <int NR, int SIGN>
__device__ __forceinline__ void block_gemv(const float* __restrict__ A, int64_t lda, int64_t pbase,
                                           int64_t ebase, int ncols, int B, const float* v,
                                           int vstride, int nr, float* red, int64_t elim,
                                           float* __restrict__ X, int64_t ldx,
                                           int64_t eshift = 0) {
    const int t = threadIdx.x;
    const int groups = NT / ncols;
    const int col = t % ncols, g = t / ncols;
    if (NR == 1) nr = 1;
    f32x4 a[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) a[r] = f32x4{0, 0, 0, 0};
    const int64_t e = ebase + 4 * col;
    const bool live = e < elim;  // elim is a multiple of 4
    if (live) {
        const float* src = A + pbase * lda + (e - eshift);  // column e - eshift of A
#pragma unroll 16
        for (int p = g; p < B; p += groups) {
            const f32x4 m = *reinterpret_cast<const f32x4*>(src + (int64_t)p * lda);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                if (r < nr) {
                    const float vp = v[r * vstride + p];
                    a[r][0] = fmaf(m[0], vp, a[r][0]);
                    a[r][1] = fmaf(m[1], vp, a[r][1]);
                    a[r][2] = fmaf(m[2], vp, a[r][2]);
                    a[r][3] = fmaf(m[3], vp, a[r][3]);
                }
            }
        }
    }
    if (groups == 1) {
        if (live) {
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                if (r < nr) {
                    f32x4* dst = reinterpret_cast<f32x4*>(X + r * ldx + e);
                    f32x4 x = *dst;
                    x[0] += SIGN * a[r][0];
                    x[1] += SIGN * a[r][1];
                    x[2] += SIGN * a[r][2];
                    x[3] += SIGN * a[r][3];
                    *dst = x;
                }
            }
        }
        return;
    }
    for (int r = 0; r < nr; ++r) {
        f32x4 ar = a[0];
#pragma unroll
        for (int rr = 1; rr < NR; ++rr)
            if (rr == r) ar = a[rr];
        *reinterpret_cast<f32x4*>(&red[g * 4 * ncols + 4 * col]) = ar;
        __syncthreads();
        if (t < 4 * ncols && ebase + t < elim) {
            float s = 0.0f;
            for (int gg = 0; gg < groups; ++gg) s += red[gg * 4 * ncols + t];
            X[r * ldx + ebase + t] += SIGN * s;
        }
        __syncthreads();
    }
}

// Solve the B x B diagonal system starting at Qoff for nr right-hand sides,
// forward (unit lower) or backward (upper), in place in X.
template <int NR, bool FWD>
__device__ void diag_solve(const float* __restrict__ A, int64_t lda, int64_t Qoff, int B,
                           const float* __restrict__ invT, float* __restrict__ X, int64_t ldx,
                           int nr, float* ybuf, float* part, float* zbuf) {
    constexpr int MAXB = gesl_cfg<NR>::MAXB;
    const int t = threadIdx.x;
    const int NQ = B / 32;
    if (NR == 1) nr = 1;
    for (int idx = t; idx < nr * B; idx += NT) {
        const int r = idx / B, e = idx % B;
        ybuf[r * MAXB + e] = X[r * ldx + Qoff + e];
    }
    __syncthreads();
    for (int step = 0; step < NQ; ++step) {
        const int s = FWD ? step : NQ - 1 - step;
        const int i = t & 31, g = t >> 5;  // 8 groups split the dot product
        // rows already solved: forward p < 32 s ; backward p >= 32 (s+1)
        const int p_lo = FWD ? 0 : 32 * (s + 1);
        const int p_hi = FWD ? 32 * s : B;
        float a[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) a[r] = 0.0f;
        const float* col = A + Qoff * lda + Qoff + 32 * s + i;
#pragma unroll 8
        for (int p = p_lo + g; p < p_hi; p += 8) {
            const float m = col[(int64_t)p * lda];
#pragma unroll
            for (int r = 0; r < NR; ++r)
                if (r < nr) a[r] = fmaf(m, ybuf[r * MAXB + p], a[r]);
        }
#pragma unroll
        for (int r = 0; r < NR; ++r)
            if (r < nr) part[(g * NR + r) * 32 + i] = a[r];
        __syncthreads();
        for (int idx = t; idx < 32 * nr; idx += NT) {
            const int r = idx >> 5, ii = idx & 31;
            float z = 0.0f;
#pragma unroll
            for (int gg = 0; gg < 8; ++gg) z += part[(gg * NR + r) * 32 + ii];
            const float yv = ybuf[r * MAXB + 32 * s + ii];
            zbuf[r * 32 + ii] = FWD ? yv + z : yv - z;
        }
        __syncthreads();
        for (int idx = t; idx < 32 * nr; idx += NT) {
            const int r = idx >> 5, ii = idx & 31;
            // y_s = X z with X^T stored: X[ii][j] = invT[j][ii]
            float y = 0.0f;
            const float* Xs = invT + s * 1024;
#pragma unroll 8
            for (int j = 0; j < 32; ++j) y = fmaf(Xs[j * 32 + ii], zbuf[r * 32 + j], y);
            ybuf[r * MAXB + 32 * s + ii] = y;
        }
        __syncthreads();
    }
    for (int idx = t; idx < nr * B; idx += NT) {
        const int r = idx / B, e = idx % B;
        X[r * ldx + Qoff + e] = ybuf[r * MAXB + e];
    }
}

// Diagonal block at Qoff with its full inverse: X[Qoff..Qoff+B) := W^T z,
// z the current entries, W[p][i] the stored inverse (SIGN -1: negated).
template <int NR, int SIGN>
__device__ void diag_apply(const float* __restrict__ W, int64_t Qoff, int B, float* __restrict__ X,
                           int64_t ldx, int nr, float* zbuf, float* red) {
    constexpr int MAXB = gesl_cfg<NR>::MAXB;
    const int t = threadIdx.x;
    if (NR == 1) nr = 1;
    for (int idx = t; idx < nr * B; idx += NT) {
        const int r = idx / B, e = idx % B;
        zbuf[r * MAXB + e] = X[r * ldx + Qoff + e];
        X[r * ldx + Qoff + e] = 0.0f;
    }
    __syncthreads();
    for (int e0 = 0; e0 < B; e0 += kDiagPass) {
        const int ncols = min(kDiagPass, B - e0) / 4;
        block_gemv<NR, SIGN>(W, B, 0, Qoff + e0, ncols, B, zbuf, MAXB, nr, red, Qoff + B, X, ldx,
                             Qoff);
    }
}

template <int NR>
struct gesl_smem {
    float ybuf[NR * gesl_cfg<NR>::MAXB];  // diagonal system being solved
    float vs[NR * gesl_cfg<NR>::MAXB];    // the solved block applied by this step
    float part[8 * NR * 32];
    float zbuf[NR * 32];
    float red[1024];
};

// Forward step q (q = -1 .. N-2): apply solved block q, solve block q+1.
template <int NR>
__global__ __launch_bounds__(NT) void gesl_fwd_kernel(const float* __restrict__ A, int64_t lda,
                                                      int64_t n, int B, int q,
                                                      const float* __restrict__ inv,
                                                      int64_t inv_stride,
                                                      const float* __restrict__ binv,
                                                      float* __restrict__ X, int64_t ldx, int nr) {
    constexpr int MAXB = gesl_cfg<NR>::MAXB, CH = gesl_cfg<NR>::CH;
    __shared__ gesl_smem<NR> sm;
    const int t = threadIdx.x;
    const int NQ = B / 32;
    const int64_t qoff = (int64_t)q * B;
    if (q >= 0) {
        for (int idx = t; idx < nr * B; idx += NT) {
            const int r = idx / B, e = idx % B;
            sm.vs[r * MAXB + e] = X[r * ldx + qoff + e];
        }
        __syncthreads();
    }
    if (blockIdx.x == 0) {
        const int64_t Q = (int64_t)(q + 1) * B;
        if (q >= 0) {
            // entries of block q+1: x_e += sum_p A[qoff+p][e] y_p
            for (int e0 = 0; e0 < B; e0 += kDiagPass) {
                const int ncols = min(kDiagPass, B - e0) / 4;
                block_gemv<NR, 1>(A, lda, qoff, Q + e0, ncols, B, sm.vs, MAXB, nr, sm.red, n, X,
                                  ldx);
            }
            __syncthreads();
        }
        if (binv)
            diag_apply<NR, 1>(binv + (int64_t)(q + 1) * 2 * B * B, Q, B, X, ldx, nr, sm.ybuf,
                              sm.red);
        else
            diag_solve<NR, true>(A, lda, Q, B, inv + (int64_t)(q + 1) * inv_stride + NQ * 1024, X,
                                 ldx, nr, sm.ybuf, sm.part, sm.zbuf);
        return;
    }
    // entries [(q+2)B, n) in chunks of CH
    const int64_t e0 = (int64_t)(q + 2) * B + (int64_t)(blockIdx.x - 1) * CH;
    block_gemv<NR, 1>(A, lda, qoff, e0, CH / 4, B, sm.vs, MAXB, nr, sm.red, n, X, ldx);
}

// Backward step q (q = N .. 1): apply solved block q (if q < N), solve q-1.
template <int NR>
__global__ __launch_bounds__(NT) void gesl_bwd_kernel(const float* __restrict__ A, int64_t lda,
                                                      int B, int q, int N,
                                                      const float* __restrict__ inv,
                                                      int64_t inv_stride,
                                                      const float* __restrict__ binv,
                                                      float* __restrict__ X, int64_t ldx, int nr) {
    constexpr int MAXB = gesl_cfg<NR>::MAXB, CH = gesl_cfg<NR>::CH;
    __shared__ gesl_smem<NR> sm;
    const int t = threadIdx.x;
    const int NQ = B / 32;
    const int64_t qoff = (int64_t)q * B;
    if (q < N) {
        for (int idx = t; idx < nr * B; idx += NT) {
            const int r = idx / B, e = idx % B;
            sm.vs[r * MAXB + e] = X[r * ldx + qoff + e];
        }
        __syncthreads();
    }
    if (blockIdx.x == 0) {
        const int64_t Q = (int64_t)(q - 1) * B;
        if (q < N) {
            for (int e0 = 0; e0 < B; e0 += kDiagPass) {
                const int ncols = min(kDiagPass, B - e0) / 4;
                block_gemv<NR, -1>(A, lda, qoff, Q + e0, ncols, B, sm.vs, MAXB, nr, sm.red, Q + B,
                                   X, ldx);
            }
            __syncthreads();
        }
        if (binv)
            diag_apply<NR, -1>(binv + (int64_t)(q - 1) * 2 * B * B + (int64_t)B * B, Q, B, X, ldx,
                               nr, sm.ybuf, sm.red);
        else
            diag_solve<NR, false>(A, lda, Q, B, inv + (int64_t)(q - 1) * inv_stride + 2 * NQ * 1024,
                                  X, ldx, nr, sm.ybuf, sm.part, sm.zbuf);
        return;
    }
    // entries [0, (q-1)B)
    const int64_t e0 = (int64_t)(blockIdx.x - 1) * CH;
    const int64_t lim = (int64_t)(q - 1) * B;
    block_gemv<NR, -1>(A, lda, qoff, e0, CH / 4, B, sm.vs, MAXB, nr, sm.red, lim, X, ldx);
}

template <int NR>
hipError_t gesl_sweeps(const float* A, int64_t lda, int64_t n, int B, const float* inv,
                       int64_t inv_stride, const float* binv, float* X, int64_t ldx, int nr,
                       hipStream_t s) {
    constexpr int CH = gesl_cfg<NR>::CH;
    if (B > gesl_cfg<NR>::MAXB) return hipErrorInvalidValue;
    const int N = (int)(n / B);
    for (int q = -1; q <= N - 2; ++q) {
        const int64_t rest = n - (int64_t)(q + 2) * B;
        const int grid = 1 + (q >= 0 ? (int)((rest + CH - 1) / CH) : 0);
        gesl_fwd_kernel<NR><<<grid, NT, 0, s>>>(A, lda, n, B, q, inv, inv_stride, binv, X, ldx, nr);
    }
    for (int q = N; q >= 1; --q) {
        const int64_t rest = (int64_t)(q - 1) * B;
        const int grid = 1 + (q < N ? (int)((rest + CH - 1) / CH) : 0);
        gesl_bwd_kernel<NR><<<grid, NT, 0, s>>>(A, lda, B, q, N, inv, inv_stride, binv, X, ldx,
                                                nr);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Split sweeps (one right-hand side, full diagonal-block inverses): per block
// step a wide launch applies the solved block to every entry still to solve
// -- the next diagonal block included, spread over workgroups of CH entries
// -- and then one 1024-thread workgroup applies the next block's inverse, each
// thread streaming its rows in a single batch of loads.  The fused sweep
// kernels above leave workgroup 0 four dependent GEMV passes per step (the
// next block's update, then its inverse, each over two passes of 256
// threads); this puts two short launches on the critical path instead.
// ---------------------------------------------------------------------------
constexpr int NTD = 1024;

// x_e += SIGN sum_p A[qoff + p][e] X[qoff + p] for e in [e_begin, e_end)
template <int SIGN>
__global__ __launch_bounds__(NT) void gesl_apply1_kernel(const float* __restrict__ A, int64_t lda,
                                                        int B, int64_t qoff, int64_t e_begin,
                                                        int64_t e_end, float* __restrict__ X) {
    constexpr int CH = gesl_cfg<1>::CH;
    __shared__ float vs[512];
    __shared__ float red[1024];
    const int t = threadIdx.x;
    for (int i = t; i < B; i += NT) vs[i] = X[qoff + i];
    __syncthreads();
    block_gemv<1, SIGN>(A, lda, qoff, e_begin + (int64_t)blockIdx.x * CH, CH / 4, B, vs, 0, 1,
                        red, e_end, X, 0);
}

// X[Qoff + i] := SIGN sum_p W[p][i] X[Qoff + p], W the B x B inverse (row-major)
template <int SIGN>
__global__ __launch_bounds__(NTD) void gesl_diag1_kernel(const float* __restrict__ W, int B,
                                                        int64_t Qoff, float* __restrict__ X) {
    __shared__ float z[512];
    __shared__ float red[4096];
    const int t = threadIdx.x;
    for (int i = t; i < B; i += NTD) z[i] = X[Qoff + i];
    __syncthreads();
    const int cols = B / 4, groups = NTD / cols;  // float4 columns; row groups
    const int c = t % cols, g = t / cols;
    // rows g, g + groups, ..: at most B / groups = B^2 / 4096 <= 16 (B <= 256),
    // all loads in flight before the first is used
    f32x4 w[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int p = g + i * groups;
        w[i] = p < B ? *reinterpret_cast<const f32x4*>(W + (int64_t)p * B + 4 * c)
                     : f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    f32x4 a = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int p = g + i * groups;
        if (p < B) {
            const float zp = z[p];
            a[0] = fmaf(w[i][0], zp, a[0]);
            a[1] = fmaf(w[i][1], zp, a[1]);
            a[2] = fmaf(w[i][2], zp, a[2]);
            a[3] = fmaf(w[i][3], zp, a[3]);
        }
    }
    *reinterpret_cast<f32x4*>(&red[g * B + 4 * c]) = a;  // groups * B = 4096
    __syncthreads();
    if (t < B) {
        float s = 0.0f;
        for (int gg = 0; gg < groups; ++gg) s += red[gg * B + t];
        X[Qoff + t] = SIGN * s;
    }
}

hipError_t gesl_sweeps_split(const float* A, int64_t lda, int64_t n, int B, const float* binv,
                             float* X, hipStream_t s) {
    constexpr int CH = gesl_cfg<1>::CH;
    const int N = (int)(n / B);
    const int64_t BB = (int64_t)B * B;
    // forward: block 0, then per q the update by block q and the solve of q+1
    gesl_diag1_kernel<1><<<1, NTD, 0, s>>>(binv, B, 0, X);
    for (int q = 0; q + 1 < N; ++q) {
        const int64_t e0 = (int64_t)(q + 1) * B;
        gesl_apply1_kernel<1><<<(int)((n - e0 + CH - 1) / CH), NT, 0, s>>>(A, lda, B,
                                                                          (int64_t)q * B, e0, n, X);
        gesl_diag1_kernel<1><<<1, NTD, 0, s>>>(binv + (int64_t)(q + 1) * 2 * BB, B, e0, X);
    }
    // backward: block N-1, then per q the update of the entries before block q
    // and the solve of q-1 (the stored U inverses are negated)
    gesl_diag1_kernel<-1><<<1, NTD, 0, s>>>(binv + (int64_t)(N - 1) * 2 * BB + BB, B,
                                            (int64_t)(N - 1) * B, X);
    for (int q = N - 1; q >= 1; --q) {
        const int64_t e1 = (int64_t)q * B;
        gesl_apply1_kernel<-1><<<(int)((e1 + CH - 1) / CH), NT, 0, s>>>(A, lda, B, e1, 0, e1, X);
        gesl_diag1_kernel<-1><<<1, NTD, 0, s>>>(binv + (int64_t)(q - 1) * 2 * BB + BB, B,
                                                (int64_t)(q - 1) * B, X);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_gesl_multi(const float* A, int64_t lda, int64_t n, int B, const float* inv,
                             int64_t inv_stride, float* X, int64_t ldx, int nrhs, hipStream_t s,
                             const float* binv) {
    if (B > 512 || B < 32 || B % 32 || n % B || nrhs < 1 || (nrhs > 1 && (ldx < n || ldx % 4)))
        return hipErrorInvalidValue;
    for (int r0 = 0; r0 < nrhs; r0 += 16) {
        const int nr = std::min(16, nrhs - r0);
        float* Xr = X + (int64_t)r0 * ldx;
        hipError_t e;
        if (nr == 1 && binv && B <= 256)
            e = gesl_sweeps_split(A, lda, n, B, binv, Xr, s);
        else if (nr == 1)
            e = gesl_sweeps<1>(A, lda, n, B, inv, inv_stride, binv, Xr, ldx, 1, s);
        else if (nr <= 4)
            e = gesl_sweeps<4>(A, lda, n, B, inv, inv_stride, binv, Xr, ldx, nr, s);
        else
            e = gesl_sweeps<16>(A, lda, n, B, inv, inv_stride, binv, Xr, ldx, nr, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_gesl(const float* A, int64_t lda, int64_t n, int B, const float* inv,
                       int64_t inv_stride, float* x, hipStream_t s, const float* binv) {
    return launch_gesl_multi(A, lda, n, B, inv, inv_stride, x, n, 1, s, binv);
}
