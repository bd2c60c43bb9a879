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
#include <cstdlib>

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

// ---------------------------------------------------------------------------
// Single-launch substitution (one right-hand side, whole diagonal-block
// inverses).  The entries are cut into slices of S = min(64, B); slice s
// belongs to workgroup s mod G, which keeps its running sum z_s in LDS for the
// whole launch.  Forward sweep, block step q ascending: every slice beyond
// block q takes block q's contribution z_s += A[qB .. qB+B)[slice]^T y_q as
// soon as y_q is known; a slice of block q+1, once complete, publishes its
// share of y_{q+1} = W_L(q+1)^T z_{q+1}: the product of the slice's S rows of
// the inverse with z_s, a B-vector.  y_{q+1} is the sum of the T = B / S
// shares, which every reader forms itself in slice order (so all see the same
// values): one publication per block on the critical path, not a gather of
// z followed by a publication of y.  The backward sweep mirrors it
// (descending, U and -W_U), each slice starting from its entries of y; the
// owners of a block finally write their entries of x.
// Publication: values are written with agent-scope coherent stores, every
// wave drains its stores, then the slice's flag is set (1 forward, 2
// backward; a small kernel clears the flags before each launch).  A waiter
// polls the T flags of the block it needs, bounded: past the limit the call
// fails with status bit 2 and every workgroup leaves.  The factor tile of a
// contribution and the inverse rows of a share do not depend on the awaited
// values, so they are loaded before the wait and only the flag latency and
// the arithmetic stay on the critical path.
// ---------------------------------------------------------------------------
constexpr int kMaxOwn = 16;            // slices per workgroup (LDS capacity)
constexpr int kMaxT = 4;               // shares per block (B / S, B <= 256)
constexpr long kGeslSpinLimit = 1L << 22;

__device__ __forceinline__ void pub_f(float* p, float v) {
    __hip_atomic_store(reinterpret_cast<unsigned*>(p), __float_as_uint(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float get_f(const float* p) {
    return __uint_as_float(__hip_atomic_load(reinterpret_cast<const unsigned*>(p),
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

__global__ void gesl_persist_reset_kernel(unsigned* flags, int64_t nflags, int* status) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nflags;
         i += (int64_t)gridDim.x * blockDim.x)
        flags[i] = 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) *status = 0;
}

template <int B>
__global__ __launch_bounds__(NT) void gesl_persist_kernel(const float* __restrict__ A, int64_t lda,
                                                          int64_t n, const float* __restrict__ binv,
                                                          float* __restrict__ x, GeslWs ws,
                                                          long spin_limit) {
    constexpr int S = B < 64 ? B : 64;  // slice width
    constexpr int T = B / S;            // slices per block
    constexpr int CG = S / 4;           // float4 column groups of a slice
    constexpr int RG = NT / CG;         // row groups
    constexpr int RPT = B / RG;         // rows of a tile per thread
    static_assert(RPT >= 1 && RPT * RG == B, "tile split");
    static_assert(B <= NT && T <= kMaxT, "one share column per thread");
    __shared__ float zown[kMaxOwn * S];  // running sums of the owned slices
    __shared__ float vbuf[B];            // the awaited block vector
    __shared__ float red[RG * S];
    __shared__ int abort_flag;
    const int t = threadIdx.x, G = gridDim.x, g = blockIdx.x;
    const int NS = (int)(n / S), N = (int)(n / B);
    const int own = (NS - g + G - 1) / G;  // slices g, g + G, ...
    const int cg = t % CG, rg = t / CG;
    unsigned* flag = ws.flags;

    for (int k = 0; k < own; ++k)
        for (int c = t; c < S; c += NT) zown[k * S + c] = x[(int64_t)(g + k * G) * S + c];

    f32x4 tile[RPT];
    // rows p = rg + RG i (stride ld) of base, the 4 columns of group cg
    auto load_tile = [&](const float* base, int64_t ld) {
#pragma unroll
        for (int i = 0; i < RPT; ++i)
            tile[i] = *reinterpret_cast<const f32x4*>(base + (int64_t)(rg + RG * i) * ld + 4 * cg);
    };
    // out[c] += sign * sum_p tile[p][c] vbuf[p]
    auto gemv = [&](float* out, float sign) {
        f32x4 a = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
            const float v = vbuf[rg + RG * i];
            a[0] = fmaf(tile[i][0], v, a[0]);
            a[1] = fmaf(tile[i][1], v, a[1]);
            a[2] = fmaf(tile[i][2], v, a[2]);
            a[3] = fmaf(tile[i][3], v, a[3]);
        }
        *reinterpret_cast<f32x4*>(&red[rg * S + 4 * cg]) = a;
        __syncthreads();
        if (t < S) {
            float sum = 0.0f;
#pragma unroll 8
            for (int r = 0; r < RG; ++r) sum += red[r * S + t];
            out[t] = out[t] + sign * sum;
        }
        __syncthreads();
    };
    // wait until the T flags from f on are >= target; true: the call failed
    auto wait = [&](const unsigned* f, unsigned target) {
        if (t == 0) {
            int ab = 0;
            long spins = 0;
            for (int m = 0; m < T && !ab; ++m)
                while (__hip_atomic_load(f + m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                       target) {
                    if (++spins > spin_limit ||
                        (__hip_atomic_load(ws.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
                         2)) {
                        atomicOr(ws.status, 2);
                        ab = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            abort_flag = ab;
        }
        __syncthreads();
        return abort_flag != 0;
    };
    // block q's vector: the sum of its T shares, in slice order
    auto fetch_sum = [&](int q, const float* shares, int lo, int len, float* dst) {
        const float* src = shares + (int64_t)q * T * B;
        for (int p = t; p < len; p += NT) {
            float v = 0.0f;
#pragma unroll
            for (int j = 0; j < T; ++j) v += get_f(src + j * B + lo + p);
            dst[p] = v;
        }
        __syncthreads();
    };
    // slice s (owned slice k, block Q): its share sign * W[slice rows]^T z_s,
    // published with the slice's flag; the inverse rows are loaded first
    // (share_load), before the wait for the slice's last contribution
    float wrow[S];
    auto share_load = [&](int Q, int s, bool fwd) {
        const float* w = binv + (int64_t)Q * 2 * B * B + (fwd ? 0 : (int64_t)B * B) +
                         (int64_t)(s - Q * T) * S * B;
        if (t < B) {
#pragma unroll
            for (int r = 0; r < S; ++r) wrow[r] = w[(int64_t)r * B + t];
        }
    };
    auto share = [&](int k, int s, bool fwd) {
        const float* z = zown + k * S;
        float acc = 0.0f;
        if (t < B) {
#pragma unroll
            for (int r = 0; r < S; ++r) acc = fmaf(wrow[r], z[r], acc);
            pub_f((fwd ? ws.pf : ws.pb) + (int64_t)s * B + t, fwd ? acc : -acc);
        }
        __builtin_amdgcn_s_waitcnt(0);  // this wave's stores have completed
        __syncthreads();
        if (t == 0)
            __hip_atomic_store(flag + s, fwd ? 1u : 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    // block q's contribution to owned slice k (= s)
    auto contrib = [&](int q, int k, int s, bool fwd) {
        load_tile(A + (int64_t)q * B * lda + (int64_t)s * S, lda);
        if (wait(flag + (int64_t)q * T, fwd ? 1u : 2u)) return true;
        fetch_sum(q, fwd ? ws.pf : ws.pb, 0, B, vbuf);
        gemv(zown + k * S, fwd ? 1.0f : -1.0f);
        return false;
    };
    __syncthreads();

    // forward: L y = b
    if (g < T) {
        share_load(0, g, true);
        share(0, g, true);
    }
    for (int q = 0; q + 1 < N; ++q)
        for (int k = 0; k < own; ++k) {
            const int s = g + k * G, blk = s / T;
            if (blk <= q) continue;
            if (blk == q + 1) share_load(q + 1, s, true);
            if (contrib(q, k, s, true)) return;
            if (blk == q + 1) share(k, s, true);
        }
    // backward: U x = y; each owned slice starts at its entries of y,
    // last block first
    for (int k = own - 1; k >= 0; --k) {
        const int s = g + k * G, blk = s / T;
        if (blk == N - 1) share_load(N - 1, s, false);
        if (wait(flag + (int64_t)blk * T, 1u)) return;
        fetch_sum(blk, ws.pf, (s - blk * T) * S, S, zown + k * S);
        if (blk == N - 1) share(k, s, false);
    }
    for (int q = N - 1; q >= 1; --q)
        for (int k = own - 1; k >= 0; --k) {
            const int s = g + k * G, blk = s / T;
            if (blk >= q) continue;
            if (blk == q - 1) share_load(q - 1, s, false);
            if (contrib(q, k, s, false)) return;
            if (blk == q - 1) share(k, s, false);
        }
    // x: the owners of a block sum its backward shares
    for (int k = 0; k < own; ++k) {
        const int s = g + k * G, blk = s / T;
        if (wait(flag + (int64_t)blk * T, 2u)) return;
        fetch_sum(blk, ws.pb, (s - blk * T) * S, S, zown + k * S);
        for (int c = t; c < S; c += NT) x[(int64_t)s * S + c] = zown[k * S + c];
    }
}

// Workgroups of the single-launch substitution of order n at block B, or 0
// when it does not apply (then the per-step launches solve).  At most half of
// what the device holds at once, so that the launch stays co-resident beside
// other work on the device.
int64_t persist_groups(int64_t n, int B) {
    if (B < 32 || B > 256 || (B & (B - 1)) || n % B) return 0;
    const int S = std::min(64, B);
    const int64_t NS = n / S;
    static int cap[4] = {0, 0, 0, 0};
    const int bi = B == 32 ? 0 : B == 64 ? 1 : B == 128 ? 2 : 3;
    if (!cap[bi]) {
        const void* k = B == 32    ? reinterpret_cast<const void*>(gesl_persist_kernel<32>)
                        : B == 64  ? reinterpret_cast<const void*>(gesl_persist_kernel<64>)
                        : B == 128 ? reinterpret_cast<const void*>(gesl_persist_kernel<128>)
                                   : reinterpret_cast<const void*>(gesl_persist_kernel<256>);
        int per_cu = 0, dev = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, NT, 0) != hipSuccess ||
            hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
            (void)hipGetLastError();
            return 0;
        }
        cap[bi] = std::max(1, per_cu * cus / 2);
    }
    int64_t G = std::min<int64_t>(NS, cap[bi]);
    if (const char* e = getenv("HPL_GESL_GROUPS"))
        G = std::max<int64_t>(1, std::min<int64_t>(G, atoll(e)));
    // a block's slices on distinct workgroups; the owned slices fit the LDS
    if (G < B / S || (NS + G - 1) / G > kMaxOwn) return 0;
    return G;
}

hipError_t gesl_persist(const float* A, int64_t lda, int64_t n, int B, const float* binv,
                        float* x, const GeslWs& ws, int64_t G, hipStream_t s) {
    const int64_t NS = n / std::min(64, B);
    gesl_persist_reset_kernel<<<(unsigned)std::min<int64_t>((NS + 255) / 256, 64), 256, 0,
                                s>>>(ws.flags, NS, ws.status);
    const unsigned grid = (unsigned)G;
    switch (B) {
        case 32:
            gesl_persist_kernel<32><<<grid, NT, 0, s>>>(A, lda, n, binv, x, ws, kGeslSpinLimit);
            break;
        case 64:
            gesl_persist_kernel<64><<<grid, NT, 0, s>>>(A, lda, n, binv, x, ws, kGeslSpinLimit);
            break;
        case 128:
            gesl_persist_kernel<128><<<grid, NT, 0, s>>>(A, lda, n, binv, x, ws, kGeslSpinLimit);
            break;
        default:
            gesl_persist_kernel<256><<<grid, NT, 0, s>>>(A, lda, n, binv, x, ws, kGeslSpinLimit);
            break;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_gesl_multi(const float* A, int64_t lda, int64_t n, int B, const float* inv,
                             int64_t inv_stride, float* X, int64_t ldx, int nrhs, hipStream_t s,
                             const float* binv, const GeslWs* ws) {
    if (B > 512 || B < 32 || B % 32 || n % B || nrhs < 1 || (nrhs > 1 && (ldx < n || ldx % 4)))
        return hipErrorInvalidValue;
    for (int r0 = 0; r0 < nrhs; r0 += 16) {
        const int nr = std::min(16, nrhs - r0);
        float* Xr = X + (int64_t)r0 * ldx;
        hipError_t e;
        const int64_t G = nr == 1 && binv && ws && (lda & 3) == 0 ? persist_groups(n, B) : 0;
        if (G > 0)
            e = gesl_persist(A, lda, n, B, binv, Xr, *ws, G, s);
        else if (nr == 1 && binv && B <= 256)
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
                       int64_t inv_stride, float* x, hipStream_t s, const float* binv,
                       const GeslWs* ws) {
    return launch_gesl_multi(A, lda, n, B, inv, inv_stride, x, n, 1, s, binv, ws);
}

int64_t gesl_persist_groups(int64_t n, int B) { return persist_groups(n, B); }

size_t gesl_ws_bytes(int64_t n) {
    // forward and backward shares (B per slice of S: at most kMaxT n floats
    // each), one flag per slice of >= 32 entries, status
    return (size_t)(2 * kMaxT * n) * sizeof(float) + (size_t)(n / 32 + 64) * sizeof(unsigned);
}

GeslWs gesl_ws_carve(void* base, int64_t n) {
    GeslWs w;
    float* f = static_cast<float*>(base);
    w.pf = f;
    w.pb = f + kMaxT * n;
    w.flags = reinterpret_cast<unsigned*>(f + 2 * kMaxT * n);
    w.status = reinterpret_cast<int*>(w.flags + n / 32);
    return w;
}
