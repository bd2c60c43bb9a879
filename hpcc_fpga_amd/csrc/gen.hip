// Device-side input generation and FP64 residual/norms.
//
// Generation: the counter-based generator of the sized workloads (element
// (r,c) of the row-major storage is a pure function of (seed, r*n+c)), so a
// 16 GiB matrix is produced in place at HBM speed.  The reference generates
// on the host (LINPACK/src/host/linpack_benchmark.hpp:211-297); its mt19937
// variant is sequential and stays on the host.
//
// Residual: r = b - M x with M = A^T, i.e. r_i = b_i - sum_j A[j][i] x_j,
// accumulated in FP64 from the FP32 matrix (coalesced rows of A).  The sum over
// j is split in fixed chunks whose partial sums are added in a fixed order, so
// the result is bitwise reproducible.
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace {

__global__ void gen_offdiag_kernel(float* __restrict__ A, int64_t lda, int64_t n, uint64_t key) {
    for (int64_t r = blockIdx.y; r < n; r += gridDim.y)
        for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
             c += (int64_t)gridDim.x * blockDim.x)
            A[r * lda + c] = hpl_u24(key, (uint64_t)(r * n + c)) - 0.5f;
}

// mode 0: M(i,i) += n, b_i = sum_j M(i,j) (double, j increasing);
// mode 2: the same without the diagonal term (no diagonal dominance)
__global__ void gen_mode0_kernel(float* __restrict__ A, int64_t lda, int64_t n,
                                 float* __restrict__ b, double* __restrict__ b64,
                                 double* __restrict__ xt, int dominant) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (dominant) A[i * lda + i] += (float)n;
    double s = 0.0;
    for (int64_t j = 0; j < n; ++j) s += (double)A[j * lda + i];
    if (b) b[i] = (float)s;
    if (b64) b64[i] = s;
    if (xt) xt[i] = 1.0;
}

// mode 1, part 1: M(i,i) = d * sum_{j!=i} |M(i,j)|, x_true_i = 2u - 1
__global__ void gen_mode1_diag_kernel(float* __restrict__ A, int64_t lda, int64_t n, double d,
                                      uint64_t xkey, double* __restrict__ xt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int64_t j = 0; j < n; ++j)
        if (j != i) s += fabs((double)A[j * lda + i]);
    A[i * lda + i] = (float)(d * s);
    xt[i] = (double)(2.0f * hpl_u24(xkey, (uint64_t)i) - 1.0f);
}

// mode 1, part 2: b = M x_true in double
__global__ void gen_mode1_rhs_kernel(const float* __restrict__ A, int64_t lda, int64_t n,
                                     const double* __restrict__ xt, float* __restrict__ b,
                                     double* __restrict__ b64) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int64_t j = 0; j < n; ++j) s = fma((double)A[j * lda + i], xt[j], s);
    if (b) b[i] = (float)s;
    if (b64) b64[i] = s;
}

constexpr int RCH = 256;  // rows of A per residual chunk

// partial[c][i] = sum_{j in chunk c} A[j][i] * x_j ; pabs[c][i] = sum |A[j][i]|
template <typename TX>
__global__ void resid_partial_kernel(const float* __restrict__ A, int64_t lda, int64_t n,
                                     const TX* __restrict__ x, double* __restrict__ partial,
                                     double* __restrict__ pabs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c = blockIdx.y;
    if (i >= n) return;
    const int64_t j0 = c * RCH, j1 = min(n, j0 + RCH);
    double s = 0.0, sa = 0.0;
    int64_t j = j0;
    // 16 rows' loads in flight at a time, then the sums in row order
    for (; j + 16 <= j1; j += 16) {
        float av[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) av[u] = A[(j + u) * lda + i];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const double a = (double)av[u];
            s = fma(a, (double)x[j + u], s);
            sa += fabs(a);
        }
    }
    for (; j < j1; ++j) {
        const double a = (double)A[j * lda + i];
        s = fma(a, (double)x[j], s);
        sa += fabs(a);
    }
    partial[c * n + i] = s;
    if (pabs) pabs[c * n + i] = sa;
}

__global__ void zero4_kernel(double* p) {
    if (threadIdx.x < 4) p[threadIdx.x] = 0.0;
}

__device__ __forceinline__ void atomic_max_nonneg(double* p, double v) {
    atomicMax(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v));
}

// r_i = b_i - sum_c partial[c][i]; norms into out[0..3] (must start at 0)
template <typename TX, typename TB>
__global__ void resid_final_kernel(int64_t n, int64_t nch, const double* __restrict__ partial,
                                   const double* __restrict__ pabs, const TX* __restrict__ x,
                                   const TB* __restrict__ b, double* __restrict__ r,
                                   double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double rv = 0, av = 0, xv = 0, bv = 0;
    if (i < n) {
        double s = 0.0, sa = 0.0;
        for (int64_t c = 0; c < nch; ++c) {
            s += partial[c * n + i];
            if (pabs) sa += pabs[c * n + i];
        }
        const double ri = (double)b[i] - s;
        if (r) r[i] = ri;
        rv = fabs(ri); av = sa; xv = fabs((double)x[i]); bv = fabs((double)b[i]);
    }
    // wave max then one atomic per wave
    for (int o = 32; o > 0; o >>= 1) {
        rv = fmax(rv, __shfl_xor(rv, o));
        av = fmax(av, __shfl_xor(av, o));
        xv = fmax(xv, __shfl_xor(xv, o));
        bv = fmax(bv, __shfl_xor(bv, o));
    }
    if ((threadIdx.x & 63) == 0 && out) {
        atomic_max_nonneg(out + 0, rv);
        atomic_max_nonneg(out + 1, av);
        atomic_max_nonneg(out + 2, xv);
        atomic_max_nonneg(out + 3, bv);
    }
}

}  // namespace

hipError_t launch_generate(float* A, int64_t lda, int64_t n, int mode, uint64_t seed, double d,
                           float* b, double* b64, double* xt, hipStream_t s) {
    const uint64_t key = hpl_stream_key(seed, 0);
    dim3 g1((unsigned)std::min<int64_t>((n + 255) / 256, 64), (unsigned)std::min<int64_t>(n, 8192));
    gen_offdiag_kernel<<<g1, 256, 0, s>>>(A, lda, n, key);
    const unsigned g = (unsigned)((n + 255) / 256);
    if (mode == 0 || mode == 2) {
        gen_mode0_kernel<<<g, 256, 0, s>>>(A, lda, n, b, b64, xt, mode == 0);
    } else {
        if (!xt) return hipErrorInvalidValue;
        gen_mode1_diag_kernel<<<g, 256, 0, s>>>(A, lda, n, d, hpl_stream_key(seed, 1), xt);
        gen_mode1_rhs_kernel<<<g, 256, 0, s>>>(A, lda, n, xt, b, b64);
    }
    return hipGetLastError();
}

int64_t residual_chunks(int64_t n) { return (n + RCH - 1) / RCH; }

template <typename TX, typename TB>
static hipError_t launch_residual_t(const float* A, int64_t lda, int64_t n, const TX* x,
                                    const TB* b, double* work, double* r, double* out4,
                                    bool with_norm, hipStream_t s) {
    const int64_t nch = residual_chunks(n);
    double* partial = work;
    double* pabs = with_norm ? work + nch * n : nullptr;
    dim3 g1((unsigned)((n + 255) / 256), (unsigned)nch);
    resid_partial_kernel<TX><<<g1, 256, 0, s>>>(A, lda, n, x, partial, pabs);
    if (out4) zero4_kernel<<<1, 64, 0, s>>>(out4);  // a kernel, not a memset node
    resid_final_kernel<TX, TB><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, nch, partial, pabs,
                                                                           x, b, r, out4);
    return hipGetLastError();
}

hipError_t launch_residual_f32x(const float* A, int64_t lda, int64_t n, const float* x,
                                const float* b, double* work, double* r, double* out4,
                                hipStream_t s) {
    return launch_residual_t<float, float>(A, lda, n, x, b, work, r, out4, true, s);
}

namespace {
__global__ void d2f_scaled_kernel(const double* __restrict__ r, float* __restrict__ x, int64_t n,
                                  double scale) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = (float)(r[i] * scale);
}
__global__ void axpy_f2d_kernel(double* __restrict__ x, const float* __restrict__ d, int64_t n,
                                double scale, int assign) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = assign ? scale * (double)d[i] : fma(scale, (double)d[i], x[i]);
}
}  // namespace

hipError_t launch_d2f_scaled(const double* r, float* x, int64_t n, double scale, hipStream_t s) {
    d2f_scaled_kernel<<<(unsigned)std::min<int64_t>((n + 255) / 256, 4096), 256, 0, s>>>(r, x, n,
                                                                                        scale);
    return hipGetLastError();
}

hipError_t launch_axpy_f2d(double* x, const float* d, int64_t n, double scale, int assign,
                           hipStream_t s) {
    axpy_f2d_kernel<<<(unsigned)std::min<int64_t>((n + 255) / 256, 4096), 256, 0, s>>>(
        x, d, n, scale, assign);
    return hipGetLastError();
}

hipError_t launch_residual_f64x(const float* A, int64_t lda, int64_t n, const double* x,
                                const double* b, double* work, double* r, double* out4,
                                bool with_norm, hipStream_t s) {
    return launch_residual_t<double, double>(A, lda, n, x, b, work, r, out4, with_norm, s);
}

// ---------------------------------------------------------------------------
// Several right-hand sides: R[r] = B[r] - M X[r], column r at offset r * ld.
// The matrix is read once for all of them; partial sums per chunk of rows
// (partial[(c * nr + r) * n + i]) are added in a fixed order as above.
namespace {
constexpr int kMaxMultiRhs = 16;

__global__ void resid_partial_multi_kernel(const float* __restrict__ A, int64_t lda, int64_t n,
                                           const double* __restrict__ X, int64_t ldx, int nr,
                                           double* __restrict__ partial,
                                           double* __restrict__ pabs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c = blockIdx.y;
    if (i >= n) return;
    const int64_t j0 = c * RCH, j1 = min(n, j0 + RCH);
    double s[kMaxMultiRhs];
#pragma unroll
    for (int r = 0; r < kMaxMultiRhs; ++r) s[r] = 0.0;
    double sa = 0.0;
    int64_t j = j0;
    // 16 rows' loads in flight at a time, then the sums in row order
    for (; j + 16 <= j1; j += 16) {
        float av[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) av[u] = A[(j + u) * lda + i];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const double a = (double)av[u];
#pragma unroll
            for (int r = 0; r < kMaxMultiRhs; ++r)
                if (r < nr) s[r] = fma(a, X[r * ldx + j + u], s[r]);
            sa += fabs(a);
        }
    }
    for (; j < j1; ++j) {
        const double a = (double)A[j * lda + i];
#pragma unroll
        for (int r = 0; r < kMaxMultiRhs; ++r)
            if (r < nr) s[r] = fma(a, X[r * ldx + j], s[r]);
        sa += fabs(a);
    }
#pragma unroll
    for (int r = 0; r < kMaxMultiRhs; ++r)
        if (r < nr) partial[(c * nr + r) * n + i] = s[r];
    if (pabs) pabs[c * n + i] = sa;
}

__global__ void zero_kernel(double* p, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0.0;
}

// grid.y = right-hand side; out[4 r + 0..3] = ||r||, ||M||, ||x||, ||b|| (inf)
__global__ void resid_final_multi_kernel(int64_t n, int64_t nch, int nr,
                                         const double* __restrict__ partial,
                                         const double* __restrict__ pabs,
                                         const double* __restrict__ X, int64_t ldx,
                                         const double* __restrict__ Bv, int64_t ldb,
                                         double* __restrict__ R, int64_t ldr,
                                         double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    double rv = 0, av = 0, xv = 0, bv = 0;
    if (i < n) {
        double s = 0.0, sa = 0.0;
        for (int64_t c = 0; c < nch; ++c) {
            s += partial[(c * nr + r) * n + i];
            if (pabs) sa += pabs[c * n + i];
        }
        const double bi = Bv[r * ldb + i];
        const double ri = bi - s;
        if (R) R[r * ldr + i] = ri;
        rv = fabs(ri); av = sa; xv = fabs(X[r * ldx + i]); bv = fabs(bi);
    }
    for (int o = 32; o > 0; o >>= 1) {
        rv = fmax(rv, __shfl_xor(rv, o));
        av = fmax(av, __shfl_xor(av, o));
        xv = fmax(xv, __shfl_xor(xv, o));
        bv = fmax(bv, __shfl_xor(bv, o));
    }
    if ((threadIdx.x & 63) == 0) {
        atomic_max_nonneg(out + 4 * r + 0, rv);
        atomic_max_nonneg(out + 4 * r + 1, av);
        atomic_max_nonneg(out + 4 * r + 2, xv);
        atomic_max_nonneg(out + 4 * r + 3, bv);
    }
}

struct ColScales {
    double v[kMaxMultiRhs];
};

// Xs[r][i] = (float)(R[r][i] * scale[r])
__global__ void d2f_multi_kernel(const double* __restrict__ R, int64_t ldr, float* __restrict__ Xs,
                                 int64_t ldxs, int64_t n, ColScales sc,
                                 const int* __restrict__ perm) {
    const int r = blockIdx.y;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        Xs[r * ldxs + i] = (float)(R[r * ldr + (perm ? (int64_t)perm[i] : i)] * sc.v[r]);
}

// X[r][i] = scale[r] * D[r][i] (assign) or X[r][i] += scale[r] * D[r][i]
__global__ void axpy_multi_kernel(double* __restrict__ X, int64_t ldx, const float* __restrict__ D,
                                  int64_t ldd, int64_t n, ColScales sc, int assign) {
    const int r = blockIdx.y;
    const double a = sc.v[r];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double d = (double)D[r * ldd + i];
        X[r * ldx + i] = assign ? a * d : fma(a, d, X[r * ldx + i]);
    }
}
}  // namespace

int64_t residual_multi_work(int64_t n, int nrhs) {
    return residual_chunks(n) * n * (nrhs + 1);  // doubles
}

hipError_t launch_residual_multi(const float* A, int64_t lda, int64_t n, const double* X,
                                 int64_t ldx, const double* Bv, int64_t ldb, int nrhs,
                                 double* work, double* R, int64_t ldr, double* out,
                                 hipStream_t s) {
    if (nrhs < 1 || nrhs > kMaxMultiRhs) return hipErrorInvalidValue;
    const int64_t nch = residual_chunks(n);
    double* partial = work;
    double* pabs = work + nch * n * nrhs;
    dim3 g1((unsigned)((n + 255) / 256), (unsigned)nch);
    resid_partial_multi_kernel<<<g1, 256, 0, s>>>(A, lda, n, X, ldx, nrhs, partial, pabs);
    zero_kernel<<<1, 64, 0, s>>>(out, 4 * nrhs);  // a kernel, not a memset node
    dim3 g2((unsigned)((n + 255) / 256), (unsigned)nrhs);
    resid_final_multi_kernel<<<g2, 256, 0, s>>>(n, nch, nrhs, partial, pabs, X, ldx, Bv, ldb, R,
                                                ldr, out);
    return hipGetLastError();
}

hipError_t launch_d2f_multi(const double* R, int64_t ldr, float* Xs, int64_t ldxs, int64_t n,
                            int nrhs, const double* scale, hipStream_t s, const int* perm) {
    if (nrhs < 1 || nrhs > kMaxMultiRhs) return hipErrorInvalidValue;
    ColScales sc{};
    for (int r = 0; r < nrhs; ++r) sc.v[r] = scale[r];
    dim3 g((unsigned)std::min<int64_t>((n + 255) / 256, 1024), (unsigned)nrhs);
    d2f_multi_kernel<<<g, 256, 0, s>>>(R, ldr, Xs, ldxs, n, sc, perm);
    return hipGetLastError();
}

hipError_t launch_axpy_multi(double* X, int64_t ldx, const float* D, int64_t ldd, int64_t n,
                             int nrhs, const double* scale, int assign, hipStream_t s) {
    if (nrhs < 1 || nrhs > kMaxMultiRhs) return hipErrorInvalidValue;
    ColScales sc{};
    for (int r = 0; r < nrhs; ++r) sc.v[r] = scale[r];
    dim3 g((unsigned)std::min<int64_t>((n + 255) / 256, 1024), (unsigned)nrhs);
    axpy_multi_kernel<<<g, 256, 0, s>>>(X, ldx, D, ldd, n, sc, assign);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Identity extension of an n x n system to npad x npad: the added rows and
// columns are zero except for ones on the diagonal, so the factor of the
// leading block is unchanged and the added unknowns solve to zero.
namespace {
__global__ void pad_identity_kernel(float* __restrict__ A, int64_t lda, int64_t n, int64_t npad) {
    for (int64_t r = blockIdx.y; r < npad; r += gridDim.y) {
        const int64_t c0 = r < n ? n : 0;  // rows < n: only the added columns
        for (int64_t c = c0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < npad;
             c += (int64_t)gridDim.x * blockDim.x)
            A[r * lda + c] = r == c ? 1.0f : 0.0f;
    }
}
__global__ void fill_f64_kernel(double* __restrict__ p, int64_t count, double v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
}  // namespace

hipError_t launch_pad_identity(float* A, int64_t lda, int64_t n, int64_t npad, hipStream_t s) {
    if (npad <= n) return hipSuccess;
    dim3 g((unsigned)std::min<int64_t>((npad + 255) / 256, 32),
           (unsigned)std::min<int64_t>(npad, 4096));
    pad_identity_kernel<<<g, 256, 0, s>>>(A, lda, n, npad);
    return hipGetLastError();
}

hipError_t launch_fill_f64(double* p, int64_t count, double v, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    fill_f64_kernel<<<(unsigned)std::min<int64_t>((count + 255) / 256, 1024), 256, 0, s>>>(p, count,
                                                                                          v);
    return hipGetLastError();
}
