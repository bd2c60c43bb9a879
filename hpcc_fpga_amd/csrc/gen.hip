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

// mode 0: M(i,i) += n, b_i = sum_j M(i,j) (double, j increasing)
__global__ void gen_mode0_kernel(float* __restrict__ A, int64_t lda, int64_t n,
                                 float* __restrict__ b, double* __restrict__ b64,
                                 double* __restrict__ xt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    A[i * lda + i] += (float)n;
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
    for (int64_t j = j0; j < j1; ++j) {
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
    if (mode == 0) {
        gen_mode0_kernel<<<g, 256, 0, s>>>(A, lda, n, b, b64, xt);
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
