// FP64 vector kernels of the GMRES-based refinement (api.hip, hpl_solve_gmres):
// matrix-vector product with the untouched FP32 matrix, block dot products
// against the Krylov basis and basis combinations.  Every reduction adds its
// partial sums in a fixed order, so results are bitwise reproducible.
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace {

constexpr int RCH = 256;  // rows of A per partial sum (as the residual kernels)
constexpr int DOT_T = 256;

// partial[c][i] = sum_{j in chunk c} M(i, j) x_j  with M(i, j) = A[j][i]
__global__ void matvec_partial_kernel(const float* __restrict__ A, int64_t lda, int64_t n,
                                      const double* __restrict__ x, double* __restrict__ partial) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c = blockIdx.y;
    if (i >= n) return;
    const int64_t j0 = c * RCH, j1 = min(n, j0 + RCH);
    double s = 0.0;
    int64_t j = j0;
    for (; j + 16 <= j1; j += 16) {
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = A[(j + u) * lda + i];
#pragma unroll
        for (int u = 0; u < 16; ++u) s = fma((double)v[u], x[j + u], s);
    }
    for (; j < j1; ++j) s = fma((double)A[j * lda + i], x[j], s);
    partial[c * n + i] = s;
}

__global__ void sum_partials_kernel(const double* __restrict__ partial, int64_t nch, int64_t n,
                                    double* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int64_t c = 0; c < nch; ++c) s += partial[c * n + i];
    y[i] = s;
}

// part[k][b] = sum over the entries of block b of V[k][e] * w[e], k < kc
__global__ __launch_bounds__(DOT_T) void multi_dot_partial_kernel(const double* __restrict__ V,
                                                                  int64_t ldv, int kc,
                                                                  const double* __restrict__ w,
                                                                  int64_t n, int64_t per_block,
                                                                  double* __restrict__ part) {
    __shared__ double red[DOT_T / 64];
    const int k = blockIdx.y;
    const int64_t e0 = (int64_t)blockIdx.x * per_block, e1 = min(n, e0 + per_block);
    double s = 0.0;
    for (int64_t e = e0 + threadIdx.x; e < e1; e += DOT_T) s = fma(V[k * ldv + e], w[e], s);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < DOT_T / 64; ++q) t += red[q];
        part[(int64_t)k * gridDim.x + blockIdx.x] = t;
    }
    (void)kc;
}

__global__ void multi_dot_final_kernel(const double* __restrict__ part, int kc, int nb,
                                       double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kc) return;
    double t = 0.0;
    for (int b = 0; b < nb; ++b) t += part[(int64_t)k * nb + b];
    out[k] = t;
}

// y[e] = alpha * y[e] + beta * sum_{k < kc} c[k] * V[k][e]   (c on the device)
__global__ void multi_axpy_kernel(double* __restrict__ y, double alpha, double beta,
                                  const double* __restrict__ V, int64_t ldv, int kc,
                                  const double* __restrict__ c, int64_t n) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int k = 0; k < kc; ++k) s = fma(c[k], V[k * ldv + e], s);
        y[e] = fma(beta, s, alpha * y[e]);
    }
}

__global__ void scale_copy_kernel(double* __restrict__ dst, const double* __restrict__ src,
                                  double a, int64_t n) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        dst[e] = a * src[e];
}

int64_t dot_blocks(int64_t n) { return std::min<int64_t>(256, (n + 4095) / 4096); }

}  // namespace

int64_t gmres_work_doubles(int64_t n, int kmax) {
    return std::max<int64_t>(residual_chunks(n) * n, (int64_t)kmax * dot_blocks(n));
}

hipError_t launch_matvec_f64(const float* A, int64_t lda, int64_t n, const double* x, double* y,
                             double* work, hipStream_t s) {
    const int64_t nch = residual_chunks(n);
    dim3 g1((unsigned)((n + 255) / 256), (unsigned)nch);
    matvec_partial_kernel<<<g1, 256, 0, s>>>(A, lda, n, x, work);
    sum_partials_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(work, nch, n, y);
    return hipGetLastError();
}

hipError_t launch_multi_dot(const double* V, int64_t ldv, int kc, const double* w, int64_t n,
                            double* out, double* work, hipStream_t s) {
    if (kc <= 0) return hipSuccess;
    const int64_t nb = dot_blocks(n);
    const int64_t per = (n + nb - 1) / nb;
    dim3 g((unsigned)nb, (unsigned)kc);
    multi_dot_partial_kernel<<<g, DOT_T, 0, s>>>(V, ldv, kc, w, n, per, work);
    multi_dot_final_kernel<<<(kc + 63) / 64, 64, 0, s>>>(work, kc, (int)nb, out);
    return hipGetLastError();
}

hipError_t launch_multi_axpy(double* y, double alpha, double beta, const double* V, int64_t ldv,
                             int kc, const double* c, int64_t n, hipStream_t s) {
    multi_axpy_kernel<<<(unsigned)std::min<int64_t>((n + 255) / 256, 2048), 256, 0, s>>>(
        y, alpha, beta, V, ldv, kc, c, n);
    return hipGetLastError();
}

hipError_t launch_scale_copy(double* dst, const double* src, double a, int64_t n, hipStream_t s) {
    scale_copy_kernel<<<(unsigned)std::min<int64_t>((n + 255) / 256, 2048), 256, 0, s>>>(dst, src,
                                                                                        a, n);
    return hipGetLastError();
}
