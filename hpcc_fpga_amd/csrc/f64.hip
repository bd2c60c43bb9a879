// FP64 LINPACK path: the reference built with DATA_TYPE=double (the "_DP"
// build, LINPACK/src/host/linpack_data.cpp:166-229 with HOST_DATA_TYPE
// double).  Same algorithm and storage as the FP32 path (row-major a,
// M = a^T, dgefa encoding without pivoting); the kernels:
//   * lu_f64_kernel: the 64 x 64 diagonal block in the registers of one wave
//     (lane c holds stored column c), pivots and multipliers broadcast with
//     v_readlane; the operation order of gefa_ref_nopvt with contraction off,
//     so the block is bit-identical to it;
//   * strip_f64_kernel: the top strip (one thread per column of the strip,
//     the closed form of test_kernel_communication.cpp:776-789) and the left
//     strip (one thread per row, :624-632) with the 64 values of the vector in
//     registers, coefficients from LDS; the left strip is also written
//     transposed (k-major) for the update;
//   * update_f64_kernel: C += Lt^T T on v_mfma_f64_16x16x4f64, 128 x 128
//     workgroup tiles of four 64 x 64 wave tiles, k-slabs of 16 through
//     double-buffered LDS, XCD-aware tile order as the FP32 update;
//   * gesl: per 64-block a one-wave diagonal solve (block staged in LDS) and a
//     column-parallel update of the rest of x (the factor streamed once per
//     sweep);
//   * residual r = b - M x with per-chunk partial sums added in fixed order.
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>

typedef double f64x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int FB = 64;  // FP64 block size

__device__ __forceinline__ double bcast64(double v, int lane) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, lane);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), lane);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// ---------------------------------------------------------------- diagonal LU
// Stored row r of the block is M's column r; lane c holds d[r] = D[r][c].
//   for gk: D[gk][gk] = -1/D[gk][gk]; D[gk][c] *= D[gk][gk] (c > gk);
//           D[r][c] += D[gk][c] * D[r][gk] (r > gk, c > gk)
__global__ __launch_bounds__(64) void lu_f64_kernel(double* __restrict__ A, int64_t lda,
                                                    int64_t off, int* __restrict__ status) {
#pragma clang fp contract(off)
    const int c = threadIdx.x;
    double* base = A + off * lda + off;
    double d[FB];
#pragma unroll
    for (int r = 0; r < FB; ++r) d[r] = base[(int64_t)r * lda + c];
    bool bad = false;
#pragma unroll
    for (int gk = 0; gk < FB; ++gk) {
        const double piv = bcast64(d[gk], gk);
        const double inv = -1.0 / piv;
        bad |= (piv == 0.0) || !__builtin_isfinite(inv);
        d[gk] = (c > gk) ? d[gk] * inv : ((c == gk) ? inv : d[gk]);
#pragma unroll
        for (int r = gk + 1; r < FB; ++r) {
            const double s = bcast64(d[r], gk);
            if (c > gk) d[r] = d[r] + d[gk] * s;
        }
    }
#pragma unroll
    for (int r = 0; r < FB; ++r) base[(int64_t)r * lda + c] = d[r];
    if (bad && c == 0 && status) atomicMax(status, 1);
}

// ---------------------------------------------------------------- strips
// Workgroup of 256 threads; blocks [0, gtop) take top-strip columns, the rest
// left-strip rows, 256 vectors each.  D (the factored diagonal block) in LDS.
__global__ __launch_bounds__(256) void strip_f64_kernel(double* __restrict__ A, int64_t lda,
                                                        int64_t off, int64_t m, int gtop,
                                                        double* __restrict__ Lt, int64_t ldl) {
#pragma clang fp contract(off)
    __shared__ double Ds[FB][FB + 1];
    const int t = threadIdx.x;
    const double* D = A + off * lda + off;
    for (int e = t; e < FB * FB; e += 256) Ds[e / FB][e % FB] = D[(int64_t)(e / FB) * lda + e % FB];
    __syncthreads();
    const bool top = (int)blockIdx.x < gtop;
    const int64_t v = (int64_t)(top ? blockIdx.x : blockIdx.x - gtop) * 256 + t;
    if (v >= m) return;
    double x[FB];
    if (top) {
        // x = column off+FB+v of stored rows off..off+FB-1:
        //   for k: x_k *= D[k][k]; x_j += x_k * D[j][k] (j > k)
        double* col = A + off * lda + off + FB + v;
#pragma unroll
        for (int k = 0; k < FB; ++k) x[k] = col[(int64_t)k * lda];
#pragma unroll
        for (int k = 0; k < FB; ++k) {
            x[k] = x[k] * Ds[k][k];
#pragma unroll
            for (int j = k + 1; j < FB; ++j) x[j] = x[j] + x[k] * Ds[j][k];
        }
#pragma unroll
        for (int k = 0; k < FB; ++k) col[(int64_t)k * lda] = x[k];
    } else {
        // x = stored row off+FB+v, columns off..off+FB-1:
        //   for k: x_i += x_k * D[k][i] (i > k)
        double* row = A + (off + FB + v) * lda + off;
#pragma unroll
        for (int k = 0; k < FB; k += 2) {
            const double2 p = *reinterpret_cast<const double2*>(row + k);
            x[k] = p.x;
            x[k + 1] = p.y;
        }
#pragma unroll
        for (int k = 0; k < FB; ++k)
#pragma unroll
            for (int i = k + 1; i < FB; ++i) x[i] = x[i] + x[k] * Ds[k][i];
#pragma unroll
        for (int k = 0; k < FB; k += 2) *reinterpret_cast<double2*>(row + k) = double2{x[k], x[k + 1]};
#pragma unroll
        for (int k = 0; k < FB; ++k) Lt[(int64_t)k * ldl + v] = x[k];
    }
}

// ---------------------------------------------------------------- update
// C[r][c] += sum_k At[k][r] * Bt[k][c].  MFMA operand of lane l: A[l%16][l/16],
// B[l/16][l%16]; accumulator register q of lane l holds D[l/16 + 4q][l%16].
constexpr int UBM = 128, UBN = 128, UBK = 16, UNT = 512;
constexpr int UPAD = 16;  // LDS row = 144 doubles: the 4 k-rows of a read fall in distinct bank halves

__device__ __forceinline__ void tile_coords64(int bid, int nblk, int tiles_m, int tiles_n, int& tm,
                                              int& tn) {
    int pid = bid;
    if ((nblk & 7) == 0) pid = (bid & 7) * (nblk >> 3) + (bid >> 3);
    constexpr int GROUP = 8;
    const int per_group = GROUP * tiles_n;
    const int g = pid / per_group;
    const int first_m = g * GROUP;
    const int gsize = min(GROUP, tiles_m - first_m);
    const int in_g = pid - g * per_group;
    tm = first_m + in_g % gsize;
    tn = in_g / gsize;
}

// Slabs are staged with global_load_lds: each wave instruction moves one
// 1 KiB row of a slab (16 B per lane) to a wave-uniform LDS row; the padded
// LDS row keeps the four k-rows of an operand read in distinct bank halves.
// The accumulators start from C, so the epilogue is a plain store.  Eight
// waves of 64 x 32 per 128 x 128 tile, at most 128 VGPRs (four waves per
// SIMD, two workgroups per CU): 56.4-56.6 TFLOP/s at M = N = 16384-24576,
// K = 384, against 53.6-54.0 with four waves of 64 x 64 (two per SIMD); the
// DMA row pointers are formed once and advanced by a slab per step (1.9 of
// those 2.6 TFLOP/s).
__global__ __launch_bounds__(UNT, 4) void update_f64_kernel(double* __restrict__ C, int64_t ldc,
                                                            const double* __restrict__ At,
                                                            int64_t lda,
                                                            const double* __restrict__ Bt,
                                                            int64_t ldb, int M, int N, int K) {
    constexpr int FM = 4, FN = 2;  // 16 x 16 blocks per 64 x 32 wave tile
    constexpr int LDS_ROW = UBM + UPAD;
    constexpr int ROWS_PER_WAVE = UBK / (UNT / 64);  // slab rows each wave moves (2)
    __shared__ double As[2][UBK][LDS_ROW];
    __shared__ double Bs[2][UBK][LDS_ROW];
    const int tiles_m = (M + UBM - 1) / UBM, tiles_n = (N + UBN - 1) / UBN;
    int tm, tn;
    tile_coords64(blockIdx.x, tiles_m * tiles_n, tiles_m, tiles_n, tm, tn);
    const int m0 = tm * UBM, n0 = tn * UBN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;  // 2 x 4 waves
    const int l16 = lane & 15, lq = lane >> 4;
    const bool full = (m0 + UBM <= M) && (n0 + UBN <= N);
    // lane l moves doubles 2l, 2l+1 of a row; clamped so edge tiles read in
    // bounds (the clamped columns only reach rows/columns of C never stored)
    const int ca = min(2 * lane, M - m0 - 2), cb = min(2 * lane, N - n0 - 2);

    const double* pa0 = At + (int64_t)(wave * ROWS_PER_WAVE) * lda + m0 + ca;
    const double* pb0 = Bt + (int64_t)(wave * ROWS_PER_WAVE) * ldb + n0 + cb;
    auto dma_slab = [&](int k0, int buf) {
        const double* pa = pa0 + (int64_t)k0 * lda;
        const double* pb = pb0 + (int64_t)k0 * ldb;
#pragma unroll
        for (int s = 0; s < ROWS_PER_WAVE; ++s) {
            const int row = wave * ROWS_PER_WAVE + s;
            __builtin_amdgcn_global_load_lds(
                pa + s * lda, (__attribute__((address_space(3))) void*)(&As[buf][row][0]), 16, 0, 0);
            __builtin_amdgcn_global_load_lds(
                pb + s * ldb, (__attribute__((address_space(3))) void*)(&Bs[buf][row][0]), 16, 0, 0);
        }
    };
    const int rbase = m0 + wm * 64, cbase = n0 + wn * 32;
    f64x4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = rbase + i * 16 + lq + 4 * q;
                const int col = cbase + j * 16 + l16;
                acc[i][j][q] = (full || (row < M && col < N)) ? C[(int64_t)row * ldc + col] : 0.0;
            }

    dma_slab(0, 0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const int nk = K / UBK;
    for (int kc = 0; kc < nk; ++kc) {
        const int buf = kc & 1;
        const bool more = kc + 1 < nk;
        if (more) dma_slab((kc + 1) * UBK, buf ^ 1);
#pragma unroll
        for (int kq = 0; kq < UBK / 4; ++kq) {
            double a[FM], b[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) a[i] = As[buf][kq * 4 + lq][wm * 64 + i * 16 + l16];
#pragma unroll
            for (int j = 0; j < FN; ++j) b[j] = Bs[buf][kq * 4 + lq][wn * 32 + j * 16 + l16];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (more) {
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
    }
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = rbase + i * 16 + lq + 4 * q;
                const int col = cbase + j * 16 + l16;
                if (full || (row < M && col < N)) C[(int64_t)row * ldc + col] = acc[i][j][q];
            }
}

// ---------------------------------------------------------------- substitution
// forward, block q (64 entries from k0): x_i += x_k * a[k][i] (k < i, inside
// the block), the block staged in LDS, one wave; then x_i += sum_{k in block}

// The 64 x 64 diagonal block at (k0, k0) into S, one column per lane, with
// all 64 loads of a lane in flight before the first is stored.
__device__ __forceinline__ void load_block64(const double* __restrict__ A, int64_t lda, int64_t k0,
                                             double (*S)[FB + 1]) {
    const int l = threadIdx.x;
    double v[FB];
#pragma unroll
    for (int r = 0; r < FB; ++r) v[r] = A[(k0 + r) * lda + k0 + l];
#pragma unroll
    for (int r = 0; r < FB; ++r) S[r][l] = v[r];
    __syncthreads();
}

// x_k * a[k][i] for i >= k0 + 64.
__global__ __launch_bounds__(64) void gesl_diag_fwd_f64(const double* __restrict__ A, int64_t lda,
                                                        int64_t k0, double* __restrict__ x) {
#pragma clang fp contract(off)
    __shared__ double S[FB][FB + 1];
    const int l = threadIdx.x;
    load_block64(A, lda, k0, S);
    double xi = x[k0 + l];
    for (int k = 0; k < FB - 1; ++k) {
        const double xk = bcast64(xi, k);
        if (l > k) xi = xi + xk * S[k][l];
    }
    x[k0 + l] = xi;
}

// backward, block at k0: for k = k0+63 .. k0: s = x_k * a[k][k]; x_k = -s;
// x_i += s * a[k][i] (k0 <= i < k)
__global__ __launch_bounds__(64) void gesl_diag_bwd_f64(const double* __restrict__ A, int64_t lda,
                                                        int64_t k0, double* __restrict__ x) {
#pragma clang fp contract(off)
    __shared__ double S[FB][FB + 1];
    const int l = threadIdx.x;
    load_block64(A, lda, k0, S);
    double xi = x[k0 + l];
    for (int k = FB - 1; k >= 0; --k) {
        const double s = bcast64(xi, k) * S[k][k];
        if (l == k) xi = -s;
        if (l < k) xi = xi + s * S[k][l];
    }
    x[k0 + l] = xi;
}

// x_i += sum over the block's k of s_k * a[k][i] for i in [i0, i1), in
// gesl_ref_nopvt's order: forward s_k = x_k, k increasing; backward the block
// holds x_k = -s_k, so s_k = -x_k, k decreasing.
template <bool BWD>
__global__ __launch_bounds__(256) void gesl_gemv_f64(const double* __restrict__ A, int64_t lda,
                                                     int64_t k0, int64_t i0, int64_t i1,
                                                     double* __restrict__ x) {
#pragma clang fp contract(off)
    __shared__ double xs[FB];
    if (threadIdx.x < FB) xs[threadIdx.x] = BWD ? -x[k0 + threadIdx.x] : x[k0 + threadIdx.x];
    __syncthreads();
    const int64_t i = i0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= i1) return;
    // the block's 64 factor entries of column i, all in flight at once
    double a[FB];
#pragma unroll
    for (int u = 0; u < FB; ++u) a[u] = A[(k0 + (BWD ? FB - 1 - u : u)) * lda + i];
    double acc = x[i];
#pragma unroll
    for (int u = 0; u < FB; ++u) {
        const int k = BWD ? FB - 1 - u : u;
        acc = acc + xs[k] * a[u];
    }
    x[i] = acc;
}

// ---------------------------------------------------------------- residual
// partial[c][i] = sum_{j in chunk c} M(i, j) x_j, M(i, j) = A[j][i]
constexpr int RCH64 = 256;
__global__ void resid_partial_f64(const double* __restrict__ A, int64_t lda, int64_t n,
                                  const double* __restrict__ x, double* __restrict__ part) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c = blockIdx.y;
    if (i >= n) return;
    const int64_t j0 = c * RCH64, j1 = min(n, j0 + RCH64);
    double s = 0.0, a = 0.0;
    for (int64_t j = j0; j < j1; ++j) {
        const double v = A[j * lda + i];
        s = fma(v, x[j], s);
        a += fabs(v);
    }
    part[(2 * c) * n + i] = s;
    part[(2 * c + 1) * n + i] = a;
}

// out = (||b - Mx||inf, ||M||inf, ||x||inf, ||b||inf), one workgroup
__global__ __launch_bounds__(256) void resid_final_f64(const double* __restrict__ part,
                                                       int64_t nch, int64_t n,
                                                       const double* __restrict__ x,
                                                       const double* __restrict__ b,
                                                       double* __restrict__ out) {
    __shared__ double red[4][256];
    double m[4] = {0, 0, 0, 0};
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        double s = 0.0, a = 0.0;
        for (int64_t c = 0; c < nch; ++c) {
            s += part[(2 * c) * n + i];
            a += part[(2 * c + 1) * n + i];
        }
        m[0] = fmax(m[0], fabs(b[i] - s));
        m[1] = fmax(m[1], a);
        m[2] = fmax(m[2], fabs(x[i]));
        m[3] = fmax(m[3], fabs(b[i]));
    }
    for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = m[q];
    __syncthreads();
    if (threadIdx.x < 4) {
        double v = 0.0;
        for (int u = 0; u < 256; ++u) v = fmax(v, red[threadIdx.x][u]);
        out[threadIdx.x] = v;
    }
}

__global__ void f32_to_f64_kernel(const float* __restrict__ s, int64_t lds, double* __restrict__ d,
                                  int64_t ldd, int64_t n) {
    for (int64_t r = blockIdx.y; r < n; r += gridDim.y)
        for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
             c += (int64_t)gridDim.x * blockDim.x)
            d[r * ldd + c] = (double)s[r * lds + c];
}

}  // namespace

int f64_block() { return FB; }

hipError_t launch_lu_f64(double* A, int64_t lda, int64_t off, int* status, hipStream_t s) {
    lu_f64_kernel<<<1, 64, 0, s>>>(A, lda, off, status);
    return hipGetLastError();
}

hipError_t launch_strips_f64(double* A, int64_t lda, int64_t off, int64_t m, double* Lt,
                             int64_t ldl, hipStream_t s) {
    if (m <= 0) return hipSuccess;
    const int g = (int)((m + 255) / 256);
    strip_f64_kernel<<<2 * g, 256, 0, s>>>(A, lda, off, m, g, Lt, ldl);
    return hipGetLastError();
}

hipError_t launch_update_f64(double* C, int64_t ldc, const double* Lt, int64_t ldl,
                             const double* T, int64_t ldt, int M, int N, int K, hipStream_t s) {
    if (M <= 0 || N <= 0) return hipSuccess;
    if (K % UBK || (M & 1) || (N & 1)) return hipErrorInvalidValue;
    const int tiles = ((M + UBM - 1) / UBM) * ((N + UBN - 1) / UBN);
    update_f64_kernel<<<tiles, UNT, 0, s>>>(C, ldc, Lt, ldl, T, ldt, M, N, K);
    return hipGetLastError();
}

hipError_t launch_gesl_f64(const double* A, int64_t lda, int64_t n, double* x, hipStream_t s) {
    if (n % FB) return hipErrorInvalidValue;
    for (int64_t k0 = 0; k0 < n; k0 += FB) {
        gesl_diag_fwd_f64<<<1, 64, 0, s>>>(A, lda, k0, x);
        const int64_t i0 = k0 + FB;
        if (i0 < n)
            gesl_gemv_f64<false><<<(unsigned)((n - i0 + 255) / 256), 256, 0, s>>>(A, lda, k0, i0,
                                                                                  n, x);
    }
    for (int64_t k0 = n - FB; k0 >= 0; k0 -= FB) {
        gesl_diag_bwd_f64<<<1, 64, 0, s>>>(A, lda, k0, x);
        if (k0 > 0)
            gesl_gemv_f64<true><<<(unsigned)((k0 + 255) / 256), 256, 0, s>>>(A, lda, k0, 0, k0, x);
    }
    return hipGetLastError();
}

int64_t residual_f64_work(int64_t n) { return 2 * ((n + RCH64 - 1) / RCH64) * n; }

hipError_t launch_residual_f64(const double* A, int64_t lda, int64_t n, const double* x,
                               const double* b, double* work, double* out4, hipStream_t s) {
    const int64_t nch = (n + RCH64 - 1) / RCH64;
    resid_partial_f64<<<dim3((unsigned)((n + 255) / 256), (unsigned)nch), 256, 0, s>>>(A, lda, n, x,
                                                                                    work);
    resid_final_f64<<<1, 256, 0, s>>>(work, nch, n, x, b, out4);
    return hipGetLastError();
}

hipError_t launch_f32_to_f64(const float* src, int64_t lds, double* dst, int64_t ldd, int64_t n,
                             hipStream_t s) {
    dim3 g((unsigned)std::min<int64_t>((n + 255) / 256, 64), (unsigned)std::min<int64_t>(n, 8192));
    f32_to_f64_kernel<<<g, 256, 0, s>>>(src, lds, dst, ldd, n);
    return hipGetLastError();
}
