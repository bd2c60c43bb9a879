// Trailing-matrix update ("inner_update_mm", LINPACK/src/device/hpl_torus_PCIE.cl:711-865).
//
// C[r][c] += sum_k Lt[k][r] * T[k][c]   for r < M, c < N, k < K (K = block B)
// C is the row-major trailing matrix (ldc), Lt the transposed left strip
// written by the panel solve (k-major, ldl) and T the top strip as it lies in
// A (k-major, ldt).  With the multipliers stored negated this is exactly the
// reference's C += Left * Top (alpha = beta = +1).
//
// The whole trailing matrix of one step is one launch.  Workgroup tile
// 128 x 128, four wave64s in 2 x 2, wave tile 64 x 64 = 2 x 2 blocks of
// v_mfma_f32_32x32x2_f32.  Both operands are k-major, so a k-slab of each is a
// contiguous row segment: float4 global loads, conflict-free LDS rows, and the
// MFMA operand of lane l is LDS[k0 + l/32][m0 + l%32].  LDS is double
// buffered (one barrier per k-slab) with the next slab prefetched into
// registers while the current one is multiplied.  The C tile is loaded into
// the accumulators before the k-loop and stored once after it.
//
// Tile order is XCD-aware: consecutive workgroup ids are dealt round-robin to
// the 8 XCDs, so ids are remapped to give each XCD a contiguous run of tiles,
// walked in groups of 8 tile-rows so a run shares Lt rows and T columns in
// that XCD's L2.
#include "common.hpp"
#include "kernels.hpp"

namespace {

constexpr int BM = 128, BN = 128, BK = 16, NT = 256;

__device__ __forceinline__ void tile_coords(int bid, int nblk, int tiles_m, int tiles_n, int& tm,
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

__global__ __launch_bounds__(NT, 2) void gemm_update_f32_kernel(
    float* __restrict__ C, int64_t ldc, const float* __restrict__ Lt, int64_t ldl,
    const float* __restrict__ T, int64_t ldt, int M, int N, int K) {
    __shared__ float As[2][BK][BM];
    __shared__ float Bs[2][BK][BN];

    const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
    int tm, tn;
    tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int lr = lane & 31, lh = lane >> 5;
    const bool full = (m0 + BM <= M) && (n0 + BN <= N);

    // global -> register slab mapping: 2 float4 of A and 2 of B per thread
    int gk[2], gc[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int f = tid + NT * s;
        gk[s] = f / (BM / 4);
        gc[s] = (f % (BM / 4)) * 4;
    }
    f32x4 ra[2], rb[2];
    auto load_slab = [&](int k0) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int64_t krow = k0 + gk[s];
            if (full || m0 + gc[s] < M)
                ra[s] = *reinterpret_cast<const f32x4*>(Lt + krow * ldl + m0 + gc[s]);
            else
                ra[s] = f32x4{0, 0, 0, 0};
            if (full || n0 + gc[s] < N)
                rb[s] = *reinterpret_cast<const f32x4*>(T + krow * ldt + n0 + gc[s]);
            else
                rb[s] = f32x4{0, 0, 0, 0};
        }
    };
    auto store_slab = [&](int buf) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            *reinterpret_cast<f32x4*>(&As[buf][gk[s]][gc[s]]) = ra[s];
            *reinterpret_cast<f32x4*>(&Bs[buf][gk[s]][gc[s]]) = rb[s];
        }
    };

    // accumulators initialised with the C tile
    f32x16 acc[2][2];
    const int rbase = m0 + wm * 64, cbase = n0 + wn * 64;
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = rbase + mi * 32 + acc_row(r, lane);
                const int col = cbase + ni * 32 + lr;
                acc[mi][ni][r] = (full || (row < M && col < N)) ? C[(int64_t)row * ldc + col] : 0.0f;
            }

    load_slab(0);
    store_slab(0);
    __syncthreads();
    const int nk = K / BK;
    for (int kc = 0; kc < nk; ++kc) {
        const int buf = kc & 1;
        if (kc + 1 < nk) load_slab((kc + 1) * BK);
#pragma unroll
        for (int kk = 0; kk < BK / 2; ++kk) {
            const int kr = 2 * kk + lh;
            const float a0 = As[buf][kr][wm * 64 + lr];
            const float a1 = As[buf][kr][wm * 64 + 32 + lr];
            const float b0 = Bs[buf][kr][wn * 64 + lr];
            const float b1 = Bs[buf][kr][wn * 64 + 32 + lr];
            acc[0][0] = mfma32(a0, b0, acc[0][0]);
            acc[0][1] = mfma32(a0, b1, acc[0][1]);
            acc[1][0] = mfma32(a1, b0, acc[1][0]);
            acc[1][1] = mfma32(a1, b1, acc[1][1]);
        }
        if (kc + 1 < nk) store_slab(buf ^ 1);
        __syncthreads();
    }

#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = rbase + mi * 32 + acc_row(r, lane);
                const int col = cbase + ni * 32 + lr;
                if (full || (row < M && col < N)) C[(int64_t)row * ldc + col] = acc[mi][ni][r];
            }
}

}  // namespace

hipError_t launch_gemm_update_f32(float* C, int64_t ldc, const float* Lt, int64_t ldl,
                                  const float* T, int64_t ldt, int M, int N, int K,
                                  hipStream_t s) {
    if (M <= 0 || N <= 0) return hipSuccess;
    if (K % BK != 0 || (M % 4) || (N % 4)) return hipErrorInvalidValue;
    const int tiles = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
    gemm_update_f32_kernel<<<tiles, NT, 0, s>>>(C, ldc, Lt, ldl, T, ldt, M, N, K);
    return hipGetLastError();
}
