// Trailing-matrix update ("inner_update_mm", LINPACK/src/device/hpl_torus_PCIE.cl:711-865).
//
// C[r][c] += sum_k Lt[k][r] * T[k][c]   for r < M, c < N, k < K
// C is the row-major trailing matrix (ldc), Lt the transposed left strips
// written by the panel solves (k-major, ldl) and T the top strips as they lie
// in A (k-major, ldt).  With the multipliers stored negated this is exactly
// the reference's C += Left * Top (alpha = beta = +1).  K is a multiple of the
// block size: the schedule may apply several steps' updates in one launch.
//
// Tile 128 x 128 with 16-deep k slabs and eight wave64s of 32 x 64 (two
// v_mfma_f32_32x32x2_f32 blocks each), 32 KiB of LDS per workgroup, four
// waves per SIMD, slabs staged by LDS DMA.  Measured on MI355X against nine
// other tile shapes at M = N = 16384 and 32256, K = 256 and 512: 111-113
// TFLOP/s at K = 256 and 121 at K = 512 with register staging (the 32-deep
// version of the same tile: 88-103 and 103-119); DMA staging adds 2-4%
// (114.5-115.6 and 121.6-125.4) with bit-identical results.
#include "gemm_kernel.hpp"
#include "kernels.hpp"

hipError_t launch_gemm_update_f32(float* C, int64_t ldc, const float* Lt, int64_t ldl,
                                  const float* T, int64_t ldt, int M, int N, int K,
                                  hipStream_t s) {
    return gemm_f32::launch<128, 128, 16, 32, 64, 2, false, false, true>(C, ldc, Lt, ldl, T, ldt, M, N, K, s);
}

hipError_t launch_gemm_update_f32_2(float* C, int64_t ldc, const float* Lt, int64_t ldl,
                                    const float* T, int64_t ldt, int M, int N, float* C2,
                                    const float* Lt2, const float* T2, int M2, int N2, int K,
                                    hipStream_t s) {
    return gemm_f32::launch2<128, 128, 16, 32, 64, 2, false, false, true>(
        C, ldc, Lt, ldl, T, ldt, M, N, K, gemm_f32::Region{C2, Lt2, T2, M2, N2}, s);
}
