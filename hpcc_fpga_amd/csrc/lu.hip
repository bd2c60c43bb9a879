// Diagonal-block LU ("lu", LINPACK/src/device/hpl_torus_PCIE.cl:211-446).
//
// Factorises the B x B diagonal block D = A[off.., off..] (row-major, lda) in
// place into the dgefa encoding without pivoting, i.e. gefa_ref_nopvt on the
// block (LINPACK/src/host/linpack_data.cpp:166-196):
//   for gk: D[gk][gk] = -1/D[gk][gk]; D[gk][c] *= D[gk][gk] (c > gk);
//           D[r][c] += D[gk][c] * D[r][gk]  (r > gk, c > gk).
// Every element sees the same multiply and add in the same order as the
// reference loop, so the block result is bit-identical to it.
//
// Blocks up to 128 are handled here; larger diagonal blocks are factorised by
// the host schedule as a 2 x 2 recursion on top of this kernel.
//
// One workgroup of NT threads holds the block in registers: thread t owns
// column c = t % B and rows rg + RG*i (rg = t / B, RG = NT / B).  Each step
// publishes the pivot column and pivot row of the NEXT step through LDS
// (double-buffered), so a step costs one workgroup barrier.
#include "common.hpp"
#include "kernels.hpp"

namespace {

template <int B, int NT>
__global__ __launch_bounds__(NT) void lu_diag_kernel(float* __restrict__ A, int64_t lda,
                                                      int64_t off, int* __restrict__ status) {
    // separate multiply and add, rounded as the reference loop rounds them
#pragma clang fp contract(off)
    constexpr int RG = NT / B;  // row groups
    constexpr int R = B / RG;   // rows per thread
    static_assert(RG * B == NT && R * RG == B, "bad LU shape");
    // colbuf[p][rg*R + i] = D[rg + RG*i][pivot column]; rowbuf[p][c] = D[pivot row][c]
    __shared__ __attribute__((aligned(16))) float colbuf[2][B];
    __shared__ float rowbuf[2][B];

    const int t = threadIdx.x;
    const int c = t % B;
    const int rg = t / B;
    float* base = A + off * lda + off;

    float d[R];
#pragma unroll
    for (int i = 0; i < R; ++i) d[i] = base[(int64_t)(rg + RG * i) * lda + c];

    // publish step 0: column 0 and row 0
    if (c == 0) {
#pragma unroll
        for (int i = 0; i < R; ++i) colbuf[0][rg * R + i] = d[i];
    }
    if (rg == 0) rowbuf[0][c] = d[0];

    bool bad = false;
#pragma unroll
    for (int i = 0; i < R; ++i) {
        for (int rr = 0; rr < RG; ++rr) {
            const int gk = i * RG + rr;
            const int p = gk & 1;
            __syncthreads();
            const float piv = colbuf[p][rr * R + i];
            const float inv = -1.0f / piv;
            bad |= !(__builtin_isfinite(inv)) || piv == 0.0f;
            const float rv = rowbuf[p][c] * inv;  // scaled pivot-row element (c > gk)
            if (rg == rr) d[i] = (c > gk) ? rv : ((c == gk) ? inv : d[i]);
            if (c > gk) {
                // rows r > gk: register i when rg > rr, every register above i.
                // Chunks of 4 with float4 LDS reads; the scheduling barrier
                // keeps the compiler from hoisting every read (register cap).
#pragma unroll
                for (int c4 = (i / 4) * 4; c4 < R; c4 += 4) {
                    const f32x4 cv = *reinterpret_cast<const f32x4*>(&colbuf[p][rg * R + c4]);
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int ii = c4 + e;
                        if (ii > i || (ii == i && rg > rr)) d[ii] += rv * cv[e];
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            // publish step gk+1 into the other buffers
            const int nk = gk + 1;
            if (nk < B) {
                if (c == nk) {
                    if (rg > rr) colbuf[1 - p][rg * R + i] = d[i];
#pragma unroll
                    for (int ii = i + 1; ii < R; ++ii) colbuf[1 - p][rg * R + ii] = d[ii];
                }
                // owner of row nk: rg == nk % RG, register nk / RG
                if (rr + 1 < RG) {
                    if (rg == rr + 1) rowbuf[1 - p][c] = d[i];
                } else {
                    if (i + 1 < R && rg == 0) rowbuf[1 - p][c] = d[(i + 1 < R) ? i + 1 : i];
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < R; ++i) base[(int64_t)(rg + RG * i) * lda + c] = d[i];
    if (bad && status) atomicMax(status, 1);
}

}  // namespace

hipError_t launch_lu_diag(float* A, int64_t lda, int64_t off, int B, int* status,
                          hipStream_t s) {
    switch (B) {
        case 32: lu_diag_kernel<32, 256><<<1, 256, 0, s>>>(A, lda, off, status); break;
        case 64: lu_diag_kernel<64, 256><<<1, 256, 0, s>>>(A, lda, off, status); break;
        case 128: lu_diag_kernel<128, 512><<<1, 512, 0, s>>>(A, lda, off, status); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
