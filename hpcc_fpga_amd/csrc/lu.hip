// Diagonal-block LU ("lu", LINPACK/src/device/hpl_torus_PCIE.cl:211-446).
//
// Factorises the B x B diagonal block D = A[off.., off..] (row-major, lda) in
// place into the dgefa encoding without pivoting, i.e. gefa_ref_nopvt on the
// block (LINPACK/src/host/linpack_data.cpp:166-196):
//   for gk: D[gk][gk] = -1/D[gk][gk]; D[gk][c] *= D[gk][gk] (c > gk);
//           D[r][c] += D[gk][c] * D[r][gk]  (r > gk, c > gk).
// Every element receives the same multiplies and adds in the same order as in
// the reference loop (no FMA contraction), so the result is bit-identical.
//
// The block lives in LDS and is processed in panels of 32 pivot rows:
//   1. wave 0 eliminates the 32 panel rows (all columns right of the panel
//      start) with the panel held in registers; pivots and multipliers are
//      broadcast with v_readlane, so the 32 steps need no barrier;
//   2. every row below the panel is solved against the panel's 32 x 32
//      diagonal block (the "left" closed form), one row per thread;
//   3. the trailing part receives the panel's 32 rank-1 updates in k order,
//      4 x 4 register tiles per thread.
// Three workgroup barriers per panel instead of one or two per pivot.
// Blocks up to 128 are handled here; larger diagonal blocks are factorised by
// the host schedule as a 2 x 2 recursion on top of this kernel.
#include "common.hpp"
#include "kernels.hpp"

namespace {

constexpr int NT = 256;

__device__ __forceinline__ float bcast(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

template <int B>
__global__ __launch_bounds__(NT) void lu_diag_kernel(float* __restrict__ A, int64_t lda,
                                                      int64_t off, int* __restrict__ status) {
#pragma clang fp contract(off)
    constexpr int LD = B + 1;  // padded LDS row
    __shared__ float Ds[B * LD];
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    float* base = A + off * lda + off;

    // the block as float4s, every load issued before the first is used (one
    // HBM round trip instead of one per element group); off is a multiple of
    // 32 and lda of 4, so the rows are 16-byte aligned
    constexpr int V = B * B / 4 / NT;
    static_assert(V >= 1 && V * NT * 4 == B * B, "block must split evenly over the threads");
    {
        float4 v[V];
#pragma unroll
        for (int i = 0; i < V; ++i) {
            const int e = t + i * NT, r = e / (B / 4), c = (e % (B / 4)) * 4;
            v[i] = *reinterpret_cast<const float4*>(base + (int64_t)r * lda + c);
        }
#pragma unroll
        for (int i = 0; i < V; ++i) {
            const int e = t + i * NT, r = e / (B / 4), c = (e % (B / 4)) * 4;
            Ds[r * LD + c] = v[i].x;
            Ds[r * LD + c + 1] = v[i].y;
            Ds[r * LD + c + 2] = v[i].z;
            Ds[r * LD + c + 3] = v[i].w;
        }
    }
    __syncthreads();

    bool bad = false;
    for (int P0 = 0; P0 < B; P0 += 32) {
        // ---- 1. panel rows [P0, P0+32), columns [P0, B): wave 0, registers
        if (wave == 0) {
            const int c0 = P0 + lane, c1 = P0 + 64 + lane;
            const bool h0 = c0 < B, h1 = c1 < B;
            float d0[32], d1[32];
#pragma unroll
            for (int r = 0; r < 32; ++r) {
                d0[r] = h0 ? Ds[(P0 + r) * LD + c0] : 0.0f;
                d1[r] = h1 ? Ds[(P0 + r) * LD + c1] : 0.0f;
            }
#pragma unroll
            for (int j = 0; j < 32; ++j) {
                const float piv = bcast(d0[j], j);
                const float inv = -1.0f / piv;
                bad |= (piv == 0.0f) || !__builtin_isfinite(inv);
                // scale pivot row j right of the diagonal; diagonal := -1/p
                d0[j] = (lane > j) ? d0[j] * inv : ((lane == j) ? inv : d0[j]);
                d1[j] = d1[j] * inv;
                // the multipliers of column j (lane j's rows) reach every lane
                // through ds_bpermute: vector results, so the 31 - j
                // broadcasts pipeline without scalar registers or the
                // readlane hazards
#pragma unroll
                for (int r = j + 1; r < 32; ++r) {
                    const float s = __shfl(d0[r], j);
                    const float u = d0[r] + d0[j] * s;
                    d0[r] = (lane > j) ? u : d0[r];  // a select, not a branch
                    d1[r] = d1[r] + d1[j] * s;
                }
                // finish this step's updates here: the compiler would
                // otherwise delay those of d1 (off the pivot chain) to the
                // end and keep every broadcast value alive until then
#pragma unroll
                for (int r = 0; r < 32; ++r) asm volatile("" : "+v"(d0[r]), "+v"(d1[r]));
            }
#pragma unroll
            for (int r = 0; r < 32; ++r) {
                if (h0) Ds[(P0 + r) * LD + c0] = d0[r];
                if (h1) Ds[(P0 + r) * LD + c1] = d1[r];
            }
        }
        __syncthreads();
        const int R0 = P0 + 32;  // first row / column after the panel
        const int W = B - R0;    // rows below and columns right of the panel
        if (W == 0) break;
        // ---- 2. rows below: X[j][P0+i] += X[j][P0+k] * D[P0+k][P0+i], i > k
        if (t < W) {
            const int j = R0 + t;
            float x[32];
#pragma unroll
            for (int i = 0; i < 32; ++i) x[i] = Ds[j * LD + P0 + i];
#pragma unroll
            for (int k = 0; k < 32; ++k)
#pragma unroll
                for (int i = k + 1; i < 32; ++i)
                    x[i] = x[i] + x[k] * Ds[(P0 + k) * LD + P0 + i];
#pragma unroll
            for (int i = 0; i < 32; ++i) Ds[j * LD + P0 + i] = x[i];
        }
        __syncthreads();
        // ---- 3. trailing W x W: D[j][i] += D[P0+k][i] * D[j][P0+k], k = 0..31
        const int tiles = (W / 4) * (W / 4);
        for (int tt = t; tt < tiles; tt += NT) {
            const int tj = R0 + (tt / (W / 4)) * 4, ti = R0 + (tt % (W / 4)) * 4;
            float acc[4][4];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = Ds[(tj + a) * LD + ti + b];
            for (int k = 0; k < 32; ++k) {
                float lv[4], tv[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) lv[a] = Ds[(tj + a) * LD + P0 + k];
#pragma unroll
                for (int b = 0; b < 4; ++b) tv[b] = Ds[(P0 + k) * LD + ti + b];
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) acc[a][b] = acc[a][b] + tv[b] * lv[a];
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) Ds[(tj + a) * LD + ti + b] = acc[a][b];
        }
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < V; ++i) {
        const int e = t + i * NT, r = e / (B / 4), c = (e % (B / 4)) * 4;
        *reinterpret_cast<float4*>(base + (int64_t)r * lda + c) =
            float4{Ds[r * LD + c], Ds[r * LD + c + 1], Ds[r * LD + c + 2], Ds[r * LD + c + 3]};
    }
    if (bad && status) atomicMax(status, 1);
}

}  // namespace

hipError_t launch_lu_diag(float* A, int64_t lda, int64_t off, int B, int* status,
                          hipStream_t s) {
    // 16-byte row vectors
    if ((lda & 3) || (off & 3) || (reinterpret_cast<uintptr_t>(A) & 15))
        return hipErrorInvalidValue;
    switch (B) {
        case 32: lu_diag_kernel<32><<<1, NT, 0, s>>>(A, lda, off, status); break;
        case 64: lu_diag_kernel<64><<<1, NT, 0, s>>>(A, lda, off, status); break;
        case 128: lu_diag_kernel<128><<<1, NT, 0, s>>>(A, lda, off, status); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
