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
//   A. wave 0 factorises the panel's 32 x 32 diagonal sub-block, one column
//      per lane; each step's pivot and multipliers come from the pivot
//      column's lane by v_readlane (scalar operands, no LDS round trip), so
//      the 32 steps need no barrier;
//   B. every column right of the sub-block (the panel's top strip) and every
//      row below it (its left strip) is solved against the factorised
//      sub-block in closed form, one vector per thread, the sub-block's
//      coefficients read as 16-byte LDS broadcasts from an aligned copy;
//   C. the trailing part receives the panel's 32 rank-1 updates in k order,
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
    // the factorised 32 x 32 sub-block of the current panel, row-major (S)
    // and transposed (St), rows 16-byte aligned for vector broadcasts
    __shared__ __attribute__((aligned(16))) float S[32 * 32];
    __shared__ __attribute__((aligned(16))) float St[32 * 32];
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
        // ---- A. sub-block rows and columns [P0, P0+32): lane c % 32 holds
        // column P0 + c.  Step j: row j right of the diagonal is scaled by
        // -1/p; then D[r][c] += D[j][c] * D[r][j] (r, c > j), D[r][j] being
        // lane j's (final) value of row r.
        if (wave == 0) {
            const int c = lane & 31;
            float d[32];
#pragma unroll
            for (int r = 0; r < 32; ++r) d[r] = Ds[(P0 + r) * LD + P0 + c];
#pragma unroll
            for (int j = 0; j < 32; ++j) {
                // the lane's column, opaque to the compiler per step: the
                // step's lane masks are formed here instead of all 64 being
                // hoisted out of the panel loop and spilled
                int cj = c;
                asm volatile("" : "+v"(cj));
                const float piv = bcast(d[j], j);
                const float inv = -1.0f / piv;
                bad |= (piv == 0.0f) || !__builtin_isfinite(inv);
                d[j] = (cj > j) ? d[j] * inv : ((cj == j) ? inv : d[j]);
                // the rows' products, then their sums, then the selects: three
                // passes of independent operations (row by row, the compiler
                // chains each row's three through one register)
                float u[32];
#pragma unroll
                for (int r = j + 1; r < 32; ++r) u[r] = d[j] * bcast(d[r], j);
#pragma unroll
                for (int r = j + 1; r < 32; ++r) u[r] = d[r] + u[r];
#pragma unroll
                for (int r = j + 1; r < 32; ++r) d[r] = (cj > j) ? u[r] : d[r];
            }
            if (lane < 32) {
#pragma unroll
                for (int r = 0; r < 32; ++r) {
                    Ds[(P0 + r) * LD + P0 + c] = d[r];
                    S[r * 32 + c] = d[r];
                    St[c * 32 + r] = d[r];
                }
            }
        }
        __syncthreads();
        const int R0 = P0 + 32;  // first row / column after the panel
        const int W = B - R0;    // rows below and columns right of the panel
        if (W == 0) break;
        // ---- B. top strip (rows P0.., column R0 + t) and left strip (row
        // R0 + t - W, columns P0..), one vector per thread, in step order
        for (int v = t; v < 2 * W; v += NT) {
            float x[32];
            if (v < W) {
                // top: step j scales x_j by D[j][j] (= -1/p), then
                // x_r += x_j * D[r][j] for r > j
                const int cc = R0 + v;
#pragma unroll
                for (int r = 0; r < 32; ++r) x[r] = Ds[(P0 + r) * LD + cc];
#pragma unroll
                for (int j = 0; j < 32; ++j) {
                    x[j] = x[j] * S[j * 32 + j];
#pragma unroll
                    for (int r4 = (j + 1) / 4; r4 < 8; ++r4) {
                        const float4 sv = *reinterpret_cast<const float4*>(&St[j * 32 + 4 * r4]);
                        const float s4[4] = {sv.x, sv.y, sv.z, sv.w};
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            if (4 * r4 + q > j) x[4 * r4 + q] = x[4 * r4 + q] + x[j] * s4[q];
                    }
                }
#pragma unroll
                for (int r = 0; r < 32; ++r) Ds[(P0 + r) * LD + cc] = x[r];
            } else {
                // left: x_i += x_k * D[k][i] for i > k, k in order
                const int rr = R0 + v - W;
#pragma unroll
                for (int i = 0; i < 32; ++i) x[i] = Ds[rr * LD + P0 + i];
#pragma unroll
                for (int k = 0; k < 32; ++k) {
#pragma unroll
                    for (int i4 = (k + 1) / 4; i4 < 8; ++i4) {
                        const float4 sv = *reinterpret_cast<const float4*>(&S[k * 32 + 4 * i4]);
                        const float s4[4] = {sv.x, sv.y, sv.z, sv.w};
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            if (4 * i4 + q > k) x[4 * i4 + q] = x[4 * i4 + q] + x[k] * s4[q];
                    }
                }
#pragma unroll
                for (int i = 0; i < 32; ++i) Ds[rr * LD + P0 + i] = x[i];
            }
        }
        __syncthreads();
        // ---- C. trailing W x W: D[j][i] += D[P0+k][i] * D[j][P0+k], k = 0..31
        const int tiles = (W / 4) * (W / 4);
        for (int tt = t; tt < tiles; tt += NT) {
            const int tj = R0 + (tt / (W / 4)) * 4, ti = R0 + (tt % (W / 4)) * 4;
            float acc[4][4];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = Ds[(tj + a) * LD + ti + b];
#pragma unroll 4
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
