// Panel triangular solves: "top_update" and "left_update"
// (LINPACK/src/device/hpl_torus_PCIE.cl:452-582 and :588-709).
//
// After the diagonal block D (row-major, B x B) of step k is factorised, the
// row-major strip right of it ("top" blocks, M's L-panel) and the strip below
// it ("left" blocks, M's U-panel) are solved.  Both are a forward substitution
// applied to many independent vectors v of length B:
//     y_i = (v_i + sum_{p<i} C[i][p] y_p) * s_i,   i = 0..B-1
//   top : vectors = columns of the strip, C[i][p] = D[i][p], s_i = D[i][i]
//         (closed form of LINPACK/tests/test_kernel_communication.cpp:776-789)
//   left: vectors = rows of the strip,    C[i][p] = D[p][i], s_i = 1
//         (closed form of test_kernel_communication.cpp:624-632)
// The left strip is also written transposed (k-major) for the trailing GEMM,
// as the reference's left_update writes its transposed copy
// (hpl_torus_PCIE.cl:697-708).
//
// Mixed precision: the solved strips can also be written as BF16/FP16 copies
// (k-contiguous) or as FP8 E4M3 copies with one power-of-two scale per block
// of 32 vectors x B (the block's amax maps to at most 448, so the scaling is
// exact and nothing saturates).
//
// Blocking: rows are taken 32 at a time.  For block q the contribution of the
// already solved rows is a 32 x 32q by 32q x 32 product on MFMA
// (v_mfma_f32_32x32x2_f32), and the 32 x 32 diagonal system is applied as its
// precomputed inverse G_q = (S^-1 - C_qq)^-1, also on MFMA.  One wavefront
// owns 32 vectors for all B rows, so no workgroup barrier is needed.
#include "common.hpp"
#include "kernels.hpp"

#include <type_traits>

namespace {

// ----------------------------------------------------------------------------
// prep: transposed copy of D (coefficients of the top solve become coalesced)
// and the 32 x 32 inverses of the diagonal sub-systems, stored transposed:
//   inv + 0*NQ*1024 : G^T for top   (S^-1 - C)^-1 with C = D lower, s = diag D
//   inv + 1*NQ*1024 : G^T for left  (I - C)^-1 with C[i][p] = D[p][i]
//   inv + 2*NQ*1024 : H^T for the back substitution of GESL, H = U_qq^-1 with
//                     U[i][p] = D[p][i] (p > i), 1/u_ii = -D[i][i]
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(64) void prep_kernel(const float* __restrict__ A, int64_t lda,
                                                  int64_t off, int B, float* __restrict__ DT,
                                                  float* __restrict__ inv, int64_t inv_stride) {
    __shared__ float S[32][33];
    const int NQ = B / 32;
    const int t = threadIdx.x;
    int blk = blockIdx.x;
    const float* D = A + off * lda + off;
    inv += (int64_t)blockIdx.y * inv_stride;
    if (gridDim.y > 1) D += (int64_t)blockIdx.y * (B * lda + B);  // all-blocks mode
    // a 32 x 32 sub-block into S: four float4 per thread, all issued before
    // the first is used (rows are 16-byte aligned: off is a multiple of 32,
    // lda of 4)
    auto load_sub = [&](int r0, int c0) {
        float4 v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int e = t + 64 * k, i = e / 8, j = (e % 8) * 4;
            v[k] = *reinterpret_cast<const float4*>(D + (int64_t)(r0 + i) * lda + c0 + j);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int e = t + 64 * k, i = e / 8, j = (e % 8) * 4;
            S[i][j] = v[k].x;
            S[i][j + 1] = v[k].y;
            S[i][j + 2] = v[k].z;
            S[i][j + 3] = v[k].w;
        }
    };
    if (blk >= 3 * NQ) {
        if (DT == nullptr) return;
        blk -= 3 * NQ;
        const int tr = blk / NQ, tc = blk % NQ;
        load_sub(tr * 32, tc * 32);
        __syncthreads();
        for (int e = t; e < 1024; e += 64) {
            const int j = e / 32, i = e % 32;
            DT[(int64_t)(tc * 32 + j) * B + tr * 32 + i] = S[i][j];
        }
        return;
    }
    const int type = blk / NQ, q = blk % NQ;
    load_sub(q * 32, q * 32);
    __syncthreads();
    // Lane j (mod 32) computes column j of the inverse in registers, x[i] =
    // X[i][j], with the same fmaf sequence per element as the textbook loops
    // (a from 0, p = j .. i - 1 increasing; backward for type 2): the terms
    // outside that range are computed and discarded by a select, so the
    // accumulator stays +0 until the first real term.  The S reads have
    // lane-independent addresses (LDS broadcasts) and do not depend on the
    // chain, so only the fmaf chain is serial.
    const int j = t & 31;
    float x[32];
    if (type == 0) {
#pragma unroll
        for (int i = 0; i < 32; ++i) {
            float a = 0.0f;
#pragma unroll
            for (int p = 0; p < i; ++p) {
                const float f = fmaf(S[i][p], x[p], a);
                a = p >= j ? f : a;
            }
            const float sii = S[i][i];
            x[i] = i > j ? a * sii : (i == j ? sii : 0.0f);
        }
    } else if (type == 1) {
#pragma unroll
        for (int i = 0; i < 32; ++i) {
            float a = 0.0f;
#pragma unroll
            for (int p = 0; p < i; ++p) {
                const float f = fmaf(S[p][i], x[p], a);
                a = p >= j ? f : a;
            }
            x[i] = i > j ? a : (i == j ? 1.0f : 0.0f);
        }
    } else {
#pragma unroll
        for (int i = 31; i >= 0; --i) {
            float a = 0.0f;
#pragma unroll
            for (int p = i + 1; p < 32; ++p) {
                const float f = fmaf(S[p][i], x[p], a);
                a = p <= j ? f : a;
            }
            const float sii = S[i][i];
            x[i] = i < j ? a * sii : (i == j ? -sii : 0.0f);
        }
    }
    // out[j][i] = X[i][j]: lane j writes its 32 values as 16-byte vectors
    if (t < 32) {
        float* out = inv + (int64_t)(type * NQ + q) * 1024 + j * 32;
#pragma unroll
        for (int i = 0; i < 32; i += 4)
            *reinterpret_cast<float4*>(out + i) = float4{x[i], x[i + 1], x[i + 2], x[i + 3]};
    }
}

// ----------------------------------------------------------------------------
// The solve.  Launch: one wave per workgroup, grid.x = ceil(m/32) vectors
// groups for the top strip followed by the same count for the left strip.
// ----------------------------------------------------------------------------
// LP: element type of the optional low-precision copies for the mixed
// precision trailing update (float = none): lpT[c][k] (top strip, column c)
// and lpL[r][k] (left strip, row r), k-contiguous with leading dimension ldlp.
template <int B, typename LP>
__global__ __launch_bounds__(64) void trsm_kernel(float* __restrict__ A, int64_t lda, int64_t off,
                                                  int64_t m, const float* __restrict__ DT,
                                                  const float* __restrict__ inv,
                                                  float* __restrict__ Lt, int64_t ldt,
                                                  int ngroups_top, LP* __restrict__ lpT,
                                                  LP* __restrict__ lpL, int64_t ldlp,
                                                  float* __restrict__ scT,
                                                  float* __restrict__ scL) {
    constexpr bool kFp8 = std::is_same<LP, fp8e4m3>::value;
    constexpr int NQ = B / 32;
    __shared__ float Y[B][32];
    __shared__ float SZ[32][33];  // staging tile of the left strip, then z_q
    auto& Stg = SZ;
    auto& Z = SZ;
    const int lane = threadIdx.x;
    const int lr = lane & 31, lh = lane >> 5;
    const bool top = (int)blockIdx.x < ngroups_top;
    const int64_t v0 = (int64_t)(top ? blockIdx.x : blockIdx.x - ngroups_top) * 32;
    const float* D = A + off * lda + off;
    // coefficient matrix read as X[p][i]: top -> D^T, left -> D
    const float* Xc = top ? DT : D;
    const int64_t ldx = top ? (int64_t)B : lda;
    const float* G = inv + (top ? 0 : NQ * 1024);
    float* stripT = A + off * lda + off + B;          // top strip: rows off.., cols off+B..
    float* stripL = A + (off + B) * lda + off;        // left strip: rows off+B.., cols off..

    // prefetch of the next block's right-hand sides (registers)
    float vn[16];
    f32x4 sn[4];
    auto fetch_v = [&](int q) {
        if (top) {
#pragma unroll
            for (int r = 0; r < 16; ++r)
                vn[r] = stripT[(int64_t)(q * 32 + acc_row(r, lane)) * lda + v0 + lr];
        } else {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int vr = (lane >> 3) + 8 * s, c4 = (lane & 7) * 4;
                sn[s] = *reinterpret_cast<const f32x4*>(stripL + (v0 + vr) * lda + q * 32 + c4);
            }
        }
    };
    fetch_v(0);

    constexpr int PF = 8;  // coefficient prefetch depth (k-pairs)
    for (int q = 0; q < NQ; ++q) {
        f32x16 acc;
        // ---- v_q in accumulator layout (row i = acc_row, vector = lr)
        if (top) {
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] = vn[r];
        } else {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int vr = (lane >> 3) + 8 * s, c4 = (lane & 7) * 4;
                Stg[vr][c4 + 0] = sn[s][0];
                Stg[vr][c4 + 1] = sn[s][1];
                Stg[vr][c4 + 2] = sn[s][2];
                Stg[vr][c4 + 3] = sn[s][3];
            }
            wave_lds_sync();
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] = Stg[lr][acc_row(r, lane)];
        }
        if (q + 1 < NQ) fetch_v(q + 1);
        // inverse of the diagonal sub-system, loaded early
        const float* Gq = G + q * 1024;
        float g[16];
#pragma unroll
        for (int kp = 0; kp < 16; ++kp) g[kp] = Gq[(2 * kp + lh) * 32 + lr];
        // ---- z_q = v_q + C[q rows][0:32q] * Y[0:32q]   (16q k-pairs)
        const int cnt = 16 * q;
        const float* xc = Xc + (int64_t)lh * ldx + q * 32 + lr;
        float an[PF];
#pragma unroll
        for (int j = 0; j < PF; ++j) an[j] = (j < cnt) ? xc[(int64_t)(2 * j) * ldx] : 0.0f;
        for (int kp = 0; kp < cnt; kp += PF) {
            float ac[PF];
#pragma unroll
            for (int j = 0; j < PF; ++j) ac[j] = an[j];
#pragma unroll
            for (int j = 0; j < PF; ++j) {
                const int kn = kp + PF + j;
                if (kn < cnt) an[j] = xc[(int64_t)(2 * kn) * ldx];
            }
#pragma unroll
            for (int j = 0; j < PF; ++j) acc = mfma32(ac[j], Y[2 * (kp + j) + lh][lr], acc);
        }
        // ---- y_q = G_q z_q
        wave_lds_sync();  // staging reads done before z overwrites the tile
#pragma unroll
        for (int r = 0; r < 16; ++r) Z[acc_row(r, lane)][lr] = acc[r];
        wave_lds_sync();
        f32x16 y = {0};
#pragma unroll
        for (int kp = 0; kp < 16; ++kp) y = mfma32(g[kp], Z[2 * kp + lh][lr], y);
#pragma unroll
        for (int r = 0; r < 16; ++r) Y[q * 32 + acc_row(r, lane)][lr] = y[r];
        // ---- low-precision copy: registers 4g..4g+3 are 4 consecutive k
        if constexpr (!std::is_same<LP, float>::value && !kFp8) {
            LP* dst = (top ? lpT : lpL) + (v0 + lr) * ldlp + q * 32 + 4 * lh;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                struct alignas(8) P4 { LP v[4]; } p4;
#pragma unroll
                for (int e = 0; e < 4; ++e) p4.v[e] = static_cast<LP>(y[4 * g + e]);
                *reinterpret_cast<P4*>(dst + 8 * g) = p4;
            }
        }
        // ---- write back
        if (top) {
#pragma unroll
            for (int r = 0; r < 16; ++r)
                stripT[(int64_t)(q * 32 + acc_row(r, lane)) * lda + v0 + lr] = y[r];
        } else {
            if (Lt)
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    Lt[(int64_t)(q * 32 + acc_row(r, lane)) * ldt + v0 + lr] = y[r];
            wave_lds_sync();
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int vr = (lane >> 3) + 8 * s, c4 = (lane & 7) * 4;
                f32x4 v;
                v[0] = Y[q * 32 + c4 + 0][vr];
                v[1] = Y[q * 32 + c4 + 1][vr];
                v[2] = Y[q * 32 + c4 + 2][vr];
                v[3] = Y[q * 32 + c4 + 3][vr];
                *reinterpret_cast<f32x4*>(stripL + (v0 + vr) * lda + q * 32 + c4) = v;
            }
        }
        wave_lds_sync();
    }
    // ---- FP8 copy with one power-of-two scale per block of 32 vectors x B:
    // lane (lr, lh) quantises vector lr, k in [lh B/2, (lh+1) B/2)
    if constexpr (kFp8) {
        constexpr int KH = B / 2;
        float am = 0.0f;
#pragma unroll 8
        for (int k = lh * KH; k < (lh + 1) * KH; ++k) am = fmaxf(am, fabsf(Y[k][lr]));
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) am = fmaxf(am, __shfl_xor(am, o));
        const int e = fp8_scale_exp(am);
        const float inv = exp2i(-e);
        uint8_t* dst = reinterpret_cast<uint8_t*>(top ? lpT : lpL) + (v0 + lr) * ldlp + lh * KH;
#pragma unroll
        for (int c = 0; c < KH; c += 16) {
            uint32_t w[4];
#pragma unroll
            for (int q4 = 0; q4 < 4; ++q4) {
                const int k = lh * KH + c + 4 * q4;
                w[q4] = fp8_e4m3_code(Y[k][lr] * inv) | (fp8_e4m3_code(Y[k + 1][lr] * inv) << 8) |
                        (fp8_e4m3_code(Y[k + 2][lr] * inv) << 16) |
                        (fp8_e4m3_code(Y[k + 3][lr] * inv) << 24);
            }
            *reinterpret_cast<uint4*>(dst + c) = make_uint4(w[0], w[1], w[2], w[3]);
        }
        if (lane == 0) (top ? scT : scL)[v0 / 32] = exp2i(e);
    }
}

// ----------------------------------------------------------------------------
// The same solve with the solved blocks kept in registers (B <= 256): no LDS
// for Y, so the kernel fits beside the trailing-update GEMM on a CU (the LDS
// version needs B*32*4 bytes per wave).  The k order of each MFMA pair
// follows the accumulator layout: register r of a lane holds row acc_row(r)
// of its block, so the lane supplies that row's y as the B operand and the
// matching coefficient as the A operand -- both z = v + C y and y = G z are
// computed from registers.
template <int B, typename LP>
__global__ __launch_bounds__(64) void trsm_reg_kernel(float* __restrict__ A, int64_t lda,
                                                      int64_t off, int64_t m,
                                                      const float* __restrict__ DT,
                                                      const float* __restrict__ inv,
                                                      float* __restrict__ Lt, int64_t ldt,
                                                      int ngroups_top, LP* __restrict__ lpT,
                                                      LP* __restrict__ lpL, int64_t ldlp,
                                                      float* __restrict__ scT,
                                                      float* __restrict__ scL) {
    constexpr int NQ = B / 32;
    constexpr bool kFp8 = std::is_same<LP, fp8e4m3>::value;
    constexpr bool kLp = !std::is_same<LP, float>::value && !kFp8;
    __shared__ float Tt[32][33];  // transposes of the left strip
    const int lane = threadIdx.x;
    const int lr = lane & 31;
    const bool top = (int)blockIdx.x < ngroups_top;
    const int64_t v0 = (int64_t)(top ? blockIdx.x : blockIdx.x - ngroups_top) * 32;
    const float* D = A + off * lda + off;
    const float* Xc = top ? DT : D;  // coefficients read as X[p][i]
    const int64_t ldx = top ? (int64_t)B : lda;
    const float* G = inv + (top ? 0 : NQ * 1024);
    float* stripT = A + off * lda + off + B;
    float* stripL = A + (off + B) * lda + off;

    f32x16 Y[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        f32x16 acc;
        if (top) {
#pragma unroll
            for (int r = 0; r < 16; ++r)
                acc[r] = stripT[(int64_t)(q * 32 + acc_row(r, lane)) * lda + v0 + lr];
        } else {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int vr = (lane >> 3) + 8 * s, c4 = (lane & 7) * 4;
                const f32x4 v =
                    *reinterpret_cast<const f32x4*>(stripL + (v0 + vr) * lda + q * 32 + c4);
                Tt[vr][c4 + 0] = v[0];
                Tt[vr][c4 + 1] = v[1];
                Tt[vr][c4 + 2] = v[2];
                Tt[vr][c4 + 3] = v[3];
            }
            wave_lds_sync();
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] = Tt[lr][acc_row(r, lane)];
            wave_lds_sync();
        }
        // z = v + sum over solved blocks pb < q of C[q][pb] y_pb
        float c[2][16];
        if (q > 0) {
#pragma unroll
            for (int r = 0; r < 16; ++r)
                c[0][r] = Xc[(int64_t)acc_row(r, lane) * ldx + q * 32 + lr];
        }
#pragma unroll
        for (int pb = 0; pb < q; ++pb) {
            if (pb + 1 < q) {
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    c[(pb + 1) & 1][r] =
                        Xc[(int64_t)((pb + 1) * 32 + acc_row(r, lane)) * ldx + q * 32 + lr];
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) acc = mfma32(c[pb & 1][r], Y[pb][r], acc);
        }
        // y = G_q z  (G[i][k] stored at Gq[k * 32 + i])
        const float* Gq = G + q * 1024;
        f32x16 y = {0};
#pragma unroll
        for (int r = 0; r < 16; ++r) y = mfma32(Gq[acc_row(r, lane) * 32 + lr], acc[r], y);
        Y[q] = y;
        if constexpr (kLp) {
            LP* dst = (top ? lpT : lpL) + (v0 + lr) * ldlp + q * 32 + 4 * (lane >> 5);
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                struct alignas(8) P4 { LP v[4]; } p4;
#pragma unroll
                for (int e = 0; e < 4; ++e) p4.v[e] = static_cast<LP>(y[4 * g + e]);
                *reinterpret_cast<P4*>(dst + 8 * g) = p4;
            }
        }
        if (top) {
#pragma unroll
            for (int r = 0; r < 16; ++r)
                stripT[(int64_t)(q * 32 + acc_row(r, lane)) * lda + v0 + lr] = y[r];
        } else {
            if (Lt)
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    Lt[(int64_t)(q * 32 + acc_row(r, lane)) * ldt + v0 + lr] = y[r];
#pragma unroll
            for (int r = 0; r < 16; ++r) Tt[acc_row(r, lane)][lr] = y[r];
            wave_lds_sync();
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int vr = (lane >> 3) + 8 * s, c4 = (lane & 7) * 4;
                f32x4 v;
                v[0] = Tt[c4 + 0][vr];
                v[1] = Tt[c4 + 1][vr];
                v[2] = Tt[c4 + 2][vr];
                v[3] = Tt[c4 + 3][vr];
                *reinterpret_cast<f32x4*>(stripL + (v0 + vr) * lda + q * 32 + c4) = v;
            }
            wave_lds_sync();
        }
    }
    // FP8 copy: one power-of-two scale for the 32 vectors x B of this wave;
    // registers 4g..4g+3 are 4 consecutive k, stored as one 32-bit word
    if constexpr (kFp8) {
        float am = 0.0f;
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int r = 0; r < 16; ++r) am = fmaxf(am, fabsf(Y[q][r]));
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) am = fmaxf(am, __shfl_xor(am, o));
        const int e = fp8_scale_exp(am);
        const float sinv = exp2i(-e);
        uint8_t* dst = reinterpret_cast<uint8_t*>(top ? lpT : lpL) + (v0 + lr) * ldlp +
                       4 * (lane >> 5);
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const uint32_t w = fp8_e4m3_code(Y[q][4 * g] * sinv) |
                                   (fp8_e4m3_code(Y[q][4 * g + 1] * sinv) << 8) |
                                   (fp8_e4m3_code(Y[q][4 * g + 2] * sinv) << 16) |
                                   (fp8_e4m3_code(Y[q][4 * g + 3] * sinv) << 24);
                *reinterpret_cast<uint32_t*>(dst + q * 32 + 8 * g) = w;
            }
        if (lane == 0) (top ? scT : scL)[v0 / 32] = exp2i(e);
    }
}

// ----------------------------------------------------------------------------
// Full inverses of the diagonal blocks for the substitution.  The strip
// recurrences above, run on unit vectors, give for diagonal block k (with
// L unit lower and U upper as the substitution uses them, M terms):
//   left recurrence on e_r :  y = L^-1 e_r      -> OutL[r][i] = (L^-1)(i, r)
//   top  recurrence on e_c :  y = -(U^T)^-1 e_c -> OutU[i][c] = -(U^-1)(c, i)
// stored per block at out + k * 2 * B * B (OutL, then OutU), both row-major
// B x B, so the substitution applies them as coalesced GEMVs.  Launch: grid
// (B/32 vector groups, 2 types, blocks), one wave each; coefficients are read
// straight from the factor (the transposed copy DT of the panel is gone).
template <int B>
__global__ __launch_bounds__(64) void block_inverse_kernel(const float* __restrict__ A,
                                                           int64_t lda,
                                                           const float* __restrict__ inv,
                                                           int64_t inv_stride,
                                                           float* __restrict__ out) {
    constexpr int NQ = B / 32;
    const int lane = threadIdx.x;
    const int lr = lane & 31;
    const int v0 = blockIdx.x * 32;
    const bool top = blockIdx.y == 0;
    const int64_t k = blockIdx.z;
    const float* D = A + k * B * lda + k * B;
    // coefficient X[p][i]: top C[i][p] = D[i][p], left C[i][p] = D[p][i]
    const int64_t sp = top ? 1 : lda, si = top ? lda : 1;
    const float* G = inv + k * inv_stride + (top ? 0 : NQ * 1024);
    float* o = out + k * 2 * (int64_t)B * B + (top ? (int64_t)B * B : 0);

    // Every load is issued one block ahead of its use (the 32 x 32 inverse of
    // the next block and the first coefficient block of the next column
    // block while this one multiplies), and the stores' layout is selected by
    // address, not by a branch, so the unrolled loop stays one stretch of code
    // the compiler can schedule.
    f32x16 Y[NQ];
    float g[2][16], c[2][16];
    int cb = 0;  // buffer of c holding the next product's coefficients
#pragma unroll
    for (int r = 0; r < 16; ++r) g[0][r] = G[acc_row(r, lane) * 32 + lr];
    if (NQ > 1) {
#pragma unroll
        for (int r = 0; r < 16; ++r) c[0][r] = D[(int64_t)acc_row(r, lane) * sp + (int64_t)(32 + lr) * si];
    }
    const int64_t ostep_i = top ? B : 1, ostep_v = top ? 1 : B;  // o offset per element / vector
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        if (q + 1 < NQ) {
#pragma unroll
            for (int r = 0; r < 16; ++r)
                g[(q + 1) & 1][r] = G[(q + 1) * 1024 + acc_row(r, lane) * 32 + lr];
        }
        f32x16 acc;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = (q * 32 + acc_row(r, lane) == v0 + lr) ? 1.0f : 0.0f;
#pragma unroll
        for (int pb = 0; pb < q; ++pb) {
            // the coefficients of the next product: block pb + 1 of this
            // column block, or block 0 of the next one
            const int nq = pb + 1 < q ? q : q + 1, np = pb + 1 < q ? pb + 1 : 0;
            if (nq < NQ) {
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    c[cb ^ 1][r] = D[(int64_t)(np * 32 + acc_row(r, lane)) * sp +
                                     (int64_t)(nq * 32 + lr) * si];
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) acc = mfma32(c[cb][r], Y[pb][r], acc);
            cb ^= 1;
        }
        f32x16 y = {0};
#pragma unroll
        for (int r = 0; r < 16; ++r) y = mfma32(g[q & 1][r], acc[r], y);
        Y[q] = y;
        // element i = q*32 + acc_row(r) of vector v0 + lr
#pragma unroll
        for (int r = 0; r < 16; ++r)
            o[(int64_t)(q * 32 + acc_row(r, lane)) * ostep_i + (int64_t)(v0 + lr) * ostep_v] = y[r];
    }
}


// ----------------------------------------------------------------------------
// Strip solves of the BF16 factorisation as products with the diagonal
// block's full inverses.  The top and left recurrences above are linear maps
// y = G v with G lower triangular (G_top = OutU, G_left = OutL transposed, as
// block_inverse_kernel writes them), so each strip is one product over
// independent 32 x 32 blocks, y_q = sum_{p <= q} G_qp v_p: no chain from block
// to block, every load of v issued before the first MFMA.  The product runs on
// v_mfma_f32_32x32x16_bf16 with each operand split as x = hi + lo (hi =
// bf16(x), lo = bf16(x - hi), both exact to about 2^-17 |x|) and
// G v = G_hi v_hi + G_hi v_lo + G_lo v_hi.
//
// Operand layout of one 32 x 32 block product (two MFMAs of K = 16): lane
// (i, h) = (l % 32, l / 32) supplies G[i][16 kb + 8 h + e] and, for vector
// i, v[16 kb + 8 h + e], e = 0..7.  tinv_pack_kernel writes G in that layout
// once per step: block (q, p), p <= q, at index q (q + 1) / 2 + p, 2048
// halves each -- [kb][hi, lo][lane][8].
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

__device__ __forceinline__ void split_bf16(const float (&x)[8], bf16x8_t& hi, bf16x8_t& lo) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const __bf16 h = static_cast<__bf16>(x[e]);
        hi[e] = h;
        lo[e] = static_cast<__bf16>(x[e] - static_cast<float>(h));
    }
}

template <int B>
__global__ __launch_bounds__(64) void tinv_pack_kernel(const float* __restrict__ binv,
                                                       bf16x8_t* __restrict__ pk) {
    constexpr int NQ = B / 32, NB = NQ * (NQ + 1) / 2;
    const int bi = blockIdx.x;
    int q = 0;
    while ((q + 1) * (q + 2) / 2 <= bi) ++q;
    const int p = bi - q * (q + 1) / 2;
    const bool top = blockIdx.y == 0;
    const int lane = threadIdx.x, i = lane & 31, h = lane >> 5;
    const float* OutL = binv;
    const float* OutU = binv + B * B;
    bf16x8_t* dst = pk + (int64_t)((top ? 0 : NB) + bi) * 256;
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        float x[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = p * 32 + 16 * kb + 8 * h + e, r = q * 32 + i;
            x[e] = top ? OutU[r * B + k] : OutL[k * B + r];
        }
        bf16x8_t hi, lo;
        split_bf16(x, hi, lo);
        dst[(kb * 2 + 0) * 64 + lane] = hi;
        dst[(kb * 2 + 1) * 64 + lane] = lo;
    }
}

// One wave per 32 vectors: grid.x = ngroups_top top groups, then as many left
// groups.
template <int B, typename LP>
__global__ __launch_bounds__(64) void trsm_inv_kernel(float* __restrict__ A, int64_t lda,
                                                      int64_t off,
                                                      const bf16x8_t* __restrict__ pk,
                                                      int ngroups_top, LP* __restrict__ lpT,
                                                      LP* __restrict__ lpL, int64_t ldlp) {
    constexpr int NQ = B / 32, NB = NQ * (NQ + 1) / 2;
    const int lane = threadIdx.x;
    const int lr = lane & 31, lh = lane >> 5;
    const bool top = (int)blockIdx.x < ngroups_top;
    const int64_t v0 = (int64_t)(top ? blockIdx.x : blockIdx.x - ngroups_top) * 32;
    float* stripT = A + off * lda + off + B;
    float* stripL = A + (off + B) * lda + off;
    const bf16x8_t* G = pk + (top ? 0 : NB * 256) + lane;

    // v of vector v0 + lr, k = 32 p + 16 kb + 8 lh + e, split into halves
    bf16x8_t vh[NQ][2], vl[NQ][2];
    {
        float x[NQ][2][8];
#pragma unroll
        for (int p = 0; p < NQ; ++p)
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) {
                const int k0 = p * 32 + 16 * kb + 8 * lh;
                if (top) {
#pragma unroll
                    for (int e = 0; e < 8; ++e) x[p][kb][e] = stripT[(int64_t)(k0 + e) * lda + v0 + lr];
                } else {
                    const float* src = stripL + (v0 + lr) * lda + k0;
                    const f32x4 a = *reinterpret_cast<const f32x4*>(src);
                    const f32x4 b = *reinterpret_cast<const f32x4*>(src + 4);
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        x[p][kb][e] = a[e];
                        x[p][kb][4 + e] = b[e];
                    }
                }
            }
#pragma unroll
        for (int p = 0; p < NQ; ++p)
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) split_bf16(x[p][kb], vh[p][kb], vl[p][kb]);
    }

#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        f32x16 acc = {0};
#pragma unroll
        for (int p = 0; p <= q; ++p) {
            const bf16x8_t* g = G + (int64_t)(q * (q + 1) / 2 + p) * 256;
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) {
                const bf16x8_t ghi = g[(kb * 2 + 0) * 64], glo = g[(kb * 2 + 1) * 64];
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(glo, vh[p][kb], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ghi, vl[p][kb], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ghi, vh[p][kb], acc, 0, 0, 0);
            }
        }
        // register r: element i = q * 32 + acc_row(r) of vector v0 + lr
        LP* dst = (top ? lpT : lpL) + (v0 + lr) * ldlp + q * 32 + 4 * lh;
#pragma unroll
        for (int g4 = 0; g4 < 4; ++g4) {
            struct alignas(8) P4 {
                LP v[4];
            } p4;
#pragma unroll
            for (int e = 0; e < 4; ++e) p4.v[e] = static_cast<LP>(acc[4 * g4 + e]);
            *reinterpret_cast<P4*>(dst + 8 * g4) = p4;
        }
        if (top) {
#pragma unroll
            for (int r = 0; r < 16; ++r)
                stripT[(int64_t)(q * 32 + acc_row(r, lane)) * lda + v0 + lr] = acc[r];
        } else {
            float* row = stripL + (v0 + lr) * lda + q * 32 + 4 * lh;
#pragma unroll
            for (int g4 = 0; g4 < 4; ++g4) {
                f32x4 v;
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = acc[4 * g4 + e];
                *reinterpret_cast<f32x4*>(row + 8 * g4) = v;
            }
        }
    }
}

}  // namespace

hipError_t launch_block_inverse(const float* A, int64_t lda, int64_t n, int B, const float* inv,
                                int64_t inv_stride, float* out, hipStream_t s) {
    if (n % B) return hipErrorInvalidValue;
    const int NQ = B / 32;
    dim3 grid(NQ, 2, (unsigned)(n / B));
    switch (B) {
        case 32: block_inverse_kernel<32><<<grid, 64, 0, s>>>(A, lda, inv, inv_stride, out); break;
        case 64: block_inverse_kernel<64><<<grid, 64, 0, s>>>(A, lda, inv, inv_stride, out); break;
        case 128: block_inverse_kernel<128><<<grid, 64, 0, s>>>(A, lda, inv, inv_stride, out); break;
        case 256: block_inverse_kernel<256><<<grid, 64, 0, s>>>(A, lda, inv, inv_stride, out); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_prep(const float* A, int64_t lda, int64_t off, int B, float* DT, float* inv,
                       hipStream_t s) {
    if ((lda & 3) || (off & 3) || (reinterpret_cast<uintptr_t>(A) & 15))
        return hipErrorInvalidValue;  // 16-byte row vectors
    const int NQ = B / 32;
    dim3 grid(3 * NQ + (DT ? NQ * NQ : 0), 1);
    prep_kernel<<<grid, 64, 0, s>>>(A, lda, off, B, DT, inv, 0);
    return hipGetLastError();
}

hipError_t launch_prep_all(const float* A, int64_t lda, int B, int nblocks, float* inv,
                           int64_t inv_stride, hipStream_t s) {
    if ((lda & 3) || (reinterpret_cast<uintptr_t>(A) & 15)) return hipErrorInvalidValue;
    const int NQ = B / 32;
    dim3 grid(3 * NQ, nblocks);
    prep_kernel<<<grid, 64, 0, s>>>(A, lda, 0, B, nullptr, inv, inv_stride);
    return hipGetLastError();
}

template <typename LP>
static hipError_t launch_trsm_t(float* A, int64_t lda, int64_t off, int64_t m, int B,
                                const float* DT, const float* inv, float* Lt, int64_t ldt,
                                LP* lpT, LP* lpL, int64_t ldlp, hipStream_t s,
                                float* scT = nullptr, float* scL = nullptr) {
    const int groups = (int)(m / 32);
    dim3 grid(2 * groups);
#define HPL_TRSM_CASE(BB, KERNEL)                                                             \
    case BB:                                                                                  \
        KERNEL<BB, LP><<<grid, 64, 0, s>>>(A, lda, off, m, DT, inv, Lt, ldt, groups, lpT, lpL,  \
                                           ldlp, scT, scL);                                   \
        break;
    switch (B) {
        HPL_TRSM_CASE(32, trsm_reg_kernel)
        HPL_TRSM_CASE(64, trsm_reg_kernel)
        HPL_TRSM_CASE(128, trsm_reg_kernel)
        HPL_TRSM_CASE(256, trsm_reg_kernel)
        HPL_TRSM_CASE(512, trsm_kernel)
        default: return hipErrorInvalidValue;
    }
#undef HPL_TRSM_CASE
    return hipGetLastError();
}

hipError_t launch_trsm_top(float* A, int64_t lda, int64_t off, int64_t m, int B, const float* DT,
                           const float* inv, hipStream_t s) {
    if (m <= 0) return hipSuccess;
    if (m % 32) return hipErrorInvalidValue;
    const int groups = (int)(m / 32);
#define HPL_TRSM_TOP(BB)                                                                      \
    case BB:                                                                                  \
        trsm_reg_kernel<BB, float><<<groups, 64, 0, s>>>(A, lda, off, m, DT, inv, nullptr, 0, \
                                                         groups, nullptr, nullptr, 0, nullptr, \
                                                         nullptr);                            \
        break;
    switch (B) {
        HPL_TRSM_TOP(32)
        HPL_TRSM_TOP(64)
        HPL_TRSM_TOP(128)
        HPL_TRSM_TOP(256)
        default: return hipErrorInvalidValue;
    }
#undef HPL_TRSM_TOP
    return hipGetLastError();
}

hipError_t launch_trsm(float* A, int64_t lda, int64_t off, int64_t m, int B, const float* DT,
                       const float* inv, float* Lt, int64_t ldt, hipStream_t s, int lp_prec,
                       void* lpT, void* lpL, int64_t ldlp, float* scT, float* scL) {
    if (m <= 0) return hipSuccess;
    if (m % 32) return hipErrorInvalidValue;
    switch (lp_prec) {
        case 0:
            return launch_trsm_t<float>(A, lda, off, m, B, DT, inv, Lt, ldt, nullptr, nullptr, 0, s);
        case 1:
            return launch_trsm_t<__bf16>(A, lda, off, m, B, DT, inv, Lt, ldt,
                                         static_cast<__bf16*>(lpT), static_cast<__bf16*>(lpL),
                                         ldlp, s);
        case 2:
            return launch_trsm_t<_Float16>(A, lda, off, m, B, DT, inv, Lt, ldt,
                                           static_cast<_Float16*>(lpT),
                                           static_cast<_Float16*>(lpL), ldlp, s);
        case 3:
            if (!scT || !scL || ldlp % 16) return hipErrorInvalidValue;
            return launch_trsm_t<fp8e4m3>(A, lda, off, m, B, DT, inv, Lt, ldt,
                                          static_cast<fp8e4m3*>(lpT), static_cast<fp8e4m3*>(lpL),
                                          ldlp, s, scT, scL);
        default:
            return hipErrorInvalidValue;
    }
}

int64_t tinv_floats(int B) {
    const int NQ = B / 32;
    // OutL and OutU, then the packed halves (2048 per block, 2 bytes each)
    return 2 * (int64_t)B * B + (int64_t)NQ * (NQ + 1) * 2048 / 2;
}

hipError_t launch_tinv_prepare(const float* A, int64_t lda, int64_t off, int B, const float* inv,
                               float* tinv, float* out, hipStream_t s) {
    if (B % 32 || B > 256 || B < 32 || (reinterpret_cast<uintptr_t>(tinv) & 15))
        return hipErrorInvalidValue;
    const int NQ = B / 32;
    const float* D = A + off * lda + off;
    bf16x8_t* pk = reinterpret_cast<bf16x8_t*>(tinv + 2 * (int64_t)B * B);
    float* o = out ? out : tinv;
    dim3 g1(NQ, 2, 1), g2(NQ * (NQ + 1) / 2, 2);
    switch (B) {
#define HPL_TINV(BB)                                                                     \
    case BB:                                                                             \
        block_inverse_kernel<BB><<<g1, 64, 0, s>>>(D, lda, inv, 0, o);                   \
        tinv_pack_kernel<BB><<<g2, 64, 0, s>>>(o, pk);                                   \
        break;
        HPL_TINV(32)
        HPL_TINV(64)
        HPL_TINV(128)
        HPL_TINV(256)
#undef HPL_TINV
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_trsm_inv(float* A, int64_t lda, int64_t off, int64_t m, int B,
                           const float* tinv, hipStream_t s, int lp_prec, void* lpT, void* lpL,
                           int64_t ldlp) {
    if (m <= 0) return hipSuccess;
    if (m % 32 || (lda & 3) || (off & 3) || (ldlp & 3) || (lp_prec != 1 && lp_prec != 2))
        return hipErrorInvalidValue;
    const int groups = (int)(m / 32);
    const bf16x8_t* pk = reinterpret_cast<const bf16x8_t*>(tinv + 2 * (int64_t)B * B);
    dim3 grid(2 * groups);
#define HPL_TRSM_INV(BB)                                                                      \
    case BB:                                                                                  \
        if (lp_prec == 1)                                                                     \
            trsm_inv_kernel<BB, __bf16><<<grid, 64, 0, s>>>(A, lda, off, pk, groups,           \
                                                            static_cast<__bf16*>(lpT),         \
                                                            static_cast<__bf16*>(lpL), ldlp);  \
        else                                                                                  \
            trsm_inv_kernel<BB, _Float16><<<grid, 64, 0, s>>>(A, lda, off, pk, groups,         \
                                                              static_cast<_Float16*>(lpT),     \
                                                              static_cast<_Float16*>(lpL),     \
                                                              ldlp);                           \
        break;
    switch (B) {
        HPL_TRSM_INV(32)
        HPL_TRSM_INV(64)
        HPL_TRSM_INV(128)
        HPL_TRSM_INV(256)
        default: return hipErrorInvalidValue;
    }
#undef HPL_TRSM_INV
    return hipGetLastError();
}
