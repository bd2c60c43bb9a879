// FP32 MFMA trailing-update kernel template (included by gemm.hip).
//
// C[r][c] += sum_k At[k][r] * Bt[k][c]  (r < M, c < N, k < K), all row-major,
// both operands k-major.  Workgroup tile BM x BN made of (BM/WM) x (BN/WN)
// wave64 tiles; a wave tile WM x WN is (WM/32) x (WN/32) blocks of
// v_mfma_f32_32x32x2_f32.  k-slabs of BK are staged through double-buffered
// LDS (row segments: float4 global loads, conflict-free LDS rows, MFMA operand
// of lane l = LDS[k0 + l/32][m0 + l%32]); the next slab is prefetched into
// registers while the current one is multiplied.  MFMA operands are read one
// k-pair ahead into a second register set, and the hand-over to the next slab
// (LDS write + the single barrier per slab) is issued before the slab's last
// MFMAs, so neither LDS latency nor the barrier stalls the MFMA pipe.  The accumulators start at
// zero and the C tile, prefetched into registers two slabs before the end, is
// added in the epilogue, so its HBM latency overlaps the MFMA work (CPRE);
// without CPRE the C tile is loaded into the accumulators up front (half the
// registers, for large wave tiles).
#pragma once

#include "common.hpp"

namespace gemm_f32 {

// XCD-aware tile order: workgroup ids are dealt round-robin to the 8 XCDs, so
// ids are remapped to give each XCD a contiguous run of tiles, walked in
// groups of 8 tile-rows so a run shares operand rows in that XCD's L2.
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

// A second, independent region of the same shape class (same leading
// dimensions and K) served by the same launch: workgroups past the first
// region's tiles take the second region's.  Empty when M = 0.
struct Region {
    float* C;
    const float* At;
    const float* Bt;
    int M, N;
};

// DMA: stage the slabs with global_load_lds (16 bytes per lane, written to LDS
// lane-linearly at a wave-uniform base), which needs no staging registers and
// no LDS write instructions; each wave moves whole 1 KiB pieces of a slab.
template <int BM, int BN, int BK, int WM, int WN, int MINB, bool CPRE, bool TRANS = false,
          bool DMA = false>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64, MINB) void update_kernel(
    float* __restrict__ C, int64_t ldc, const float* __restrict__ At, int64_t lda,
    const float* __restrict__ Bt, int64_t ldb, int M, int N, int K, Region r2) {
    constexpr int WAVES_M = BM / WM, WAVES_N = BN / WN;
    constexpr int NT = WAVES_M * WAVES_N * 64;
    constexpr int FM = WM / 32, FN = WN / 32;  // MFMA blocks per wave
    constexpr int A4 = BK * BM / 4 / NT;       // float4 per thread per slab (A)
    constexpr int B4 = BK * BN / 4 / NT;
    static_assert(A4 >= 1 && B4 >= 1 && A4 * NT * 4 == BK * BM && B4 * NT * 4 == BK * BN,
                  "slab must split evenly over the threads");
    __shared__ float As[2][BK][BM];
    __shared__ float Bs[2][BK][BN];

    int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
    int bid = blockIdx.x, nblk = tiles_m * tiles_n;
    if (bid >= nblk) {  // second region
        bid -= nblk;
        nblk = gridDim.x - nblk;
        C = r2.C;
        At = r2.At;
        Bt = r2.Bt;
        M = r2.M;
        N = r2.N;
        tiles_m = (M + BM - 1) / BM;
        tiles_n = (N + BN - 1) / BN;
    }
    int tm, tn;
    tile_coords(bid, nblk, tiles_m, tiles_n, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WAVES_M, wn = wave / WAVES_M;
    const int lr = lane & 31, lh = lane >> 5;
    const bool full = (m0 + BM <= M) && (n0 + BN <= N);

    f32x4 ra[DMA ? 1 : A4], rb[DMA ? 1 : B4];
    // DMA: float4 index f = tid + NT * s of a slab lands at LDS float 4 * f,
    // i.e. wave w's lanes fill floats [256 (w + NT/64 s), +256) in order
    auto dma_slab = [&](int k0, int buf) {
#pragma unroll
        for (int s = 0; s < A4; ++s) {
            const int f = tid + NT * s;
            const int kk = f / (BM / 4), cc = min((f % (BM / 4)) * 4, M - m0 - 4);
            __builtin_amdgcn_global_load_lds(
                At + (int64_t)(k0 + kk) * lda + m0 + cc,
                (__attribute__((address_space(3))) void*)(&As[buf][0][0] + 256 * (wave + (NT / 64) * s)),
                16, 0, 0);
        }
#pragma unroll
        for (int s = 0; s < B4; ++s) {
            const int f = tid + NT * s;
            const int kk = f / (BN / 4), cc = min((f % (BN / 4)) * 4, N - n0 - 4);
            __builtin_amdgcn_global_load_lds(
                Bt + (int64_t)(k0 + kk) * ldb + n0 + cc,
                (__attribute__((address_space(3))) void*)(&Bs[buf][0][0] + 256 * (wave + (NT / 64) * s)),
                16, 0, 0);
        }
    };
    auto load_slab = [&](int k0) {
#pragma unroll
        for (int s = 0; s < A4; ++s) {
            const int f = tid + NT * s;
            const int kk = f / (BM / 4), cc = (f % (BM / 4)) * 4;
            ra[s] = (full || m0 + cc < M)
                        ? *reinterpret_cast<const f32x4*>(At + (int64_t)(k0 + kk) * lda + m0 + cc)
                        : f32x4{0, 0, 0, 0};
        }
#pragma unroll
        for (int s = 0; s < B4; ++s) {
            const int f = tid + NT * s;
            const int kk = f / (BN / 4), cc = (f % (BN / 4)) * 4;
            rb[s] = (full || n0 + cc < N)
                        ? *reinterpret_cast<const f32x4*>(Bt + (int64_t)(k0 + kk) * ldb + n0 + cc)
                        : f32x4{0, 0, 0, 0};
        }
    };
    auto store_slab = [&](int buf) {
#pragma unroll
        for (int s = 0; s < A4; ++s) {
            const int f = tid + NT * s;
            *reinterpret_cast<f32x4*>(&As[buf][f / (BM / 4)][(f % (BM / 4)) * 4]) = ra[s];
        }
#pragma unroll
        for (int s = 0; s < B4; ++s) {
            const int f = tid + NT * s;
            *reinterpret_cast<f32x4*>(&Bs[buf][f / (BN / 4)][(f % (BN / 4)) * 4]) = rb[s];
        }
    };

    const int rbase = m0 + wm * WM, cbase = n0 + wn * WN;
    // accumulator register r of block (i, j) -> C element (row, col).  TRANS
    // multiplies T^T Lt^T instead, so the 4 registers of a group are 4
    // consecutive columns of one row (float4 C traffic).
    auto c_row = [&](int i, int r) {
        return TRANS ? rbase + i * 32 + lr : rbase + i * 32 + acc_row(r, lane);
    };
    auto c_col = [&](int j, int r) {
        return TRANS ? cbase + j * 32 + acc_row(r, lane) : cbase + j * 32 + lr;
    };
    f32x16 acc[FM][FN];
    float cpre[CPRE ? FM : 1][CPRE ? FN : 1][16];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = c_row(i, r);
                const int col = c_col(j, r);
                acc[i][j][r] = (!CPRE && (full || (row < M && col < N)))
                                   ? C[(int64_t)row * ldc + col]
                                   : 0.0f;
            }
    auto load_c = [&]() {
        if constexpr (CPRE) {
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int row = c_row(i, r);
                        const int col = c_col(j, r);
                        cpre[i][j][r] = (full || (row < M && col < N))
                                            ? C[(int64_t)row * ldc + col]
                                            : 0.0f;
                    }
        }
    };

    // operand registers, double-buffered across k-pairs
    float a[2][FM], b[2][FN];
    auto read_ops = [&](int buf, int kk, int slot) {
        const int kr = 2 * kk + lh;
#pragma unroll
        for (int i = 0; i < FM; ++i) a[slot][i] = As[buf][kr][wm * WM + i * 32 + lr];
#pragma unroll
        for (int j = 0; j < FN; ++j) b[slot][j] = Bs[buf][kr][wn * WN + j * 32 + lr];
    };

    auto mfmas = [&](int slot) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
                acc[i][j] = TRANS ? mfma32(b[slot][j], a[slot][i], acc[i][j])
                                  : mfma32(a[slot][i], b[slot][j], acc[i][j]);
    };

    if constexpr (DMA) {
        dma_slab(0, 0);
        __builtin_amdgcn_s_waitcnt(0);
    } else {
        load_slab(0);
        store_slab(0);
    }
    __syncthreads();
    const int nk = K / BK;
    constexpr int KP = BK / 2;  // k-pairs per slab
    const int kc_c = nk >= 2 ? nk - 2 : 0;  // slab during which C is prefetched
    read_ops(0, 0, 0);
    for (int kc = 0; kc < nk; ++kc) {
        const int buf = kc & 1;
        const bool more = kc + 1 < nk;
        // the last slab reloads itself (in bounds, unused) to stay branch-free
        if constexpr (DMA) {
            if (more) dma_slab((kc + 1) * BK, buf ^ 1);
        } else {
            load_slab((more ? kc + 1 : kc) * BK);
        }
        if (kc == kc_c) load_c();
        // k-pairs 0 .. KP-2: read the next operands, then multiply the current
#pragma unroll
        for (int kk = 0; kk + 1 < KP; ++kk) {
            read_ops(buf, kk + 1, (kk & 1) ^ 1);
            mfmas(kk & 1);
        }
#pragma unroll
        for (int kk = 0; kk + 1 < KP; ++kk) {
            __builtin_amdgcn_sched_group_barrier(0x100, FM + FN > 2 ? 2 : 1, 0);  // DS read
            __builtin_amdgcn_sched_group_barrier(0x008, FM * FN, 0);              // MFMA
        }
        // last k-pair: hand over to the next slab while its MFMAs run
        if (more) {
            if constexpr (DMA)
                __builtin_amdgcn_s_waitcnt(0);
            else
                store_slab(buf ^ 1);
            __syncthreads();
            read_ops(buf ^ 1, 0, KP & 1);
        }
        mfmas((KP - 1) & 1);
    }

#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = c_row(i, r);
                const int col = c_col(j, r);
                if (full || (row < M && col < N)) {
                    if constexpr (CPRE)
                        C[(int64_t)row * ldc + col] = cpre[i][j][r] + acc[i][j][r];
                    else
                        C[(int64_t)row * ldc + col] = acc[i][j][r];
                }
            }
}

// One launch over region (C, At, Bt, M, N) and the optional second region r2.
template <int BM, int BN, int BK, int WM, int WN, int MINB, bool CPRE, bool TRANS = false,
          bool DMA = false>
inline hipError_t launch2(float* C, int64_t ldc, const float* At, int64_t lda, const float* Bt,
                          int64_t ldb, int M, int N, int K, Region r2, hipStream_t s) {
    if (r2.M <= 0 || r2.N <= 0) r2.M = r2.N = 0;
    if (M <= 0 || N <= 0) {  // only the second region (or nothing)
        if (r2.M == 0) return hipSuccess;
        C = r2.C;
        At = r2.At;
        Bt = r2.Bt;
        M = r2.M;
        N = r2.N;
        r2.M = r2.N = 0;
    }
    if (K % BK != 0 || (M % 4) || (N % 4) || (r2.M % 4) || (r2.N % 4)) return hipErrorInvalidValue;
    constexpr int NT = (BM / WM) * (BN / WN) * 64;
    const int tiles = ((M + BM - 1) / BM) * ((N + BN - 1) / BN) +
                      ((r2.M + BM - 1) / BM) * ((r2.N + BN - 1) / BN);
    update_kernel<BM, BN, BK, WM, WN, MINB, CPRE, TRANS, DMA>
        <<<tiles, NT, 0, s>>>(C, ldc, At, lda, Bt, ldb, M, N, K, r2);
    return hipGetLastError();
}

template <int BM, int BN, int BK, int WM, int WN, int MINB, bool CPRE, bool TRANS = false,
          bool DMA = false>
inline hipError_t launch(float* C, int64_t ldc, const float* At, int64_t lda, const float* Bt,
                         int64_t ldb, int M, int N, int K, hipStream_t s) {
    return launch2<BM, BN, BK, WM, WN, MINB, CPRE, TRANS, DMA>(C, ldc, At, lda, Bt, ldb, M, N, K,
                                                          Region{nullptr, nullptr, nullptr, 0, 0},
                                                          s);
}

}  // namespace gemm_f32
