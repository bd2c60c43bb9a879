// Mixed-precision trailing update: BF16, FP16 or scaled FP8 operands, FP32
// accumulation.
//
// C[r][c] += sum_k L[r][k] * T[c][k]     (FP32 C, 16-bit L and T)
// L and T are the low-precision copies of the left and top strips written by
// the panel solve, both k-contiguous ("NT"): L[r][k] is row r of M's U-panel,
// T[c][k] row c of M's L-panel.  A step's copy is a slab of Bblk values per
// row; the schedule may apply several steps at once (K = steps * Bblk), their
// slabs sitting `slot` elements apart.
//
// v_mfma_f32_32x32x16_{bf16,f16}: lane l supplies row (l % 32) of the A block
// and column (l % 32) of the B block, 8 consecutive k values (8 * (l / 32) ..
// + 7) each -- one 16-byte LDS read per operand.  LDS rows are padded by 16
// bytes so the eight lanes of a read phase hit distinct banks.  The FP32 C
// tile is the dominant traffic (8 bytes per element per pass), so the kernel
// is sized for streaming: 128 x 128 tiles, 4 waves of 64 x 64, 64-deep k slabs
// double-buffered in LDS, C loaded into the accumulators up front.
#include "common.hpp"
#include "kernels.hpp"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

namespace {

template <typename T>
struct lp_traits;
template <>
struct lp_traits<__bf16> {
    typedef bf16x8 vec;
    static __device__ __forceinline__ f32x16 mfma(vec a, vec b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};
template <>
struct lp_traits<_Float16> {
    typedef f16x8 vec;
    static __device__ __forceinline__ f32x16 mfma(vec a, vec b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};

constexpr int BM = 128, BN = 128, BK = 64, NT = 256, WM = 64, WN = 64;

// Optional second region served by the same launch (workgroups past the
// first region's tiles); empty when M = 0.
struct LpRegion {
    float* C;
    const void* L;
    const void* T;
    const float* scL;
    const float* scT;
    int M, N;
};

// The region of this workgroup, with its tile index and tile count.
template <typename TE>
struct RegionSel {
    float* C;
    const TE* L;
    const TE* T;
    const float* scL;
    const float* scT;
    int M, N, bid, nblk;
};
template <typename TE>
__device__ __forceinline__ RegionSel<TE> pick_region(float* C, const TE* L, const TE* Tt,
                                                     const float* scL, const float* scT, int M,
                                                     int N, const LpRegion& r2) {
    RegionSel<TE> r{C, L, Tt, scL, scT, M, N, (int)blockIdx.x,
                    ((M + BM - 1) / BM) * ((N + BN - 1) / BN)};
    if (r.bid >= r.nblk) {
        r.bid -= r.nblk;
        r.nblk = gridDim.x - r.nblk;
        r.C = r2.C;
        r.L = static_cast<const TE*>(r2.L);
        r.T = static_cast<const TE*>(r2.T);
        r.scL = r2.scL;
        r.scT = r2.scT;
        r.M = r2.M;
        r.N = r2.N;
    }
    return r;
}

// Normalise the pair of regions (first non-empty one first); false when
// both are empty.
bool order_regions(float*& C, const void*& L, const void*& Tt, const float*& scL,
                   const float*& scT, int& M, int& N, LpRegion& r2) {
    if (r2.M <= 0 || r2.N <= 0) r2.M = r2.N = 0;
    if (M <= 0 || N <= 0) {
        if (r2.M == 0) return false;
        C = r2.C;
        L = r2.L;
        Tt = r2.T;
        scL = r2.scL;
        scT = r2.scT;
        M = r2.M;
        N = r2.N;
        r2.M = r2.N = 0;
    }
    return true;
}

int region_tiles(int M, int N) { return ((M + BM - 1) / BM) * ((N + BN - 1) / BN); }

constexpr int LDR = BK + 8;  // padded LDS row (elements)

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

template <typename T>
__global__ __launch_bounds__(NT, 2) void update_lp_kernel(
    float* __restrict__ C, int64_t ldc, const T* __restrict__ L, int64_t l_ld, int64_t l_slot,
    const T* __restrict__ Tt, int64_t t_ld, int64_t t_slot, int M, int N, int K, int Bblk,
    LpRegion r2) {
    typedef typename lp_traits<T>::vec vec;
    __shared__ __attribute__((aligned(16))) T As[2][BM * LDR];
    __shared__ __attribute__((aligned(16))) T Bs[2][BN * LDR];

    const auto rs = pick_region<T>(C, L, Tt, nullptr, nullptr, M, N, r2);
    C = rs.C;
    L = rs.L;
    Tt = rs.T;
    M = rs.M;
    N = rs.N;
    const int bid = rs.bid, nblk = rs.nblk;
    const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
    int tm, tn;
    tile_coords(bid, nblk, tiles_m, tiles_n, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int lr = lane & 31, lh = lane >> 5;
    const bool full = (m0 + BM <= M) && (n0 + BN <= N);

    // global -> LDS: 4 chunks of 16 bytes per operand per thread
    u16x8 ra[4], rb[4];
    auto load_slab = [&](int k0) {
        const int64_t ks = k0 / Bblk, ko = k0 % Bblk;
        const T* lb = L + ks * l_slot + ko;
        const T* tb = Tt + ks * t_slot + ko;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int f = tid + NT * s, row = f >> 3, ch = (f & 7) * 8;
            ra[s] = (full || m0 + row < M)
                        ? *reinterpret_cast<const u16x8*>(lb + (int64_t)(m0 + row) * l_ld + ch)
                        : u16x8{0};
            rb[s] = (full || n0 + row < N)
                        ? *reinterpret_cast<const u16x8*>(tb + (int64_t)(n0 + row) * t_ld + ch)
                        : u16x8{0};
        }
    };
    auto store_slab = [&](int buf) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int f = tid + NT * s, row = f >> 3, ch = (f & 7) * 8;
            *reinterpret_cast<u16x8*>(&As[buf][row * LDR + ch]) = ra[s];
            *reinterpret_cast<u16x8*>(&Bs[buf][row * LDR + ch]) = rb[s];
        }
    };

    f32x16 acc[2][2];
    const int rbase = m0 + wm * WM, cbase = n0 + wn * WN;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = rbase + i * 32 + acc_row(r, lane);
                const int col = cbase + j * 32 + lr;
                acc[i][j][r] = (full || (row < M && col < N)) ? C[(int64_t)row * ldc + col] : 0.0f;
            }

    load_slab(0);
    store_slab(0);
    __syncthreads();
    const int nk = K / BK;
    for (int kc = 0; kc < nk; ++kc) {
        const int buf = kc & 1;
        const bool more = kc + 1 < nk;
        load_slab((more ? kc + 1 : kc) * BK);
#pragma unroll
        for (int ks = 0; ks < BK / 16; ++ks) {
            vec a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                a[i] = *reinterpret_cast<const vec*>(
                    &As[buf][(wm * WM + i * 32 + lr) * LDR + ks * 16 + 8 * lh]);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                b[j] = *reinterpret_cast<const vec*>(
                    &Bs[buf][(wn * WN + j * 32 + lr) * LDR + ks * 16 + 8 * lh]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = lp_traits<T>::mfma(a[i], b[j], acc[i][j]);
        }
        if (more) store_slab(buf ^ 1);
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = rbase + i * 32 + acc_row(r, lane);
                const int col = cbase + j * 32 + lr;
                if (full || (row < M && col < N)) C[(int64_t)row * ldc + col] = acc[i][j][r];
            }
}

template <typename T>
hipError_t launch_t(float* C, int64_t ldc, const void* L, int64_t l_ld, int64_t l_slot,
                    const void* Tt, int64_t t_ld, int64_t t_slot, int M, int N, int K, int Bblk,
                    hipStream_t s, LpRegion r2) {
    const float* scL = nullptr;
    const float* scT = nullptr;
    if (!order_regions(C, L, Tt, scL, scT, M, N, r2)) return hipSuccess;
    if (K % BK || Bblk % BK || (l_ld % 8) || (t_ld % 8)) return hipErrorInvalidValue;
    const int tiles = region_tiles(M, N) + region_tiles(r2.M, r2.N);
    update_lp_kernel<T><<<tiles, NT, 0, s>>>(C, ldc, static_cast<const T*>(L), l_ld, l_slot,
                                            static_cast<const T*>(Tt), t_ld, t_slot, M, N, K, Bblk,
                                            r2);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FP8 E4M3 operands with per-block power-of-two scales.
//
// v_mfma_f32_32x32x16_fp8_fp8: lane l supplies row (l % 32) of the A block and
// column (l % 32) of the B block, 8 consecutive k (8 * (l / 32) .. + 7) as 8
// bytes.  A step (panel) of Bblk k values carries one scale per 32 rows of L
// and per 32 columns of T, so each 32 x 32 accumulator tile of a step has the
// single factor scL * scT: the step is accumulated unscaled in `part` and
// added to the FP32 result with that factor when the step ends.
constexpr int BK8 = 64;          // bytes (= k values) per slab
constexpr int LDR8 = BK8 + 16;   // padded LDS row, bytes

__global__ __launch_bounds__(NT, 2) void update_fp8_kernel(
    float* __restrict__ C, int64_t ldc, const uint8_t* __restrict__ L, int64_t l_ld,
    int64_t l_slot, const uint8_t* __restrict__ Tt, int64_t t_ld, int64_t t_slot,
    const float* __restrict__ scL, const float* __restrict__ scT, int64_t sc_slot, int M, int N,
    int K, int Bblk, LpRegion r2) {
    __shared__ __attribute__((aligned(16))) uint8_t As[2][BM * LDR8];
    __shared__ __attribute__((aligned(16))) uint8_t Bs[2][BN * LDR8];

    const auto rs = pick_region<uint8_t>(C, L, Tt, scL, scT, M, N, r2);
    C = rs.C;
    L = rs.L;
    Tt = rs.T;
    scL = rs.scL;
    scT = rs.scT;
    M = rs.M;
    N = rs.N;
    const int bid = rs.bid, nblk = rs.nblk;
    const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
    int tm, tn;
    tile_coords(bid, nblk, tiles_m, tiles_n, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int lr = lane & 31, lh = lane >> 5;
    const bool full = (m0 + BM <= M) && (n0 + BN <= N);

    // global -> LDS: 2 chunks of 16 bytes per operand per thread
    uint4 ra[2], rb[2];
    auto load_slab = [&](int k0) {
        const int64_t ks = k0 / Bblk, ko = k0 % Bblk;
        const uint8_t* lb = L + ks * l_slot + ko;
        const uint8_t* tb = Tt + ks * t_slot + ko;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int f = tid + NT * s, row = f >> 2, ch = (f & 3) * 16;
            ra[s] = (full || m0 + row < M)
                        ? *reinterpret_cast<const uint4*>(lb + (int64_t)(m0 + row) * l_ld + ch)
                        : make_uint4(0, 0, 0, 0);
            rb[s] = (full || n0 + row < N)
                        ? *reinterpret_cast<const uint4*>(tb + (int64_t)(n0 + row) * t_ld + ch)
                        : make_uint4(0, 0, 0, 0);
        }
    };
    auto store_slab = [&](int buf) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int f = tid + NT * s, row = f >> 2, ch = (f & 3) * 16;
            *reinterpret_cast<uint4*>(&As[buf][row * LDR8 + ch]) = ra[s];
            *reinterpret_cast<uint4*>(&Bs[buf][row * LDR8 + ch]) = rb[s];
        }
    };

    f32x16 acc[2][2], part[2][2];
    const int rbase = m0 + wm * WM, cbase = n0 + wn * WN;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = rbase + i * 32 + acc_row(r, lane);
                const int col = cbase + j * 32 + lr;
                acc[i][j][r] = (full || (row < M && col < N)) ? C[(int64_t)row * ldc + col] : 0.0f;
                part[i][j][r] = 0.0f;
            }

    load_slab(0);
    store_slab(0);
    __syncthreads();
    const int nk = K / BK8, per_step = Bblk / BK8;
    for (int kc = 0; kc < nk; ++kc) {
        const int buf = kc & 1;
        const bool more = kc + 1 < nk;
        load_slab((more ? kc + 1 : kc) * BK8);
#pragma unroll
        for (int ks = 0; ks < BK8 / 16; ++ks) {
            long a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                a[i] = *reinterpret_cast<const long*>(
                    &As[buf][(wm * WM + i * 32 + lr) * LDR8 + ks * 16 + 8 * lh]);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                b[j] = *reinterpret_cast<const long*>(
                    &Bs[buf][(wn * WN + j * 32 + lr) * LDR8 + ks * 16 + 8 * lh]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    part[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(a[i], b[j], part[i][j],
                                                                           0, 0, 0);
        }
        if ((kc + 1) % per_step == 0) {
            // end of a step: scale its tiles into the result
            const int64_t p = kc / per_step;
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int rg = (rbase + i * 32) / 32;
                const float sl = rbase + i * 32 < M ? scL[p * sc_slot + rg] : 0.0f;
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const int cg = (cbase + j * 32) / 32;
                    const float f = cbase + j * 32 < N ? sl * scT[p * sc_slot + cg] : 0.0f;
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        acc[i][j][r] = fmaf(f, part[i][j][r], acc[i][j][r]);
                        part[i][j][r] = 0.0f;
                    }
                }
            }
        }
        if (more) store_slab(buf ^ 1);
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = rbase + i * 32 + acc_row(r, lane);
                const int col = cbase + j * 32 + lr;
                if (full || (row < M && col < N)) C[(int64_t)row * ldc + col] = acc[i][j][r];
            }
}

}  // namespace

hipError_t launch_gemm_update_fp8(float* C, int64_t ldc, const void* L, int64_t l_ld,
                                  int64_t l_slot, const void* Tt, int64_t t_ld, int64_t t_slot,
                                  const float* scL, const float* scT, int64_t sc_slot, int M,
                                  int N, int K, int Bblk, hipStream_t s, const LpRegion2* r2in) {
    LpRegion r2{nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0};
    if (r2in) r2 = LpRegion{r2in->C, r2in->L, r2in->T, r2in->scL, r2in->scT, r2in->M, r2in->N};
    if (!order_regions(C, L, Tt, scL, scT, M, N, r2)) return hipSuccess;
    if (K % Bblk || Bblk % BK8 || (l_ld % 16) || (t_ld % 16) || (M % 32) || (N % 32) ||
        (r2.M % 32) || (r2.N % 32))
        return hipErrorInvalidValue;
    const int tiles = region_tiles(M, N) + region_tiles(r2.M, r2.N);
    update_fp8_kernel<<<tiles, NT, 0, s>>>(C, ldc, static_cast<const uint8_t*>(L), l_ld, l_slot,
                                           static_cast<const uint8_t*>(Tt), t_ld, t_slot, scL, scT,
                                           sc_slot, M, N, K, Bblk, r2);
    return hipGetLastError();
}

hipError_t launch_gemm_update_lp(int prec, float* C, int64_t ldc, const void* L, int64_t l_ld,
                                 int64_t l_slot, const void* Tt, int64_t t_ld, int64_t t_slot,
                                 int M, int N, int K, int Bblk, hipStream_t s,
                                 const LpRegion2* r2in) {
    LpRegion r2{nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0};
    if (r2in) r2 = LpRegion{r2in->C, r2in->L, r2in->T, nullptr, nullptr, r2in->M, r2in->N};
    if (prec == 1)
        return launch_t<__bf16>(C, ldc, L, l_ld, l_slot, Tt, t_ld, t_slot, M, N, K, Bblk, s, r2);
    if (prec == 2)
        return launch_t<_Float16>(C, ldc, L, l_ld, l_slot, Tt, t_ld, t_slot, M, N, K, Bblk, s, r2);
    return hipErrorInvalidValue;
}
