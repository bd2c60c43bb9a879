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

#include <algorithm>
#include <cstdlib>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

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

// ---------------------------------------------------------------------------
// Wide tiles for the 16-bit update: TBM x TBN workgroup tiles (256 x 256) of
// four 128 x 128 wave tiles, one workgroup per CU.  A wave tile of 4 x 4
// blocks reads 8 operand vectors per 16 MFMAs, half the LDS traffic per MFMA
// of the 64 x 64 wave tiles above, and a 256-wide tile halves the operand
// traffic from L2 per flop.  Slabs of BK k values arrive by LDS DMA
// (global_load_lds, 16 bytes per lane, written lane-linearly) into NBUF
// buffers, NBUF - 1 slabs ahead of the MFMAs, so the DMA latency is covered by
// NBUF - 1 slabs of MFMA work.  Chunk c (16 bytes) of row r sits at position
// c ^ f(r), chosen so that the 32 rows an MFMA operand read touches at one k
// offset fall in distinct banks whether the LDS serves 8 or 16 lanes of a
// 16-byte read per cycle: f(r) = (r & 7) ^ ((r >> 3) & 1) for 8 chunks per row,
// ((r >> 1) ^ (r >> 2)) & 3 for 4.  The launcher sends only regions whose
// sides are multiples of the tile, so every tile is whole.
template <int CPR>
__device__ __forceinline__ int swz(int r) {
    if constexpr (CPR == 8) return (r & 7) ^ ((r >> 3) & 1);
    else return ((r >> 1) ^ (r >> 2)) & 3;
}

// A raw buffer resource over the bytes from p on (stride 0, range 2^31 - 1;
// dword 3 as for every raw gfx9 buffer: 32-bit data format).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void* p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, 0x7fffffff,
                                             0x00020000);
}

// N 16-byte-per-lane loads into LDS from rows of a raw buffer at `base`:
// load j reads byte offset voff + j * jst of each lane and fills the 64
// 16-byte LDS chunks at lds + j * lds_st.  (Buffer resources live in plain
// device functions: held in a lambda of a kernel template, they cost the
// kernel its host-side launch stub.)
template <int N>
__device__ __forceinline__ void dma_rows(const void* base, uint32_t voff, uint32_t jst,
                                         u32x4* lds, int lds_st) {
    const __amdgpu_buffer_rsrc_t r = raw_rsrc(base);
#pragma unroll
    for (int j = 0; j < N; ++j)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            r, (__attribute__((address_space(3))) void*)(lds + j * lds_st), 16, voff, j * jst, 0,
            0);
}

// Two 4-byte-per-lane loads into the LDS words sink[0..63] and sink[64..127]
// (read by nobody): they bring the 128-byte lines at voff and voff + st of
// each lane into L2.
__device__ __forceinline__ void prefetch_rows(const void* base, uint32_t voff, uint32_t st,
                                              unsigned* sink) {
    const __amdgpu_buffer_rsrc_t r = raw_rsrc(base);
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)sink, 4,
                                             voff, 0, 0, 0);
    __builtin_amdgcn_raw_ptr_buffer_load_lds(
        r, (__attribute__((address_space(3))) void*)(sink + 64), 4, voff, st, 0, 0);
}

// s_waitcnt immediate that waits until at most n vector memory operations and
// at most lgkm LDS / scalar memory operations of this wave are outstanding
// (expcnt not waited for)
constexpr int vmcnt_le(int n, int lgkm = 15) {
    return (n & 15) | ((n >> 4) << 14) | (7 << 4) | (lgkm << 8);
}

template <typename T, int TBM, int TBN, int BK, int NBUF>
__global__ __launch_bounds__((TBM / 128) * (TBN / 128) * 64, 1) void update_lpw_kernel(
    float* __restrict__ C, int64_t ldc, const T* __restrict__ L, int64_t l_ld, int64_t l_slot,
    const T* __restrict__ Tt, int64_t t_ld, int64_t t_slot, int M, int N, int K, int Bblk,
    LpRegion r2) {
    typedef typename lp_traits<T>::vec vec;
    constexpr int WM = 128, WN = 128, WAVES_M = TBM / WM, WAVES = WAVES_M * (TBN / WN);
    constexpr int FM = WM / 32, FN = WN / 32;
    constexpr int CPR = BK / 8;      // 16-byte chunks per slab row
    constexpr int RPI = 64 / CPR;    // rows per DMA instruction
    static_assert(CPR == 4 || CPR == 8, "slabs of 32 or 64 k values");
    static_assert(WAVES == 4, "the DMA row mapping assumes four waves");
    constexpr int IA = TBM / (RPI * WAVES), IB = TBN / (RPI * WAVES);  // DMA per wave per slab
    constexpr int KS = BK / 16;      // MFMA k-steps per slab
    __shared__ __attribute__((aligned(16))) u32x4 As[NBUF][TBM * CPR];
    __shared__ __attribute__((aligned(16))) u32x4 Bs[NBUF][TBN * CPR];

    // A grid smaller than the tile count makes the workgroups persistent:
    // workgroup b takes tiles b, b + G, b + 2G, ... (G = gridDim.x, a multiple
    // of 8, so every tile of a workgroup maps to the same XCD's share).
    const int tiles1 = (M / TBM) * (N / TBN);
    const int total = tiles1 + (r2.M > 0 ? (r2.M / TBM) * (r2.N / TBN) : 0);
    float* const C1 = C;
    const T* const L1 = L;
    const T* const T1 = Tt;
    const int M1 = M, N1 = N;
    for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
    if (tile != (int)blockIdx.x) __syncthreads();  // every wave done with the LDS slabs
    C = C1;
    L = L1;
    Tt = T1;
    M = M1;
    N = N1;
    int tiles_m = M / TBM, tiles_n = N / TBN;
    int bid = tile, nblk = tiles1;
    if (bid >= nblk) {  // second region
        bid -= nblk;
        nblk = total - tiles1;
        C = r2.C;
        L = static_cast<const T*>(r2.L);
        Tt = static_cast<const T*>(r2.T);
        M = r2.M;
        N = r2.N;
        tiles_m = M / TBM;
        tiles_n = N / TBN;
    }
    int tm, tn;
    tile_coords(bid, nblk, tiles_m, tiles_n, tm, tn);
    const int m0 = tm * TBM, n0 = tn * TBN;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar LDS addresses
    const int wm = wave % WAVES_M, wn = wave / WAVES_M;
    const int lr = lane & 31, lh = lane >> 5;

    // DMA instruction j of this wave moves slab rows j * 4 * RPI + rl, chunk
    // (lane % CPR) ^ f(row); f only depends on the low row bits, which j does
    // not change, so the per-lane part of the address is one 32-bit byte
    // offset (a VGPR, fixed for the tile) and the rest -- slab, row block j --
    // a scalar base: each DMA is a few scalar instructions and one load.
    // Slab kc lies in strip slot kc / spb (spb = Bblk / BK, a power of two).
    const int rl = wave * RPI + lane / CPR;
    const int ce = ((lane % CPR) ^ swz<CPR>(rl)) * 8;  // element offset in the row
    const uint32_t voffA = (uint32_t)((rl * (int)l_ld + ce) * (int)sizeof(T));
    const uint32_t voffB = (uint32_t)((rl * (int)t_ld + ce) * (int)sizeof(T));
    const uint32_t jstA = (uint32_t)(WAVES * RPI * l_ld * (int64_t)sizeof(T));  // row block j
    const uint32_t jstB = (uint32_t)(WAVES * RPI * t_ld * (int64_t)sizeof(T));
    const int spb_sh = __builtin_ctz((unsigned)(Bblk / BK)), spb_m = Bblk / BK - 1;
    // The slab's rows as raw buffers (base = the slab's first row of the
    // tile): buffer loads into LDS take the base and the row block's offset in
    // scalar registers and the lane's offset in one VGPR.  (Unlike the FLAT
    // form they count in vmcnt only, so the compiler can wait for just the
    // LDS reads a k-step needs.)
    auto slab_base = [&](const T* base, int kcs, int64_t slot, int64_t ld, int r0) {
        const int64_t ks = kcs >> spb_sh, ko = (int64_t)(kcs & spb_m) * BK;
        return base + ks * slot + ko + (int64_t)r0 * ld;
    };
    auto dma_slab = [&](int kcs, int buf) {
        dma_rows<IA>(slab_base(L, kcs, l_slot, l_ld, m0), voffA, jstA, &As[buf][wave * 64],
                     WAVES * 64);
        dma_rows<IB>(slab_base(Tt, kcs, t_slot, t_ld, n0), voffB, jstB, &Bs[buf][wave * 64],
                     WAVES * 64);
    };

    // L2 prefetch of the slab NBUF slabs ahead: one 4-byte DMA per 128-byte
    // row segment of that slab (waves 0-1 the L rows, waves 2-3 the T rows,
    // two rows per lane) into a scratch area of LDS nobody reads, so that the
    // slab's own DMA, issued a slab later, finds its rows in L2.  A DMA
    // returns no register, so the prefetch never holds up the MFMAs; it only
    // counts in vmcnt, which the waits below allow for.
    static_assert(TBM == 256 && TBN == 256, "the prefetch mapping assumes 256 x 256 tiles");
    __shared__ __attribute__((aligned(16))) unsigned pf_sink[4][2][64];
    const int pf_r = (wave & 1) * 128 + lane;  // the lane's first row
    const uint32_t pf_off = (uint32_t)(pf_r * (wave < 2 ? l_ld : t_ld) * (int64_t)sizeof(T));
    const uint32_t pf_st = (uint32_t)(64 * (wave < 2 ? l_ld : t_ld) * (int64_t)sizeof(T));
    auto prefetch_slab = [&](int kcs) {
        prefetch_rows(wave < 2 ? slab_base(L, kcs, l_slot, l_ld, m0)
                               : slab_base(Tt, kcs, t_slot, t_ld, n0),
                      pf_off, pf_st, &pf_sink[wave][0][0]);
    };
    constexpr int PFN = 2;  // prefetch DMAs per wave and slab

    const int nk = K / BK;
    // the first NBUF - 1 slabs go out before the loads of C
#pragma unroll
    for (int p = 0; p + 1 < NBUF; ++p)
        if (p < nk) dma_slab(p, p);

    // The MFMAs take T as their A operand and L as their B operand, so each
    // 32 x 32 accumulator block holds a block of C transposed: register
    // 4 q + e of block (i, j) is C row rbase + 32 i + lane % 32, column
    // cbase + 32 j + 8 q + 4 (lane / 32) + e.  Four consecutive registers are
    // four consecutive floats of a row, so C moves as 16-byte vectors: one
    // row pointer per i, the 16 vectors of a row at immediate offsets.  (The
    // products and their k order are those of the 128-tile kernel, which puts
    // L first; the accumulators start from C in both, so both give the same
    // bits.)  The row stride, read back from LDS, lives in a VGPR, so the row
    // offsets are formed per use rather than kept as scalars.
    const int rbase = m0 + wm * WM, cbase = n0 + wn * WN;
    __shared__ int sh_ldc;
    if (tid == 0) sh_ldc = (int)ldc;
    __syncthreads();
    const int ldc_v = sh_ldc;
    float* cb = C + (int64_t)(rbase + lr) * ldc + cbase + 4 * lh;
    f32x16 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        const float* p = cb + (int64_t)(i * 32) * ldc_v;
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const f32x4 v = *reinterpret_cast<const f32x4*>(p + j * 32 + 8 * q);
#pragma unroll
                for (int e = 0; e < 4; ++e) acc[i][j][4 * q + e] = v[e];
            }
    }

    u32x4 a[2][FM], b[2][FN];
    const int sl = swz<CPR>(lr);  // f of every row this lane reads
    auto read_ops = [&](int buf, int ks, int slot) {
        const int c = (2 * ks + lh) ^ sl;
#pragma unroll
        for (int i = 0; i < FM; ++i) a[slot][i] = As[buf][(wm * WM + i * 32 + lr) * CPR + c];
#pragma unroll
        for (int j = 0; j < FN; ++j) b[slot][j] = Bs[buf][(wn * WN + j * 32 + lr) * CPR + c];
    };
    auto mfmas = [&](int slot) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
                acc[i][j] = lp_traits<T>::mfma(__builtin_bit_cast(vec, b[slot][j]),
                                               __builtin_bit_cast(vec, a[slot][i]), acc[i][j]);
    };

    __builtin_amdgcn_s_waitcnt(0);  // C and the first slab
    __syncthreads();
    read_ops(0, 0, 0);
    for (int kc = 0; kc < nk; ++kc) {
        const int buf = kc % NBUF;
        // DMA of slab kc + NBUF - 1 into the buffer of slab kc - 1 (every
        // wave finished reading it before the barrier that ended that slab)
        // and the L2 prefetch of slab kc + NBUF (clamped to the last slab).
        // (In a block of its own: issued in the same block as the LDS reads,
        // the DMA makes the compiler wait for it before the first read.)
        const bool dma_now = kc + NBUF - 1 < nk;
        if (dma_now) {
            dma_slab(kc + NBUF - 1, (kc + NBUF - 1) % NBUF);
            prefetch_slab(min(kc + NBUF, nk - 1));
        }
#pragma unroll
        for (int ks = 0; ks + 1 < KS; ++ks) {
            read_ops(buf, ks + 1, (ks & 1) ^ 1);
            mfmas(ks & 1);
        }
        // Each k-step's LDS reads go out between its predecessor's MFMAs,
        // after the first of them: the wait for a step's operands then sits
        // in front of that step's first MFMA, a whole step after the reads
        // were issued.  (While a DMA is in flight the compiler can only wait
        // for all LDS reads at once, so reads issued right before that wait
        // would stall the MFMAs for their full latency.)
#pragma unroll
        for (int ks = 0; ks + 1 < KS; ++ks) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
#pragma unroll
            for (int q = 0; q < FM + FN; ++q) {
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
            }
            __builtin_amdgcn_sched_group_barrier(0x008, FM * FN - 1 - (FM + FN), 0);
        }
        if (kc + 1 < nk) {
            // slab kc + 1 landed -- at most the DMA of the slabs after it and
            // this slab's prefetch (issued after them) still in flight -- and
            // this wave's LDS reads of slab kc are done; then a bare barrier
            // (no fence: that would also wait for the prefetch)
            if (dma_now)
                __builtin_amdgcn_s_waitcnt(vmcnt_le((NBUF - 2) * (IA + IB) + PFN, 0));
            else
                __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_s_barrier();
            read_ops((kc + 1) % NBUF, 0, KS & 1);
        }
        mfmas((KS - 1) & 1);
    }
    __builtin_amdgcn_s_waitcnt(0);

#pragma unroll
    for (int i = 0; i < FM; ++i) {
        float* p = cb + (int64_t)(i * 32) * ldc_v;
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                f32x4 v;
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = acc[i][j][4 * q + e];
                *reinterpret_cast<f32x4*>(p + j * 32 + 8 * q) = v;
            }
    }
    }  // tiles
}

// The wide-tile kernel is the 16-bit update; HPL_LP_TILES=128 selects the
// 128 x 128 one (for measurements).
// Slab depth of the wide kernel: 64 k values in two buffers (default) or 32
// in four (HPL_LP_SLAB=32, three slabs of DMA ahead instead of one).  At
// M = N = 16384 the first runs at 950 / 1109 / 729 TFLOP/s for K = 3072 /
// 6144 / 1536 (928 / 993 / 710 without the L2 prefetch); the second ran at
// 713 / 761 / 570: the extra barriers of the shorter slabs cost more than the
// deeper prefetch gains.
int lp_wide_slab() {
    static const int k = [] {
        const char* e = getenv("HPL_LP_SLAB");
        return e && atoi(e) == 32 ? 32 : 64;
    }();
    return k;
}

bool lp_wide() {
    static const int w = [] {
        const char* e = getenv("HPL_LP_TILES");
        return e && atoi(e) == 128 ? 0 : 1;
    }();
    return w != 0;
}

template <typename T>
hipError_t launch_t(float* C, int64_t ldc, const void* L, int64_t l_ld, int64_t l_slot,
                    const void* Tt, int64_t t_ld, int64_t t_slot, int M, int N, int K, int Bblk,
                    hipStream_t s, LpRegion r2, int max_groups) {
    const float* scL = nullptr;
    const float* scT = nullptr;
    if (!order_regions(C, L, Tt, scL, scT, M, N, r2)) return hipSuccess;
    if (K % BK || Bblk % BK || (l_ld % 8) || (t_ld % 8)) return hipErrorInvalidValue;
    constexpr int W = 256;
    // the wide kernel moves C as 16-byte row vectors
    const bool c_vec = (ldc % 4) == 0 && (reinterpret_cast<uintptr_t>(C) & 15) == 0 &&
                       (reinterpret_cast<uintptr_t>(r2.C) & 15) == 0;
    // Regions of fewer than 128 wide tiles (half the CUs) take the 128 x 128
    // tiles: four times the workgroups, so a small update on the critical
    // path (a diagonal block, the last row and column blocks) ends sooner.
    // Both kernels give the same bits.
    const int wide_tiles = (M / W) * (N / W) + (r2.M > 0 ? (r2.M / W) * (r2.N / W) : 0);
    if (lp_wide() && c_vec && M % W == 0 && N % W == 0 && r2.M % W == 0 && r2.N % W == 0 &&
        wide_tiles >= 128 && (Bblk & (Bblk - 1)) == 0) {
        // at most max_groups persistent workgroups (rounded down to a
        // multiple of 8), which leaves the other CUs to concurrent launches
        int tiles = wide_tiles;
        if (max_groups > 0 && max_groups < tiles) tiles = std::max(8, max_groups & ~7);
        if (lp_wide_slab() == 64)
            update_lpw_kernel<T, W, W, 64, 2><<<tiles, (W / 128) * (W / 128) * 64, 0, s>>>(
                C, ldc, static_cast<const T*>(L), l_ld, l_slot, static_cast<const T*>(Tt), t_ld,
                t_slot, M, N, K, Bblk, r2);
        else
            update_lpw_kernel<T, W, W, 32, 4><<<tiles, (W / 128) * (W / 128) * 64, 0, s>>>(
                C, ldc, static_cast<const T*>(L), l_ld, l_slot, static_cast<const T*>(Tt), t_ld,
                t_slot, M, N, K, Bblk, r2);
        return hipGetLastError();
    }
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
                                 const LpRegion2* r2in, int max_groups) {
    LpRegion r2{nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0};
    if (r2in) r2 = LpRegion{r2in->C, r2in->L, r2in->T, nullptr, nullptr, r2in->M, r2in->N};
    if (prec == 1)
        return launch_t<__bf16>(C, ldc, L, l_ld, l_slot, Tt, t_ld, t_slot, M, N, K, Bblk, s, r2,
                                max_groups);
    if (prec == 2)
        return launch_t<_Float16>(C, ldc, L, l_ld, l_slot, Tt, t_ld, t_slot, M, N, K, Bblk, s, r2,
                                  max_groups);
    return hipErrorInvalidValue;
}
