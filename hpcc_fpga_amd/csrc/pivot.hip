// LU with partial pivoting (the reference's gefa_ref / gesl_ref,
// LINPACK/src/host/linpack_data.cpp:72-156, selected by --uniform).
//
// Layout.  The matrix is stored row-major as a[j*lda + i] = M(i, j), so a row
// exchange of M touches two scattered elements of every stored row.  The
// pivoted factorisation therefore runs on the transpose X (X[i][j] = M(i, j),
// obtained in place by transpose_kernel), where an exchange of M rows moves
// two contiguous rows of X, and transposes back at the end.  On X the panel is
// B contiguous floats of each row, the trailing update is the FP32 MFMA GEMM
// with T = X's U-panel rows and Lt = the transposed L-panel, and the U-panel
// solve is the register TRSM with a unit-diagonal copy of the diagonal block.
//
// Panel.  One persistent launch factorises the m x B panel (m = n - k0):
// workgroup g keeps rows [k0 + gR, k0 + gR + R) of the panel in LDS for all B
// columns.  Column c needs the whole column's maximum, so every column ends in
// one grid-wide barrier:
//   * each workgroup publishes its best candidate as a 64-bit key
//     (|value| bits << 32 | ~position: max = largest magnitude, then smallest
//     position, the first-index rule of gefa_ref's strict '<', :80-86) with
//     one atomicMax, and writes its candidate row and (owner of row k) row k
//     to double-buffered slots;
//   * after the barrier every workgroup reads the winning key, copies the
//     winning row, exchanges rows k and p where it owns them, and applies
//     gefa_ref's arithmetic to its rows: multiplier l = M(i,k) * (-1.0/u) in
//     double rounded to float (:96-98), update M(i,j) += l * u_j as a float
//     product and a float sum (:100-105).
// Whole rows move (all B columns), so inside the panel the multipliers follow
// LAPACK's convention (every later exchange applied); the exchanges of the
// other columns follow as one gather of the moved rows.  A factor with only
// one panel is bit-identical to gefa_ref after the exchanges of the earlier
// columns are undone (a data movement done on download).
//
// The barrier spins with a bound: a workgroup that waits longer than about a
// second raises the status and releases every waiter, so every wave exits
// (the host sizes the grid to what is co-resident, so this is an error path).
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace {

constexpr int PNT = 256;                // threads of the panel kernel
constexpr unsigned kAbort = 1u << 30;   // added to the barrier counter on abort
constexpr long kSpinLimit = 1L << 21;   // default polls (each >= ~0.5 us) before abort

// ---------------------------------------------------------------- transpose
// In-place transpose of an n x n matrix: workgroup (J, I), I <= J, exchanges
// tiles (I, J) and (J, I) through LDS (a diagonal tile is transposed alone).
__global__ __launch_bounds__(256) void transpose_kernel(float* __restrict__ A, int64_t ld,
                                                        int64_t n) {
    __shared__ float P[64][65];
    __shared__ float Q[64][65];
    const int I = blockIdx.y, J = blockIdx.x;
    if (I > J) return;
    const int t = threadIdx.x;
    const int64_t i0 = (int64_t)I * 64, j0 = (int64_t)J * 64;
    for (int e = t; e < 4096; e += 256) {
        const int r = e >> 6, c = e & 63;
        if (i0 + r < n && j0 + c < n) P[r][c] = A[(i0 + r) * ld + j0 + c];
        if (I != J && j0 + r < n && i0 + c < n) Q[r][c] = A[(j0 + r) * ld + i0 + c];
    }
    __syncthreads();
    for (int e = t; e < 4096; e += 256) {
        const int r = e >> 6, c = e & 63;
        if (I != J) {
            if (i0 + r < n && j0 + c < n) A[(i0 + r) * ld + j0 + c] = Q[c][r];
            if (j0 + r < n && i0 + c < n) A[(j0 + r) * ld + i0 + c] = P[c][r];
        } else if (i0 + r < n && j0 + c < n) {
            A[(i0 + r) * ld + j0 + c] = P[c][r];
        }
    }
}

// ---------------------------------------------------------------- panel
__device__ __forceinline__ unsigned long long pivot_key(float v, int64_t pos, int64_t k) {
    const float a = fabsf(v);
    // gefa_ref never moves to a NaN (max_val < NaN is false); a NaN on the
    // diagonal keeps row k (every comparison against it is false)
    if (a != a) return pos == k ? ~0ULL : 0ULL;
    return ((unsigned long long)__float_as_uint(a) << 32) |
           (unsigned long long)(0xFFFFFFFFu - (unsigned)pos);
}
__device__ __forceinline__ int64_t key_position(unsigned long long key, int64_t k) {
    return key == ~0ULL ? k : (int64_t)(0xFFFFFFFFu - (unsigned)(key & 0xFFFFFFFFull));
}
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned lo = __shfl_xor((unsigned)(v & 0xFFFFFFFFull), o);
        const unsigned hi = __shfl_xor((unsigned)(v >> 32), o);
        const unsigned long long w = ((unsigned long long)hi << 32) | lo;
        v = w > v ? w : v;
    }
    return v;
}

// LDS: tile[R][B + 1] (column B holds the row's position of origin as int
// bits) | urow[B + 1] (the pivot row of the current column) | lcol[R] (the
// multipliers of the current column).
//
// Barrier of column c: thread 0 of every workgroup folds its key into the
// column's key (atomicMax), waits for that to complete, and counts the
// workgroup in on one counter; every workgroup polls the counter until all G
// have arrived (an abort adds kAbort).  What workgroups exchange -- the
// candidate rows and row k -- is written and read with agent-scope coherent
// accesses, and every wave drains its stores before the arrival, so no L2
// writeback or invalidate is needed (those serialise across the device: with
// them a barrier of 256 workgroups measured 18 us, without 6 us).
__device__ __forceinline__ void publish(float* p, float v) {
    __hip_atomic_store(reinterpret_cast<unsigned*>(p), __float_as_uint(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float fetch(const float* p) {
    return __uint_as_float(__hip_atomic_load(reinterpret_cast<const unsigned*>(p), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT));
}

template <int B>
__global__ __launch_bounds__(PNT) void panel_pivot_kernel(
    float* __restrict__ X, int64_t ldx, int64_t n, int64_t k0, int R, float* __restrict__ pub,
    float* __restrict__ krow, unsigned long long* __restrict__ keys, unsigned* __restrict__ ctr,
    float* __restrict__ Lt, int64_t ldl, float* __restrict__ diagU, int* __restrict__ ipvt,
    int* __restrict__ src, int* __restrict__ status, long spin_limit, unsigned phantom) {
#pragma clang fp contract(off)
    constexpr int LDT = B + 1;
    constexpr int TPC = PNT / B;  // threads per column in the update
    extern __shared__ __attribute__((aligned(16))) float sm[];
    float* tile = sm;
    float* urow = sm + (size_t)R * LDT;
    float* lcol = urow + LDT;
    __shared__ unsigned long long wkey[PNT / 64];
    __shared__ int abort_flag;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int G = gridDim.x, g = blockIdx.x;
    const int64_t p0 = k0 + (int64_t)g * R;  // position of local row 0
    const int rows = (int)min<int64_t>((int64_t)R, n - p0);

    // the workgroup's rows of the panel as float4s, 16 per thread in flight at
    // a time (ldx and k0 are multiples of 4, so the rows are 16-byte aligned)
    {
        constexpr int Q = B / 4;
        const int nv = rows * Q;
        for (int e0 = t; e0 < nv; e0 += 16 * PNT) {
            float4 v[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int e = e0 + u * PNT;
                if (e < nv)
                    v[u] = *reinterpret_cast<const float4*>(X + (p0 + e / Q) * ldx + k0 +
                                                            (e % Q) * 4);
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int e = e0 + u * PNT;
                if (e < nv) {
                    float* d = tile + (e / Q) * LDT + (e % Q) * 4;
                    d[0] = v[u].x;
                    d[1] = v[u].y;
                    d[2] = v[u].z;
                    d[3] = v[u].w;
                }
            }
        }
    }
    for (int r = t; r < rows; r += PNT) tile[r * LDT + B] = __int_as_float((int)(p0 + r));
    __syncthreads();

    // Column c's candidate search, the publication of the candidate row and
    // of row k, and the arrival at column c's barrier.  With upd (every
    // column after the first) the workgroup has applied the previous column's
    // rank-1 update to column c only: the published rows get it on the fly
    // for their columns beyond c (the same float product and sum), and the
    // workgroup updates those columns after arriving, while the other
    // workgroups arrive (a split barrier).
    unsigned long long best = 0;
    auto search_arrive = [&](int c, bool upd) {
        const int64_t k = k0 + c;
        const int buf = c & 1;
        const int64_t rk = k - p0;
        best = 0;
        for (int r = (int)max<int64_t>(0, rk) + t; r < rows; r += PNT)
            best = max(best, pivot_key(tile[r * LDT + c], p0 + r, k));
        best = wave_max_u64(best);
        if (lane == 0) wkey[wave] = best;
        __syncthreads();
        best = wkey[0];
#pragma unroll
        for (int w = 1; w < PNT / 64; ++w) best = max(best, wkey[w]);
        auto cur = [&](int r, int j) {
            float v = tile[r * LDT + j];
            if (upd && j > c && j < B) v = v + lcol[r] * urow[j];
            return v;
        };
        if (best != 0) {
            float* slot = pub + ((int64_t)buf * G + g) * LDT;
            const int rl = (int)(key_position(best, k) - p0);
            for (int j = t; j <= B; j += PNT) publish(slot + j, cur(rl, j));
        }
        if (rk >= 0 && rk < rows)
            for (int j = t; j <= B; j += PNT) publish(krow + buf * LDT + j, cur((int)rk, j));
        __builtin_amdgcn_s_waitcnt(0);  // this wave's stores have completed
        __syncthreads();
        if (t == 0) {
            if (best != 0) {
                atomicMax(keys + c, best);
                __builtin_amdgcn_s_waitcnt(0);
            }
            __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    };
    // wait until all G workgroups have arrived at column c's barrier; true
    // when the panel was aborted (a wait past the spin limit)
    auto wait_all = [&](int c) {
        if (t == 0) {
            // phantom > 0 (a test setting) waits for arrivals that never come
            const unsigned target = (unsigned)(c + 1) * (unsigned)G + phantom;
            unsigned v;
            long spins = 0;
            while ((v = __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) <
                   target) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > spin_limit) {
                    atomicOr(status, 2);
                    v = __hip_atomic_fetch_add(ctr, kAbort, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT) +
                        kAbort;
                    break;
                }
            }
            abort_flag = v >= kAbort;
        }
        __syncthreads();
        return abort_flag != 0;
    };

    search_arrive(0, false);
    for (int c = 0; c < B; ++c) {
        if (wait_all(c)) return;
        const int64_t k = k0 + c;
        const int buf = c & 1;
        const int64_t rk = k - p0;  // local row of position k (may be outside)
        // ---- pivot, exchange
        const unsigned long long kk =
            __hip_atomic_load(keys + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int64_t p = key_position(kk, k);
        if (p < k || p >= n) {  // cannot happen with a complete barrier; never index with it
            if (t == 0) atomicOr(status, 4);
            p = k;
        }
        const int gw = (int)((p - k0) / R);
        const float* wrow = pub + ((int64_t)buf * G + gw) * LDT;
        for (int j = t; j <= B; j += PNT) urow[j] = fetch(wrow + j);
        __syncthreads();
        if (p != k) {
            if (rk >= 0 && rk < rows)
                for (int j = t; j <= B; j += PNT) tile[rk * LDT + j] = urow[j];
            const int64_t rp = p - p0;
            if (rp >= 0 && rp < rows)
                for (int j = t; j <= B; j += PNT) tile[rp * LDT + j] = fetch(krow + buf * LDT + j);
        }
        const float piv = urow[c];
        if (g == 0 && t == 0) {
            ipvt[k] = (int)p;
            diagU[k] = piv;
            if (!(piv != 0.0f) || !__builtin_isfinite(piv)) atomicOr(status, 1);
        }
        __syncthreads();
        // ---- multipliers of the rows below position k (gefa_ref: double
        // reciprocal and product, rounded once), then the rank-1 update as a
        // float product and a float sum
        const int rs1 = (int)max<int64_t>(0, k + 1 - p0);
        const double inv = -1.0 / (double)piv;
        for (int r = rs1 + t; r < rows; r += PNT) {
            const float l = (float)((double)tile[r * LDT + c] * inv);
            tile[r * LDT + c] = l;
            lcol[r] = l;
        }
        __syncthreads();
        if (c + 1 == B) break;
        // column c+1 first, one row per thread: the next search reads it
        {
            const float u1 = urow[c + 1];
            for (int r = rs1 + t; r < rows; r += PNT)
                tile[r * LDT + c + 1] = tile[r * LDT + c + 1] + lcol[r] * u1;
        }
        __syncthreads();
        search_arrive(c + 1, true);
        // the other columns, one column per TPC threads
        const int j = t % B;
        if (j > c + 1) {
            const float uj = urow[j];
            int r = rs1 + t / B;
            for (; r + 7 * TPC < rows; r += 8 * TPC) {
                float v[8], lv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    v[u] = tile[(r + u * TPC) * LDT + j];
                    lv[u] = lcol[r + u * TPC];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) tile[(r + u * TPC) * LDT + j] = v[u] + lv[u] * uj;
            }
            for (; r < rows; r += TPC) tile[r * LDT + j] = tile[r * LDT + j] + lcol[r] * uj;
        }
        __syncthreads();
    }
    // ---- write back: the panel (-1/u_kk on the diagonal, the encoding of
    // the unpivoted path), the transposed multipliers below the block for the
    // trailing update, and each position's row of origin
    for (int e = t; e < rows * B; e += PNT) {
        const int r = e / B, c = e % B;
        float v = tile[r * LDT + c];
        if (p0 + r == k0 + c) v = (float)(-1.0 / (double)v);
        X[(p0 + r) * ldx + k0 + c] = v;
    }
    for (int c = 0; c < B; ++c)
        for (int r = t; r < rows; r += PNT)
            if (p0 + r >= k0 + B) Lt[(int64_t)c * ldl + p0 + r] = tile[r * LDT + c];
    for (int r = t; r < rows; r += PNT) src[p0 + r] = __float_as_int(tile[r * LDT + B]);
}

__global__ void panel_reset_kernel(unsigned long long* keys, int B, unsigned* ctr) {
    for (int i = threadIdx.x; i < B; i += blockDim.x) keys[i] = 0ULL;
    if (threadIdx.x == 0) *ctr = 0u;
}

// ---------------------------------------------------------------- exchanges
// The row moves of panel k0: every position t in [k0, k0+B) and every pivot
// position beyond it (once) whose row of origin src[t] differs from t.
// list[2i] = t, list[2i+1] = src[t]; *count = number of moves.  One workgroup.
// The moves index rows by the panel's outputs (ipvt, src), so they are taken
// only from a panel that completed: if the panel's barrier gave up (status
// bits 2 and 4) ipvt and src may be stale, and no move is listed.  Every
// position read or listed is checked against [k0, n) as well; one outside it
// raises status bit 4 and empties the list.
__global__ __launch_bounds__(256) void move_list_kernel(const int* __restrict__ ipvt,
                                                        const int* __restrict__ src, int64_t k0,
                                                        int64_t n, int B, int* __restrict__ list,
                                                        int* __restrict__ count,
                                                        int* __restrict__ status) {
    __shared__ int piv[256];
    __shared__ int cnt, bad;
    const int c = threadIdx.x;
    // the panel kernel ran before on this stream: status is final here and
    // the same for every thread
    if (*status & 6) {
        if (c == 0) *count = 0;
        return;
    }
    if (c == 0) {
        cnt = 0;
        bad = 0;
    }
    auto inside = [&](int64_t v) { return v >= k0 && v < n; };
    if (c < B) piv[c] = ipvt[k0 + c];
    __syncthreads();
    if (c < B) {
        const int t1 = (int)(k0 + c);
        const int s1 = src[t1];
        if (!inside(s1)) {
            bad = 1;
        } else if (s1 != t1) {
            const int i = atomicAdd(&cnt, 1);
            list[2 * i] = t1;
            list[2 * i + 1] = s1;
        }
        const int t2 = piv[c];
        if (t2 < t1 || t2 >= n) {
            bad = 1;
        } else {
            bool first = t2 >= k0 + B;
            for (int q = 0; q < c && first; ++q) first = piv[q] != t2;
            if (first) {
                const int s2 = src[t2];
                if (!inside(s2)) {
                    bad = 1;
                } else if (s2 != t2) {
                    const int i = atomicAdd(&cnt, 1);
                    list[2 * i] = t2;
                    list[2 * i + 1] = s2;
                }
            }
        }
    }
    __syncthreads();
    if (c == 0) {
        *count = bad ? 0 : cnt;
        if (bad) atomicOr(status, 4);
    }
}

// Y[i][col] = X[src_i][col] over the columns outside [k0, k0+B) (4 per thread)
__global__ __launch_bounds__(256) void move_gather_kernel(const float* __restrict__ X, int64_t ldx,
                                                          int64_t n, int64_t k0, int B,
                                                          const int* __restrict__ list,
                                                          const int* __restrict__ count,
                                                          float* __restrict__ Y, int64_t ldy) {
    const int i = blockIdx.y;
    if (i >= *count) return;
    const int64_t s = list[2 * i + 1];
    const int64_t col = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 4;
    if (col >= n || (col >= k0 && col < k0 + B)) return;
    *reinterpret_cast<f32x4*>(Y + i * ldy + col) =
        *reinterpret_cast<const f32x4*>(X + s * ldx + col);
}

__global__ __launch_bounds__(256) void move_scatter_kernel(float* __restrict__ X, int64_t ldx,
                                                           int64_t n, int64_t k0, int B,
                                                           const int* __restrict__ list,
                                                           const int* __restrict__ count,
                                                           const float* __restrict__ Y,
                                                           int64_t ldy) {
    const int i = blockIdx.y;
    if (i >= *count) return;
    const int64_t d = list[2 * i];
    const int64_t col = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 4;
    if (col >= n || (col >= k0 && col < k0 + B)) return;
    *reinterpret_cast<f32x4*>(X + d * ldx + col) =
        *reinterpret_cast<const f32x4*>(Y + i * ldy + col);
}

// perm[t_i] = perm[src_i] for the moves of one panel (one workgroup).  The
// old values pass through LDS, so every read has completed before the
// barrier and before any write.
__global__ __launch_bounds__(256) void move_perm_kernel(int* __restrict__ perm,
                                                        const int* __restrict__ list,
                                                        const int* __restrict__ count) {
    __shared__ int v[512];
    const int cnt = *count;
    for (int i = threadIdx.x; i < cnt; i += 256) v[i] = perm[list[2 * i + 1]];
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += 256) perm[list[2 * i]] = v[i];
}

__global__ void iota_kernel(int* __restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (int)i;
}

// D[i][j] = X(k0+i, k0+j) below the diagonal, 1 on it, 0 above (B x B)
__global__ void unit_lower_kernel(const float* __restrict__ X, int64_t ldx, int64_t k0, int B,
                                  float* __restrict__ D) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < B * B; e += gridDim.x * blockDim.x) {
        const int i = e / B, j = e % B;
        D[e] = j < i ? X[(k0 + i) * ldx + k0 + j] : (j == i ? 1.0f : 0.0f);
    }
}

// x[i] = b[perm[i]]
__global__ void gather_f32_kernel(float* __restrict__ x, const float* __restrict__ b,
                                  const int* __restrict__ perm, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = b[perm[i]];
}

}  // namespace

// ---------------------------------------------------------------- launchers
hipError_t launch_transpose_square(float* A, int64_t ld, int64_t n, hipStream_t s) {
    const unsigned T = (unsigned)((n + 63) / 64);
    transpose_kernel<<<dim3(T, T), 256, 0, s>>>(A, ld, n);
    return hipGetLastError();
}

int panel_pivot_rows(int B) { return B >= 256 ? 128 : 256; }

size_t panel_pivot_lds(int B) {
    const int R = panel_pivot_rows(B);
    return ((size_t)R * (B + 1) + (B + 1) + R) * sizeof(float);
}

int panel_pivot_groups(int64_t m, int B) {
    const int R = panel_pivot_rows(B);
    return (int)((m + R - 1) / R);
}

hipError_t panel_pivot_prepare(int B, int* max_groups) {
    const size_t lds = panel_pivot_lds(B);
    const void* k = nullptr;
    switch (B) {
        case 32: k = reinterpret_cast<const void*>(panel_pivot_kernel<32>); break;
        case 64: k = reinterpret_cast<const void*>(panel_pivot_kernel<64>); break;
        case 128: k = reinterpret_cast<const void*>(panel_pivot_kernel<128>); break;
        case 256: k = reinterpret_cast<const void*>(panel_pivot_kernel<256>); break;
        default: return hipErrorInvalidValue;
    }
    hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, PNT, lds);
    if (e != hipSuccess) return e;
    int dev = 0, cus = 0;
    e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    *max_groups = per_cu * cus;
    return hipSuccess;
}

hipError_t launch_panel_pivot(float* X, int64_t ldx, int64_t n, int64_t k0, int B, float* pub,
                              float* krow, unsigned long long* keys, unsigned* ctr, float* Lt,
                              int64_t ldl, float* diagU, int* ipvt, int* src, int* status,
                              int max_groups, hipStream_t s, long spin_limit) {
    const int R = panel_pivot_rows(B);
    // spin_limit < 0: give up after -spin_limit polls, waiting for one more
    // workgroup than the grid holds, so that every barrier gives up (the
    // error path, exercised from the host with every kernel exiting)
    const unsigned phantom = spin_limit < 0 ? 1u : 0u;
    spin_limit = spin_limit == 0 ? kSpinLimit : (spin_limit < 0 ? -spin_limit : spin_limit);
    const int64_t m = n - k0;
    const int G = panel_pivot_groups(m, B);
    if (G < 1 || G > max_groups) return hipErrorInvalidConfiguration;
    // 16-byte row vectors
    if ((ldx & 3) || (k0 & 3) || (reinterpret_cast<uintptr_t>(X) & 15)) return hipErrorInvalidValue;
    panel_reset_kernel<<<1, 256, 0, s>>>(keys, B, ctr);
    const size_t lds = panel_pivot_lds(B);
#define HPL_PANEL(BB)                                                                           \
    case BB:                                                                                    \
        panel_pivot_kernel<BB><<<G, PNT, lds, s>>>(X, ldx, n, k0, R, pub, krow, keys, ctr, Lt, \
                                                   ldl, diagU, ipvt, src, status, spin_limit,   \
                                                   phantom);                                    \
        break;
    switch (B) {
        HPL_PANEL(32)
        HPL_PANEL(64)
        HPL_PANEL(128)
        HPL_PANEL(256)
        default: return hipErrorInvalidValue;
    }
#undef HPL_PANEL
    return hipGetLastError();
}

hipError_t launch_row_moves(float* X, int64_t ldx, int64_t n, int64_t k0, int B, const int* ipvt,
                            const int* src, int* list, int* count, float* Y, int64_t ldy,
                            int* perm, int* status, hipStream_t s) {
    if (B < 1 || B > 256 || !status) return hipErrorInvalidValue;
    move_list_kernel<<<1, 256, 0, s>>>(ipvt, src, k0, n, B, list, count, status);
    dim3 g((unsigned)((n + 1023) / 1024), (unsigned)(2 * B));
    move_gather_kernel<<<g, 256, 0, s>>>(X, ldx, n, k0, B, list, count, Y, ldy);
    move_scatter_kernel<<<g, 256, 0, s>>>(X, ldx, n, k0, B, list, count, Y, ldy);
    move_perm_kernel<<<1, 256, 0, s>>>(perm, list, count);
    return hipGetLastError();
}

hipError_t launch_iota(int* p, int64_t n, hipStream_t s) {
    iota_kernel<<<(unsigned)std::min<int64_t>((n + 255) / 256, 1024), 256, 0, s>>>(p, n);
    return hipGetLastError();
}

hipError_t launch_unit_lower(const float* X, int64_t ldx, int64_t k0, int B, float* D,
                             hipStream_t s) {
    unit_lower_kernel<<<(unsigned)((B * B + 255) / 256), 256, 0, s>>>(X, ldx, k0, B, D);
    return hipGetLastError();
}

hipError_t launch_gather_f32(float* x, const float* b, const int* perm, int64_t n, hipStream_t s) {
    gather_f32_kernel<<<(unsigned)std::min<int64_t>((n + 255) / 256, 1024), 256, 0, s>>>(x, b,
                                                                                       perm, n);
    return hipGetLastError();
}
