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
// one grid-wide exchange:
//   * each workgroup publishes its best candidate as a 64-bit key
//     (|value| bits << 32 | ~position: max = largest magnitude, then smallest
//     position, the first-index rule of gefa_ref's strict '<', :80-86), the
//     candidate's values at columns c and c+1, its whole candidate row and
//     (owner of row k) row k and its values at c and c+1;
//   * every workgroup polls all G key slots; once each holds column c's key,
//     it knows the winner and the values at c and c+1 of the winning row and
//     of row k, forms the multipliers of column c and column c+1 and searches
//     column c+1 while the two whole rows it needs are still in flight;
//   * it applies gefa_ref's arithmetic to its rows: multiplier l = M(i,k) *
//     (-1.0/u) in double rounded to float (:96-98), update M(i,j) += l * u_j as
//     a float product and a float sum (:100-105).
// Every word exchanged is 64 bits, (tag << 32) | payload, with tag = the
// global column + 1, written and read with single-copy-atomic agent-scope
// accesses.  A reader accepts a word only when it carries the column's tag,
// so a writer never waits for its stores to complete and no counter, fence
// or atomic read-modify-write is needed: one store-to-load trip per column
// plus the fetch of the winning row.  Slots alternate by column parity: a
// slot of column c is rewritten at column c+2, after every workgroup has
// arrived at column c+1, which each does only after reading column c's
// words.  The host zeroes the slots before each factorisation, so tags are
// unique since the last clear.
// Whole rows move (all B columns), so inside the panel the multipliers follow
// LAPACK's convention (every later exchange applied); the exchanges of the
// other columns follow as one gather of the moved rows.  A factor with only
// one panel is bit-identical to gefa_ref after the exchanges of the earlier
// columns are undone (a data movement done on download).
//
// Every wait spins with a bound: a workgroup that waits longer than about a
// second raises the status, marks its key slots with the abort tag (which
// every waiter takes as a release) and exits, so every wave exits (the host
// sizes the grid to what is co-resident, so this is an error path).
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace {

constexpr int PNT = 512;                   // threads of the panel kernel
constexpr int kMaxPanelGroups = 4 * PNT;   // key slots one workgroup polls
constexpr int kSlotWords = 8;              // words of a key slot (6 used)
constexpr unsigned kAbortTag = 0xFFFFFFFFu;
constexpr long kSpinLimit = 1L << 21;      // default polls (each >= ~0.5 us) before abort

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

// 64-bit exchange words: (tag << 32) | payload
__device__ __forceinline__ void put_tagged(unsigned long long* p, unsigned payload, unsigned tag) {
    __hip_atomic_store(p, ((unsigned long long)tag << 32) | payload, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long get_tagged(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned word_tag(unsigned long long w) { return (unsigned)(w >> 32); }
__device__ __forceinline__ unsigned word_payload(unsigned long long w) { return (unsigned)w; }

// LDS: tile[R][B + 1] (column B holds the row's position of origin as int
// bits) | urow[B + 1] (the pivot row of the current column) | lcol[R] (the
// multipliers of the current column).
//
// Exchange buffers (64-bit words, see above): pub[2][G][B+1] candidate rows,
// krow[2][B+1] row k, kslot[2][gslots][kSlotWords] key slots: words 0/1 the
// key's low/high half, 2/3 the candidate's values at columns c and c+1, 4/5
// row k's (written by its owner only).
template <int B>
__global__ __launch_bounds__(PNT) void panel_pivot_kernel(
    float* __restrict__ X, int64_t ldx, int64_t n, int64_t k0, int R,
    unsigned long long* __restrict__ pub, unsigned long long* __restrict__ krow,
    unsigned long long* __restrict__ kslot, int gslots, float* __restrict__ Lt, int64_t ldl,
    float* __restrict__ diagU, int* __restrict__ ipvt, int* __restrict__ src,
    int* __restrict__ status, long spin_limit, unsigned phantom) {
#pragma clang fp contract(off)
    constexpr int LDT = B + 1;
    constexpr int TPC = PNT / B;  // threads per column in the update
    constexpr int NF = (B + PNT) / PNT;  // row words per thread (B + 1 of them)
    extern __shared__ __attribute__((aligned(16))) float sm[];
    float* tile = sm;
    float* urow = sm + (size_t)R * LDT;
    float* lcol = urow + LDT;
    __shared__ unsigned long long skey[PNT / 64], wkey[PNT / 64];
    __shared__ float wval[PNT / 64][2];
    __shared__ float kval[2];
    __shared__ int abort_flag;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int G = gridDim.x, g = blockIdx.x;
    const int64_t p0 = k0 + (int64_t)g * R;  // position of local row 0
    const int rows = (int)min<int64_t>((int64_t)R, n - p0);
    const size_t kstride = (size_t)gslots * kSlotWords;
    if (t == 0) abort_flag = 0;

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

    // the workgroup's best candidate of column c (rows at positions >= k)
    auto search = [&](int c) {
        const int64_t k = k0 + c;
        unsigned long long b = 0;
        for (int r = (int)max<int64_t>(0, k - p0) + t; r < rows; r += PNT)
            b = max(b, pivot_key(tile[r * LDT + c], p0 + r, k));
        b = wave_max_u64(b);
        if (lane == 0) skey[wave] = b;
        __syncthreads();
        b = skey[0];
#pragma unroll
        for (int w = 1; w < PNT / 64; ++w) b = max(b, skey[w]);
        return b;
    };
    // Publish column c's candidate `best`: the whole row, row k when this
    // workgroup owns it, then the key slot.  With upd (every column after the
    // first) the workgroup has applied the previous column's rank-1 update to
    // column c only: the published rows get it on the fly for their columns
    // beyond c (the same float product and sum), and the workgroup updates
    // those columns after publishing, while the others arrive.
    auto publish = [&](int c, unsigned long long best, bool upd) {
        const int64_t k = k0 + c;
        const unsigned tag = (unsigned)(k + 1);
        const int buf = c & 1;
        const int64_t rk = k - p0;
        const bool own_k = rk >= 0 && rk < rows;
        auto cur = [&](int r, int j) {
            float v = tile[r * LDT + j];
            if (upd && j > c && j < B) v = v + lcol[r] * urow[j];
            return v;
        };
        const int rl = best != 0 ? (int)(key_position(best, k) - p0) : 0;
        if (best != 0) {
            unsigned long long* slot = pub + ((int64_t)buf * G + g) * LDT;
            for (int j = t; j <= B; j += PNT) put_tagged(slot + j, __float_as_uint(cur(rl, j)), tag);
        }
        if (own_k)
            for (int j = t; j <= B; j += PNT)
                put_tagged(krow + buf * LDT + j, __float_as_uint(cur((int)rk, j)), tag);
        if (t == 0) {
            unsigned long long* w = kslot + (size_t)buf * kstride + (size_t)g * kSlotWords;
            if (best != 0) {
                put_tagged(w + 2, __float_as_uint(cur(rl, c)), tag);
                put_tagged(w + 3, c + 1 < B ? __float_as_uint(cur(rl, c + 1)) : 0u, tag);
            }
            if (own_k) {
                put_tagged(w + 4, __float_as_uint(cur((int)rk, c)), tag);
                put_tagged(w + 5, c + 1 < B ? __float_as_uint(cur((int)rk, c + 1)) : 0u, tag);
            }
            put_tagged(w + 0, (unsigned)best, tag);
            put_tagged(w + 1, (unsigned)(best >> 32), tag);
        }
    };
    // give up: raise the status, mark this workgroup's key slots (both
    // parities) with the abort tag so that every waiter is released
    auto give_up = [&]() {
        if (t == 0) {
            atomicOr(status, 2);
            for (int b2 = 0; b2 < 2; ++b2) {
                unsigned long long* w = kslot + (size_t)b2 * kstride + (size_t)g * kSlotWords;
                put_tagged(w + 0, 0u, kAbortTag);
                put_tagged(w + 1, 0u, kAbortTag);
            }
        }
    };
    // Column c's exchange: poll the key slots of all G groups (and the
    // phantom ones of the error-path test, which nobody writes) until each
    // carries column c's tag.  Returns the winning key, the winner's values at
    // c and c+1 in *u0, *u1; row k's land in kval.  False in *ok when the wait
    // gave up or another workgroup did.  wkey / wval are next written after
    // several barriers, so they need no barrier after the reads here.
    auto exchange = [&](int c, bool* ok, float* u0, float* u1) {
        const unsigned tag = (unsigned)(k0 + c + 1);
        const unsigned long long* ks = kslot + (size_t)(c & 1) * kstride;
        const int GW = G + (int)phantom;
        const int gk = c / R;  // the group holding position k
        unsigned long long mx = 0;
        float m0 = 0.0f, m1 = 0.0f;
        bool fail = false;
        for (int q = t; q < GW && !fail; q += PNT) {
            const unsigned long long* w = ks + (size_t)q * kSlotWords;
            const bool want_k = q == gk;
            unsigned long long v[6];
            long spins = 0;
            for (;;) {
#pragma unroll
                for (int i = 0; i < 6; ++i) v[i] = get_tagged(w + i);
                bool have = word_tag(v[0]) == tag && word_tag(v[1]) == tag;
                if (have) {
                    const unsigned long long key =
                        ((unsigned long long)word_payload(v[1]) << 32) | word_payload(v[0]);
                    if (key != 0) have = word_tag(v[2]) == tag && word_tag(v[3]) == tag;
                }
                if (want_k) have = have && word_tag(v[4]) == tag && word_tag(v[5]) == tag;
                if (have) break;
                if (word_tag(v[0]) == kAbortTag || word_tag(v[1]) == kAbortTag ||
                    ++spins > spin_limit) {
                    fail = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (fail) break;
            const unsigned long long key =
                ((unsigned long long)word_payload(v[1]) << 32) | word_payload(v[0]);
            if (key > mx) {
                mx = key;
                m0 = __uint_as_float(word_payload(v[2]));
                m1 = __uint_as_float(word_payload(v[3]));
            }
            if (want_k) {
                kval[0] = __uint_as_float(word_payload(v[4]));
                kval[1] = __uint_as_float(word_payload(v[5]));
            }
        }
        if (fail) abort_flag = 1;
        // the maximum key with its values, over the wave, then the workgroup
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned lo = __shfl_xor((unsigned)mx, o);
            const unsigned hi = __shfl_xor((unsigned)(mx >> 32), o);
            const float o0 = __shfl_xor(m0, o), o1 = __shfl_xor(m1, o);
            const unsigned long long om = ((unsigned long long)hi << 32) | lo;
            if (om > mx) {
                mx = om;
                m0 = o0;
                m1 = o1;
            }
        }
        if (lane == 0) {
            wkey[wave] = mx;
            wval[wave][0] = m0;
            wval[wave][1] = m1;
        }
        __syncthreads();
        *ok = abort_flag == 0;
        int bw = 0;
        mx = wkey[0];
#pragma unroll
        for (int w = 1; w < PNT / 64; ++w)
            if (wkey[w] > mx) {
                mx = wkey[w];
                bw = w;
            }
        *u0 = wval[bw][0];
        *u1 = wval[bw][1];
        return mx;
    };

    publish(0, search(0), false);
    for (int c = 0; c < B; ++c) {
        bool ok = true;
        float piv = 0.0f, u1 = 0.0f;
        const unsigned long long win = exchange(c, &ok, &piv, &u1);
        if (!ok) {
            give_up();
            return;
        }
        const int64_t k = k0 + c;
        const unsigned tag = (unsigned)(k + 1);
        const int buf = c & 1;
        const int64_t rk = k - p0;  // local row of position k (may be outside)
        int64_t p = key_position(win, k);
        if (p < k || p >= n) {  // cannot happen with a complete exchange; never index with it
            if (t == 0) atomicOr(status, 4);
            p = k;
        }
        const int gw = (int)((p - k0) / R);
        const int64_t rp = p - p0;
        const bool swap = p != k;
        const bool own_k = swap && rk >= 0 && rk < rows;
        const bool own_p = swap && rp >= 0 && rp < rows;
        // the whole winning row (and row k where position p is ours) start
        // now and are checked after the search of column c+1
        const unsigned long long* wrow = pub + ((int64_t)buf * G + gw) * LDT;
        const unsigned long long* krw = krow + buf * LDT;
        unsigned long long fw[NF], fk[NF];
#pragma unroll
        for (int i = 0; i < NF; ++i) {
            const int j = t + i * PNT;
            fw[i] = j <= B ? get_tagged(wrow + j) : 0ull;
            fk[i] = (own_p && j <= B) ? get_tagged(krw + j) : 0ull;
        }
        if (t == 0) {
            if (g == 0) {
                ipvt[k] = (int)p;
                diagU[k] = piv;
                if (!(piv != 0.0f) || !__builtin_isfinite(piv)) atomicOr(status, 1);
            }
            // the exchanged rows' values at columns c and c+1 (the rest of
            // the rows follows from the fetches)
            if (own_k) {
                tile[rk * LDT + c] = piv;
                if (c + 1 < B) tile[rk * LDT + c + 1] = u1;
            }
            if (own_p) {
                tile[rp * LDT + c] = kval[0];
                if (c + 1 < B) tile[rp * LDT + c + 1] = kval[1];
            }
        }
        __syncthreads();
        // ---- multipliers of the rows below position k (gefa_ref: double
        // reciprocal and product, rounded once)
        const int rs1 = (int)max<int64_t>(0, k + 1 - p0);
        const double inv = -1.0 / (double)piv;
        for (int r = rs1 + t; r < rows; r += PNT) {
            const float l = (float)((double)tile[r * LDT + c] * inv);
            tile[r * LDT + c] = l;
            lcol[r] = l;
        }
        __syncthreads();
        unsigned long long next = 0;
        if (c + 1 < B) {
            // column c+1 first, one row per thread: the next search reads it
            for (int r = rs1 + t; r < rows; r += PNT)
                tile[r * LDT + c + 1] = tile[r * LDT + c + 1] + lcol[r] * u1;
            __syncthreads();
            next = search(c + 1);
        }
        // ---- the fetched rows: the pivot row into urow and position k, row k
        // into position p (its columns c and c+1 are already in place)
        bool fail = false;
#pragma unroll
        for (int i = 0; i < NF; ++i) {
            const int j = t + i * PNT;
            if (j > B) continue;
            long spins = 0;
            while (word_tag(fw[i]) != tag && !fail) {
                if (++spins > spin_limit) fail = true;
                __builtin_amdgcn_s_sleep(1);
                fw[i] = get_tagged(wrow + j);
            }
            if (own_p)
                while (word_tag(fk[i]) != tag && !fail) {
                    if (++spins > spin_limit) fail = true;
                    __builtin_amdgcn_s_sleep(1);
                    fk[i] = get_tagged(krw + j);
                }
            const float uj = __uint_as_float(word_payload(fw[i]));
            urow[j] = uj;
            if (own_k) tile[rk * LDT + j] = uj;
            // (at the last column, c + 1 == B is the row-of-origin word)
            if (own_p && j != c && (j != c + 1 || j == B))
                tile[rp * LDT + j] = __uint_as_float(word_payload(fk[i]));
        }
        if (fail) abort_flag = 1;
        __syncthreads();
        if (abort_flag) {
            give_up();
            return;
        }
        if (c + 1 == B) break;
        publish(c + 1, next, true);
        __syncthreads();  // the published rows were read before they change
        // the other columns, one column per TPC threads, 16 rows in flight
        const int j = t % B;
        if (j > c + 1) {
            const float uj = urow[j];
            int r = rs1 + t / B;
            for (; r + 15 * TPC < rows; r += 16 * TPC) {
                float v[16], lv[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    v[u] = tile[(r + u * TPC) * LDT + j];
                    lv[u] = lcol[r + u * TPC];
                }
#pragma unroll
                for (int u = 0; u < 16; ++u) tile[(r + u * TPC) * LDT + j] = v[u] + lv[u] * uj;
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

__global__ void zero_words_kernel(unsigned long long* __restrict__ p, size_t words) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
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

namespace {
size_t panel_pivot_lds(int B) {
    const int R = panel_pivot_rows(B);
    return ((size_t)R * (B + 1) + (B + 1) + R) * sizeof(float);
}
// Dynamic LDS requested at least, so that a CU holds at most one panel
// workgroup: a grid no larger than the CU count then gets a CU per workgroup
// instead of two workgroups sharing one.
constexpr size_t kSpreadLds = 96 * 1024;
}  // namespace

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
    hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)std::max(lds, kSpreadLds));
    if (e != hipSuccess) return e;
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, PNT, lds);
    if (e != hipSuccess) return e;
    int dev = 0, cus = 0;
    e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    *max_groups = std::min(per_cu * cus, kMaxPanelGroups - 1);
    return hipSuccess;
}

hipError_t launch_panel_pivot(float* X, int64_t ldx, int64_t n, int64_t k0, int B,
                              unsigned long long* pub, unsigned long long* krow,
                              unsigned long long* kslot, int gslots, float* Lt, int64_t ldl,
                              float* diagU, int* ipvt, int* src, int* status, int max_groups,
                              int num_cu, hipStream_t s, long spin_limit) {
    const int R = panel_pivot_rows(B);
    // spin_limit < 0: give up after -spin_limit polls, waiting for one more
    // workgroup than the grid holds, so that every exchange gives up (the
    // error path, exercised from the host with every kernel exiting)
    const unsigned phantom = spin_limit < 0 ? 1u : 0u;
    spin_limit = spin_limit == 0 ? kSpinLimit : (spin_limit < 0 ? -spin_limit : spin_limit);
    const int64_t m = n - k0;
    const int G = panel_pivot_groups(m, B);
    if (G < 1 || G > max_groups || G + (int)phantom > gslots ||
        G + (int)phantom > kMaxPanelGroups)
        return hipErrorInvalidConfiguration;
    // 16-byte row vectors; tags are positions + 1 and must stay below the abort tag
    if ((ldx & 3) || (k0 & 3) || (reinterpret_cast<uintptr_t>(X) & 15) || n >= (1LL << 31))
        return hipErrorInvalidValue;
    size_t lds = panel_pivot_lds(B);
    if (G <= num_cu) lds = std::max(lds, kSpreadLds);
#define HPL_PANEL(BB)                                                                        \
    case BB:                                                                                 \
        panel_pivot_kernel<BB><<<G, PNT, lds, s>>>(X, ldx, n, k0, R, pub, krow, kslot, gslots, \
                                                   Lt, ldl, diagU, ipvt, src, status,          \
                                                   spin_limit, phantom);                       \
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

size_t panel_exchange_words(int64_t n, int B) {
    const int64_t groups = panel_pivot_groups(n, B);
    return 2 * (size_t)groups * (B + 1) + 2 * (size_t)(B + 1) +
           2 * (size_t)(groups + 1) * kSlotWords;
}

hipError_t launch_zero_words(unsigned long long* p, size_t words, hipStream_t s) {
    const unsigned blocks = (unsigned)std::min<size_t>((words + 255) / 256, 1024);
    if (blocks) zero_words_kernel<<<blocks, 256, 0, s>>>(p, words);
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
