// Batched solves of many independent systems of uneven order in one
// persistent launch (no reference counterpart: the reference solves one
// matrix per run; each system here follows the same algorithm -- LU without
// pivoting in the dgefa encoding, forward/back substitution as
// LINPACK/src/host/linpack_data.cpp:166-229 -- plus FP64 refinement).
//
// Work distribution: a queue of system indices ordered by decreasing size;
// every workgroup takes the next index with one atomic add and solves that
// system alone, until the queue is empty (every workgroup reaches the exit
// after at most nsys + grid claims).  Largest-first keeps the tail short.
//
// One system per workgroup, blocked right-looking LU with 32-wide steps:
//   1. wave 0 factorises the 32 x 32 diagonal block in registers (readlane
//      broadcasts, no barrier), bit-identical to gefa_ref_nopvt on the block;
//   2. the top strip (columns right of it) and the left strip (rows below)
//      get the reference's closed-form solves, one column / row per thread,
//      and are converted to FP16 (or kept FP32) in LDS;
//   3. the trailing matrix receives C += Left * Top on MFMA: 32 x 32 tiles,
//      v_mfma_f32_32x32x16_f16 (FP16 inputs, FP32 accumulation) or
//      v_mfma_f32_32x32x2_f32 (FP32 path).
// Then substitution in FP32 and classic refinement with FP64 residuals
// against the untouched FP32 matrix until ||r|| / (||M|| ||x|| + ||b||) <= tol.
//
// Storage of system s: rows of length npad (= n rounded up to 32), matrix at
// a_off[s] (row-major, M = A^T as everywhere), vectors at v_off[s].  The pad
// is an identity extension written by the generator, so it changes nothing.
#include "common.hpp"
#include "kernels.hpp"

#include <type_traits>

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

namespace {

__device__ __forceinline__ float bcast(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

template <int NT>
__device__ __forceinline__ double block_max(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double r = red[0];
    for (int i = 1; i < NT / 64; ++i) r = fmax(r, red[i]);
    return r;
}

// a[q] = sum_p A[p][q] x[p] over the 32 rows p of a block, for 4 adjacent
// columns (16-byte loads, 8 rows in flight; each column's sum in the order
// p = 0..31)
__device__ __forceinline__ void gemv4(const float* __restrict__ A, int npad,
                                      const float* __restrict__ x, f32x4& a) {
#pragma unroll
    for (int h = 0; h < 32; h += 8) {
        f32x4 v[8];
#pragma unroll
        for (int p = 0; p < 8; ++p)
            v[p] = *reinterpret_cast<const f32x4*>(A + (int64_t)(h + p) * npad);
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            const float xp = x[h + p];
            a[0] = fmaf(v[p][0], xp, a[0]);
            a[1] = fmaf(v[p][1], xp, a[1]);
            a[2] = fmaf(v[p][2], xp, a[2]);
            a[3] = fmaf(v[p][3], xp, a[3]);
        }
    }
}

// Solve M x = rhs in place on x (FP32, in LDS), factor in A (npad, dgefa
// encoding).  The diagonal blocks' solves keep all 32 factor loads of a block
// in flight; the GEMV loops read 4 adjacent columns per thread.
template <int NT>
__device__ __forceinline__ void substitute(const float* __restrict__ A, int npad,
                                           float* __restrict__ x) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    // forward: x_e += sum_p A[p][e] x_p, block by block
    for (int k0 = 0; k0 < npad; k0 += 32) {
        if (wave == 0 && lane < 32) {
            // the block's column `lane`, all 32 loads in flight before the
            // dependent chain
            float d[32];
#pragma unroll
            for (int p = 0; p < 32; ++p) d[p] = A[(int64_t)(k0 + p) * npad + k0 + lane];
            float xi = x[k0 + lane];
#pragma unroll
            for (int p = 0; p < 32; ++p) {
                const float xp = __shfl(xi, p);
                if (lane > p) xi += xp * d[p];
            }
            x[k0 + lane] = xi;
        }
        __syncthreads();
        for (int e = k0 + 32 + 4 * t; e < npad; e += 4 * NT) {
            f32x4 a = {0.0f, 0.0f, 0.0f, 0.0f};
            gemv4(A + (int64_t)k0 * npad + e, npad, x + k0, a);
            x[e] += a[0];
            x[e + 1] += a[1];
            x[e + 2] += a[2];
            x[e + 3] += a[3];
        }
        __syncthreads();
    }
    // backward: x_k = x_k / u_kk (stored -1/u_kk), x_i -= x_k u_ik (i < k)
    for (int k0 = npad - 32; k0 >= 0; k0 -= 32) {
        if (wave == 0 && lane < 32) {
            float d[32];
#pragma unroll
            for (int p = 0; p < 32; ++p) d[p] = A[(int64_t)(k0 + p) * npad + k0 + lane];
            float xi = x[k0 + lane];
#pragma unroll
            for (int p = 31; p >= 0; --p) {
                // the diagonal element (k0+p, k0+p) is lane p's d[p]
                float sc = __shfl(xi, p) * __shfl(d[p], p);
                if (lane == p) xi = -sc;
                if (lane < p) xi += sc * d[p];
            }
            x[k0 + lane] = xi;
        }
        __syncthreads();
        for (int e = 4 * t; e < k0; e += 4 * NT) {
            f32x4 a = {0.0f, 0.0f, 0.0f, 0.0f};
            gemv4(A + (int64_t)k0 * npad + e, npad, x + k0, a);
            x[e] -= a[0];
            x[e + 1] -= a[1];
            x[e + 2] -= a[2];
            x[e + 3] -= a[3];
        }
        __syncthreads();
    }
}

// LDS layout (dynamic): D[32][33] f32 (the refinement's reduction scratch
// aliases it) | Left[MAXM][40] T | TopT[MAXM][40] T
// FACTOR false: the systems were factorised by the blocked kernels below
// (their zero-pivot flags in pre_bad): substitution and refinement only, in a
// kernel sized for three waves per SIMD (its registers no longer hold the
// factorisation's tiles), so several systems share a CU.
template <int NT, bool F16, bool FACTOR>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(FACTOR ? 1 : 3))) void
batch_solve_kernel(
    const BatchSystem* __restrict__ sys, const int* __restrict__ order, int nsys,
    int* __restrict__ counter, float* __restrict__ A, const float* __restrict__ A0,
    const double* __restrict__ b64, double* __restrict__ x64, float* __restrict__ xs,
    double* __restrict__ rs, int max_iter, double tol_per_n, BatchResult* __restrict__ out,
    int maxm, const int* __restrict__ pre_bad) {
    typedef typename std::conditional<F16, _Float16, float>::type T;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    constexpr int LDT = 40;  // padded row of the strip copies (elements)
    float* D = reinterpret_cast<float*>(smem);
    T* Left = reinterpret_cast<T*>(smem + 32 * 33 * 4);
    T* TopT = Left + (size_t)maxm * LDT;
    double* red = reinterpret_cast<double*>(smem);  // used after the factorisation only
    int* claim = reinterpret_cast<int*>(D + 32);  // D[0][32]: padding, never a matrix element

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int lr = lane & 31, lh = lane >> 5;
    for (;;) {
        __syncthreads();
        if (t == 0) *claim = atomicAdd(counter, 1);
        __syncthreads();
        const int q = *claim;
        __syncthreads();
        if (q >= nsys) break;
        const int s = order[q];
        const int npad = sys[s].npad;
        float* a = A + sys[s].a_off;
        bool bad = false;

        // ------------------------------------------------------------ factor
        if constexpr (!FACTOR) bad = pre_bad[s] != 0;
        for (int k0 = 0; k0 < (FACTOR ? npad : 0); k0 += 32) {
            if (wave == 0) {
                // diagonal block: lane c < 32 holds column k0+c, rows 0..31
                float d[32];
#pragma unroll
                for (int r = 0; r < 32; ++r)
                    d[r] = lr == lane ? a[(int64_t)(k0 + r) * npad + k0 + lr] : 0.0f;
#pragma unroll
                for (int j = 0; j < 32; ++j) {
                    const float piv = bcast(d[j], j);
                    const float inv = -1.0f / piv;
                    bad |= (piv == 0.0f) || !__builtin_isfinite(inv);
                    d[j] = (lane > j) ? __fmul_rn(d[j], inv) : ((lane == j) ? inv : d[j]);
#pragma unroll
                    for (int r = j + 1; r < 32; ++r) {
                        const float sv = bcast(d[r], j);
                        if (lane > j) d[r] = __fadd_rn(d[r], __fmul_rn(d[j], sv));
                    }
                }
                if (lane < 32) {
#pragma unroll
                    for (int r = 0; r < 32; ++r) {
                        a[(int64_t)(k0 + r) * npad + k0 + lane] = d[r];
                        D[r * 33 + lane] = d[r];
                    }
                }
            }
            __syncthreads();
            const int r0 = k0 + 32, m = npad - r0;
            if (m > 0) {
                // top strip: columns c >= r0 of rows k0..k0+31 (closed form top)
                for (int c = t; c < m; c += NT) {
                    float x[32];
#pragma unroll
                    for (int i = 0; i < 32; ++i) x[i] = a[(int64_t)(k0 + i) * npad + r0 + c];
#pragma unroll
                    for (int kk = 0; kk < 32; ++kk) {
                        x[kk] = __fmul_rn(x[kk], D[kk * 33 + kk]);
#pragma unroll
                        for (int j = kk + 1; j < 32; ++j)
                            x[j] = __fadd_rn(x[j], __fmul_rn(x[kk], D[j * 33 + kk]));
                    }
#pragma unroll
                    for (int i = 0; i < 32; ++i) {
                        a[(int64_t)(k0 + i) * npad + r0 + c] = x[i];
                        TopT[(size_t)c * LDT + i] = (T)x[i];
                    }
                }
                // left strip: rows r >= r0, columns k0..k0+31 (closed form left)
                for (int r = t; r < m; r += NT) {
                    float x[32];
                    float* row = a + (int64_t)(r0 + r) * npad + k0;
#pragma unroll
                    for (int i = 0; i < 32; i += 4) {
                        const f32x4 v = *reinterpret_cast<const f32x4*>(row + i);
                        x[i] = v[0]; x[i + 1] = v[1]; x[i + 2] = v[2]; x[i + 3] = v[3];
                    }
#pragma unroll
                    for (int kk = 0; kk < 32; ++kk)
#pragma unroll
                        for (int i = kk + 1; i < 32; ++i)
                            x[i] = __fadd_rn(x[i], __fmul_rn(x[kk], D[kk * 33 + i]));
#pragma unroll
                    for (int i = 0; i < 32; i += 4)
                        *reinterpret_cast<f32x4*>(row + i) = f32x4{x[i], x[i + 1], x[i + 2], x[i + 3]};
#pragma unroll
                    for (int i = 0; i < 32; ++i) Left[(size_t)r * LDT + i] = (T)x[i];
                }
                __syncthreads();
                // trailing update on MFMA, 32 x 32 tiles over the waves; a wave
                // takes TPW tiles at a time so their loads are in flight together
                // (the update streams the trailing matrix through L2 and is
                // bound by load latency, not by the MFMAs)
                constexpr int TPW = 4;
                const int tiles = m / 32, ntile = tiles * tiles;
                for (int tt0 = wave * TPW; tt0 < ntile; tt0 += (NT / 64) * TPW) {
                    f32x16 acc[TPW];
                    float* cbase[TPW];
#pragma unroll
                    for (int u = 0; u < TPW; ++u) {
                        const int tt = min(tt0 + u, ntile - 1);  // clamped: in bounds, unused
                        cbase[u] = a + (int64_t)(r0 + (tt / tiles) * 32) * npad + r0 +
                                   (tt % tiles) * 32;
#pragma unroll
                        for (int rr = 0; rr < 16; ++rr)
                            acc[u][rr] = cbase[u][(int64_t)acc_row(rr, lane) * npad + lr];
                    }
#pragma unroll
                    for (int u = 0; u < TPW; ++u) {
                        const int tt = min(tt0 + u, ntile - 1);
                        const int ti = tt / tiles, tj = tt % tiles;
                        if constexpr (F16) {
#pragma unroll
                            for (int ks = 0; ks < 2; ++ks) {
                                const f16x8 av = *reinterpret_cast<const f16x8*>(
                                    &Left[(size_t)(ti * 32 + lr) * LDT + ks * 16 + 8 * lh]);
                                const f16x8 bv = *reinterpret_cast<const f16x8*>(
                                    &TopT[(size_t)(tj * 32 + lr) * LDT + ks * 16 + 8 * lh]);
                                acc[u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv, acc[u], 0,
                                                                               0, 0);
                            }
                        } else {
#pragma unroll
                            for (int kp = 0; kp < 16; ++kp) {
                                const float av = Left[(size_t)(ti * 32 + lr) * LDT + 2 * kp + lh];
                                const float bv = TopT[(size_t)(tj * 32 + lr) * LDT + 2 * kp + lh];
                                acc[u] = mfma32(av, bv, acc[u]);
                            }
                        }
                    }
#pragma unroll
                    for (int u = 0; u < TPW; ++u) {
                        if (tt0 + u >= ntile) break;  // wave-uniform
#pragma unroll
                        for (int rr = 0; rr < 16; ++rr)
                            cbase[u][(int64_t)acc_row(rr, lane) * npad + lr] = acc[u][rr];
                    }
                }
            }
            __syncthreads();
        }

        // ------------------------------------------------- solve + refinement
        const int64_t vo = sys[s].v_off;
        const float* a0 = A0 + sys[s].a_off;
        // the FP32 working vector of the substitutions lives in LDS (in the
        // dead strip copies; maxm >= 32 makes room for npad floats there)
        float* x32 = reinterpret_cast<float*>(TopT);
        double* x = x64 + vo;
        double* r = rs + vo;
        const double* b = b64 + vo;
        for (int i = t; i < npad; i += NT) x32[i] = (float)b[i];
        __syncthreads();
        substitute<NT>(a, npad, x32);
        for (int i = t; i < npad; i += NT) x[i] = (double)x32[i];
        __syncthreads();
        double anorm = 0.0, berr = 0.0;
        // x is staged in LDS (the strip copies are dead after the
        // factorisation) so the residual's inner loop streams only the matrix:
        // each thread takes 4 adjacent rows i (16-byte loads of the stored
        // transpose), 8 j at a time in flight; every sum still runs over j in
        // increasing order
        double* xsh = reinterpret_cast<double*>(Left);
        int it = 0;
        for (;; ++it) {
            double rmax = 0.0, amax = 0.0, xmax = 0.0, bmax = 0.0;
            for (int i = t; i < npad; i += NT) xsh[i] = x[i];
            __syncthreads();
            for (int i0 = 4 * t; i0 < npad; i0 += 4 * NT) {
                double acc[4] = {0.0, 0.0, 0.0, 0.0}, sa[4] = {0.0, 0.0, 0.0, 0.0};
                for (int j0 = 0; j0 < npad; j0 += 8) {
                    f32x4 v4[8];
#pragma unroll
                    for (int jj = 0; jj < 8; ++jj)
                        v4[jj] = *reinterpret_cast<const f32x4*>(a0 + (int64_t)(j0 + jj) * npad + i0);
#pragma unroll
                    for (int jj = 0; jj < 8; ++jj) {
                        const double xj = xsh[j0 + jj];
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const double v = (double)v4[jj][q];
                            acc[q] = fma(v, xj, acc[q]);
                            sa[q] += fabs(v);
                        }
                    }
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int i = i0 + q;
                    const double ri = b[i] - acc[q];
                    r[i] = ri;
                    rmax = fmax(rmax, fabs(ri));
                    amax = fmax(amax, sa[q]);
                    xmax = fmax(xmax, fabs(x[i]));
                    bmax = fmax(bmax, fabs(b[i]));
                }
            }
            rmax = block_max<NT>(rmax, red);
            if (it == 0) anorm = block_max<NT>(amax, red);
            xmax = block_max<NT>(xmax, red);
            bmax = block_max<NT>(bmax, red);
            berr = rmax / (anorm * xmax + bmax);
            if (berr <= tol_per_n * sys[s].n || it >= max_iter || !(rmax > 0.0)) break;
            for (int i = t; i < npad; i += NT) x32[i] = (float)(r[i] / rmax);
            __syncthreads();
            substitute<NT>(a, npad, x32);
            for (int i = t; i < npad; i += NT) x[i] = fma(rmax, (double)x32[i], x[i]);
            __syncthreads();
        }
        if (t == 0) {
            out[s].iterations = it;
            out[s].backward_error = berr;
            out[s].status = bad ? 1 : 0;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- blocked
// The same factorisation, one 32-wide step at a time over every system of a
// bucket still longer than the step: the first `act` systems of its
// largest-first order.  One launch factorises the diagonal blocks, one solves
// the strips, one applies the trailing updates, so a large system's step
// spreads over many CUs instead of one workgroup.  Each element sees the
// persistent kernel's operations in the same order, so the factors are
// bit-identical.  The strips' copies (type T) go to global scratch, row
// stride 32, at Lg / Tg + j * sstride for the j-th active system; its
// diagonal block to Dg + j * 1056 (32 x 33).

// one wave per system: the diagonal block of step k0
__global__ __launch_bounds__(64) void bstep_diag_kernel(const BatchSystem* __restrict__ sys,
                                                        const int* __restrict__ order, int k0,
                                                        float* __restrict__ A,
                                                        const float* __restrict__ S,
                                                        float* __restrict__ Dg,
                                                        int* __restrict__ bad_out) {
    const int j = blockIdx.x, s = order[j];
    const int npad = sys[s].npad;
    float* a = A + sys[s].a_off;
    const float* sa = S + sys[s].a_off;
    const int lane = threadIdx.x, lr = lane & 31;
    bool bad = false;
    float d[32];
#pragma unroll
    for (int r = 0; r < 32; ++r) d[r] = lr == lane ? sa[(int64_t)(k0 + r) * npad + k0 + lr] : 0.0f;
#pragma unroll
    for (int jj = 0; jj < 32; ++jj) {
        const float piv = bcast(d[jj], jj);
        const float inv = -1.0f / piv;
        bad |= (piv == 0.0f) || !__builtin_isfinite(inv);
        d[jj] = (lane > jj) ? __fmul_rn(d[jj], inv) : ((lane == jj) ? inv : d[jj]);
#pragma unroll
        for (int r = jj + 1; r < 32; ++r) {
            const float sv = bcast(d[r], jj);
            if (lane > jj) d[r] = __fadd_rn(d[r], __fmul_rn(d[jj], sv));
        }
    }
    if (lane < 32) {
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            a[(int64_t)(k0 + r) * npad + k0 + lane] = d[r];
            Dg[(size_t)j * 1056 + r * 33 + lane] = d[r];
        }
    }
    // the first step sets the flag, later ones only raise it
    if (lane == 0 && (k0 == 0 || bad)) bad_out[s] = bad ? 1 : 0;
}

// the strips of step k0: 256 columns of the top strip and 256 rows of the
// left strip per workgroup
template <bool F16>
__global__ __launch_bounds__(256) void bstep_strips_kernel(const BatchSystem* __restrict__ sys,
                                                           const int* __restrict__ order, int k0,
                                                           float* __restrict__ A,
                                                           const float* __restrict__ S,
                                                           const float* __restrict__ Dg,
                                                           void* __restrict__ Lv,
                                                           void* __restrict__ Tv,
                                                           size_t sstride) {
    typedef typename std::conditional<F16, _Float16, float>::type T;
    const int j = blockIdx.y, s = order[j];
    const int npad = sys[s].npad, r0 = k0 + 32, m = npad - r0;
    // the coefficients are the same for every thread: read from global at
    // wave-uniform addresses, they arrive through scalar loads into SGPRs
    // instead of one LDS read per multiply-add
    const float* __restrict__ D = Dg + (size_t)j * 1056;
    float* a = A + sys[s].a_off;
    const float* sa = S + sys[s].a_off;
    T* Left = static_cast<T*>(Lv) + (size_t)j * sstride;
    T* TopT = static_cast<T*>(Tv) + (size_t)j * sstride;
    const int c = blockIdx.x * 256 + threadIdx.x;
    if ((int)blockIdx.x * 256 >= m || c >= m) return;
    {  // top strip: column r0 + c of rows k0..k0+31 (closed form top)
        float x[32];
#pragma unroll
        for (int i = 0; i < 32; ++i) x[i] = sa[(int64_t)(k0 + i) * npad + r0 + c];
#pragma unroll
        for (int kk = 0; kk < 32; ++kk) {
            x[kk] = __fmul_rn(x[kk], D[kk * 33 + kk]);
#pragma unroll
            for (int jj = kk + 1; jj < 32; ++jj) x[jj] = __fadd_rn(x[jj], __fmul_rn(x[kk], D[jj * 33 + kk]));
        }
#pragma unroll
        for (int i = 0; i < 32; ++i) {
            a[(int64_t)(k0 + i) * npad + r0 + c] = x[i];
            TopT[(size_t)c * 32 + i] = (T)x[i];
        }
    }
    {  // left strip: row r0 + c, columns k0..k0+31 (closed form left)
        float x[32];
        float* row = a + (int64_t)(r0 + c) * npad + k0;
        const float* srow = sa + (int64_t)(r0 + c) * npad + k0;
#pragma unroll
        for (int i = 0; i < 32; i += 4) {
            const f32x4 v = *reinterpret_cast<const f32x4*>(srow + i);
            x[i] = v[0];
            x[i + 1] = v[1];
            x[i + 2] = v[2];
            x[i + 3] = v[3];
        }
#pragma unroll
        for (int kk = 0; kk < 32; ++kk)
#pragma unroll
            for (int i = kk + 1; i < 32; ++i) x[i] = __fadd_rn(x[i], __fmul_rn(x[kk], D[kk * 33 + i]));
#pragma unroll
        for (int i = 0; i < 32; i += 4)
            *reinterpret_cast<f32x4*>(row + i) = f32x4{x[i], x[i + 1], x[i + 2], x[i + 3]};
#pragma unroll
        for (int i = 0; i < 32; ++i) Left[(size_t)c * 32 + i] = (T)x[i];
    }
}

// acc += the step's rank-32 product on tile (ti, tj) of its trailing matrix
// (strip rows ti*32.. of Left, tj*32.. of TopT)
template <bool F16>
__device__ __forceinline__ void bstep_tile_mfma(f32x16& acc, const void* Lv, const void* Tv,
                                                int ti, int tj, int lr, int lh) {
    typedef typename std::conditional<F16, _Float16, float>::type T;
    const T* Left = static_cast<const T*>(Lv);
    const T* TopT = static_cast<const T*>(Tv);
    if constexpr (F16) {
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const f16x8 av =
                *reinterpret_cast<const f16x8*>(&Left[(size_t)(ti * 32 + lr) * 32 + ks * 16 + 8 * lh]);
            const f16x8 bv =
                *reinterpret_cast<const f16x8*>(&TopT[(size_t)(tj * 32 + lr) * 32 + ks * 16 + 8 * lh]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv, acc, 0, 0, 0);
        }
    } else {
#pragma unroll
        for (int kp = 0; kp < 16; ++kp) {
            const float av = Left[(size_t)(ti * 32 + lr) * 32 + 2 * kp + lh];
            const float bv = TopT[(size_t)(tj * 32 + lr) * 32 + 2 * kp + lh];
            acc = mfma32(av, bv, acc);
        }
    }
}

// Trailing update of step k0, one 32 x 32 tile per wave.  MODE 0: every tile
// of the step's trailing matrix.  MODE 1: its first tile row and column only
// (the next step's diagonal block and strips).  MODE 2: the other tiles, with
// step k0's update and then step k0+32's (strips Lv2 / Tv2, tile indices one
// less): the same products in the same order as two MODE 0 passes, since the
// FP32 accumulator is what a pass would store and reload.
template <bool F16, int MODE>
__global__ __launch_bounds__(256) void bstep_update_kernel(
    const BatchSystem* __restrict__ sys, const int* __restrict__ order, int k0,
    float* __restrict__ A, const float* __restrict__ S, const void* __restrict__ Lv,
    const void* __restrict__ Tv, const void* __restrict__ Lv2, const void* __restrict__ Tv2,
    size_t sstride) {
    typedef typename std::conditional<F16, _Float16, float>::type T;
    const int j = blockIdx.y, s = order[j];
    const int npad = sys[s].npad, r0 = k0 + 32, m = npad - r0;
    const int tiles = m / 32;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tt = blockIdx.x * 4 + wave;
    int ti, tj;
    if constexpr (MODE == 0) {
        if (tt >= tiles * tiles) return;  // wave-uniform; no barrier in this kernel
        ti = tt / tiles;
        tj = tt % tiles;
    } else if constexpr (MODE == 1) {
        if (tt >= 2 * tiles - 1) return;
        ti = tt < tiles ? 0 : tt - tiles + 1;
        tj = tt < tiles ? tt : 0;
    } else {
        const int t2 = tiles - 1;
        if (tt >= t2 * t2) return;
        ti = 1 + tt / t2;
        tj = 1 + tt % t2;
    }
    const int lr = lane & 31, lh = lane >> 5;
    const int64_t toff = sys[s].a_off + (int64_t)(r0 + ti * 32) * npad + r0 + tj * 32;
    float* cbase = A + toff;
    const float* sbase = S + toff;
    f32x16 acc;
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) acc[rr] = sbase[(int64_t)acc_row(rr, lane) * npad + lr];
    bstep_tile_mfma<F16>(acc, static_cast<const T*>(Lv) + (size_t)j * sstride,
                         static_cast<const T*>(Tv) + (size_t)j * sstride, ti, tj, lr, lh);
    if constexpr (MODE == 2)
        bstep_tile_mfma<F16>(acc, static_cast<const T*>(Lv2) + (size_t)j * sstride,
                             static_cast<const T*>(Tv2) + (size_t)j * sstride, ti - 1, tj - 1, lr,
                             lh);
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) cbase[(int64_t)acc_row(rr, lane) * npad + lr] = acc[rr];
}

// Generator: system s, row-major npad x npad; M(i,j) = U[-0.5,0.5) for
// i, j < n (counter-based, key of (seed, s)), M(i,i) += n, identity on the
// pad; b = M * 1 in FP64 on the first n entries, 0 on the pad.
__global__ void batch_generate_kernel(const BatchSystem* __restrict__ sys, int nsys, uint64_t seed,
                                      float* __restrict__ A, double* __restrict__ b64) {
    const int s = blockIdx.x;
    if (s >= nsys) return;
    const int n = sys[s].n, npad = sys[s].npad;
    float* a = A + sys[s].a_off;
    const uint64_t key = hpl_stream_key(seed, 2 + (uint64_t)s);
    for (int64_t e = threadIdx.x; e < (int64_t)npad * npad; e += blockDim.x) {
        const int r = (int)(e / npad), c = (int)(e % npad);
        float v;
        if (r < n && c < n)
            v = hpl_u24(key, (uint64_t)r * n + c) - 0.5f + (r == c ? (float)n : 0.0f);
        else
            v = (r == c) ? 1.0f : 0.0f;
        a[e] = v;
    }
    __syncthreads();
    double* b = b64 + sys[s].v_off;
    for (int i = threadIdx.x; i < npad; i += blockDim.x) {
        double acc = 0.0;
        if (i < n)
            for (int j = 0; j < n; ++j) acc += (double)a[(int64_t)j * npad + i];
        b[i] = acc;
    }
}

__global__ void reset_counter_kernel(int* c) {
    if (threadIdx.x == 0) *c = 0;
}

}  // namespace

hipError_t batch_prepare() {
    const int lds = (int)kBatchLdsLimit;
    const void* ks[] = {reinterpret_cast<const void*>(batch_solve_kernel<128, true, true>),
                        reinterpret_cast<const void*>(batch_solve_kernel<128, false, true>),
                        reinterpret_cast<const void*>(batch_solve_kernel<256, true, true>),
                        reinterpret_cast<const void*>(batch_solve_kernel<256, false, true>),
                        reinterpret_cast<const void*>(batch_solve_kernel<512, true, true>),
                        reinterpret_cast<const void*>(batch_solve_kernel<512, false, true>),
                        reinterpret_cast<const void*>(batch_solve_kernel<128, true, false>),
                        reinterpret_cast<const void*>(batch_solve_kernel<128, false, false>),
                        reinterpret_cast<const void*>(batch_solve_kernel<256, true, false>),
                        reinterpret_cast<const void*>(batch_solve_kernel<256, false, false>),
                        reinterpret_cast<const void*>(batch_solve_kernel<512, true, false>),
                        reinterpret_cast<const void*>(batch_solve_kernel<512, false, false>)};
    for (const void* k : ks) {
        hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

size_t batch_lds_bytes(int maxm, bool f16) {
    const size_t es = f16 ? 2 : 4;
    return 32 * 33 * 4 + 2 * (size_t)maxm * 40 * es;
}

hipError_t launch_batch_generate(const BatchSystem* sys, int nsys, uint64_t seed, float* A,
                                 double* b64, hipStream_t s) {
    batch_generate_kernel<<<nsys, 256, 0, s>>>(sys, nsys, seed, A, b64);
    return hipGetLastError();
}

hipError_t launch_batch_step(bool f16, const BatchSystem* sys, const int* order, int k0,
                             int act_diag, int act, int mmax, float* A, float* Dg, int* bad,
                             void* Lg, void* Tg, size_t sstride, hipStream_t s) {
    if (act_diag < 1 || act > act_diag || act > 65535) return hipErrorInvalidValue;
    bstep_diag_kernel<<<act_diag, 64, 0, s>>>(sys, order, k0, A, A, Dg, bad);
    if (act > 0 && mmax > 0) {
        if (mmax % 32 || (size_t)mmax * 32 > sstride) return hipErrorInvalidValue;
        const dim3 gs((unsigned)((mmax + 255) / 256), (unsigned)act);
        const int tiles = mmax / 32;
        const dim3 gu((unsigned)((tiles * tiles + 3) / 4), (unsigned)act);
        if (f16) {
            bstep_strips_kernel<true><<<gs, 256, 0, s>>>(sys, order, k0, A, A, Dg, Lg, Tg, sstride);
            bstep_update_kernel<true, 0><<<gu, 256, 0, s>>>(sys, order, k0, A, A, Lg, Tg, nullptr,
                                                            nullptr, sstride);
        } else {
            bstep_strips_kernel<false><<<gs, 256, 0, s>>>(sys, order, k0, A, A, Dg, Lg, Tg,
                                                          sstride);
            bstep_update_kernel<false, 0><<<gu, 256, 0, s>>>(sys, order, k0, A, A, Lg, Tg, nullptr,
                                                             nullptr, sstride);
        }
    }
    return hipGetLastError();
}

hipError_t launch_batch_step2(bool f16, const BatchSystem* sys, const int* order, int k0,
                              int act_diag, int act, int act2, int mmax, float* A,
                              const float* A0, float* Dg, int* bad, void* Lg, void* Tg, void* Lg2,
                              void* Tg2, size_t sstride, hipStream_t s) {
    if (act_diag < 1 || act > act_diag || act2 > act || act > 65535) return hipErrorInvalidValue;
    if (act > 0 && (mmax < 32 || mmax % 32 || (size_t)mmax * 32 > sstride))
        return hipErrorInvalidValue;
    const int tiles = mmax / 32;
    // step k0's reads: the original matrices at the first step (every element
    // is written there, so A needs no restore), A after
    const float* S = (k0 == 0 && A0) ? A0 : A;
    bstep_diag_kernel<<<act_diag, 64, 0, s>>>(sys, order, k0, A, S, Dg, bad);
    if (act == 0) return hipGetLastError();
    const dim3 gs((unsigned)((mmax + 255) / 256), (unsigned)act);
    const dim3 gn((unsigned)((2 * tiles - 1 + 3) / 4), (unsigned)act);
    if (f16) {
        bstep_strips_kernel<true><<<gs, 256, 0, s>>>(sys, order, k0, A, S, Dg, Lg, Tg, sstride);
        bstep_update_kernel<true, 1><<<gn, 256, 0, s>>>(sys, order, k0, A, S, Lg, Tg, nullptr,
                                                        nullptr, sstride);
    } else {
        bstep_strips_kernel<false><<<gs, 256, 0, s>>>(sys, order, k0, A, S, Dg, Lg, Tg, sstride);
        bstep_update_kernel<false, 1><<<gn, 256, 0, s>>>(sys, order, k0, A, S, Lg, Tg, nullptr,
                                                         nullptr, sstride);
    }
    // step k0 + 32 on the systems longer than it (its block and strips were
    // written to A by the update above)
    bstep_diag_kernel<<<act, 64, 0, s>>>(sys, order, k0 + 32, A, A, Dg, bad);
    if (act2 > 0 && tiles > 1) {
        const dim3 gs2((unsigned)((mmax - 32 + 255) / 256), (unsigned)act2);
        const dim3 gw((unsigned)(((tiles - 1) * (tiles - 1) + 3) / 4), (unsigned)act2);
        if (f16) {
            bstep_strips_kernel<true><<<gs2, 256, 0, s>>>(sys, order, k0 + 32, A, A, Dg, Lg2, Tg2,
                                                          sstride);
            bstep_update_kernel<true, 2><<<gw, 256, 0, s>>>(sys, order, k0, A, S, Lg, Tg, Lg2,
                                                            Tg2, sstride);
        } else {
            bstep_strips_kernel<false><<<gs2, 256, 0, s>>>(sys, order, k0 + 32, A, A, Dg, Lg2,
                                                           Tg2, sstride);
            bstep_update_kernel<false, 2><<<gw, 256, 0, s>>>(sys, order, k0, A, S, Lg, Tg, Lg2,
                                                             Tg2, sstride);
        }
    }
    return hipGetLastError();
}

hipError_t launch_batch_solve(int threads, bool f16, int grid, const BatchSystem* sys,
                              const int* order, int nsys, int* counter, float* A, const float* A0,
                              const double* b64, double* x64, float* xs, double* rs, int max_iter,
                              double tol_per_n, BatchResult* out, int maxm, hipStream_t s,
                              const int* pre_bad) {
    const size_t lds = batch_lds_bytes(maxm, f16);
    if (lds > kBatchLdsLimit) return hipErrorInvalidValue;
    reset_counter_kernel<<<1, 64, 0, s>>>(counter);
#define HPL_BATCH(NTV, F)                                                                   \
    do {                                                                                    \
        if (pre_bad)                                                                        \
            batch_solve_kernel<NTV, F, false><<<grid, NTV, lds, s>>>(                       \
                sys, order, nsys, counter, A, A0, b64, x64, xs, rs, max_iter, tol_per_n, out, \
                maxm, pre_bad);                                                             \
        else                                                                                \
            batch_solve_kernel<NTV, F, true><<<grid, NTV, lds, s>>>(                        \
                sys, order, nsys, counter, A, A0, b64, x64, xs, rs, max_iter, tol_per_n, out, \
                maxm, nullptr);                                                             \
    } while (0)
    if (threads == 512) {
        if (f16) HPL_BATCH(512, true); else HPL_BATCH(512, false);
    } else if (threads == 256) {
        if (f16) HPL_BATCH(256, true); else HPL_BATCH(256, false);
    } else if (threads == 128) {
        if (f16) HPL_BATCH(128, true); else HPL_BATCH(128, false);
    } else {
        return hipErrorInvalidValue;
    }
#undef HPL_BATCH
    return hipGetLastError();
}
