// LINPACK oracle -- TEST INFRASTRUCTURE ONLY.
//
// This file is a CPU restatement of the reference's LINPACK host path. It is
// the checker for the HIP library, never part of it: only tests/, the smoke
// check and the CPU-baseline leg of bench.py may load it.  Nothing in the
// product links, loads or calls it.
//
// Storage convention (same as the reference): the n x n matrix is stored
// row-major, A[row*lda + col], and factorised as the column-major matrix
// M = A^T, i.e. M(i,j) = A[j*lda + i].  The system solved is M x = b.
// The factor encoding is LINPACK dgefa without pivoting: strictly-lower part
// of M holds the NEGATED multipliers -l_ik, the strictly-upper part holds
// u_kj and the diagonal holds -1/u_kk.
//
// Every function cites the reference file:line it restates.  Pinning of each
// function (what checks it besides itself) is listed in tests/test_oracle.py.
//
// Build: oracle/Makefile (g++ -O3 -ffp-contract=off: no FMA contraction, so
// the float loops round exactly like the reference compiled for plain x86-64).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <random>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// ---------------------------------------------------------------------------
// A1: reference generator, LINPACK/src/host/linpack_benchmark.hpp:211-297,
// single rank (P = Q = 1, so every MPI reduction is the identity and the
// global row/column index maps collapse to the local ones, :251-253).
//   std::mt19937 gen(rank); uniform_real_distribution<double>(0,1) -> float
//   (:222-238); diagonally dominant: diag := 0, then float sequential row sum
//   of the row-major row -> diag (:243-272); b_j = float sequential column
//   sum (:288-295); norma/normb are running maxima (:236,268,294).
// ---------------------------------------------------------------------------
void orc_generate_ref(float* A, float* b, int64_t n, int64_t lda,
                      uint32_t seed, int diag_dominant,
                      float* norma_out, float* normb_out) {
    std::mt19937 gen(seed);
    std::uniform_real_distribution<> dis(0.0, 1.0);
    float norma = 0.0f, normb = 0.0f;
    for (int64_t j = 0; j < n; j++) {
        for (int64_t i = 0; i < n; i++) {
            float temp = dis(gen);
            A[lda * j + i] = temp;
            norma = (temp > norma) ? temp : norma;
        }
    }
    if (diag_dominant) {
        for (int64_t r = 0; r < n; r++) {
            A[lda * r + r] = 0.0f;
            float row_sum = 0.0f;
            for (int64_t i = 0; i < n; i++) row_sum += A[lda * r + i];
            norma = (row_sum > norma) ? row_sum : norma;
            A[lda * r + r] = row_sum;
        }
    }
    for (int64_t j = 0; j < n; j++) {
        float col_sum = 0.0f;
        for (int64_t i = 0; i < n; i++) col_sum += A[lda * i + j];
        b[j] = col_sum;
        normb = (b[j] > normb) ? b[j] : normb;
    }
    if (norma_out) *norma_out = norma;
    if (normb_out) *normb_out = normb;
}

// The same generator as the reference builds it with DATA_TYPE=double
// (HOST_DATA_TYPE double, linpack_benchmark.hpp:222-295): the U(0,1) draws
// are kept in double, the row and column sums are double sums.
void orc_generate_ref_f64(double* A, double* b, int64_t n, int64_t lda, uint32_t seed,
                          int diag_dominant) {
    std::mt19937 gen(seed);
    std::uniform_real_distribution<> dis(0.0, 1.0);
    for (int64_t j = 0; j < n; j++)
        for (int64_t i = 0; i < n; i++) A[lda * j + i] = dis(gen);
    if (diag_dominant) {
        for (int64_t r = 0; r < n; r++) {
            A[lda * r + r] = 0.0;
            double row_sum = 0.0;
            for (int64_t i = 0; i < n; i++) row_sum += A[lda * r + i];
            A[lda * r + r] = row_sum;
        }
    }
    for (int64_t j = 0; j < n; j++) {
        double col_sum = 0.0;
        for (int64_t i = 0; i < n; i++) col_sum += A[lda * i + j];
        b[j] = col_sum;
    }
}

// ---------------------------------------------------------------------------
// Counter-based generator for the sized workloads (no reference counterpart:
// the reference has only the mt19937 generator above).  Element k of stream s
// under seed S is  mix64(mix64(S ^ (s << 56)) + (k + 1) * 0x9E3779B97F4A7C15),
// mix64 being the splitmix64 finaliser (Steele, Lea, Flood 2014).  The top 24
// bits give u in [0,1) exactly representable in float.  Because every element
// is a pure function of (S, s, k) the device generator reproduces it bit for
// bit in any order.
// ---------------------------------------------------------------------------
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static inline float u24(uint64_t key, uint64_t k) {
    uint64_t z = mix64(key + (k + 1) * 0x9E3779B97F4A7C15ULL);
    return (float)(uint32_t)(z >> 40) * (1.0f / 16777216.0f);
}
static inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
    return mix64(seed ^ (stream << 56));
}

// Off-diagonal values of the sized workloads: A[r*lda+c] = u24(r*n+c) - 0.5,
// Uniform[-0.5,0.5) on a 2^-24 grid.
float orc_hash_uniform(uint64_t seed, uint64_t stream, uint64_t k) {
    return u24(stream_key(seed, stream), k);
}

// mode 0 ("add_n"): M(i,i) += n (float add), b = M*1, row sums of M taken in
//                   double in increasing column order and rounded to float.
// mode 2 ("uniform"): as mode 0 without the diagonal term: every entry is
//                   U[-0.5,0.5), no diagonal dominance (the pivoting workload).
// mode 1 ("scaled"): M(i,i) = d * sum_{j!=i} |M(i,j)| (double, rounded to
//                   float), x_true_i = 2*u24(stream 1, i) - 1, b = M x_true
//                   in double (b64) and rounded to float (b).
// Storage is the reference's: M(i,j) = A[j*lda + i].
void orc_generate_sized(float* A, float* b, double* b64, double* xtrue,
                        int64_t n, int64_t lda, uint64_t seed, int mode,
                        double d) {
    const uint64_t key = stream_key(seed, 0);
    for (int64_t r = 0; r < n; r++)
        for (int64_t c = 0; c < n; c++)
            A[lda * r + c] = u24(key, (uint64_t)(r * n + c)) - 0.5f;
    if (mode == 0 || mode == 2) {
        if (mode == 0)
            for (int64_t i = 0; i < n; i++) A[lda * i + i] += (float)n;
        for (int64_t i = 0; i < n; i++) {
            double s = 0.0;
            for (int64_t j = 0; j < n; j++) s += (double)A[lda * j + i];
            if (b) b[i] = (float)s;
            if (b64) b64[i] = s;
            if (xtrue) xtrue[i] = 1.0;
        }
    } else {
        for (int64_t i = 0; i < n; i++) {
            double s = 0.0;
            for (int64_t j = 0; j < n; j++)
                if (j != i) s += std::fabs((double)A[lda * j + i]);
            A[lda * i + i] = (float)(d * s);
        }
        const uint64_t xkey = stream_key(seed, 1);
        std::vector<double> xt(n);
        for (int64_t i = 0; i < n; i++)
            xt[i] = (double)(2.0f * u24(xkey, (uint64_t)i) - 1.0f);
        for (int64_t i = 0; i < n; i++) {
            double s = 0.0;
            for (int64_t j = 0; j < n; j++) s = std::fma((double)A[lda * j + i], xt[j], s);
            if (b) b[i] = (float)s;
            if (b64) b64[i] = s;
            if (xtrue) xtrue[i] = xt[i];
        }
    }
}

// System s of a batch: the sized generator (mode 0) on stream 2 + s, stored
// with leading dimension npad >= n and an identity extension on the pad;
// b = M * 1 in double (0 on the pad).
void orc_generate_batch_system(float* A, double* b64, int64_t n, int64_t npad, uint64_t seed,
                               uint64_t s) {
    const uint64_t key = stream_key(seed, 2 + s);
    for (int64_t r = 0; r < npad; r++)
        for (int64_t c = 0; c < npad; c++) {
            float v;
            if (r < n && c < n) {
                v = u24(key, (uint64_t)(r * n + c)) - 0.5f;
                if (r == c) v += (float)n;
            } else {
                v = (r == c) ? 1.0f : 0.0f;
            }
            A[r * npad + c] = v;
        }
    for (int64_t i = 0; i < npad; i++) {
        double acc = 0.0;
        if (i < n)
            for (int64_t j = 0; j < n; j++) acc += (double)A[j * npad + i];
        b64[i] = acc;
    }
}

// ---------------------------------------------------------------------------
// A8: gefa_ref_nopvt, LINPACK/src/host/linpack_data.cpp:166-196.
// ---------------------------------------------------------------------------
void orc_gefa_nopvt(float* a, int64_t n, int64_t lda) {
    for (int64_t k = 0; k < n; k++) {
        a[k * lda + k] = -1.0 / a[k * lda + k];
        for (int64_t i = k + 1; i < n; i++) a[k * lda + i] *= a[k * lda + k];
        for (int64_t j = k + 1; j < n; j++)
            for (int64_t i = k + 1; i < n; i++)
                a[j * lda + i] += a[k * lda + i] * a[j * lda + k];
    }
}

// Same loop nest with the independent columns j > k of step k spread over
// threads.  Not the reference (which is single threaded): a labelled parallel
// restatement used only as a second CPU timing.
void orc_gefa_nopvt_omp(float* a, int64_t n, int64_t lda, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    for (int64_t k = 0; k < n; k++) {
        a[k * lda + k] = -1.0 / a[k * lda + k];
        for (int64_t i = k + 1; i < n; i++) a[k * lda + i] *= a[k * lda + k];
#pragma omp parallel for schedule(static)
        for (int64_t j = k + 1; j < n; j++)
            for (int64_t i = k + 1; i < n; i++)
                a[j * lda + i] += a[k * lda + i] * a[j * lda + k];
    }
}

// ---------------------------------------------------------------------------
// A9: gesl_ref_nopvt, LINPACK/src/host/linpack_data.cpp:199-229.
// ---------------------------------------------------------------------------
void orc_gesl_nopvt(const float* a, float* b, int64_t n, int64_t lda) {
    std::vector<float> t(b, b + n);
    for (int64_t k = 0; k < n - 1; k++)
        for (int64_t i = k + 1; i < n; i++) t[i] += t[k] * a[lda * k + i];
    for (int64_t k = n - 1; k >= 0; k--) {
        float scale = t[k] * a[lda * k + k];
        t[k] = -scale;
        for (int64_t i = 0; i < k; i++) t[i] += scale * a[lda * k + i];
    }
    std::memcpy(b, t.data(), sizeof(float) * n);
}

// ---------------------------------------------------------------------------
// gefa_ref, LINPACK/src/host/linpack_data.cpp:72-119: LU with partial pivoting
// by rows of M (LINPACK dgefa).  Column k: the pivot is the first index of the
// largest |M(i,k)|, i >= k (strict '<' keeps the earlier index on ties, :80-86);
// rows k and p are exchanged in columns k..n-1 only (:88-92), so the
// multipliers of earlier columns keep the row order of their own step; the
// multipliers are M(i,k) * (-1.0 / M(k,k)) evaluated in double (the double
// literal promotes, :96-98) and rounded to float; the update is a float
// product added in float (:100-105).  The diagonal keeps u_kk.  ipvt[k] = p,
// ipvt[n-1] = n-1.
// ---------------------------------------------------------------------------
void orc_gefa_pvt(float* a, int64_t n, int64_t lda, int32_t* ipvt) {
    for (int64_t i = 0; i < n; i++) ipvt[i] = (int32_t)i;
    for (int64_t k = 0; k < n - 1; k++) {
        float max_val = std::fabs(a[k * lda + k]);
        int64_t p = k;
        for (int64_t i = k + 1; i < n; i++) {
            if (max_val < std::fabs(a[k * lda + i])) {
                p = i;
                max_val = std::fabs(a[k * lda + i]);
            }
        }
        for (int64_t i = k; i < n; i++) std::swap(a[i * lda + k], a[i * lda + p]);
        ipvt[k] = (int32_t)p;
        const double inv = -1.0 / (double)a[k * lda + k];
        for (int64_t i = k + 1; i < n; i++) a[k * lda + i] = (float)((double)a[k * lda + i] * inv);
        for (int64_t j = k + 1; j < n; j++)
            for (int64_t i = k + 1; i < n; i++)
                a[j * lda + i] += a[k * lda + i] * a[j * lda + k];
    }
}

// gesl_ref, LINPACK/src/host/linpack_data.cpp:121-156: forward elimination
// with the row exchanges of gefa_ref applied step by step, then back
// substitution dividing by u_kk.
void orc_gesl_pvt(const float* a, float* b, const int32_t* ipvt, int64_t n, int64_t lda) {
    std::vector<float> t(b, b + n);
    for (int64_t k = 0; k < n - 1; k++) {
        if (ipvt[k] != k) std::swap(t[k], t[ipvt[k]]);
        for (int64_t i = k + 1; i < n; i++) t[i] += t[k] * a[lda * k + i];
    }
    for (int64_t k = n - 1; k >= 0; k--) {
        t[k] = t[k] / a[lda * k + k];
        for (int64_t i = 0; i < k; i++) t[i] -= t[k] * a[lda * k + i];
    }
    std::memcpy(b, t.data(), sizeof(float) * n);
}

// FP64 oracle: the same two algorithms (linpack_data.cpp:166-229) in double,
// used as the "FP64 LU oracle" for forward and backward errors.
void orc_gefa_nopvt_f64(double* a, int64_t n, int64_t lda) {
    for (int64_t k = 0; k < n; k++) {
        a[k * lda + k] = -1.0 / a[k * lda + k];
        for (int64_t i = k + 1; i < n; i++) a[k * lda + i] *= a[k * lda + k];
        for (int64_t j = k + 1; j < n; j++)
            for (int64_t i = k + 1; i < n; i++)
                a[j * lda + i] += a[k * lda + i] * a[j * lda + k];
    }
}
void orc_gesl_nopvt_f64(const double* a, double* b, int64_t n, int64_t lda) {
    std::vector<double> t(b, b + n);
    for (int64_t k = 0; k < n - 1; k++)
        for (int64_t i = k + 1; i < n; i++) t[i] += t[k] * a[lda * k + i];
    for (int64_t k = n - 1; k >= 0; k--) {
        double scale = t[k] * a[lda * k + k];
        t[k] = -scale;
        for (int64_t i = 0; i < k; i++) t[i] += scale * a[lda * k + i];
    }
    std::memcpy(b, t.data(), sizeof(double) * n);
}

// ---------------------------------------------------------------------------
// Block kernels as closed forms, on one B x B row-major block (ld = ldx).
// A5 top_update (hpl_torus_PCIE.cl:452-582), closed form restated from
//   LINPACK/tests/test_kernel_communication.cpp:776-789 (and :1060-1072):
//   for k: row k *= LU(k,k); rows j>k += row k * LU(j,k).
// A6 left_update (hpl_torus_PCIE.cl:588-709), closed form restated from
//   test_kernel_communication.cpp:624-632 (and :1028-1037):
//   for k: all rows j: X(j,i) += X(j,k) * LU(k,i) for i > k.
// A7 inner_update_mm (hpl_torus_PCIE.cl:711-865): C += Left * Top (row-major
//   products, both operands as stored).
// ---------------------------------------------------------------------------
void orc_top_update(float* x, int64_t ldx, const float* lu, int64_t ldl,
                    int64_t B) {
    for (int64_t k = 0; k < B; k++) {
        for (int64_t i = 0; i < B; i++) x[k * ldx + i] *= lu[k * ldl + k];
        for (int64_t j = k + 1; j < B; j++)
            for (int64_t i = 0; i < B; i++)
                x[j * ldx + i] += x[k * ldx + i] * lu[j * ldl + k];
    }
}
void orc_left_update(float* x, int64_t ldx, const float* lu, int64_t ldl,
                     int64_t B) {
    for (int64_t k = 0; k < B; k++)
        for (int64_t j = 0; j < B; j++)
            for (int64_t i = k + 1; i < B; i++)
                x[j * ldx + i] += x[j * ldx + k] * lu[k * ldl + i];
}
void orc_inner_update(float* c, int64_t ldc, const float* left, int64_t ldlft,
                      const float* top, int64_t ldt, int64_t B) {
    for (int64_t i = 0; i < B; i++)
        for (int64_t j = 0; j < B; j++) {
            float s = c[i * ldc + j];
            for (int64_t k = 0; k < B; k++)
                s += left[i * ldlft + k] * top[k * ldt + j];
            c[i * ldc + j] = s;
        }
}

// Blocked right-looking GEFA composed of the four block operations in the
// order of execution_pcie.hpp:177-641 (lu -> top/left -> inner), on one rank.
// Used to check that the block decomposition reproduces gefa_ref_nopvt.
void orc_gefa_blocked(float* a, int64_t n, int64_t lda, int64_t B) {
    const int64_t N = n / B;
    for (int64_t k = 0; k < N; k++) {
        float* d = a + k * B * lda + k * B;
        // lu: gefa_ref_nopvt on the diagonal block (hpl_torus_PCIE.cl:211-446)
        orc_gefa_nopvt(d, B, lda);
        for (int64_t j = k + 1; j < N; j++)
            orc_top_update(a + k * B * lda + j * B, lda, d, lda, B);
        for (int64_t i = k + 1; i < N; i++)
            orc_left_update(a + i * B * lda + k * B, lda, d, lda, B);
        for (int64_t i = k + 1; i < N; i++)
            for (int64_t j = k + 1; j < N; j++)
                orc_inner_update(a + i * B * lda + j * B, lda,
                                 a + i * B * lda + k * B, lda,
                                 a + k * B * lda + j * B, lda, B);
    }
}

// ---------------------------------------------------------------------------
// Mixed-precision trailing update (no reference counterpart; the reference's
// inner_update_mm is FP32, hpl_torus_PCIE.cl:711-865).  Rounding of a value to
// a binary format with `frac` fraction bits, minimum normal exponent `emin`
// and largest finite value `maxv`, round to nearest with ties to even (the
// default rounding mode of nearbyint), subnormals included; values beyond
// maxv saturate (only FP8 can reach it, and its block scale prevents that).
static double round_format(double x, int frac, int emin, double maxv) {
    if (x == 0.0 || !std::isfinite(x)) return x;
    int e = std::ilogb(x);
    if (e < emin) e = emin;
    const double q = std::ldexp(std::nearbyint(std::ldexp(x, frac - e)), e - frac);
    return std::fabs(q) > maxv ? std::copysign(maxv, q) : q;
}

double orc_round_lp(double x, int prec) {
    switch (prec) {
        case 1: return round_format(x, 7, -126, 3.3895313892515355e38);  // bfloat16
        case 2: return round_format(x, 10, -14, 65504.0);                // IEEE binary16
        case 3: return round_format(x, 3, -6, 448.0);                    // OCP E4M3
        default: return x;
    }
}

// Power-of-two block scale of FP8: 2^e with e the smallest integer such that
// amax <= 448 * 2^e (1 for an all-zero block).
double orc_fp8_scale(double amax) {
    if (!(amax > 0.0)) return 1.0;
    int e = (int)std::floor(std::log2(amax / 448.0));
    while (amax > 448.0 * std::ldexp(1.0, e)) ++e;
    while (amax <= 448.0 * std::ldexp(1.0, e - 1)) --e;
    return std::ldexp(1.0, e);
}

// Quantise-dequantise `count` vectors of length B (vector v at x + v*stride_v,
// element k at + k*stride_k) as the trailing update sees them: BF16/FP16
// rounding, or FP8 with one scale per group of 32 consecutive vectors.
static void quantise_vectors(const float* x, int64_t stride_v, int64_t stride_k, int64_t count,
                             int64_t B, int prec, double* out /* count x B */) {
    for (int64_t g = 0; g < count; g += 32) {
        const int64_t ge = std::min<int64_t>(count, g + 32);
        double scale = 1.0;
        if (prec == 3) {
            double amax = 0.0;
            for (int64_t v = g; v < ge; v++)
                for (int64_t k = 0; k < B; k++)
                    amax = std::max(amax, std::fabs((double)x[v * stride_v + k * stride_k]));
            scale = orc_fp8_scale(amax);
        }
        for (int64_t v = g; v < ge; v++)
            for (int64_t k = 0; k < B; k++)
                out[v * B + k] =
                    scale * orc_round_lp((double)x[v * stride_v + k * stride_k] / scale, prec);
    }
}

// Blocked factorisation as orc_gefa_blocked, with the trailing update taken
// from low-precision copies of the solved strips (prec 1 BF16, 2 FP16, 3 FP8)
// and accumulated exactly (double) before rounding to FP32 once per step.
void orc_gefa_blocked_lp(float* a, int64_t n, int64_t lda, int64_t B, int prec) {
    const int64_t N = n / B;
    std::vector<double> Lq, Tq;
    for (int64_t k = 0; k < N; k++) {
        float* d = a + k * B * lda + k * B;
        orc_gefa_nopvt(d, B, lda);
        for (int64_t j = k + 1; j < N; j++)
            orc_top_update(a + k * B * lda + j * B, lda, d, lda, B);
        for (int64_t i = k + 1; i < N; i++)
            orc_left_update(a + i * B * lda + k * B, lda, d, lda, B);
        const int64_t o = (k + 1) * B, m = n - o;
        if (m <= 0) break;
        // top strip: vectors = columns c (element k at row k*B + kk);
        // left strip: vectors = rows r (element at column k*B + kk)
        Tq.assign(m * B, 0.0);
        Lq.assign(m * B, 0.0);
        quantise_vectors(a + k * B * lda + o, 1, lda, m, B, prec, Tq.data());
        quantise_vectors(a + o * lda + k * B, lda, 1, m, B, prec, Lq.data());
        for (int64_t r = 0; r < m; r++)
            for (int64_t c = 0; c < m; c++) {
                double s = (double)a[(o + r) * lda + o + c];
                const double* lr = &Lq[r * B];
                const double* tc = &Tq[c * B];
                for (int64_t kk = 0; kk < B; kk++) s += lr[kk] * tc[kk];
                a[(o + r) * lda + o + c] = (float)s;
            }
    }
}

// ---------------------------------------------------------------------------
// A11: validation metrics.
// Reference residual (linpack_benchmark.hpp:400-417, DISTRIBUTED_VALIDATION
// branch): resid = max|x_i - 1|, residn = resid / (n * normb * eps_f32).
// ---------------------------------------------------------------------------
void orc_residn_ref(const float* x, int64_t n, float normb, double* resid,
                    double* residn) {
    double r = 0.0;
    for (int64_t i = 0; i < n; i++) {
        double v = std::abs(x[i] - 1);
        r = (r > v) ? r : v;
    }
    float eps = std::numeric_limits<float>::epsilon();
    *resid = r;
    *residn = r / (static_cast<double>(n) * normb * eps);
}

// Norm-wise quantities in double on the float matrix (M = A^T):
//   out[0] = ||b - M x||_inf, out[1] = ||M||_inf, out[2] = ||x||_inf,
//   out[3] = ||b||_inf.
void orc_norms(const float* A, int64_t n, int64_t lda, const double* x,
               const double* b, double* out) {
    std::vector<double> r(b, b + n), rowabs(n, 0.0);
    for (int64_t j = 0; j < n; j++) {
        const float* col = A + j * lda;  // M(:, j)
        const double xj = x[j];
        for (int64_t i = 0; i < n; i++) {
            r[i] -= (double)col[i] * xj;
            rowabs[i] += std::fabs((double)col[i]);
        }
    }
    double rn = 0, an = 0, xn = 0, bn = 0;
    for (int64_t i = 0; i < n; i++) {
        rn = std::max(rn, std::fabs(r[i]));
        an = std::max(an, rowabs[i]);
        xn = std::max(xn, std::fabs(x[i]));
        bn = std::max(bn, std::fabs(b[i]));
    }
    out[0] = rn; out[1] = an; out[2] = xn; out[3] = bn;
}

}  // extern "C"
