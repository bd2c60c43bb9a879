// LINPACK benchmark host with the HIP execution backend.
//
// Mirrors the reference's LinpackBenchmark (LINPACK/src/host/linpack_benchmark.hpp:50-578),
// LinpackProgramSettings / LinpackData (LINPACK/src/host/linpack_data.hpp:48-214,
// linpack_data.cpp:37-65) and the backend entry point
// execution::pcie::calculate (LINPACK/src/host/execution_types/execution_pcie.hpp:53-728),
// single rank.  The HIP backend factorises AND solves on the GPU: after
// calculate(), data.A holds the dgefa-encoded factor and data.b the solution
// x, the state the reference reaches after its device GEFA plus its host
// distributed_gesl_nopvt_ref (linpack_benchmark.hpp:306-324).
#pragma once

#include <omp.h>
#include <stdlib.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <limits>
#include <random>
#include <type_traits>

// The element type, as the reference's HOST_DATA_TYPE: float by default,
// double in the DATA_TYPE=double build (Linpack_hip_dp, -DHOST_DATA_TYPE=double),
// whose factorisation and solve run on the FP64 calls of the C ABI.
#ifndef HOST_DATA_TYPE
#define HOST_DATA_TYPE float
#endif

#include "../../include/hpl_hip.h"
#include "hpcc_base.hpp"

namespace linpack {

using hpcc_base::CommunicationType;
using real_t = HOST_DATA_TYPE;
constexpr bool kDoublePrecision = std::is_same<real_t, double>::value;

// Error convention: a non-zero hpl_status becomes an exception, as ASSERT_CL
// does in the reference (shared/setup/fpga_setup.cpp:112-122).
class HipException : public std::runtime_error {
   public:
    HipException(hpl_status st, const std::string& msg)
        : std::runtime_error(std::string(hpl_status_string(st)) + ": " + msg), status(st) {}
    hpl_status status;
};

// Device context of the HIP backend (the TContext of the benchmark).
class HipContext {
   public:
    explicit HipContext(int device) {
        hpl_status st = hpl_create(device, &ctx_);
        if (st != HPL_SUCCESS) throw HipException(st, "could not open HIP device");
    }
    ~HipContext() { hpl_destroy(ctx_); }
    HipContext(const HipContext&) = delete;
    HipContext& operator=(const HipContext&) = delete;
    hpl_ctx* get() const { return ctx_; }
    std::string name() const {
        char buf[256] = {0};
        if (hpl_device_name(ctx_, buf, sizeof buf) != HPL_SUCCESS) return "unknown";
        return buf;
    }

   private:
    hpl_ctx* ctx_ = nullptr;
};

#define ASSERT_HPL(ctx, expr)                                                     \
    do {                                                                          \
        hpl_status st_ = (expr);                                                  \
        if (st_ != HPL_SUCCESS) throw HipException(st_, hpl_last_error(ctx));     \
    } while (0)

class LinpackProgramSettings : public hpcc_base::BaseSettings {
   public:
    unsigned matrixSize;  // n = m * 2^b (values)
    unsigned blockSize;   // 2^b
    bool isDiagonallyDominant;
    bool isEmulationKernel;
    int torus_row = 0, torus_col = 0, torus_width, torus_height;
    std::string generator;  // "ref" (reference generator) or "sized"
    uint64_t seed;

    explicit LinpackProgramSettings(const hpcc_base::Options::Result& r)
        : hpcc_base::BaseSettings(r),
          matrixSize((unsigned)(r.integer("m") * (1LL << r.integer("b")))),
          blockSize((unsigned)(1u << r.integer("b"))),
          isDiagonallyDominant(r.count("uniform") == 0),
          isEmulationKernel(r.count("emulation") > 0),
          torus_width((int)r.integer("p")),
          generator(r.str("generator")),
          seed((uint64_t)r.integer("seed")) {
        const int mpi_size = 1;  // single rank
        if (torus_width <= 0 || mpi_size % torus_width != 0)
            throw std::runtime_error("MPI size not dividable by P=" + std::to_string(torus_width) +
                                     "!");
        torus_height = mpi_size / torus_width;
        if (generator != "ref" && generator != "sized")
            throw std::runtime_error("unknown generator '" + generator + "' (ref or sized)");
    }

    std::map<std::string, std::string> getSettingsMap() override {
        auto map = hpcc_base::BaseSettings::getSettingsMap();
        map["Matrix Size"] = std::to_string(matrixSize);
        map["Block Size"] = std::to_string(blockSize);
        map["Emulate"] = isEmulationKernel ? "Yes" : "No";
        map["Diagonally Dominant"] = isDiagonallyDominant ? "Yes" : "No";
        map["Data Type"] = kDoublePrecision ? "cl_double" : "cl_float";
        map["FPGA Torus"] = "P=" + std::to_string(torus_width) + ", Q=" + std::to_string(torus_height);
        map["Generator"] = generator;
        return map;
    }
};

class LinpackData {
   public:
    real_t* A = nullptr;
    real_t* b = nullptr;
    int* ipvt = nullptr;
    size_t matrix_width, matrix_height;
    real_t norma = 0, normb = 0;
    LinpackData(size_t width, size_t height) : matrix_width(width), matrix_height(height) {
        if (posix_memalign(reinterpret_cast<void**>(&A), 4096, width * height * sizeof(real_t)) ||
            posix_memalign(reinterpret_cast<void**>(&b), 4096, width * sizeof(real_t)) ||
            posix_memalign(reinterpret_cast<void**>(&ipvt), 4096, height * sizeof(int)))
            throw std::bad_alloc();
    }
    ~LinpackData() {
        free(A);
        free(b);
        free(ipvt);
    }
    LinpackData(const LinpackData&) = delete;
    LinpackData& operator=(const LinpackData&) = delete;
};

using Settings = hpcc_base::ExecutionSettings<LinpackProgramSettings, HipContext>;

namespace execution {
namespace hip {

// The HIP counterpart of execution::pcie::calculate: upload once (outside the
// timers, as execution_pcie.hpp:123-127), then per repetition restore the
// matrix on the device and time GEFA and GESL with HIP events.  Leaves the
// factor in data.A and the solution x in data.b.  A --uniform matrix is
// factorised with partial pivoting (gefa_ref, linpack_data.cpp:72-119): data.A
// then holds gefa_ref's layout and data.ipvt its pivots.
inline std::map<std::string, std::vector<double>> calculate(const Settings& config,
                                                             LinpackData& data) {
    hpl_ctx* c = config.context->get();
    const uint64_t n = config.programSettings->matrixSize;
    const uint32_t B = config.programSettings->blockSize;
    const bool pivoting = !config.programSettings->isDiagonallyDominant;
    hpl_timing t{};
    if constexpr (kDoublePrecision) {
        // DATA_TYPE=double: the FP64 system and calls (block 64)
        ASSERT_HPL(c, hpl_upload_f64(c, reinterpret_cast<double*>(data.A), n, n,
                                     reinterpret_cast<double*>(data.b), &t));
        std::vector<double> gefa, gesl;
        for (unsigned rep = 0; rep < config.programSettings->numRepetitions; ++rep) {
            if (rep > 0) ASSERT_HPL(c, hpl_restore_f64(c));
            ASSERT_HPL(c, hpl_gefa_dev_f64(c, &t));
            ASSERT_HPL(c, hpl_gesl_dev_f64(c, &t));
            gefa.push_back(t.gefa_s);
            gesl.push_back(t.gesl_s);
        }
        ASSERT_HPL(c, hpl_download_f64(c, reinterpret_cast<double*>(data.A), n,
                                       reinterpret_cast<double*>(data.b), &t));
        return {{"gefa", gefa}, {"gesl", gesl}};
    }
    ASSERT_HPL(c, hpl_keep_original(c, 1));
    ASSERT_HPL(c, hpl_set_pivoting(c, pivoting ? 1 : 0));
    ASSERT_HPL(c, hpl_upload(c, reinterpret_cast<float*>(data.A), n, n,
                             reinterpret_cast<float*>(data.b), &t));
    std::vector<double> gefa, gesl;
    for (unsigned rep = 0; rep < config.programSettings->numRepetitions; ++rep) {
        if (rep > 0) ASSERT_HPL(c, hpl_restore(c));
        ASSERT_HPL(c, hpl_gefa_dev(c, B, HPL_PREC_FP32, &t));
        ASSERT_HPL(c, hpl_gesl_dev(c, &t));
        gefa.push_back(t.gefa_s);
        gesl.push_back(t.gesl_s);
    }
    ASSERT_HPL(c, hpl_download(c, reinterpret_cast<float*>(data.A), n,
                               reinterpret_cast<float*>(data.b), &t));
    if (pivoting) ASSERT_HPL(c, hpl_get_pivots(c, data.ipvt));
    return {{"gefa", gefa}, {"gesl", gesl}};
}

}  // namespace hip

namespace cpu {

// The CPU backend (communication type "CPU", cpu_only in
// shared/include/communication_types.hpp:58,88, which the reference's LINPACK
// declares but has no calculate for, linpack_benchmark.hpp:306-320): the
// reference's own host algorithms, gefa_ref_nopvt / gesl_ref_nopvt
// (LINPACK/src/host/linpack_data.cpp:166-229) or, for a --uniform matrix,
// gefa_ref / gesl_ref (:72-156).  Within an elimination step the columns
// right of the pivot are independent, so they are spread over OpenMP threads;
// every element still receives the reference's operations in the reference's
// order (no contraction on the default x86-64 target), so the factor and the
// solution are the reference loops' bit for bit.

inline void gefa_nopvt(real_t* a, int64_t n, int64_t lda) {
    for (int64_t k = 0; k < n; ++k) {
        real_t* ck = a + k * lda;
        ck[k] = -1.0 / ck[k];
        for (int64_t i = k + 1; i < n; ++i) ck[i] *= ck[k];
#pragma omp parallel for schedule(static) if (n - k > 128)
        for (int64_t j = k + 1; j < n; ++j) {
            real_t* cj = a + j * lda;
            const real_t s = cj[k];
            for (int64_t i = k + 1; i < n; ++i) cj[i] += ck[i] * s;
        }
    }
}

inline void gesl_nopvt(const real_t* a, real_t* b, int64_t n, int64_t lda) {
    for (int64_t k = 0; k + 1 < n; ++k)
        for (int64_t i = k + 1; i < n; ++i) b[i] += b[k] * a[lda * k + i];
    for (int64_t k = n - 1; k >= 0; --k) {
        const real_t scale = b[k] * a[lda * k + k];
        b[k] = -scale;
        for (int64_t i = 0; i < k; ++i) b[i] += scale * a[lda * k + i];
    }
}

// Partial pivoting: column k takes the first row of largest |M(i,k)|.
inline void gefa_pvt(real_t* a, int64_t n, int64_t lda, int* ipvt) {
    for (int64_t i = 0; i < n; ++i) ipvt[i] = (int)i;
    for (int64_t k = 0; k + 1 < n; ++k) {
        real_t* ck = a + k * lda;
        real_t max_val = std::abs(ck[k]);
        int64_t p = k;
        for (int64_t i = k + 1; i < n; ++i)
            if (max_val < std::abs(ck[i])) {
                p = i;
                max_val = std::abs(ck[i]);
            }
        for (int64_t i = k; i < n; ++i) std::swap(a[i * lda + k], a[i * lda + p]);
        ipvt[k] = (int)p;
        for (int64_t i = k + 1; i < n; ++i) ck[i] *= -1.0 / ck[k];
#pragma omp parallel for schedule(static) if (n - k > 128)
        for (int64_t j = k + 1; j < n; ++j) {
            real_t* cj = a + j * lda;
            const real_t s = cj[k];
            for (int64_t i = k + 1; i < n; ++i) cj[i] += ck[i] * s;
        }
    }
}

inline void gesl_pvt(const real_t* a, real_t* b, const int* ipvt, int64_t n, int64_t lda) {
    for (int64_t k = 0; k + 1 < n; ++k) {
        if (ipvt[k] != k) std::swap(b[k], b[ipvt[k]]);
        for (int64_t i = k + 1; i < n; ++i) b[i] += b[k] * a[lda * k + i];
    }
    for (int64_t k = n - 1; k >= 0; --k) {
        b[k] = b[k] / a[lda * k + k];
        for (int64_t i = 0; i < k; ++i) b[i] -= b[k] * a[lda * k + i];
    }
}

// Per repetition: factorise a fresh copy of the generated matrix, then solve
// a fresh copy of b, each timed with the host clock.  Leaves the factor in
// data.A, the pivots in data.ipvt and the solution x in data.b, as the HIP
// backend does.
inline std::map<std::string, std::vector<double>> calculate(const Settings& config,
                                                             LinpackData& data) {
    const int64_t n = config.programSettings->matrixSize;
    const bool pivoting = !config.programSettings->isDiagonallyDominant;
    const std::vector<real_t> a0(data.A, data.A + n * n), b0(data.b, data.b + n);
    std::vector<double> gefa, gesl;
    for (unsigned rep = 0; rep < config.programSettings->numRepetitions; ++rep) {
        std::copy(a0.begin(), a0.end(), data.A);
        std::copy(b0.begin(), b0.end(), data.b);
        const auto t0 = std::chrono::steady_clock::now();
        if (pivoting)
            gefa_pvt(data.A, n, n, data.ipvt);
        else
            gefa_nopvt(data.A, n, n);
        const auto t1 = std::chrono::steady_clock::now();
        if (pivoting)
            gesl_pvt(data.A, data.b, data.ipvt, n, n);
        else
            gesl_nopvt(data.A, data.b, n, n);
        const auto t2 = std::chrono::steady_clock::now();
        gefa.push_back(std::chrono::duration<double>(t1 - t0).count());
        gesl.push_back(std::chrono::duration<double>(t2 - t1).count());
    }
    return {{"gefa", gefa}, {"gesl", gesl}};
}

}  // namespace cpu
}  // namespace execution

class LinpackBenchmark
    : public hpcc_base::HpccBenchmark<LinpackProgramSettings, HipContext, LinpackData> {
   protected:
    std::string programName() const override { return "LINPACK"; }

    void addAdditionalParseOptions(hpcc_base::Options& options) override {
        options
            .add("m",
                 "Global matrix size in number of blocks in one dimension. Local matrix sizes "
                 "will be determined by PQ grid.",
                 "2")
            .add("b", "Log2 of the block size in number of values in one dimension", "5")
            .add("p", "Width of the FPGA grid. The heigth (Q) will be calculated from mpi_size / P.",
                 "1")
            .flag("uniform",
                  "Generate a uniform matrix instead of a diagonally dominant. This has to be "
                  "supported by the FPGA kernel!")
            .flag("emulation",
                  "Use kernel arguments for emulation. This may be necessary to simulate "
                  "persistent local memory on the FPGA")
            .add("generator",
                 "Input generator: ref (mt19937(rank), U(0,1), diagonal = row sum) or sized "
                 "(counter-based U[-0.5,0.5), A_ii += n, b = A*1)",
                 "ref")
            .add("seed", "Seed of the sized generator", "42");
    }

    std::unique_ptr<HipContext> createContext(const LinpackProgramSettings& s) override {
        if (s.communicationType != CommunicationType::hip) return nullptr;
        return std::unique_ptr<HipContext>(new HipContext(s.defaultDevice));
    }

   public:
    bool checkInputParameters() override {
        auto& s = *executionSettings->programSettings;
        if (s.communicationType == CommunicationType::hip) {
            if (kDoublePrecision && !s.isDiagonallyDominant) {
                std::cerr << "ERROR: the DATA_TYPE=double build factorises without pivoting; "
                             "--uniform needs the float build"
                          << std::endl;
                return false;
            }
            if (kDoublePrecision && s.matrixSize % 64) {
                std::cerr << "ERROR: the DATA_TYPE=double build needs a matrix size that is a "
                             "multiple of 64"
                          << std::endl;
                return false;
            }
            if (!s.isDiagonallyDominant && s.blockSize > 256) {
                std::cerr << "ERROR: partial pivoting (--uniform) needs a block size in "
                             "[32, 256] (b in 5..8)"
                          << std::endl;
                return false;
            }
            if (s.blockSize < 32 || s.blockSize > 512) {
                std::cerr << "ERROR: the HIP backend needs a block size in [32, 512] (b in 5..9)"
                          << std::endl;
                return false;
            }
        }
        return true;
    }

    // Input generation (linpack_benchmark.hpp:211-297) for one rank.
    std::unique_ptr<LinpackData> generateInputData() override {
        auto& s = *executionSettings->programSettings;
        const size_t n = s.matrixSize;
        if ((s.matrixSize / s.blockSize) % s.torus_width > 0 ||
            (s.matrixSize / s.blockSize) % s.torus_height > 0)
            throw std::runtime_error("Global matrix size must be multiple of LCM of PQ grid!");
        std::unique_ptr<LinpackData> d(new LinpackData(n, n));
        if (s.generator == "ref")
            generateReference(*d, n, s.isDiagonallyDominant);
        else
            generateSized(*d, n, s.seed, s.isDiagonallyDominant);
        for (size_t i = 0; i < n; ++i) d->ipvt[i] = (int)i;
        return d;
    }

    // mt19937(rank) draws U(0,1) as double, stored as HOST_DATA_TYPE row by
    // row; a diagonally dominant matrix gets the row sum (in HOST_DATA_TYPE)
    // on its diagonal; b is the column sum, so the solution is all ones.
    void generateReference(LinpackData& d, size_t n, bool dd) {
        std::mt19937 gen(mpi_comm_rank);
        std::uniform_real_distribution<> dis(0.0, 1.0);
        real_t* A = d.A;
        d.norma = 0;
        d.normb = 0;
        for (size_t j = 0; j < n; ++j)
            for (size_t i = 0; i < n; ++i) {
                const real_t v = dis(gen);
                A[n * j + i] = v;
                if (v > d.norma) d.norma = v;
            }
        if (dd) {
            for (size_t r = 0; r < n; ++r) {
                A[n * r + r] = 0;
                real_t sum = 0;
                for (size_t i = 0; i < n; ++i) sum += A[n * r + i];
                if (sum > d.norma) d.norma = sum;
                A[n * r + r] = sum;
            }
        }
        for (size_t j = 0; j < n; ++j) {
            real_t sum = 0;
            for (size_t i = 0; i < n; ++i) sum += A[n * i + j];
            d.b[j] = sum;
            if (sum > d.normb) d.normb = sum;
        }
    }

    // Counter-based generator of the sized workloads (the device generator
    // hpl_generate mode 0, or mode 2 without dominance, produces the same
    // bits): A = U[-0.5,0.5) on a 2^-24 grid from splitmix64 of (seed, r*n+c),
    // A_ii += n when diagonally dominant, b = row sums of M in double rounded
    // to float (the DATA_TYPE=double build widens these float values, as
    // hpl_generate_f64 does).
    static uint64_t mix64(uint64_t z) {
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    void generateSized(LinpackData& d, size_t n, uint64_t seed, bool dd) {
        const uint64_t key = mix64(seed ^ (0ULL << 56));
        real_t* A = d.A;
#pragma omp parallel for schedule(static)
        for (long long r = 0; r < (long long)n; ++r)
            for (size_t c = 0; c < n; ++c) {
                const uint64_t z = mix64(key + ((uint64_t)r * n + c + 1) * 0x9E3779B97F4A7C15ULL);
                A[r * n + c] = (real_t)((float)(uint32_t)(z >> 40) * (1.0f / 16777216.0f) - 0.5f);
            }
        if (dd)
            for (size_t i = 0; i < n; ++i) A[i * n + i] = (real_t)((float)A[i * n + i] + (float)n);
        std::vector<double> s(n, 0.0);
        const int nt = omp_get_max_threads();
        const size_t chunk = (n + nt - 1) / nt;
#pragma omp parallel for schedule(static)
        for (int t = 0; t < nt; ++t) {
            const size_t i0 = t * chunk, i1 = std::min(n, i0 + chunk);
            for (size_t j = 0; j < n; ++j)
                for (size_t i = i0; i < i1; ++i) s[i] += (double)A[j * n + i];
        }
        d.norma = 0.0f;
        d.normb = 0.0f;
        for (size_t i = 0; i < n; ++i) {
            d.b[i] = (real_t)s[i];
            if (d.b[i] > d.normb) d.normb = d.b[i];
            if (dd && A[i * n + i] > d.norma) d.norma = A[i * n + i];
        }
        if (!dd) {  // the running maximum of the generated values, as the reference keeps it
            real_t m = 0;
#pragma omp parallel for reduction(max : m)
            for (long long e = 0; e < (long long)(n * n); ++e) m = std::max(m, A[e]);
            d.norma = m;
        }
    }

    // Backend dispatch (linpack_benchmark.hpp:306-324).
    void executeKernel(LinpackData& data) override {
        auto& s = *executionSettings->programSettings;
        switch (s.communicationType) {
            case CommunicationType::hip:
                timings = execution::hip::calculate(*executionSettings, data);
                break;
            case CommunicationType::cpu_only:
                timings = execution::cpu::calculate(*executionSettings, data);
                break;
            default:
                throw std::runtime_error("No calculate method implemented for communication type " +
                                         hpcc_base::commToString(s.communicationType));
        }
    }

    // Validation (linpack_benchmark.hpp:400-469, distributed branch: data.b
    // holds x): resid = max|x_i - 1|, residn = resid / (n * normb * eps_f32).
    // The HPL scaled residual ||Mx - b|| / (eps ||M|| ||x|| n), computed in
    // FP64 on the device against the kept original, is recorded as well.
    bool validateOutput(LinpackData& data) override {
        const unsigned n = executionSettings->programSettings->matrixSize;
        double resid = 0.0;
        const double normx = data.normb;
        for (size_t i = 0; i < data.matrix_width; ++i) {
            const double v = std::abs(data.b[i] - 1);
            resid = (resid > v) ? resid : v;
        }
        const real_t eps = std::numeric_limits<real_t>::epsilon();
        const double residn = resid / (static_cast<double>(n) * normx * eps);
        errors.emplace("epsilon", eps);
        errors.emplace("residual", resid);
        errors.emplace("residual_norm", residn);
        if (executionSettings->context) {
            double nrm[4];
            hpl_ctx* c = executionSettings->context->get();
            if ((kDoublePrecision ? hpl_residual_f64(c, nrm) : hpl_residual(c, nrm)) == HPL_SUCCESS)
                errors.emplace("hpl_scaled_residual", nrm[0] / (eps * nrm[1] * nrm[2] * n));
        }
        return residn < 1;
    }

    void printError() override {
        std::cout << std::setw(ENTRY_SPACE) << " norm. residual" << std::setw(ENTRY_SPACE)
                  << " res. error" << std::setw(ENTRY_SPACE) << " mach. eps" << std::endl;
        std::cout << std::setw(ENTRY_SPACE) << errors.at("residual_norm")
                  << std::setw(ENTRY_SPACE) << errors.at("residual") << std::setw(ENTRY_SPACE)
                  << errors.at("epsilon") << std::endl;
    }

    // Metrics (linpack_benchmark.hpp:480-538).
    void collectResults() override {
        const std::vector<double>& lu = timings["gefa"];
        const std::vector<double>& sl = timings["gesl"];
        double t = 0, tlu = 0, tsl = 0;
        double tmin = std::numeric_limits<double>::max(), lu_min = tmin, sl_min = tmin;
        const double nn = executionSettings->programSettings->matrixSize;
        const double gflop_lu = ((2.0e0 * nn * nn * nn) / 3.0) / 1.0e9;
        const double gflop_sl = (2.0 * (nn * nn)) / 1.0e9;
        for (size_t i = 0; i < lu.size(); ++i) {
            const double cur = lu[i] + sl[i];
            t += cur;
            tlu += lu[i];
            tsl += sl[i];
            tmin = std::min(tmin, cur);
            lu_min = std::min(lu_min, lu[i]);
            sl_min = std::min(sl_min, sl[i]);
        }
        const double reps = (double)lu.size();
        results.emplace("t_mean", hpcc_base::HpccResult(t / reps, "s"));
        results.emplace("t_min", hpcc_base::HpccResult(tmin, "s"));
        results.emplace("tlu_mean", hpcc_base::HpccResult(tlu / reps, "s"));
        results.emplace("tlu_min", hpcc_base::HpccResult(lu_min, "s"));
        results.emplace("tsl_mean", hpcc_base::HpccResult(tsl / reps, "s"));
        results.emplace("tsl_min", hpcc_base::HpccResult(sl_min, "s"));
        results.emplace("gflops", hpcc_base::HpccResult((gflop_lu + gflop_sl) / tmin, "GFLOP/s"));
        results.emplace("gflops_lu", hpcc_base::HpccResult(gflop_lu / lu_min, "GFLOP/s"));
        results.emplace("gflops_sl", hpcc_base::HpccResult(gflop_sl / sl_min, "GFLOP/s"));
    }

    void printResults() override {
        std::cout << std::left << std::setw(ENTRY_SPACE) << " Method" << std::setw(ENTRY_SPACE)
                  << " best" << std::setw(ENTRY_SPACE) << " mean" << std::setw(ENTRY_SPACE)
                  << " GFLOPS" << std::endl;
        std::cout << std::left << std::setw(ENTRY_SPACE) << " total" << results.at("t_min")
                  << results.at("t_mean") << results.at("gflops") << std::endl;
        std::cout << std::left << std::setw(ENTRY_SPACE) << " GEFA" << results.at("tlu_min")
                  << results.at("tlu_mean") << results.at("gflops_lu") << std::endl;
        std::cout << std::left << std::setw(ENTRY_SPACE) << " GESL" << results.at("tsl_min")
                  << results.at("tsl_mean") << results.at("gflops_sl") << std::right << std::endl;
    }

    LinpackBenchmark(int argc, char** argv) { this->setupBenchmark(argc, argv); }
};

}  // namespace linpack
