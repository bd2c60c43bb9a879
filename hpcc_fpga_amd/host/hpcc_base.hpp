// Benchmark framework of the LINPACK host: lifecycle, settings, command line,
// stdout tables and JSON dump, with the behaviour of the reference's
// hpcc_base (shared/include/hpcc_benchmark.hpp, shared/include/hpcc_settings.hpp,
// shared/hpcc_settings.cpp, shared/include/communication_types.hpp).
//
// The reference builds on cxxopts, nlohmann/json and MPI, none of which is
// available here; this header carries a small option parser and JSON writer of
// its own and runs as a single rank (MPI rank 0 of 1).
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#ifndef HPCC_VERSION
#define HPCC_VERSION "2.6"
#endif
#ifndef GIT_COMMIT_HASH
#define GIT_COMMIT_HASH "unknown"
#endif
#ifndef CONFIG_TIME
#define CONFIG_TIME __DATE__ " " __TIME__
#endif

#define HLINE "-------------------------------------------------------------\n"
#define VALUE_SPACE 11
#define UNIT_SPACE 8
#define ENTRY_SPACE (VALUE_SPACE + UNIT_SPACE + 1)

namespace hpcc_base {

// ---------------------------------------------------------------------------
// Communication types (shared/include/communication_types.hpp:35-132), plus
// the HIP backend of this build.
// ---------------------------------------------------------------------------
enum class CommunicationType {
    intel_external_channels,
    pcie_mpi,
    accl,
    cpu_only,
    udp,
    aurora,
    hip,
    unsupported,
    automatic
};

inline const std::map<std::string, CommunicationType>& comm_to_str_map() {
    static const std::map<std::string, CommunicationType> m{
        {"IEC", CommunicationType::intel_external_channels},
        {"PCIE", CommunicationType::pcie_mpi},
        {"ACCL", CommunicationType::accl},
        {"CPU", CommunicationType::cpu_only},
        {"UNSUPPORTED", CommunicationType::unsupported},
        {"UDP", CommunicationType::udp},
        {"AURORA", CommunicationType::aurora},
        {"HIP", CommunicationType::hip},
        {"AUTO", CommunicationType::automatic}};
    return m;
}

inline std::string commToString(CommunicationType c) {
    for (auto& e : comm_to_str_map())
        if (e.second == c) return e.first;
    throw std::runtime_error("Communication type could not be converted to string!");
}

// AUTO matches a type name contained in the kernel file name, as the reference.
inline CommunicationType retrieveCommunicationType(const std::string& comm_name,
                                                   const std::string& kernel_filename) {
    auto it = comm_to_str_map().find(comm_name);
    if (it == comm_to_str_map().end())
        throw std::runtime_error("Communication type could not be converted from string: " +
                                 comm_name);
    if (it->second != CommunicationType::automatic) return it->second;
    for (auto& e : comm_to_str_map())
        if (e.second != CommunicationType::automatic &&
            kernel_filename.find(e.first) != std::string::npos)
            return e.second;
    throw std::runtime_error("Communication type could not be autodetected from kernel_filename: " +
                             kernel_filename);
}

// ---------------------------------------------------------------------------
// Exceptions (fpga_setup::FpgaSetupException / OpenClException counterparts).
// ---------------------------------------------------------------------------
class SetupException : public std::runtime_error {
   public:
    explicit SetupException(const std::string& m) : std::runtime_error(m) {}
};

// ---------------------------------------------------------------------------
// Minimal command-line options (the subset of cxxopts the benchmark uses).
// ---------------------------------------------------------------------------
class Options {
   public:
    struct Opt {
        std::string shortName, longName, help, def;
        bool flag;
    };
    explicit Options(std::string prog, std::string desc)
        : prog_(std::move(prog)), desc_(std::move(desc)) {}
    // spec "f,file" or "n" or "dump-json"
    Options& add(const std::string& spec, const std::string& help, const std::string& def = "",
                 bool flag = false) {
        Opt o;
        auto comma = spec.find(',');
        if (comma != std::string::npos) {
            o.shortName = spec.substr(0, comma);
            o.longName = spec.substr(comma + 1);
        } else if (spec.size() == 1) {
            o.shortName = spec;
        } else {
            o.longName = spec;
        }
        o.help = help;
        o.def = def;
        o.flag = flag;
        opts_.push_back(o);
        return *this;
    }
    Options& flag(const std::string& spec, const std::string& help) {
        return add(spec, help, "", true);
    }

    struct Result {
        std::map<std::string, std::string> values;
        std::map<std::string, int> counts;
        size_t count(const std::string& k) const {
            auto it = counts.find(k);
            return it == counts.end() ? 0 : it->second;
        }
        const std::string& str(const std::string& k) const {
            auto it = values.find(k);
            if (it == values.end()) throw std::runtime_error("Option '" + k + "' has no value");
            return it->second;
        }
        long long integer(const std::string& k) const {
            const std::string& s = str(k);
            char* end = nullptr;
            long long v = std::strtoll(s.c_str(), &end, 10);
            if (s.empty() || *end) throw std::runtime_error("Argument '" + s + "' failed to parse");
            return v;
        }
    };

    Result parse(int argc, char** argv) const {
        Result r;
        for (auto& o : opts_) {
            const std::string key = canonical(o);
            if (!o.flag && !o.def.empty()) r.values[key] = o.def;
        }
        for (int i = 1; i < argc; ++i) {
            std::string a = argv[i];
            std::string name, value;
            bool has_value = false;
            if (a.rfind("--", 0) == 0) {
                name = a.substr(2);
                auto eq = name.find('=');
                if (eq != std::string::npos) {
                    value = name.substr(eq + 1);
                    name = name.substr(0, eq);
                    has_value = true;
                }
            } else if (a.size() >= 2 && a[0] == '-') {
                name = a.substr(1, 1);
                if (a.size() > 2) {
                    value = a.substr(2);
                    has_value = true;
                }
            } else {
                throw std::runtime_error("Unexpected positional argument " + a);
            }
            const Opt* o = find(name);
            if (!o) throw std::runtime_error("Option '" + name + "' does not exist");
            const std::string key = canonical(*o);
            r.counts[key]++;
            if (o->flag) continue;
            if (!has_value) {
                if (i + 1 >= argc)
                    throw std::runtime_error("Option '" + name + "' is missing an argument");
                value = argv[++i];
            }
            r.values[key] = value;
        }
        return r;
    }

    std::string help() const {
        std::ostringstream ss;
        ss << desc_ << "\nUsage:\n  " << prog_ << " [OPTION...]\n\n";
        for (auto& o : opts_) {
            std::string n = o.shortName.empty() ? "    --" + o.longName
                                                : "-" + o.shortName +
                                                      (o.longName.empty() ? "" : ", --" + o.longName);
            if (!o.flag) n += " arg";
            ss << "  " << std::left << std::setw(28) << n << o.help;
            if (!o.def.empty()) ss << " (default: " << o.def << ")";
            ss << "\n";
        }
        return ss.str();
    }

   private:
    static std::string canonical(const Opt& o) { return o.shortName.empty() ? o.longName : o.shortName; }
    const Opt* find(const std::string& n) const {
        for (auto& o : opts_)
            if (o.shortName == n || o.longName == n) return &o;
        return nullptr;
    }
    std::string prog_, desc_;
    std::vector<Opt> opts_;
};

// ---------------------------------------------------------------------------
// Minimal JSON value + writer (the subset nlohmann::json provides for the dump).
// ---------------------------------------------------------------------------
class Json {
   public:
    enum Kind { Null, Bool, Number, Integer, String, Object, Array };
    Json() : k_(Null) {}
    Json(bool b) : k_(Bool), b_(b) {}
    Json(double d) : k_(Number), d_(d) {}
    Json(int i) : k_(Integer), i_(i) {}
    Json(long long i) : k_(Integer), i_(i) {}
    Json(const char* s) : k_(String), s_(s) {}
    Json(std::string s) : k_(String), s_(std::move(s)) {}
    static Json object() { Json j; j.k_ = Object; return j; }
    static Json array() { Json j; j.k_ = Array; return j; }
    Json& operator[](const std::string& key) {
        if (k_ == Null) k_ = Object;
        for (auto& kv : obj_)
            if (kv.first == key) return kv.second;
        obj_.emplace_back(key, Json());
        return obj_.back().second;
    }
    void push_back(Json v) {
        if (k_ == Null) k_ = Array;
        arr_.push_back(std::move(v));
    }
    std::string dump() const {
        std::ostringstream ss;
        write(ss);
        return ss.str();
    }

   private:
    static void esc(std::ostream& os, const std::string& s) {
        os << '"';
        for (unsigned char c : s) {
            switch (c) {
                case '"': os << "\\\""; break;
                case '\\': os << "\\\\"; break;
                case '\n': os << "\\n"; break;
                case '\t': os << "\\t"; break;
                case '\r': os << "\\r"; break;
                default:
                    if (c < 0x20) {
                        char b[8];
                        snprintf(b, sizeof b, "\\u%04x", c);
                        os << b;
                    } else {
                        os << c;
                    }
            }
        }
        os << '"';
    }
    void write(std::ostream& os) const {
        switch (k_) {
            case Null: os << "null"; break;
            case Bool: os << (b_ ? "true" : "false"); break;
            case Integer: os << i_; break;
            case Number: {
                if (!std::isfinite(d_)) {
                    os << "null";
                } else {
                    char b[32];
                    snprintf(b, sizeof b, "%.17g", d_);
                    os << b;
                }
                break;
            }
            case String: esc(os, s_); break;
            case Object: {
                // keys in lexicographic order, as nlohmann::json orders them
                std::vector<const std::pair<std::string, Json>*> v;
                for (auto& kv : obj_) v.push_back(&kv);
                std::sort(v.begin(), v.end(),
                          [](auto* a, auto* b) { return a->first < b->first; });
                os << '{';
                for (size_t i = 0; i < v.size(); ++i) {
                    if (i) os << ',';
                    esc(os, v[i]->first);
                    os << ':';
                    v[i]->second.write(os);
                }
                os << '}';
                break;
            }
            case Array: {
                os << '[';
                for (size_t i = 0; i < arr_.size(); ++i) {
                    if (i) os << ',';
                    arr_[i].write(os);
                }
                os << ']';
                break;
            }
        }
    }
    Kind k_;
    bool b_ = false;
    double d_ = 0;
    long long i_ = 0;
    std::string s_;
    std::vector<std::pair<std::string, Json>> obj_;
    std::vector<Json> arr_;
};

// ---------------------------------------------------------------------------
// Results and settings (hpcc_benchmark.hpp:70-92, hpcc_settings.hpp:38-154).
// ---------------------------------------------------------------------------
class HpccResult {
   public:
    double value;
    std::string unit;
    HpccResult(double v, std::string u) : value(v), unit(std::move(u)) {}
    friend std::ostream& operator<<(std::ostream& os, const HpccResult& r) {
        os << std::setw(VALUE_SPACE) << r.value << " " << std::left << std::setw(UNIT_SPACE)
           << r.unit << std::right;
        return os;
    }
};

class BaseSettings {
   public:
    unsigned numRepetitions;
    bool useMemoryInterleaving = true;
    bool skipValidation;
    int defaultPlatform;
    std::string platformString;
    int defaultDevice;
    std::string kernelFileName;
    unsigned kernelReplications;
    bool testOnly;
    std::string dumpfilePath;
    CommunicationType communicationType;

    explicit BaseSettings(const Options::Result& r)
        : numRepetitions((unsigned)r.integer("n")),
          skipValidation(r.count("skip-validation") > 0),
          defaultPlatform((int)r.integer("platform")),
          platformString(r.values.count("platform_str") ? r.str("platform_str") : ""),
          defaultDevice((int)r.integer("device")),
          kernelFileName(r.str("f")),
          kernelReplications(r.count("r") ? (unsigned)r.integer("r") : 1),
          testOnly(r.count("test") > 0),
          dumpfilePath(r.values.count("dump-json") ? r.str("dump-json") : ""),
          communicationType(retrieveCommunicationType(r.str("comm-type"), r.str("f"))) {}
    virtual ~BaseSettings() = default;

    virtual std::map<std::string, std::string> getSettingsMap() {
        return {{"Repetitions", std::to_string(numRepetitions)},
                {"Kernel Replications", std::to_string(kernelReplications)},
                {"Kernel File", kernelFileName},
                {"MPI Ranks", "None"},
                {"Test Mode", testOnly ? "Yes" : "No"},
                {"Communication Type", commToString(communicationType)}};
    }
};

// Owns the program settings and the device handle of one run
// (hpcc_settings.hpp:162-235).  TContext is the backend's device context.
template <class TSettings, class TContext>
class ExecutionSettings {
   public:
    std::unique_ptr<TSettings> programSettings;
    int device;
    std::unique_ptr<TContext> context;
    ExecutionSettings(std::unique_ptr<TSettings> s, int dev, std::unique_ptr<TContext> ctx)
        : programSettings(std::move(s)), device(dev), context(std::move(ctx)) {}
    std::string getDeviceName() const { return context ? context->name() : "None"; }
};

// ---------------------------------------------------------------------------
// Benchmark base class (hpcc_benchmark.hpp:101-755), single rank.
// ---------------------------------------------------------------------------
template <class TSettings, class TContext, class TData>
class HpccBenchmark {
   private:
    bool benchmark_setup_succeeded = false;

   protected:
    std::unique_ptr<ExecutionSettings<TSettings, TContext>> executionSettings;
    int mpi_comm_rank = 0;
    int mpi_comm_size = 1;
    std::map<std::string, std::vector<double>> timings;
    std::map<std::string, HpccResult> results;
    std::map<std::string, double> errors;
    bool validated = false;

    virtual void addAdditionalParseOptions(Options& options) = 0;
    virtual std::string programName() const = 0;
    virtual std::unique_ptr<TContext> createContext(const TSettings& settings) = 0;

   public:
    virtual std::unique_ptr<TData> generateInputData() = 0;
    virtual void executeKernel(TData& data) = 0;
    virtual bool validateOutput(TData& data) = 0;
    virtual void printError() = 0;
    virtual void collectResults() = 0;
    virtual void printResults() = 0;
    virtual bool checkInputParameters() { return true; }

    std::string programDescription() const {
        return "Implementation of the " + programName() +
               " benchmark proposed in the HPCC benchmark suite for FPGA.\nVersion: " HPCC_VERSION
               "\n";
    }

    std::unique_ptr<TSettings> parseProgramParameters(int argc, char** argv) {
        std::ostringstream ss;
        ss << programDescription() << "Config. Time: " << CONFIG_TIME << "\n"
           << "Git Commit:   " << GIT_COMMIT_HASH;
        Options options(argv[0], ss.str());
        options.add("f,file", "Kernel file name (for the HIP backend: any name, e.g. HIP)")
            .add("n", "Number of repetitions", "10")
            .flag("skip-validation",
                  "Skip the validation of the output data. This will speed up execution and "
                  "helps when working with special data types.")
            .add("device", "Index of the device that has to be used.", "-1")
            .add("platform", "Index of the platform that has to be used.", "-1")
            .add("platform_str", "Name of the platform that has to be used", "")
            .add("r", "Number of used kernel replications", "1")
            .add("comm-type", "Used communication type for inter-device communication", "AUTO")
            .add("dump-json", "dump benchmark configuration and results to this file in json format",
                 "")
            .flag("test", "Only test given configuration and skip execution and validation")
            .flag("h,help", "Print this help");
        addAdditionalParseOptions(options);
        Options::Result result;
        try {
            result = options.parse(argc, argv);
        } catch (const std::exception& e) {
            throw SetupException(
                "Input parameters could not be parsed! Use -h to show all available options. "
                "ERROR: " + std::string(e.what()));
        }
        if (result.count("h")) {
            std::cout << options.help() << std::endl;
            exit(0);
        }
        if (result.count("f") <= 0)
            throw SetupException(
                "Mandatory option is missing! Use -h to show all available options. ERROR: "
                "Kernel file must be given with option -f!");
        return std::unique_ptr<TSettings>(new TSettings(result));
    }

    void printFinalConfiguration() {
        std::cout << programDescription();
        std::cout << "Config. Time: " << CONFIG_TIME << std::endl;
        std::cout << "Git Commit:   " << GIT_COMMIT_HASH << std::endl;
        std::cout << std::endl << "Summary:" << std::endl;
        std::cout << std::left;
        for (auto& k : executionSettings->programSettings->getSettingsMap())
            std::cout << std::setw(2 * ENTRY_SPACE) << k.first << k.second << std::endl;
        std::cout << std::setw(2 * ENTRY_SPACE) << "Device" << executionSettings->getDeviceName()
                  << std::endl;
        std::cout << std::right << std::endl;
    }

    // Settings as JSON (hpcc_benchmark.hpp:460-482), including its treatment of
    // every non-integer string setting other than "FPGA Torus" as a Yes/No flag.
    Json jsonifySettingsMap(const std::map<std::string, std::string>& m) {
        Json j = Json::object();
        for (auto& item : m) {
            const std::string& v = item.second;
            char* end = nullptr;
            long val = std::strtol(v.c_str(), &end, 10);
            if (!v.empty() && end != v.c_str()) {
                j[item.first] = Json((long long)val);
            } else if (item.first == "FPGA Torus") {
                Json t = Json::object();
                auto p = v.find("P="), q = v.find("Q=");
                t["P"] = Json(std::atoi(v.c_str() + p + 2));
                t["Q"] = Json(std::atoi(v.c_str() + q + 2));
                j[item.first] = t;
            } else {
                j[item.first] = Json(v == "Yes");
            }
        }
        return j;
    }

    static std::string currentTime() {
        time_t t = std::time(nullptr);
        const tm* u = std::gmtime(&t);
        std::ostringstream oss;
        oss << std::put_time(u, "%a %b %d %T UTC %Y");
        return oss.str();
    }

    Json configurationAndResultsJson() {
        Json dump = Json::object();
        dump["name"] = Json(programName());
        dump["config_time"] = Json(CONFIG_TIME);
        dump["execution_time"] = Json(currentTime());
        dump["git_commit"] = Json(GIT_COMMIT_HASH);
        dump["version"] = Json(HPCC_VERSION);
        dump["device"] = Json(executionSettings->getDeviceName());
        dump["settings"] = jsonifySettingsMap(executionSettings->programSettings->getSettingsMap());
        Json t = Json::object();
        for (auto& kv : timings) {
            Json list = Json::array();
            for (double v : kv.second) {
                Json e = Json::object();
                e["unit"] = Json("s");
                e["value"] = Json(v);
                list.push_back(e);
            }
            t[kv.first] = list;
        }
        dump["timings"] = t;
        Json r = Json::object();
        for (auto& kv : results) {
            Json e = Json::object();
            e["unit"] = Json(kv.second.unit);
            e["value"] = Json(kv.second.value);
            r[kv.first] = e;
        }
        dump["results"] = r;
        Json er = Json::object();
        for (auto& kv : errors) er[kv.first] = Json(kv.second);
        dump["errors"] = er;
        dump["validated"] = Json(validated);
        Json env = Json::object();
        const char* ld = std::getenv("LD_LIBRARY_PATH");
        env["LD_LIBRARY_PATH"] = Json(ld ? ld : "");
        dump["environment"] = env;
        return dump;
    }

    void dumpConfigurationAndResults(const std::string& path) {
        std::ofstream fs(path);
        if (!fs.is_open()) {
            std::cout << "Unable to open file for dumping configuration and results" << std::endl;
            return;
        }
        fs << configurationAndResultsJson().dump();
    }

    bool setupBenchmark(int argc, char** argv) {
        bool success = true;
        try {
            std::unique_ptr<TSettings> s = parseProgramParameters(argc, argv);
            std::unique_ptr<TContext> ctx;
            if (!s->testOnly) ctx = createContext(*s);
            const int dev = s->defaultDevice;
            executionSettings.reset(
                new ExecutionSettings<TSettings, TContext>(std::move(s), dev, std::move(ctx)));
            if (!checkInputParameters()) {
                std::cerr << "ERROR: Input parameter check failed!" << std::endl;
                throw std::runtime_error("Input parameter check failed!");
            }
            printFinalConfiguration();
        } catch (std::exception& e) {
            std::cerr << "An error occured while setting up the benchmark: " << std::endl;
            std::cerr << "\t" << e.what() << std::endl;
            success = false;
        }
        benchmark_setup_succeeded = success;
        return success;
    }

    bool executeBenchmark() {
        if (!benchmark_setup_succeeded) {
            std::cerr << "Benchmark execution started without successfully running the benchmark "
                         "setup!"
                      << std::endl;
            return false;
        }
        if (executionSettings->programSettings->testOnly) {
            std::cout << "TEST MODE ENABLED: SKIP DATA GENERATION, EXECUTION, AND VALIDATION!"
                      << std::endl;
            std::cout << "SUCCESSFULLY parsed input parameters!" << std::endl;
            return true;
        }
        std::cout << HLINE << "Start benchmark using the given configuration. Generating data..."
                  << std::endl
                  << HLINE;
        try {
            auto gen_start = std::chrono::high_resolution_clock::now();
            std::unique_ptr<TData> data = generateInputData();
            std::chrono::duration<double> gen_time =
                std::chrono::high_resolution_clock::now() - gen_start;
            std::cout << "Generation Time: " << gen_time.count() << " s" << std::endl;
            std::cout << HLINE << "Execute benchmark kernel..." << std::endl << HLINE;
            auto exe_start = std::chrono::high_resolution_clock::now();
            executeKernel(*data);
            std::chrono::duration<double> exe_time =
                std::chrono::high_resolution_clock::now() - exe_start;
            std::cout << "Execution Time: " << exe_time.count() << " s" << std::endl;
            std::cout << HLINE << "Validate output..." << std::endl << HLINE;
            if (!executionSettings->programSettings->skipValidation) {
                auto eval_start = std::chrono::high_resolution_clock::now();
                validated = validateOutput(*data);
                printError();
                std::chrono::duration<double> eval_time =
                    std::chrono::high_resolution_clock::now() - eval_start;
                std::cout << "Validation Time: " << eval_time.count() << " s" << std::endl;
            }
            std::cout << HLINE << "Collect results..." << std::endl << HLINE;
            collectResults();
            if (!executionSettings->programSettings->dumpfilePath.empty())
                dumpConfigurationAndResults(executionSettings->programSettings->dumpfilePath);
            printResults();
            if (!validated)
                std::cerr << HLINE << "ERROR: VALIDATION OF OUTPUT DATA FAILED!" << std::endl;
            else
                std::cout << HLINE << "Validation: SUCCESS!" << std::endl;
            return validated;
        } catch (const std::exception& e) {
            std::cerr << "An error occured while executing the benchmark: " << std::endl;
            std::cerr << "\t" << e.what() << std::endl;
            return false;
        }
    }

    ExecutionSettings<TSettings, TContext>& getExecutionSettings() { return *executionSettings; }
    std::map<std::string, std::vector<double>>& getTimingsMap() { return timings; }
    const std::map<std::string, double>& getErrors() const { return errors; }

    // General setup banner (shared/setup/fpga_setup.cpp:320-344).
    HpccBenchmark() {
        std::cout << std::setprecision(5) << std::scientific;
        std::cout << HLINE << "General setup:" << std::endl
                  << "C++ high resolution clock is used." << std::endl
                  << "The clock precision seems to be "
                  << static_cast<double>(std::chrono::high_resolution_clock::period::num) /
                         std::chrono::high_resolution_clock::period::den * 10e9
                  << "ns" << std::endl
                  << HLINE;
    }
    virtual ~HpccBenchmark() = default;
};

}  // namespace hpcc_base
