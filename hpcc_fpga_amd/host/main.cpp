// LINPACK benchmark executable with the HIP backend
// (the counterpart of LINPACK/src/host/main.cpp:12-32): exit code 0 when the
// run validates, 1 otherwise.
#include "linpack.hpp"

int main(int argc, char* argv[]) {
    linpack::LinpackBenchmark bm(argc, argv);
    const bool success = bm.executeBenchmark();
    return success ? 0 : 1;
}
