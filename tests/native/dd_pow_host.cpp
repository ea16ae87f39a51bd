// Host build of the device pow (nebulasem_amd/csrc/hip/dd_math.hpp) for the
// correctness test: the same source, compiled by g++.
#include <cmath>
#include <cstdint>

#include "../../nebulasem_amd/csrc/hip/dd_math.hpp"

extern "C" void dd_pow_many(const double* x, double y, double* out, int64_t n) {
    for (int64_t i = 0; i < n; i++) out[i] = nsem::dd::pow_dd(x[i], y);
}
extern "C" void libm_pow_many(const double* x, double y, double* out, int64_t n) {
    for (int64_t i = 0; i < n; i++) out[i] = std::pow(x[i], y);
}
