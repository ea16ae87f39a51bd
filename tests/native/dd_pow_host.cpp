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
extern "C" void dd_pow_accurate_many(const double* x, double y, double* out, int64_t n) {
    for (int64_t i = 0; i < n; i++) out[i] = nsem::dd::pow_dd_accurate(x[i], y);
}
// pow_dd's first evaluation against its second: relative difference of the
// two double-double values (-1 where the first is out of its range) and
// whether the rounding test passed.
extern "C" void dd_pow_first_eval(const double* x, double y, double* rel, int32_t* passed, int64_t n) {
    using namespace nsem::dd;
    for (int64_t i = 0; i < n; i++) {
        DD e;
        rel[i] = -1.0;
        passed[i] = 0;
        if (!pow_fast(x[i], y, e)) continue;
        const DD a = exp_dd(mul_d(log_dd(x[i]), y));
        rel[i] = std::fabs((e.hi - a.hi) + (e.lo - a.lo)) / a.hi;
        const double d = (1.0 + std::fabs(y)) * 0x1p-70 * e.hi;
        passed[i] = e.hi + (e.lo + d) == e.hi && e.hi + (e.lo - d) == e.hi;
    }
}
