// Double-double arithmetic and a pow() accurate far below half an ulp.
//
// The reference evaluates its equation of state with the C library's pow
// (p = P0 (rho theta R / P0)^gamma, apps/euler/euler.cpp:211-213), whose
// result is the correctly rounded power in all but rare cases (glibc's pow
// is documented below 0.52 ulp).  A device pow with a full-ulp error would
// move p' by an ulp of ~1e5 Pa at many nodes; pow_dd instead computes x^y to
// about 2^-90 relative and rounds once, so it returns the correctly rounded
// power except where the exact value lies within ~2^-90 of a rounding
// boundary -- the same value glibc returns wherever glibc rounds correctly.
//
// Methods (all published, none table-based):
//   * error-free transformations TwoSum (Knuth, TAOCP vol. 2, 4.2.2),
//     FastTwoSum (Dekker 1971) and TwoProd by fused multiply-add;
//     double-double add and multiply as in Hida, Li & Bailey, "Library for
//     double-double and quad-double arithmetic" (2001);
//   * exp of a double-double: Cody-Waite reduction a = k ln2 + r with ln2 in
//     double-double, r scaled by 2^-8, expm1 of the scaled argument by its
//     Taylor series to degree 11 (first omitted term < 2^-125), then eight
//     expm1 doublings e(2t) - 1 = u (2 + u), which keep the small result's
//     relative accuracy, and the scaling by 2^k;
//   * log of a double: one Newton step on exp from the hardware log y0,
//     log x = y0 + log(x e^-y0), with x e^-y0 - 1 = O(2^-52) so that
//     log(1 + r) = r - r^2/2 to 2^-106;
//   * x^y = exp(y log x) with the product in double-double.
// The same source compiles for the host (the correctness test builds it
// with g++ and checks it against exact decimal arithmetic and glibc).
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NSEM_DD_FN __host__ __device__ __forceinline__
#else
#define NSEM_DD_FN inline
#endif

namespace nsem {
namespace dd {

struct DD {
    double hi, lo;
};

NSEM_DD_FN DD two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
// |a| >= |b| (or a == 0)
NSEM_DD_FN DD fast_two_sum(double a, double b) {
    const double s = a + b;
    return {s, b - (s - a)};
}
NSEM_DD_FN DD two_prod(double a, double b) {
    const double p = a * b;
    return {p, std::fma(a, b, -p)};
}
NSEM_DD_FN DD add(DD a, DD b) {
    DD s = two_sum(a.hi, b.hi);
    const DD t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}
NSEM_DD_FN DD mul(DD a, DD b) {
    DD p = two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return fast_two_sum(p.hi, p.lo);
}
NSEM_DD_FN DD mul_d(DD a, double b) {
    DD p = two_prod(a.hi, b);
    p.lo += a.lo * b;
    return fast_two_sum(p.hi, p.lo);
}

// e^a for |a.hi| <= 708 (callers check the range).
NSEM_DD_FN DD exp_dd(DD a) {
    const double kLn2Hi = 6.93147180559945286227e-01, kLn2Lo = 2.31904681384629955842e-17;
    const double k = std::nearbyint(a.hi * 1.4426950408889634);
    const DD kh = two_prod(k, kLn2Hi), kl = two_prod(k, kLn2Lo);
    DD r = add(a, DD{-kh.hi, -kh.lo});
    r = add(r, DD{-kl.hi, -kl.lo});
    r.hi *= 0.00390625;  // 2^-8, exact
    r.lo *= 0.00390625;
    // expm1(r) = r (1 + r (1/2 + r (1/3! + ... + r/11!))), |r| <= 1.36e-3
    const DD c[10] = {{0.5, 0.0},
                      {0.16666666666666666, 9.25185853854297e-18},
                      {0.041666666666666664, 2.3129646346357427e-18},
                      {0.008333333333333333, 1.1564823173178714e-19},
                      {0.001388888888888889, -5.300543954373577e-20},
                      {0.0001984126984126984, 1.7209558293420705e-22},
                      {2.48015873015873e-05, 2.1511947866775882e-23},
                      {2.7557319223985893e-06, -1.858393274046472e-22},
                      {2.755731922398589e-07, 2.3767714622250297e-23},
                      {2.505210838544172e-08, -1.448814070935912e-24}};
    DD s = c[9];
    for (int i = 8; i >= 0; i--) s = add(mul(s, r), c[i]);
    s = add(mul(s, r), DD{1.0, 0.0});
    DD u = mul(s, r);
    // e^(2t) - 1 = u (2 + u), eight times: u = e^(256 r) - 1
    for (int i = 0; i < 8; i++) u = mul(u, add(u, DD{2.0, 0.0}));
    DD e = add(DD{1.0, 0.0}, u);
    const int ki = (int)k;
    e.hi = std::ldexp(e.hi, ki);
    e.lo = std::ldexp(e.lo, ki);
    return e;
}

// log x for positive normal finite x.
NSEM_DD_FN DD log_dd(double x) {
    const double y0 = std::log(x);
    const DD t = mul_d(exp_dd(DD{-y0, 0.0}), x);  // x e^-y0 = 1 + r
    const double r = (t.hi - 1.0) + t.lo;          // t.hi - 1 is exact (Sterbenz)
    return fast_two_sum(y0, r - 0.5 * r * r);
}

// x^y rounded once from a double-double value accurate to about 2^-90.
// Outside positive normal finite x, or where y log x leaves the range of
// normal results, the C library's pow (its special cases: NaN, 0, Inf,
// negative bases, overflow, underflow).
NSEM_DD_FN double pow_dd(double x, double y) {
    if (!(x >= 2.2250738585072014e-308 && x <= 1.7976931348623157e308) || !std::isfinite(y))
        return std::pow(x, y);
    const DD t = mul_d(log_dd(x), y);
    if (!(std::fabs(t.hi) <= 700.0)) return std::pow(x, y);
    return exp_dd(t).hi;
}

// e^x rounded once from the double-double value (C library exp outside
// |x| <= 700 and for NaN).
NSEM_DD_FN double exp_cr(double x) {
    if (!(std::fabs(x) <= 700.0)) return std::exp(x);
    return exp_dd(DD{x, 0.0}).hi;
}

}  // namespace dd
}  // namespace nsem
