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
// Methods (all published):
//   * error-free transformations TwoSum (Knuth, TAOCP vol. 2, 4.2.2),
//     FastTwoSum (Dekker 1971) and TwoProd by fused multiply-add;
//     double-double add and multiply as in Hida, Li & Bailey, "Library for
//     double-double and quad-double arithmetic" (2001);
//   * exp of a double-double by Tang's table-driven method (P. T. P. Tang,
//     "Table-driven implementation of the exponential function in IEEE
//     floating-point arithmetic", ACM TOMS 15(2), 1989) carried in
//     double-double: a = (64 k + j) ln2/64 + r with ln2/64 in double-double
//     and |r| <= ln2/128, e^a = 2^k 2^(j/64) (1 + expm1 r); 2^(j/64) from a
//     64-entry double-double table (each entry the value rounded to double
//     plus the remainder rounded to double: within 2^-107 relative), expm1 r
//     by its Taylor series to degree 10 (first omitted term < 2^-108), the
//     terms of degree >= 6 (below 2^-45) in double, the others in
//     double-double;
//   * log of a double: one Newton step on exp from the hardware log y0,
//     log x = y0 + log(x e^-y0), with x e^-y0 - 1 = O(2^-52) so that
//     log(1 + r) = r - r^2/2 to 2^-106;
//   * x^y = exp(y log x) with the product in double-double;
//   * Ziv's strategy (A. Ziv, "Fast evaluation of elementary mathematical
//     functions with correctly rounded last bit", ACM TOMS 17(3), 1991): a
//     first, cheaper evaluation of the same formulas to about 2^-73
//     relative (exp_fast: the expm1 series to degree 7, only r and r^2 in
//     double-double), rounded when a test shows that every value within its
//     error bound rounds to the same double; only the rare inputs failing
//     the test (about one in 2^16) take the 2^-90 evaluation.  Where the
//     test passes both evaluations round to the same double, so pow_dd
//     returns exactly what the 2^-90 evaluation alone returns.
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
// a + b where no cancellation can occur (|b| much smaller than |a|, or the
// same sign): the sloppy addition of Hida, Li & Bailey
NSEM_DD_FN DD add_nc(DD a, DD b) {
    DD s = two_sum(a.hi, b.hi);
    s.lo += a.lo + b.lo;
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

// 2^(j/64), j = 0..63: hi = the value rounded to double, lo = the
// remainder rounded to double.
constexpr DD kExp2Tab[64] = {
    {1.0, 0.0}, {1.0108892860517005, -1.5234778603368577e-17},
    {1.0218971486541166, 5.109225028973444e-17}, {1.0330248790212284, 7.600838874027088e-18},
    {1.0442737824274138, 8.551889705537965e-17}, {1.0556451783605572, 1.759325738772092e-18},
    {1.0671404006768237, -7.899853966841582e-17}, {1.0787607977571199, -6.656660436056593e-17},
    {1.0905077326652577, -3.046782079812471e-17}, {1.102382583307841, 5.2660368715706944e-17},
    {1.1143867425958924, 1.0410278456845571e-16}, {1.1265216186082418, 5.165856758795457e-17},
    {1.1387886347566916, 8.912812676025408e-17}, {1.1511892299529827, 3.250710218863827e-17},
    {1.1637248587775775, 3.8292048369240935e-17}, {1.1763969916502812, 5.554203254218079e-17},
    {1.189207115002721, 3.982015231465646e-17}, {1.202156731452703, 6.644981499252301e-17},
    {1.215247359980469, -7.712630692681488e-17}, {1.22848053610687, -1.89878163130253e-17},
    {1.241857812073484, 4.658027591836937e-17}, {1.255380757024691, -6.7113898212968784e-18},
    {1.2690509571917332, 2.667932131342186e-18}, {1.2828700160787783, 1.713594918243561e-17},
    {1.2968395546510096, 2.5382502794888315e-17}, {1.3109612115247644, -7.181536135519454e-17},
    {1.3252366431597413, -2.8587312100388614e-17}, {1.339667524053303, 8.927282594831732e-17},
    {1.3542555469368927, 7.70094837980299e-17}, {1.3690024229745905, 9.593797919118849e-17},
    {1.383909881963832, -6.770511658794786e-17}, {1.3989796725383112, -9.614213209051323e-17},
    {1.4142135623730951, -9.667293313452913e-17}, {1.42961333839197, -1.2031642489053655e-17},
    {1.4451808069770467, -3.0237581349939873e-17}, {1.460917794180647, -5.600377186075216e-17},
    {1.4768261459394993, -3.483994556892796e-17}, {1.4929077282912648, 1.4192920154284036e-17},
    {1.5091644275934228, -1.016455327754295e-16}, {1.5255981507445384, -1.1024941712342561e-16},
    {1.5422108254079407, 7.949834809697621e-17}, {1.559004400237837, 3.7812070533575275e-17},
    {1.5759808451078865, -1.0136916471278304e-17}, {1.593142151342267, -1.0094406542311964e-16},
    {1.6104903319492543, 2.4707192569797888e-17}, {1.6280274218573478, -6.712955084707084e-17},
    {1.645755478153965, -1.0125679913674773e-16}, {1.6636765803267364, 5.8909926967131e-17},
    {1.681792830507429, 8.199010020581497e-17}, {1.7001063537185235, -8.0237193703977e-18},
    {1.718619298122478, -1.851380418263111e-17}, {1.7373338352737062, 3.164389299292957e-17},
    {1.7562521603732995, 2.960140695448873e-17}, {1.7753764925265212, 6.429731796556572e-17},
    {1.7947090750031072, 1.8227458427912087e-17}, {1.8142521755003989, -9.969531538920349e-17},
    {1.8340080864093424, 3.283107224245627e-17}, {1.8539791250833855, 9.761887490727594e-17},
    {1.8741676341103, -6.122763413004143e-17}, {1.8945759815869656, 3.4034035352165297e-17},
    {1.9152065613971474, -1.0619946056195963e-16}, {1.9360617934922943, 1.0332385960676326e-16},
    {1.9571441241754002, 8.960767791036668e-17}, {1.978456026387951, 4.0388753109278167e-17}};

// e^a for |a.hi| <= 708 (callers check the range).
NSEM_DD_FN DD exp_dd(DD a) {
    // ln2/64 in double-double, 64/ln2 rounded
    const double kL64Hi = 0.010830424696249145, kL64Lo = 3.623510646634843e-19;
    const double m = std::nearbyint(a.hi * 92.33248261689366);
    const DD mh = two_prod(m, kL64Hi), ml = two_prod(m, kL64Lo);
    DD r = add(a, DD{-mh.hi, -mh.lo});
    r = add(r, DD{-ml.hi, -ml.lo});
    // expm1(r) = r + r^2 (1/2 + r (1/3! + r (1/4! + r (1/5! + r q)))), q the
    // degree >= 6 part in double
    const double rh = r.hi;
    const double q = 1.388888888888889e-03 +
                     rh * (1.984126984126984e-04 +
                           rh * (2.48015873015873e-05 + rh * (2.7557319223985893e-06 + rh * 2.755731922398589e-07)));
    // (each partial sum is at most 0.006 of the coefficient or term it is
    // added to, so the additions cannot cancel)
    DD s = DD{q, 0.0};
    s = add_nc(DD{0.008333333333333333, 1.1564823173178714e-19}, mul(s, r));
    s = add_nc(DD{0.041666666666666664, 2.3129646346357427e-18}, mul(s, r));
    s = add_nc(DD{0.16666666666666666, 9.25185853854297e-18}, mul(s, r));
    s = add_nc(DD{0.5, 0.0}, mul(s, r));
    const DD u = add_nc(r, mul(mul(s, r), r));
    const int mi = (int)m;
    const DD t = kExp2Tab[mi & 63];
    DD e = add_nc(t, mul(t, u));
    const int ki = mi >> 6;  // floor(m / 64)
    e.hi = std::ldexp(e.hi, ki);
    e.lo = std::ldexp(e.lo, ki);
    return e;
}

// e^a to about 2^-74 relative for |a.hi| <= 700, the first evaluation of
// pow_dd.  The reduction of exp_dd with its one cancelling subtraction done
// by TwoSum (exact) and the small terms (a.lo, the low parts of m ln2/64;
// each below 2^-42) summed in double, error below 2^-92; expm1 r = r + r^2
// (1/2 + r p(r)) with p the Taylor coefficients of degrees 3-7 in double
// (first omitted term below 2^-75; r p below 2^-10, so its rounding is
// below 2^-63 and reaches the result times r^2 < 2^-15), r^2 and the
// product in double-double.
NSEM_DD_FN DD exp_fast(DD a) {
    const double kL64Hi = 0.010830424696249145, kL64Lo = 3.623510646634843e-19;
    const double m = std::nearbyint(a.hi * 92.33248261689366);
    const DD mh = two_prod(m, kL64Hi);
    const DD s0 = two_sum(a.hi, -mh.hi);
    const DD r = two_sum(s0.hi, ((a.lo - mh.lo) - m * kL64Lo) + s0.lo);
    const double rh = r.hi;
    const double p = 0.16666666666666666 +
                     rh * (0.041666666666666664 +
                           rh * (0.008333333333333333 + rh * (1.388888888888889e-03 + rh * 1.984126984126984e-04)));
    const DD s = fast_two_sum(0.5, rh * p);
    DD sq = two_prod(rh, rh);
    sq.lo += (rh + rh) * r.lo;
    const DD u = add_nc(r, mul(sq, s));
    const int mi = (int)m;
    const DD t = kExp2Tab[mi & 63];
    DD e = add_nc(t, mul(t, u));
    const int ki = mi >> 6;
    e.hi = std::ldexp(e.hi, ki);
    e.lo = std::ldexp(e.lo, ki);
    return e;
}

// log x for positive normal finite x (FAST: through exp_fast, error below
// 2^-74 absolute for 2^-1000 <= x <= 2^1000, where e^-y0 stays normal).
template <bool FAST>
NSEM_DD_FN DD log_t(double x) {
    const double y0 = std::log(x);
    const DD e = FAST ? exp_fast(DD{-y0, 0.0}) : exp_dd(DD{-y0, 0.0});
    const DD t = mul_d(e, x);              // x e^-y0 = 1 + r
    const double r = (t.hi - 1.0) + t.lo;  // t.hi - 1 is exact (Sterbenz)
    return fast_two_sum(y0, r - 0.5 * r * r);
}
NSEM_DD_FN DD log_dd(double x) { return log_t<false>(x); }

// x^y from the 2^-90 evaluation alone (pow_dd's second evaluation).
NSEM_DD_FN double pow_dd_accurate(double x, double y) {
    if (!(x >= 2.2250738585072014e-308 && x <= 1.7976931348623157e308) || !std::isfinite(y))
        return std::pow(x, y);
    const DD t = mul_d(log_dd(x), y);
    if (!(std::fabs(t.hi) <= 700.0)) return std::pow(x, y);
    return exp_dd(t).hi;
}

// The first evaluation of pow_dd: e^(y log x) as a double-double e (e.hi the
// double nearest e.hi + e.lo) with relative error below (1 + |y|) 2^-73, for
// 2^-1000 <= x <= 2^1000 and |y log x| <= 680 (below e^-680 = 2^-981 the
// low part would lose bits to the subnormal range); returns false outside
// that range.
NSEM_DD_FN bool pow_fast(double x, double y, DD& e) {
    if (!(x >= 0x1p-1000 && x <= 0x1p1000)) return false;
    const DD t = mul_d(log_t<true>(x), y);
    if (!(std::fabs(t.hi) <= 680.0)) return false;
    e = exp_fast(t);
    return true;
}

// x^y, correctly rounded except where the exact power lies within about
// 2^-90 of a rounding boundary; outside positive normal finite x, or where
// y log x leaves the range of normal results, the C library's pow (its
// special cases: NaN, 0, Inf, negative bases, overflow, underflow).  The
// cheap evaluation comes first: with its error bound taken as
// eps = (1 + |y|) 2^-71 (four times the bound above) and d = 2 eps e.hi,
// e.hi + (e.lo + d) and e.hi + (e.lo - d) both round to e.hi only if every
// value within eps e.hi of e.hi + e.lo -- the exact power among them --
// lies inside e.hi's rounding interval; e.hi is then the correctly rounded
// power, the double that the 2^-90 evaluation returns too.  Otherwise that
// evaluation decides.
NSEM_DD_FN double pow_dd(double x, double y) {
    if (std::isfinite(y)) {
        DD e;
        if (pow_fast(x, y, e)) {
            const double d = (1.0 + std::fabs(y)) * 0x1p-70 * e.hi;
            if (e.hi + (e.lo + d) == e.hi && e.hi + (e.lo - d) == e.hi) return e.hi;
        }
    }
    return pow_dd_accurate(x, y);
}

// e^x rounded once from the double-double value (C library exp outside
// |x| <= 700 and for NaN).
NSEM_DD_FN double exp_cr(double x) {
    if (!(std::fabs(x) <= 700.0)) return std::exp(x);
    return exp_dd(DD{x, 0.0}).hi;
}

}  // namespace dd
}  // namespace nsem
