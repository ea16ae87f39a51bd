// Vector, symmetric tensor and tensor values of the source-compatible field
// API (src/tensor/tensor.h, tensor.cpp: component order XX YY ZZ XY YZ XZ YX
// ZY ZX, Vector * Vector component-wise, dot products right-nested as the
// reference unrolls them, tensor rows left to right), for host and device
// code.  The evaluation order of every operation is the reference's, so
// expressions evaluated on the device round as the reference's do.
#ifndef NSEM_FIELD_TENSOR_H
#define NSEM_FIELD_TENSOR_H

#include <cmath>
#include <iostream>

#include "dd_math.hpp"
#include "types.h"

template <int N>
struct TTensor {
    Scalar P[N];
    NSEM_HD TTensor() {}
    NSEM_HD explicit TTensor(const Scalar& v) {
        for (int i = 0; i < N; i++) P[i] = v;
    }
    NSEM_HD TTensor(const Scalar& xx, const Scalar& yy, const Scalar& zz) {
        P[0] = xx;
        P[1] = yy;
        P[2] = zz;
        for (int i = 3; i < N; i++) P[i] = Scalar(0);
    }
    NSEM_HD Scalar& operator[](Int i) { return P[i]; }
    NSEM_HD const Scalar& operator[](Int i) const { return P[i]; }
    NSEM_HD TTensor operator-() const {
        TTensor r;
        for (int i = 0; i < N; i++) r.P[i] = -P[i];
        return r;
    }
#define NSEM_TT_OP(op)                                                          \
    NSEM_HD TTensor& operator op(const TTensor& q) {                            \
        for (int i = 0; i < N; i++) P[i] op q.P[i];                             \
        return *this;                                                           \
    }                                                                           \
    NSEM_HD TTensor& operator op(const Scalar& q) {                             \
        for (int i = 0; i < N; i++) P[i] op q;                                  \
        return *this;                                                           \
    }
    NSEM_TT_OP(+=)
    NSEM_TT_OP(-=)
    NSEM_TT_OP(*=)
    NSEM_TT_OP(/=)
#undef NSEM_TT_OP
};

typedef TTensor<3> Vector;
typedef TTensor<6> STensor;
typedef TTensor<9> Tensor;

// binary operators as the reference defines them from the compound ones:
// T op T and T op Scalar start from the left operand; Scalar * T and
// Scalar + T from the right one; Scalar - T and Scalar / T from T(Scalar)
#define NSEM_TT_BIN(op, cop)                                                           \
    template <int N>                                                                   \
    NSEM_HD TTensor<N> operator op(const TTensor<N>& p, const TTensor<N>& q) {         \
        TTensor<N> r = p;                                                              \
        r cop q;                                                                       \
        return r;                                                                      \
    }                                                                                  \
    template <int N>                                                                   \
    NSEM_HD TTensor<N> operator op(const TTensor<N>& p, const Scalar& q) {             \
        TTensor<N> r = p;                                                              \
        r cop q;                                                                       \
        return r;                                                                      \
    }
NSEM_TT_BIN(+, +=)
NSEM_TT_BIN(-, -=)
NSEM_TT_BIN(*, *=)
NSEM_TT_BIN(/, /=)
#undef NSEM_TT_BIN
template <int N>
NSEM_HD TTensor<N> operator+(const Scalar& p, const TTensor<N>& q) {
    TTensor<N> r = q;
    r += p;
    return r;
}
template <int N>
NSEM_HD TTensor<N> operator*(const Scalar& p, const TTensor<N>& q) {
    TTensor<N> r = q;
    r *= p;
    return r;
}
template <int N>
NSEM_HD TTensor<N> operator-(const Scalar& p, const TTensor<N>& q) {
    TTensor<N> r(p);
    r -= q;
    return r;
}
template <int N>
NSEM_HD TTensor<N> operator/(const Scalar& p, const TTensor<N>& q) {
    TTensor<N> r(p);
    r /= q;
    return r;
}

// p & q: sum of products nested from the right (p0 q0 + (p1 q1 + (... + 0)));
// a symmetric tensor counts its off-diagonal products twice
template <int N>
NSEM_HD Scalar dot_unrolled(const Scalar* p, const Scalar* q) {
    Scalar r = 0;
    for (int i = N - 1; i >= 0; i--) r = p[i] * q[i] + r;
    return r;
}
template <int N>
NSEM_HD Scalar operator&(const TTensor<N>& p, const TTensor<N>& q) {
    Scalar r = dot_unrolled<N>(p.P, q.P);
    if (N == 6) r += dot_unrolled<3>(&p.P[3], &q.P[3]);
    return r;
}
template <int N>
NSEM_HD Scalar dot(const TTensor<N>& p, const TTensor<N>& q) { return p & q; }
template <int N>
NSEM_HD Scalar magSq(const TTensor<N>& p) { return p & p; }
template <int N>
NSEM_HD Scalar mag(const TTensor<N>& p) { return sqrt(magSq(p)); }
template <int N>
NSEM_HD TTensor<N> unit(const TTensor<N>& p) {
    TTensor<N> r = p;
    const Scalar m = mag(r);
    r /= m;
    return r;
}
template <int N>
NSEM_HD TTensor<N> mul(const TTensor<N>& p, const Scalar& q) { return p * q; }
template <int N>
NSEM_HD Scalar tr(const TTensor<N>& p) { return p[0] + p[1] + p[2]; }
// deviatoric part: the trace times factor / 3 off the diagonal entries
template <int N>
NSEM_HD TTensor<N> dev(const TTensor<N>& p, const Scalar& factor) {
    TTensor<N> r = p;
    const Scalar t = tr(p) * factor / 3;
    r[0] -= t;
    r[1] -= t;
    r[2] -= t;
    return r;
}
// hydrostatic part: the trace times factor / 3 on the diagonal, the rest 0
template <int N>
NSEM_HD TTensor<N> hyd(const TTensor<N>& p, const Scalar& factor) {
    TTensor<N> r(1, 1, 1);
    const Scalar t = tr(p) * factor / 3;
    r[0] = t;
    r[1] = t;
    r[2] = t;
    return r;
}

// component-wise functions of the C library, as the reference applies them
#define NSEM_TT_FN1(name, impl)                                                \
    template <int N>                                                           \
    NSEM_HD TTensor<N> name(const TTensor<N>& p) {                             \
        TTensor<N> r;                                                          \
        for (int i = 0; i < N; i++) r.P[i] = impl(p.P[i]);                     \
        return r;                                                              \
    }
NSEM_TT_FN1(sqrt, ::sqrt)
NSEM_TT_FN1(fabs, ::fabs)
NSEM_TT_FN1(exp, nsem::dd::exp_cr)
#undef NSEM_TT_FN1
template <int N>
NSEM_HD TTensor<N> pow(const TTensor<N>& p, const Scalar& e) {
    TTensor<N> r;
    for (int i = 0; i < N; i++) r.P[i] = nsem::dd::pow_dd(p.P[i], e);
    return r;
}
template <int N>
NSEM_HD TTensor<N> max(const TTensor<N>& p, const TTensor<N>& q) {
    TTensor<N> r;
    for (int i = 0; i < N; i++) r.P[i] = (p.P[i] >= q.P[i]) ? p.P[i] : q.P[i];
    return r;
}
template <int N>
NSEM_HD TTensor<N> min(const TTensor<N>& p, const TTensor<N>& q) {
    TTensor<N> r;
    for (int i = 0; i < N; i++) r.P[i] = (p.P[i] <= q.P[i]) ? p.P[i] : q.P[i];
    return r;
}

// Scalar forms
NSEM_HD Scalar mag(const Scalar& p) { return fabs(p); }
NSEM_HD Scalar magSq(const Scalar& p) { return p * p; }
NSEM_HD Scalar dot(const Scalar& p, const Scalar& q) { return p * q; }
NSEM_HD Scalar mul(const Scalar& p, const Scalar& q) { return p * q; }
NSEM_HD Scalar sdiv(const Scalar& p, const Scalar& q) { return q ? (p / q) : 0; }
NSEM_HD Scalar max(const Scalar& p, const Scalar& q) { return (p >= q) ? p : q; }
NSEM_HD Scalar min(const Scalar& p, const Scalar& q) { return (p <= q) ? p : q; }

namespace Constants {
enum { XX, YY, ZZ, XY, YZ, XZ, YX, ZY, ZX };
const Int MAX_INT = Int(1u << 31);
const Scalar PI = Scalar(3.14159265358979323846264);
const Scalar E = Scalar(2.71828182845904523536028);
const Scalar MachineEpsilon = Scalar(1e-15);
const Scalar EqualEpsilon = Scalar(1e-7);
const Vector I_V(1, 1, 1);
const Tensor I_T(1, 1, 1);
const STensor I_ST(1, 1, 1);
}  // namespace Constants

// outer products, tensor-vector and tensor-tensor products (tensor.cpp)
NSEM_HD Tensor mul(const Vector& p, const Vector& q) {
    using namespace Constants;
    Tensor r;
    r[XX] = p[XX] * q[XX];
    r[YY] = p[YY] * q[YY];
    r[ZZ] = p[ZZ] * q[ZZ];
    r[XY] = p[XX] * q[YY];
    r[YZ] = p[YY] * q[ZZ];
    r[XZ] = p[XX] * q[ZZ];
    r[YX] = p[YY] * q[XX];
    r[ZY] = p[ZZ] * q[YY];
    r[ZX] = p[ZZ] * q[XX];
    return r;
}
NSEM_HD STensor mul(const Vector& p) {
    using namespace Constants;
    STensor r;
    r[XX] = p[XX] * p[XX];
    r[YY] = p[YY] * p[YY];
    r[ZZ] = p[ZZ] * p[ZZ];
    r[XY] = p[XX] * p[YY];
    r[YZ] = p[YY] * p[ZZ];
    r[XZ] = p[XX] * p[ZZ];
    return r;
}
// a vector times a scalar field value: the outer product's scalar case
NSEM_HD Vector mul(const Vector& p, const Scalar& q) { return p * q; }
NSEM_HD Tensor mul(const Tensor& p, const Tensor& q) {
    using namespace Constants;
    Tensor r;
    r[XX] = p[XX] * q[XX] + p[XY] * q[YX] + p[XZ] * q[ZX];
    r[XY] = p[XX] * q[XY] + p[XY] * q[YY] + p[XZ] * q[ZY];
    r[XZ] = p[XX] * q[XZ] + p[XY] * q[YZ] + p[XZ] * q[ZZ];
    r[YX] = p[YX] * q[XX] + p[YY] * q[YX] + p[YZ] * q[ZX];
    r[YY] = p[YX] * q[XY] + p[YY] * q[YY] + p[YZ] * q[ZY];
    r[YZ] = p[YX] * q[XZ] + p[YY] * q[YZ] + p[YZ] * q[ZZ];
    r[ZX] = p[ZX] * q[XX] + p[ZY] * q[YX] + p[ZZ] * q[ZX];
    r[ZY] = p[ZX] * q[XY] + p[ZY] * q[YY] + p[ZZ] * q[ZY];
    r[ZZ] = p[ZX] * q[XZ] + p[ZY] * q[YZ] + p[ZZ] * q[ZZ];
    return r;
}
// the symmetric tensors' product kept symmetric: the upper triangle of p q
NSEM_HD STensor mul(const STensor& p, const STensor& q) {
    using namespace Constants;
    STensor r;
    r[XX] = p[XX] * q[XX] + p[XY] * q[XY] + p[XZ] * q[XZ];
    r[XY] = p[XX] * q[XY] + p[XY] * q[YY] + p[XZ] * q[YZ];
    r[XZ] = p[XX] * q[XZ] + p[XY] * q[YZ] + p[XZ] * q[ZZ];
    r[YY] = p[XY] * q[XY] + p[YY] * q[YY] + p[YZ] * q[YZ];
    r[YZ] = p[XY] * q[XZ] + p[YY] * q[YZ] + p[YZ] * q[ZZ];
    r[ZZ] = p[XZ] * q[XZ] + p[YZ] * q[YZ] + p[ZZ] * q[ZZ];
    return r;
}
NSEM_HD Vector dot(const Tensor& q, const Vector& p) {
    using namespace Constants;
    Vector r;
    r[XX] = q[XX] * p[XX] + q[XY] * p[YY] + q[XZ] * p[ZZ];
    r[YY] = q[YX] * p[XX] + q[YY] * p[YY] + q[YZ] * p[ZZ];
    r[ZZ] = q[ZX] * p[XX] + q[ZY] * p[YY] + q[ZZ] * p[ZZ];
    return r;
}
NSEM_HD Vector dot(const STensor& q, const Vector& p) {
    using namespace Constants;
    Vector r;
    r[XX] = q[XX] * p[XX] + q[XY] * p[YY] + q[XZ] * p[ZZ];
    r[YY] = q[XY] * p[XX] + q[YY] * p[YY] + q[YZ] * p[ZZ];
    r[ZZ] = q[XZ] * p[XX] + q[YZ] * p[YY] + q[ZZ] * p[ZZ];
    return r;
}
NSEM_HD Tensor trn(const Tensor& p) {
    using namespace Constants;
    Tensor r = p;
    r[XY] = p[YX];
    r[YZ] = p[ZY];
    r[XZ] = p[ZX];
    r[YX] = p[XY];
    r[ZY] = p[YZ];
    r[ZX] = p[XZ];
    return r;
}
NSEM_HD STensor sym(const Tensor& p) {
    using namespace Constants;
    STensor r;
    r[XX] = p[XX];
    r[YY] = p[YY];
    r[ZZ] = p[ZZ];
    r[XY] = (p[XY] + p[YX]) / 2;
    r[YZ] = (p[YZ] + p[ZY]) / 2;
    r[XZ] = (p[XZ] + p[ZX]) / 2;
    return r;
}

// equality within a tolerance (tensor.h: absolute or relative to either)
NSEM_HD bool equal(const Scalar& p, const Scalar& q, const Scalar tol = Scalar(1e-7)) {
    const Scalar delta = mag(p - q);
    return delta <= tol || delta <= tol * mag(p) || delta <= tol * mag(q);
}
NSEM_HD bool equal(const Vector& p, const Vector& q, const Scalar tol = Scalar(1e-7)) {
    return equal(p[0], q[0], tol) && equal(p[1], q[1], tol) && equal(p[2], q[2], tol);
}

// the SYMMETRY ghost value (tensor.h): scalars copied; a vector projected
// on the face plane, (I - n n) p, and rescaled to |p|
NSEM_HD Scalar sym(const Scalar& p, const Vector& n) { return p; }
NSEM_HD Vector sym(const Vector& p, const Vector& n) {
    const Vector en = unit(n);
    const STensor A = STensor(1, 1, 1) - mul(en);
    const Vector r = dot(A, p);
    const Scalar magR = mag(r);
    if (equal(magR, Scalar(0))) return r;
    return r * (mag(p) / magR);
}
NSEM_HD STensor sym(const STensor& p, const Vector& n) {
    const Vector en = unit(n);
    const STensor A = STensor(1, 1, 1) - mul(en);
    const STensor r = mul(mul(A, p), A);
    const Scalar magR = mag(r);
    if (equal(magR, Scalar(0))) return r;
    return r * (mag(p) / magR);
}
NSEM_HD Tensor sym(const Tensor& p, const Vector& n) {
    const Vector en = unit(n);
    const Tensor A = Tensor(1, 1, 1) - mul(en, en);
    const Tensor r = mul(mul(A, p), A);
    const Scalar magR = mag(r);
    if (equal(magR, Scalar(0))) return r;
    return r * (mag(p) / magR);
}

template <int N>
std::ostream& operator<<(std::ostream& os, const TTensor<N>& p) {
    for (int i = 0; i < N; i++) os << p[i] << " ";
    return os;
}
template <int N>
std::istream& operator>>(std::istream& is, TTensor<N>& p) {
    for (int i = 0; i < N; i++) is >> p[i];
    return is;
}

#endif
