// Pointwise physics of the method-of-lines Euler residual, written once for
// every arithmetic type the kernels use: float (FP32 perturbation form),
// double (FP64), and forward-mode dual numbers Dual<K> (value + K directional
// derivatives) for the tangent-linear model and the local Jacobians of the
// adjoint.
//
// State: conserved perturbations over the hydrostatic background,
//   rho' = rho - rho_ref,  m = rho U,  Theta' = rho theta - rho_ref T0,
//   p' = p_ref expm1(gamma log1p(Theta' / (rho_ref T0))) (the exact p - p_ref).
// Flux rows (the reference's divf arguments, euler.cpp:193-258, evaluated at
// one state): F_rho = m, F_m = m (x) U + p' I, F_Theta = m theta.
// Rusanov (field.h:2927-2943): lambda = (avg|U| + avg sqrt(gamma R theta))/2,
// momentum dissipation along the normal only; SYMMETRY ghost = tangential
// projection rescaled to |U| (src/tensor/tensor.h:483-492).
#pragma once
#include <hip/hip_runtime.h>

namespace nsem {

// ------------------------------------------------------------ dual numbers
template <int K>
struct Dual {
    double v;
    double d[K];
    __device__ __forceinline__ Dual() {}
    __device__ __forceinline__ Dual(double x) : v(x) {
        for (int k = 0; k < K; k++) d[k] = 0.0;
    }
    __device__ __forceinline__ explicit operator double() const { return v; }
};

template <typename T>
struct is_dual { static constexpr bool value = false; };
template <int K>
struct is_dual<Dual<K>> { static constexpr bool value = true; };

template <int K>
__device__ __forceinline__ Dual<K> operator+(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r; r.v = a.v + b.v;
    for (int k = 0; k < K; k++) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r; r.v = a.v - b.v;
    for (int k = 0; k < K; k++) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a) {
    Dual<K> r; r.v = -a.v;
    for (int k = 0; k < K; k++) r.d[k] = -a.d[k];
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r; r.v = a.v * b.v;
    for (int k = 0; k < K; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator/(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r; r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
    for (int k = 0; k < K; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K>& operator+=(Dual<K>& a, const Dual<K>& b) { a = a + b; return a; }
template <int K>
__device__ __forceinline__ Dual<K>& operator-=(Dual<K>& a, const Dual<K>& b) { a = a - b; return a; }
template <int K>
__device__ __forceinline__ Dual<K>& operator*=(Dual<K>& a, const Dual<K>& b) { a = a * b; return a; }
template <int K>
__device__ __forceinline__ bool operator<=(const Dual<K>& a, const Dual<K>& b) { return a.v <= b.v; }

// ------------------------------------------------------------ math per type
__device__ __forceinline__ double r_expm1(double x) { return expm1(x); }
__device__ __forceinline__ float r_expm1(float x) { return expm1f(x); }
__device__ __forceinline__ double r_log1p(double x) { return log1p(x); }
__device__ __forceinline__ float r_log1p(float x) { return log1pf(x); }
__device__ __forceinline__ double r_sqrt(double x) { return sqrt(x); }
// FP32: the hardware square root (1 ulp) rather than the correctly rounded
// sequence; the FP32 form's error budget is set by its FP32 state, not by it.
__device__ __forceinline__ float r_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

template <int K>
__device__ __forceinline__ Dual<K> r_expm1(const Dual<K>& x) {
    Dual<K> r; r.v = expm1(x.v);
    const double e = r.v + 1.0;
    for (int k = 0; k < K; k++) r.d[k] = x.d[k] * e;
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> r_log1p(const Dual<K>& x) {
    Dual<K> r; r.v = log1p(x.v);
    const double s = 1.0 / (1.0 + x.v);
    for (int k = 0; k < K; k++) r.d[k] = x.d[k] * s;
    return r;
}
// |.| is not differentiable at 0: sqrt(0) takes derivative 0 (the central
// difference of the discrete map sees the same, the |h| terms cancel).
template <int K>
__device__ __forceinline__ Dual<K> r_sqrt(const Dual<K>& x) {
    Dual<K> r; r.v = sqrt(x.v);
    const double s = r.v > 0.0 ? 0.5 / r.v : 0.0;
    for (int k = 0; k < K; k++) r.d[k] = x.d[k] * s;
    return r;
}

// Division: correctly rounded in FP64 and for dual numbers; in FP32 the
// hardware reciprocal (1 ulp) times the numerator, which is far inside the
// FP32 form's error budget and a tenth of the IEEE division's instructions.
__device__ __forceinline__ double r_div(double a, double b) { return a / b; }
__device__ __forceinline__ float r_div(float a, float b) { return __fdividef(a, b); }
template <int K>
__device__ __forceinline__ Dual<K> r_div(const Dual<K>& a, const Dual<K>& b) { return a / b; }

// p' / p_ref = (1 + x)^gamma - 1 with x = Theta' / (rho_ref T0), computed as
// expm1(gamma log1p(x)) (no cancellation).  For small x the binomial series
// sum_k binom(gamma, k) x^k is used instead, a few multiply-adds in place of
// two libm calls: in FP64 to x^8 for |x| <= 1/64 (for gamma in (1, 2) the
// first omitted term is below 1e-17 of the sum) with the coefficients given
// (DgDev::ppc: uniform, so they cost no vector registers), in FP32 to x^6
// for |x| <= 1/32 (below 1e-10, under FP32 rounding).
__device__ __forceinline__ double pow1p_m1(double x, double gam, const double* c) {
    if (!c || fabs(x) > 0.015625) return expm1(gam * log1p(x));
    return x * (c[0] + x * (c[1] + x * (c[2] + x * (c[3] + x * (c[4] + x * (c[5] + x * (c[6] + x * c[7])))))));
}
__device__ __forceinline__ float pow1p_m1(float x, float gam, const double*) {
    if (fabsf(x) > 0.03125f) return expm1f(gam * log1pf(x));
    const float c1 = gam, c2 = c1 * (gam - 1.0f) * 0.5f, c3 = c2 * (gam - 2.0f) * (1.0f / 3.0f);
    const float c4 = c3 * (gam - 3.0f) * 0.25f, c5 = c4 * (gam - 4.0f) * 0.2f;
    const float c6 = c5 * (gam - 5.0f) * (1.0f / 6.0f);
    return x * (c1 + x * (c2 + x * (c3 + x * (c4 + x * (c5 + x * c6)))));
}
template <int K>
__device__ __forceinline__ Dual<K> pow1p_m1(const Dual<K>& x, const Dual<K>& gam, const double*) {
    if (fabs(x.v) > 0.015625) return r_expm1(gam * r_log1p(x));
    // the FP64 series and its derivative sum_k k binom(gamma, k) x^(k-1) dx
    // (gamma a constant here, as in every caller)
    const double g = gam.v;
    const double c1 = g, c2 = c1 * (g - 1.0) * 0.5, c3 = c2 * (g - 2.0) * (1.0 / 3.0);
    const double c4 = c3 * (g - 3.0) * 0.25, c5 = c4 * (g - 4.0) * 0.2, c6 = c5 * (g - 5.0) * (1.0 / 6.0);
    const double c7 = c6 * (g - 6.0) * (1.0 / 7.0), c8 = c7 * (g - 7.0) * 0.125;
    const double t = x.v;
    Dual<K> r;
    r.v = t * (c1 + t * (c2 + t * (c3 + t * (c4 + t * (c5 + t * (c6 + t * (c7 + t * c8)))))));
    const double dv =
        c1 + t * (2 * c2 + t * (3 * c3 + t * (4 * c4 + t * (5 * c5 + t * (6 * c6 + t * (7 * c7 + t * (8 * c8)))))));
    for (int k = 0; k < K; k++) r.d[k] = dv * x.d[k];
    return r;
}

__device__ __forceinline__ double value_of(double x) { return x; }
__device__ __forceinline__ double value_of(float x) { return x; }
template <int K>
__device__ __forceinline__ double value_of(const Dual<K>& x) { return x.v; }

// ------------------------------------------------------------ physics
template <typename R>
struct Prim {
    R rp, m[3], tp;   // perturbation conserved: rho', m, Theta'
    R u[3];           // velocity
    R thp;            // theta - T0
    R pp;             // p'
    R c;              // sqrt(gamma R theta)
    R umag;
};

template <typename R>
__device__ __forceinline__ void make_prim(Prim<R>& P, R rho_ref, R p_ref, R T0, R gam, R gR,
                                          const double* ppc = nullptr) {
    const R rho = rho_ref + P.rp;
    const R ir = r_div(R(1), rho);
    P.u[0] = P.m[0] * ir; P.u[1] = P.m[1] * ir; P.u[2] = P.m[2] * ir;
    P.thp = (P.tp - T0 * P.rp) * ir;
    P.pp = p_ref * pow1p_m1(r_div(P.tp, rho_ref * T0), gam, ppc);
    P.c = r_sqrt(gR * (T0 + P.thp));
    P.umag = r_sqrt(P.u[0] * P.u[0] + P.u[1] * P.u[1] + P.u[2] * P.u[2]);
}

// Flux row `row` (0 rho, 1-3 m_x m_y m_z, 4 Theta) at one node.
template <typename R>
__device__ __forceinline__ void flux_row(const Prim<R>& P, int row, R T0, R& f0, R& f1, R& f2) {
    if (row == 0) { f0 = P.m[0]; f1 = P.m[1]; f2 = P.m[2]; }
    else if (row < 4) {
        const int r = row - 1;
        f0 = P.m[r] * P.u[0] + (r == 0 ? P.pp : R(0));
        f1 = P.m[r] * P.u[1] + (r == 1 ? P.pp : R(0));
        f2 = P.m[r] * P.u[2] + (r == 2 ? P.pp : R(0));
    } else {
        f0 = P.m[0] * T0 + P.m[0] * P.thp;
        f1 = P.m[1] * T0 + P.m[1] * P.thp;
        f2 = P.m[2] * T0 + P.m[2] * P.thp;
    }
}

// Contravariant components (JcV^T f) in the metric order XX YY ZZ XY YZ XZ YX ZY ZX.
template <typename R>
__device__ __forceinline__ void contravariant(const R J[9], R f0, R f1, R f2, R& g0, R& g1, R& g2) {
    g0 = f0 * J[0] + f1 * J[6] + f2 * J[8];
    g1 = f0 * J[3] + f1 * J[1] + f2 * J[7];
    g2 = f0 * J[5] + f1 * J[4] + f2 * J[2];
}

// SYMMETRY ghost of a vector.
template <typename R>
__device__ __forceinline__ void sym_r(const R* p, const R* n, R* r) {
    const R imn = r_div(R(1), r_sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]));
    const R e0 = n[0] * imn, e1 = n[1] * imn, e2 = n[2] * imn;
    const R dn = e0 * p[0] + e1 * p[1] + e2 * p[2];
    r[0] = p[0] - e0 * dn; r[1] = p[1] - e1 * dn; r[2] = p[2] - e2 * dn;
    const R mr = r_sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    if (mr <= R(1e-7)) return;
    const R s = r_div(r_sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]), mr);
    r[0] *= s; r[1] *= s; r[2] *= s;
}

// Ghost state of a boundary face node from the owner state A: velocity and
// momentum by the velocity condition (SYMMETRY: tangential projection,
// DIRICHLET: the patch value with rho of A), scalars copied (zero gradient).
template <typename R>
__device__ __forceinline__ void ghost_state(const Prim<R>& A, const R nv[3], int kind, const double* uval,
                                            R rho_ref, Prim<R>& Q) {
    Q = A;
    if (kind == 1) {           // SYMMETRY
        sym_r(A.u, nv, Q.u);
        sym_r(A.m, nv, Q.m);
    } else if (kind == 2) {    // DIRICHLET
        Q.u[0] = R(uval[0]); Q.u[1] = R(uval[1]); Q.u[2] = R(uval[2]);
        const R rho = rho_ref + A.rp;
        Q.m[0] = rho * Q.u[0]; Q.m[1] = rho * Q.u[1]; Q.m[2] = rho * Q.u[2];
    }
    Q.umag = r_sqrt(Q.u[0] * Q.u[0] + Q.u[1] * Q.u[1] + Q.u[2] * Q.u[2]);
}

// Rusanov flux through a face node with area-weighted normal nv, owner side O
// and neighbour side N, fI = 1/2 inside, 0 on a boundary (the ghost side).
template <typename R>
__device__ __forceinline__ void rusanov(const Prim<R>& O, const Prim<R>& N, const R nv[3], R fI, R T0, R flux[5]) {
    const R fJ = R(1) - fI;
    const R lam = ((O.umag * fI + N.umag * fJ) + (O.c * fI + N.c * fJ)) * R(0.5);
    const R mn = r_sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    const R imn = r_div(R(1), mn);
    const R en[3] = {nv[0] * imn, nv[1] * imn, nv[2] * imn};
    // rho: avg(m).n - lambda (rho'_N - rho'_O) |n|
    flux[0] = (O.m[0] * fI + N.m[0] * fJ) * nv[0] + (O.m[1] * fI + N.m[1] * fJ) * nv[1] +
              (O.m[2] * fI + N.m[2] * fJ) * nv[2] - lam * (N.rp - O.rp) * mn;
    // momentum: avg(m (x) U + p'I).n - e_r lambda (dm . n)
    const R dmn = lam * ((N.m[0] - O.m[0]) * nv[0] + (N.m[1] - O.m[1]) * nv[1] + (N.m[2] - O.m[2]) * nv[2]);
    const R un_O = O.u[0] * nv[0] + O.u[1] * nv[1] + O.u[2] * nv[2];
    const R un_N = N.u[0] * nv[0] + N.u[1] * nv[1] + N.u[2] * nv[2];
    for (int r = 0; r < 3; r++)
        flux[1 + r] = (O.m[r] * un_O + O.pp * nv[r]) * fI + (N.m[r] * un_N + N.pp * nv[r]) * fJ - en[r] * dmn;
    // Theta: avg(m theta).n - lambda (Theta'_N - Theta'_O) |n|
    const R mnO = O.m[0] * nv[0] + O.m[1] * nv[1] + O.m[2] * nv[2];
    const R mnN = N.m[0] * nv[0] + N.m[1] * nv[1] + N.m[2] * nv[2];
    flux[4] = (mnO * T0 + mnO * O.thp) * fI + (mnN * T0 + mnN * N.thp) * fJ - lam * (N.tp - O.tp) * mn;
}

// The 12 node values a mortar face interpolates from the coarse side (and
// takes at the refined node): the normal fluxes through nv (5), rho', m (3),
// Theta', |U|, c.
template <typename R>
__device__ __forceinline__ void mortar_node_values(const Prim<R>& P, const R nv[3], R T0, R* v) {
    const R mn = P.m[0] * nv[0] + P.m[1] * nv[1] + P.m[2] * nv[2];
    const R un = P.u[0] * nv[0] + P.u[1] * nv[1] + P.u[2] * nv[2];
    v[0] = mn;
    for (int r = 0; r < 3; r++) v[1 + r] = P.m[r] * un + P.pp * nv[r];
    v[4] = mn * T0 + mn * P.thp;
    v[5] = P.rp;
    for (int c = 0; c < 3; c++) v[6 + c] = P.m[c];
    v[9] = P.tp;
    v[10] = P.umag;
    v[11] = P.c;
}

// Rusanov flux per unit face-node weight (phi) from the refined node's
// values vf and the coarse side's interpolated values vc, at the interior
// weight 1/2; owner O, neighbour N (fo: the refined element owns the face).
// The background density and Theta enter the jumps through bg (coarse
// interpolation minus the refined node's; Theta_ref = rho_ref T0).
template <typename R>
__device__ __forceinline__ void mortar_rusanov(const R* vf, const R* vc, R bg, const R nv[3], bool fo, R T0,
                                               R* phi) {
    const R drho = fo ? (vc[5] - vf[5]) + bg : (vf[5] - vc[5]) - bg;   // rho_N - rho_O
    const R dth = fo ? (vc[9] - vf[9]) + T0 * bg : (vf[9] - vc[9]) - T0 * bg;
    R dm[3];
    for (int c = 0; c < 3; c++) dm[c] = fo ? vc[6 + c] - vf[6 + c] : vf[6 + c] - vc[6 + c];
    const R lam = (((vf[10] + vc[10]) * R(0.5)) + ((vf[11] + vc[11]) * R(0.5))) * R(0.5);
    const R mn = r_sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    const R imn = r_div(R(1), mn);
    const R dmn = lam * (dm[0] * nv[0] + dm[1] * nv[1] + dm[2] * nv[2]);
    phi[0] = (vf[0] + vc[0]) * R(0.5) - lam * drho * mn;
    for (int r = 0; r < 3; r++) phi[1 + r] = (vf[1 + r] + vc[1 + r]) * R(0.5) - (nv[r] * imn) * dmn;
    phi[4] = (vf[4] + vc[4]) * R(0.5) - lam * dth * mn;
}

}  // namespace nsem
