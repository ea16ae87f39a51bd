// Operators of the field API (src/field/field.h:2880-3920): face
// interpolation, explicit divergence and gradient, the FV laplacian, source
// terms, MeshMatrix and the temporal schemes.  The node loops run on the
// device (field_ops.hip); these templates pick the helper for the value
// type and keep the reference's sequence of operations.
#ifndef NSEM_FIELD_OPS_H
#define NSEM_FIELD_OPS_H

namespace nsem_field {
template <class T> inline Scalar* S(T* p) { return reinterpret_cast<Scalar*>(p); }
template <class T> inline const Scalar* S(const T* p) { return reinterpret_cast<const Scalar*>(p); }
template <class T> constexpr int NC = (int)(sizeof(T) / sizeof(Scalar));

// fF = cF[FO] fI + cF[FN] (1 - fI) per face node (cds, field.h:2880-2893)
void cds(const Scalar* cell, int nc, Scalar* facet);
// fF = cds(F) - unit(fN) (x) lambda (q[FN] - q[FO]) (rusanov, field.h:2927-2943)
void rusanov(const Scalar* F, int ncF, const Scalar* q, int ncq, const Scalar* lam, Scalar* fF);
// r = sum over the cell's faces of +-dot(fF, fN) (div_flux<false>, field.h:3085-3117)
void div_flux(const Scalar* fF, int ncF, Scalar* r);
// r -= the weak volume term of DIVD (field.h:3443-3464)
void divf_volume(const Scalar* F, int ncF, Scalar* r);
// r = sum over the cell's faces of +-fN (x) (fF - p) (strong) or fN (x) fF
// (grad_flux, field.h:3050-3082)
void grad_flux(const Scalar* p, int nc, const Scalar* fF, bool strong, Scalar* r);
// r += the strong volume term of GRADD (field.h:3336-3357)
void grad_volume(const Scalar* p, int nc, Scalar* r);
// r = sum over the cell's faces of +-fF (sum, field.h:2990-3014)
void sum_facet(const Scalar* fF, int nc, Scalar* r);
// r -= (add: +=) x[other] a over the faces of each interior node's cell,
// a = c0[k] on the owner side, c1[k] on the neighbour side: mul
// (field.h:2369-2420) with c0 = ann, c1 = ano; its transpose mult
// (2429-2480) with c0 = ano, c1 = ann; getRHS (2483-2534) adds with ann, ano
void offdiag(const Scalar* x, int nc, const Scalar* c0, const Scalar* c1, bool add, Scalar* r);
// FV implicit convection: ano, ann per face and ap += their sums
// (div_flux_implicit, field.h:3133-3189); mu_facet only for HYBRID
void conv_coeffs(const Scalar* flux, const Scalar* mu_facet, int scheme, Scalar blend, Scalar* ano, Scalar* ann,
                 Scalar* ap);
// lap coefficients: ano = ann = fD mu (penalty or FV) or mu, ap = sum of
// the cell's couplings (field.h:3580-3612)
void lap_coeffs(const Scalar* mu_facet, bool use_fD, Scalar* ano, Scalar* ann, Scalar* ap);
// FV non-orthogonal correction: r = dot(g, K) with K = fN - fD (cC[FN] -
// cC[FO]), limited to half of ano (x[FN] - x[FO]) (field.h:3676-3699)
// (scalar fields with their vector gradient, vector fields with their
// tensor gradient; nc components of x and r)
void nonortho_flux(const Scalar* gface, const Scalar* x, const Scalar* ano, Scalar* r, int nc);
// applyImplicitBCs (field.h:2541-2582) of the field fIndex's conditions
void implicit_bcs(Scalar* ap, Scalar* ann, Scalar* Su, const Scalar* cF, int nc, Int fIndex);
// SolveT for a scalar FV system (ap, ano, ann after applyImplicitBCs, Su,
// solution x in place; solve.cpp:39-559 PCG with the DIAG preconditioner)
// by the library's conjugate gradient
void solve_fv_pcg(const Scalar* ap, const Scalar* ano, const Scalar* ann, const Scalar* Su, Scalar* x,
                  Int fIndex);
// SolveT for any FV system and every method and preconditioner: Jacobi,
// SOR, conjugate gradient (symmetric) or biconjugate gradient, with none,
// DIAG, SSOR or DILU preconditioning, in the reference's order of
// operations throughout (level-scheduled substitutions and sweeps,
// sequential-order dot products)
void solve_general(const Scalar* ap, const Scalar* ano, const Scalar* ann, const Scalar* Su, Scalar* x, int nc,
                   int flags);
// entity sizes, DG shape
bool dg();
}  // namespace nsem_field

namespace Mesh {
extern VectorFacetField fN;
}

// ------------------------------------------------------------- cds / flx --
template <class T>
MeshField<T, FACET> cds(const MeshField<T, CELL>& cF) {
    MeshField<T, FACET> f;
    nsem_field::cds(nsem_field::S(cF.dptr()), nsem_field::NC<T>, nsem_field::S(f.dptr_w()));
    return f;
}
template <class T, class A>
MeshField<T, FACET> cds(const DVExpr<T, A>& e) {
    return cds(MeshField<T, CELL>(e));
}
template <class T>
inline MeshField<T, CELL> flxc(const MeshField<T, CELL>& p) {
    return p;
}
template <class T, class A>
MeshField<T, CELL> flxc(const DVExpr<T, A>& e) {
    return MeshField<T, CELL>(e);
}
template <class T>
ScalarFacetField flx(const MeshField<T, CELL>& p) {
    MeshField<T, FACET> f = cds(p);
    ScalarFacetField r = dot(f, Mesh::fN);
    return r;
}
template <class T, class A>
ScalarFacetField flx(const DVExpr<T, A>& e) {
    return flx(MeshField<T, CELL>(e));
}
template <class T1, class T2>
MeshField<T1, FACET> rusanov(const MeshField<T1, CELL>& F, const MeshField<T2, CELL>& q,
                             const ScalarFacetField& lambdaMax) {
    MeshField<T1, FACET> f;
    nsem_field::rusanov(nsem_field::S(F.dptr()), nsem_field::NC<T1>, nsem_field::S(q.dptr()), nsem_field::NC<T2>,
                        lambdaMax.dptr(), nsem_field::S(f.dptr_w()));
    return f;
}
template <class T>
MeshField<T, CELL> sum(const MeshField<T, FACET>& fF) {
    MeshField<T, CELL> r;
    nsem_field::sum_facet(nsem_field::S(fF.dptr()), nsem_field::NC<T>, nsem_field::S(r.dptr_w()));
    return r;
}

// ------------------------------------------------------------ divergence --
template <bool strong = false, class T1, class T2 = Scalar>
auto divf(const MeshField<T1, CELL>& p, bool perunit_volume = false, const ScalarFacetField* flux = 0,
          const MeshField<T2, CELL>* q = 0, const ScalarFacetField* lambdaMax = 0) {
    static_assert(!strong, "the strong-form divergence is not part of this build");
    using R = decltype(dot(std::declval<T1>(), std::declval<Vector>()));
    MeshField<T1, FACET> fF;
    if (lambdaMax && q && Controls::convection_scheme == Controls::RUSANOV)
        fF = rusanov(p, *q, *lambdaMax);
    else if (flux == 0 || Controls::convection_scheme == Controls::CDS)
        fF = cds(p);
    else
        nsem_field::fail("divf: only the RUSANOV and CDS face fluxes are built");
    MeshField<R, CELL> r;
    nsem_field::div_flux(nsem_field::S(fF.dptr()), nsem_field::NC<T1>, nsem_field::S(r.dptr_w()));
    if (nsem_field::dg()) nsem_field::divf_volume(nsem_field::S(p.dptr()), nsem_field::NC<T1>, nsem_field::S(r.dptr_w()));
    if (perunit_volume) r = r / Mesh::cV;
    fillBCs(r);
    return r;
}
template <bool strong = false, class T, class A>
auto divf(const DVExpr<T, A>& e, bool perunit_volume = false) {
    return divf<strong>(MeshField<T, CELL>(e), perunit_volume);
}

// -------------------------------------------------------------- gradient --
template <bool strong = false, class T>
auto gradf(const MeshField<T, CELL>& p, bool perunit_volume = false) {
    using R = decltype(mul(std::declval<Vector>(), std::declval<T>()));
    MeshField<T, FACET> fF = cds(p);
    MeshField<R, CELL> r;
    const bool dg = nsem_field::dg();
    nsem_field::grad_flux(nsem_field::S(p.dptr()), nsem_field::NC<T>, nsem_field::S(fF.dptr()), dg && strong,
                          nsem_field::S(r.dptr_w()));
    if (dg) {
        static_assert(strong || !strong, "");
        if (!strong) nsem_field::fail("gradf: only the strong-form DG gradient is built");
        nsem_field::grad_volume(nsem_field::S(p.dptr()), nsem_field::NC<T>, nsem_field::S(r.dptr_w()));
    }
    if (perunit_volume) r = r / Mesh::cV;
    nsem_field::fill_bcs(nsem_field::S(r.dptr_w()), nsem_field::NC<R>, p.fIndex, nsem_field::NC<T>);
    return r;
}
template <bool strong = false, class T, class A>
auto gradf(const DVExpr<T, A>& e, bool perunit_volume = false) {
    return gradf<strong>(MeshField<T, CELL>(e), perunit_volume);
}

// ------------------------------------------------------------ the matrix --
template <class T1, class T2 = Scalar, class T3 = T1>
struct MeshMatrix {
    MeshField<T1, CELL>* cF;
    MeshField<T2, CELL> ap;
    MeshField<T2, FACET> ano;
    MeshField<T2, FACET> ann;
    MeshField<T3, CELL> Su;
    Int flags;
    enum FLAG { SYMMETRIC = 1, DIAGONAL = 2 };

    MeshMatrix() : cF(0), flags(SYMMETRIC | DIAGONAL) {
        ap = T2(0);
        ano = T2(0);
        ann = T2(0);
        Su = T3(0);
    }
    MeshMatrix(const MeshMatrix& p) : cF(p.cF), ap(p.ap), ano(p.ano), ann(p.ann), Su(p.Su), flags(p.flags) {}
    MeshMatrix(MeshField<T3, CELL>* pcF) : MeshMatrix() { cF = pcF; }
    MeshMatrix(const MeshField<T3, CELL>& p) : cF(0), flags(SYMMETRIC | DIAGONAL) {
        ap = T2(0);
        ano = T2(0);
        ann = T2(0);
        Su = p;
    }
    template <class A>
    MeshMatrix(const DVExpr<T3, A>& p) : cF(0), flags(SYMMETRIC | DIAGONAL) {
        ap = T2(0);
        ano = T2(0);
        ann = T2(0);
        Su = p;
    }
    MeshMatrix operator-() const {
        MeshMatrix r;
        r.cF = cF;
        r.flags = flags;
        r.ap = -ap;
        r.ano = -ano;
        r.ann = -ann;
        r.Su = -Su;
        return r;
    }
    MeshMatrix& operator=(const MeshMatrix& q) {
        cF = q.cF;
        flags = q.flags;
        ap = q.ap;
        ano = q.ano;
        ann = q.ann;
        Su = q.Su;
        return *this;
    }
    MeshMatrix& operator+=(const MeshMatrix& q) {
        flags &= q.flags;
        ap += q.ap;
        ano += q.ano;
        ann += q.ann;
        Su += q.Su;
        return *this;
    }
    MeshMatrix& operator-=(const MeshMatrix& q) {
        flags &= q.flags;
        ap -= q.ap;
        ano -= q.ano;
        ann -= q.ann;
        Su -= q.Su;
        return *this;
    }
    MeshMatrix& operator*=(const Scalar& q) {
        ap *= q;
        ano *= q;
        ann *= q;
        Su *= q;
        return *this;
    }
    MeshMatrix& operator/=(const Scalar& q) {
        ap /= q;
        ano /= q;
        ann /= q;
        Su /= q;
        return *this;
    }
    friend MeshMatrix operator+(const MeshMatrix& p, const MeshMatrix& q) {
        MeshMatrix r = p;
        r += q;
        return r;
    }
    friend MeshMatrix operator-(const MeshMatrix& p, const MeshMatrix& q) {
        MeshMatrix r = p;
        r -= q;
        return r;
    }
    friend MeshMatrix operator==(const MeshMatrix& p, const MeshMatrix& q) {
        MeshMatrix r = p;
        r -= q;
        return r;
    }
    // relaxation of steady problems (field.h:1802-1806)
    void Relax(Scalar UR) {
        ap /= UR;
        Su += (*cF) * ap * (1 - UR);
    }
};

typedef MeshMatrix<Scalar> ScalarCellMatrix;
typedef MeshMatrix<Vector> VectorCellMatrix;
typedef MeshMatrix<STensor> STensorCellMatrix;
typedef MeshMatrix<Tensor> TensorCellMatrix;

// A x (mul, field.h:2369-2420)
template <class T1, class T2, class T3>
MeshField<T3, CELL> mul(const MeshMatrix<T1, T2, T3>& p, const MeshField<T1, CELL>& q, const bool = false) {
    MeshField<T3, CELL> r = q * p.ap;
    if (p.flags & p.DIAGONAL) return r;
    nsem_field::offdiag(nsem_field::S(q.dptr()), nsem_field::NC<T1>, p.ann.dptr(), p.ano.dptr(), false,
                        nsem_field::S(r.dptr_w()));
    return r;
}
// A^T x (mult, field.h:2429-2480)
template <class T1, class T2, class T3>
MeshField<T3, CELL> mult(const MeshMatrix<T1, T2, T3>& p, const MeshField<T1, CELL>& q, const bool = false) {
    MeshField<T3, CELL> r = q * p.ap;
    if (p.flags & p.DIAGONAL) return r;
    nsem_field::offdiag(nsem_field::S(q.dptr()), nsem_field::NC<T1>, p.ano.dptr(), p.ann.dptr(), false,
                        nsem_field::S(r.dptr_w()));
    return r;
}
// b - (L + U) x (getRHS, field.h:2483-2534)
template <class T1, class T2, class T3>
MeshField<T3, CELL> getRHS(const MeshMatrix<T1, T2, T3>& p, const MeshField<T1, CELL>& q, const bool = false) {
    MeshField<T3, CELL> r = p.Su;
    if (p.flags & p.DIAGONAL) return r;
    nsem_field::offdiag(nsem_field::S(q.dptr()), nsem_field::NC<T1>, p.ann.dptr(), p.ano.dptr(), true,
                        nsem_field::S(r.dptr_w()));
    return r;
}
template <class T1, class T2, class T3, class A>
MeshField<T3, CELL> mul(const MeshMatrix<T1, T2, T3>& p, const DVExpr<T1, A>& q, const bool sync = false) {
    return mul(p, MeshField<T1, CELL>(q), sync);
}

// src (field.h:3711-3723)
template <class T1, class T2, class T3>
MeshMatrix<T1, T2, T3> src(MeshField<T1, CELL>& cF, const MeshField<T3, CELL>& Su, const MeshField<T2, CELL>& Sp) {
    MeshMatrix<T1, T2, T3> m;
    m.cF = &cF;
    m.flags |= (m.SYMMETRIC | m.DIAGONAL);
    m.ap = -(Sp * Mesh::cV);
    m.Su = (Su * Mesh::cV);
    m.ano = Scalar(0);
    m.ann = Scalar(0);
    return m;
}
template <class type>
MeshField<type, CELL> srcf(const MeshField<type, CELL>& Su) {
    return MeshField<type, CELL>(Su * Mesh::cV);
}

// FV laplacian (field.h:3571-3704; the DG interior-penalty branch is not
// built)
template <bool strong = false, class type>
MeshMatrix<type> lap(MeshField<type, CELL>& cF, const ScalarCellField& muc, const bool penalty = false) {
    using namespace Controls;
    if (nsem_field::dg()) nsem_field::fail("lap: the DG laplacian is not part of this build");
    MeshMatrix<type> m;
    m.cF = &cF;
    m.flags = m.SYMMETRIC;
    m.Su = type(0);
    m.ap = Scalar(0);
    {
        ScalarFacetField mu = cds(muc);
        nsem_field::lap_coeffs(mu.dptr(), true, m.ano.dptr_w(), m.ann.dptr_w(), m.ap.dptr_w());
    }
    if (nonortho_scheme != NONE) {
        MeshField<decltype(mul(std::declval<Vector>(), std::declval<type>())), FACET> g =
            cds(muc * gradf<false>(cF, true));
        MeshField<type, FACET> r;
        nsem_field::nonortho_flux(nsem_field::S(g.dptr()), nsem_field::S(cF.dptr()), m.ano.dptr(),
                                  nsem_field::S(r.dptr_w()), nsem_field::NC<type>);
        m.Su = sum(r);
    }
    (void)penalty;
    return m;
}
template <bool strong = false, class type>
MeshMatrix<type> lap(MeshField<type, CELL>& cF, const Scalar& mu, const bool penalty = false) {
    return lap<strong>(cF, ScalarCellField(mu), penalty);
}

// FV implicit convection (div, field.h:3496-3567, and the FV part of
// div_flux_implicit, 3133-3189); the deferred-correction schemes (LUD and
// after) and the DG implicit form are not built
template <bool strong = false, class type>
MeshMatrix<type> div(MeshField<type, CELL>& cF, const VectorCellField& fluxc, const ScalarFacetField& flux,
                     const ScalarFacetField* lambdaMax = 0, const ScalarCellField* muc = 0) {
    using namespace Controls;
    if (nsem_field::dg()) nsem_field::fail("div: the implicit DG convection is not part of this build");
    if (convection_scheme >= LUD)
        nsem_field::fail("div: the deferred-correction convection schemes (LUD and after) are not part of this build");
    (void)fluxc;
    (void)lambdaMax;
    MeshMatrix<type> m;
    m.cF = &cF;
    m.flags = 0;
    m.Su = type(0);
    m.ap = Scalar(0);
    ScalarFacetField mu_f(false);
    const Scalar* mup = nullptr;
    if (convection_scheme == HYBRID && muc) {
        mu_f = cds(*muc);
        mup = mu_f.dptr();
    }
    nsem_field::conv_coeffs(flux.dptr(), mup, (int)convection_scheme, blend_factor, m.ano.dptr_w(), m.ann.dptr_w(),
                            m.ap.dptr_w());
    return m;
}

// -------------------------------------------------- temporal derivative --
// ddt (field.h:3740-3841): the BDF and Runge-Kutta forms (the Adams forms
// are not built)
template <class type>
MeshMatrix<type> ddt(MeshMatrix<type>& M, const MeshField<type, CELL>& rhs, ScalarCellField* rho,
                     ScalarCellField* rho0) {
    using namespace Controls;
    MeshMatrix<type> m;
    m.cF = M.cF;
    m.flags = (m.SYMMETRIC | m.DIAGONAL);
    const Scalar dt = Controls::dt;
    auto prev = [&](Int k) -> const MeshField<type, CELL>& { return *M.cF->tstore[k]; };
    auto cFR = [&]() { return rho0 ? MeshField<type, CELL>((*M.cF) * (*rho0)) : MeshField<type, CELL>(*M.cF); };
    const Int ns = M.cF->nstored;
    if (time_scheme <= BDF6) {
        if (time_scheme >= BDF6 && ns >= 6) {
            m.ap = (-147.0 / (60.0 * dt)) * Mesh::cV;
            m.Su = ((360.0 * prev(0) - 450.0 * prev(1) + 400.0 * prev(2) - 225.0 * prev(3) + 72.0 * prev(4) -
                     10.0 * prev(5)) / 147.0) * m.ap;
        } else if (time_scheme >= BDF5 && ns >= 5) {
            m.ap = (-137.0 / (60.0 * dt)) * Mesh::cV;
            m.Su = ((300.0 * prev(0) - 300.0 * prev(1) + 200.0 * prev(2) - 75.0 * prev(3) + 12.0 * prev(4)) /
                    137.0) * m.ap;
        } else if (time_scheme >= BDF4 && ns >= 4) {
            m.ap = (-25.0 / (12.0 * dt)) * Mesh::cV;
            m.Su = ((48.0 * prev(0) - 36.0 * prev(1) + 16.0 * prev(2) - 3.0 * prev(3)) / 25.0) * m.ap;
        } else if (time_scheme >= BDF3 && ns >= 3) {
            m.ap = (-11.0 / (6.0 * dt)) * Mesh::cV;
            m.Su = ((18.0 * prev(0) - 9.0 * prev(1) + 2.0 * prev(2)) / 11.0) * m.ap;
        } else if (time_scheme >= BDF2 && ns >= 2) {
            m.ap = (-3.0 / (2.0 * dt)) * Mesh::cV;
            m.Su = ((4.0 * prev(0) - prev(1)) / 3.0) * m.ap;
        } else if (time_scheme >= BDF1 && ns >= 1) {
            m.ap = (-1.0 / dt) * Mesh::cV;
            m.Su = prev(0) * m.ap;
        }
    } else if (time_scheme >= RK1) {
        ScalarCellField mdt = dt / Mesh::cV;
        const MeshField<type, CELL>& k1 = rhs;
        m.ap = (-1.0 / dt) * Mesh::cV;
        if (time_scheme >= RK4) {
            MeshField<type, CELL> k2 = k1 - mul(M, k1 * mdt / 2);
            MeshField<type, CELL> k3 = k1 - mul(M, k2 * mdt / 2);
            MeshField<type, CELL> k4 = k1 - mul(M, k3 * mdt);
            m.Su = cFR() * m.ap + (k1 + 2 * k2 + 2 * k3 + k4) / 6;
        } else if (time_scheme >= RK3) {
            MeshField<type, CELL> k2 = k1 - mul(M, k1 * mdt / 2);
            MeshField<type, CELL> k3 = k1 - mul(M, (2 * k2 - k1) * mdt);
            m.Su = cFR() * m.ap + (k1 + 4 * k2 + k3) / 6;
        } else if (time_scheme >= RK2) {
            MeshField<type, CELL> k2 = k1 - mul(M, k1 * mdt);
            m.Su = cFR() * m.ap + (k1 + k2) / 2;
        } else {
            m.Su = cFR() * m.ap + k1;
        }
    } else {
        nsem_field::fail("ddt: the Adams-Moulton and Adams-Bashforth schemes are not part of this build");
    }
    if (rho) m.ap *= (*rho);
    m.ano = Scalar(0);
    m.ann = Scalar(0);
    return m;
}

// addTemporal (field.h:3874-3920)
template <int order, class type>
void addTemporal(MeshMatrix<type>& M, Scalar cF_UR, ScalarCellField* rho = 0, ScalarCellField* rho0 = 0) {
    using namespace Controls;
    static_assert(order == 1, "only the first time derivative is built");
    if (state == STEADY) {
        M.Relax(cF_UR);
        return;
    }
    MeshField<type, CELL> rhs = M.Su - mul(M, *(M.cF));
    if (time_scheme <= BDF6) {
        MeshField<type, CELL> cFR = rho0 ? MeshField<type, CELL>((*M.cF) * (*rho0)) : MeshField<type, CELL>(*M.cF);
        if (!(M.cF->access & STOREPREV))
            M.cF->initStore(time_scheme - BDF1 + 1, cFR);
        else if (current_step > M.cF->nstored - 1)
            M.cF->updateStore(cFR);
        M += ddt(M, rhs, rho, rho0);
    } else if (time_scheme >= RK1) {
        M = ddt(M, rhs, rho, rho0);
    } else {
        nsem_field::fail("addTemporal: the Adams schemes are not part of this build");
    }
}


// transport-equation forms (field.h:3924-3982)
template <class type>
MeshMatrix<type> diffusion(MeshField<type, CELL>& cF, const ScalarCellField& mu, Scalar cF_UR,
                           ScalarCellField* rho = 0, ScalarCellField* rho0 = 0) {
    MeshMatrix<type> M = -lap(cF, mu, true);
    M.cF = &cF;
    addTemporal<1>(M, cF_UR, rho, rho0);
    return M;
}
template <class type>
MeshMatrix<type> diffusion(MeshField<type, CELL>& cF, const ScalarCellField& mu, Scalar cF_UR,
                           const MeshField<type, CELL>& Su, const ScalarCellField& Sp, ScalarCellField* rho = 0,
                           ScalarCellField* rho0 = 0) {
    MeshMatrix<type> M = -lap(cF, mu, true) - src(cF, Su, Sp);
    M.cF = &cF;
    addTemporal<1>(M, cF_UR, rho, rho0);
    return M;
}
template <class type>
MeshMatrix<type> convection(MeshField<type, CELL>& cF, const VectorCellField& Fc, const ScalarFacetField& F,
                            Scalar cF_UR, const ScalarFacetField* lambdaMax = 0, ScalarCellField* rho = 0,
                            ScalarCellField* rho0 = 0) {
    MeshMatrix<type> M = div(cF, Fc, F, lambdaMax);
    addTemporal<1>(M, cF_UR, rho, rho0);
    return M;
}
template <class type>
MeshMatrix<type> convection(MeshField<type, CELL>& cF, const VectorCellField& Fc, const ScalarFacetField& F,
                            Scalar cF_UR, const MeshField<type, CELL>& Su, const ScalarCellField& Sp,
                            const ScalarFacetField* lambdaMax = 0, ScalarCellField* rho = 0,
                            ScalarCellField* rho0 = 0) {
    MeshMatrix<type> M = div(cF, Fc, F, lambdaMax) - src(cF, Su, Sp);
    addTemporal<1>(M, cF_UR, rho, rho0);
    return M;
}
template <class type>
MeshMatrix<type> transport(MeshField<type, CELL>& cF, const VectorCellField& Fc, const ScalarFacetField& F,
                           const ScalarCellField& mu, Scalar cF_UR, const ScalarFacetField* lambdaMax = 0,
                           ScalarCellField* rho = 0, ScalarCellField* rho0 = 0) {
    MeshMatrix<type> M = div(cF, Fc, F, lambdaMax, &mu) - lap(cF, mu);
    addTemporal<1>(M, cF_UR, rho, rho0);
    return M;
}
template <class type>
MeshMatrix<type> transport(MeshField<type, CELL>& cF, const VectorCellField& Fc, const ScalarFacetField& F,
                           const ScalarCellField& mu, Scalar cF_UR, const MeshField<type, CELL>& Su,
                           const ScalarCellField& Sp, const ScalarFacetField* lambdaMax = 0,
                           ScalarCellField* rho = 0, ScalarCellField* rho0 = 0) {
    MeshMatrix<type> M = div(cF, Fc, F, lambdaMax, &mu) - lap(cF, mu) - src(cF, Su, Sp);
    addTemporal<1>(M, cF_UR, rho, rho0);
    return M;
}

#endif
