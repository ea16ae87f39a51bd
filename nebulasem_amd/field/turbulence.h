// Turbulence model interface of the field API (src/turbulence/turbulence.h):
// the base model, which adds only the viscous stress and so makes the
// solvers laminar, with the reference's run-time selection by the
// `turbulence_model` key.  The RANS and LES models (mixing length, k-e,
// RNG k-e, realizable k-e, k-w, LES) are not part of this build: selecting
// one stops the run with a message.
#ifndef NSEM_FIELD_TURBULENCE_H
#define NSEM_FIELD_TURBULENCE_H

#include "field.h"
#include "solve.h"

struct Turbulence_Model {
    VectorCellField& U;
    VectorCellField& Fc;
    ScalarFacetField& F;
    ScalarCellField& rho;
    ScalarCellField& mu;
    Util::ParamList params;
    Int writeStress;

    Turbulence_Model(VectorCellField& tU, VectorCellField& tFc, ScalarFacetField& tF, ScalarCellField& trho,
                     ScalarCellField& tmu)
        : U(tU), Fc(tFc), F(tF), rho(trho), mu(tmu), params("turbulence"), writeStress(0) {}
    virtual ~Turbulence_Model() {}
    virtual void enroll() { params.enroll("writeStress", new Util::BoolOption(&writeStress)); }
    virtual void solve() {}
    virtual ScalarCellField getTurbVisc() { return Scalar(0); }
    // div(mu dev(grad(U)^T, 2)) per unit volume (turbulence.h:104-106)
    virtual VectorCellField getExplicitStresses() { return divf(mu * dev(trn(gradf(U, true)), 2.0), true); }
    // V = 2 mu dev(sym(grad U), 1)
    STensorCellField getViscousStress() { return 2.0 * mu * dev(sym(gradf(U, true)), 1.0); }
    virtual STensorCellField getReynoldsStress() { return STensor(0); }
    virtual ScalarCellField getK() { return Scalar(0); }

    static Int turb_model;
    static bool bneedWallDist;
    static bool needWallDist() { return bneedWallDist; }
    static void RegisterTable(Util::ParamList& params);
    static Turbulence_Model* Select(VectorCellField& U, VectorCellField& Fc, ScalarFacetField& F,
                                    ScalarCellField& rho, ScalarCellField& mu);
};

#endif
