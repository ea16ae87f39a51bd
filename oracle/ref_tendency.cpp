// Tendency dump driver for the reference build (test infrastructure only).
//
// Links against the reference library built by oracle/ref.mk.  Sets up an
// Euler problem exactly as apps/euler/euler.cpp:30-149 does (controls,
// AmrIteration's LoadMesh, field reads, hydrostatic background), replaces the
// interior node values of rho, U and theta' by a state given in a raw file,
// and evaluates at that one state the reference's own operators that
// euler.cpp:180-258 applies: lambdaMax from cds (field.h:2880-2893), the
// Rusanov divergences divf(fq, false, &F, &q, &lambdaMax) (field.h:3416-3478
// with rusanov 2927-2943, div_flux 3085-3117) of the density, momentum and
// theta fluxes, with the artificial viscosity gradf<true> terms
// (field.h:3327-3372) when `diffusion YES`.  Every flux is formed from the
// same state (the method-of-lines residual), not in the split step's
// sequence.  The divergence fields r (interior nodes) and p' are written
// out, so the method-of-lines residual L = -r / cV (+ (rho - rho_ref) g for
// the momentum) can be compared with the oracle's.
//
// usage: euler_tendency <controls> <state.bin> <out.bin>   (in the case dir)
//   state.bin: doubles rho[nb], U[3 nb] (xyz per node), theta'[nb]
//   out.bin:   doubles r_rho[nb], r_m[3 nb], r_theta[nb], p'[nb]
// The binary's name contains "euler": Solver::Initialize checks it against
// the controls' solver name (apps/utils/wrapper.cpp:32-38).
#include "solve.h"
#include "iteration.h"
#include "wrapper.h"
#include "properties.h"

#include <cstdio>
#include <vector>

int main(int argc, char* argv[]) {
    MP mp(argc, argv);
    if (argc < 4) {
        std::fprintf(stderr, "usage: euler_tendency <controls> <state.bin> <out.bin>\n");
        return 2;
    }
    Solver::Initialize(argc, argv);
    std::istream& input = Solver::input;
    Int buoyancy = 1, diffusion = 1;
    Scalar dummy_ur = 0.5;
    {
        Util::ParamList params("general");
        Fluid::enroll(params);
        params.read(input);
    }
    Util::ParamList params("euler");
    params.enroll("velocity_UR", &dummy_ur);
    params.enroll("pressure_UR", &dummy_ur);
    params.enroll("t_UR", &dummy_ur);
    params.enroll("buoyancy", new Util::BoolOption(&buoyancy));
    params.enroll("diffusion", new Util::BoolOption(&diffusion));
    Util::read_params(input, false);

    AmrIteration ait;
    ScalarCellField p("p", READWRITE);
    VectorCellField U("U", READWRITE);
    ScalarCellField T("T", READWRITE);
    ScalarCellField rho("rho", READWRITE);
    VectorCellField g(false);
    Iteration it(ait.get_step());

    using namespace Fluid;
    const Scalar R = cp - cv, p_gamma = cp / cv, iPr = 1 / Pr;
    ScalarCellField p_ref, rho_ref, gh;
    if (buoyancy) {
        g.construct("gravity", WRITE);
        g = Controls::gravity;
        gh = dot(g, Mesh::cC);
        Mesh::fixedBCs<Vector>(U, g);
        applyExplicitBCs(g, true);
        p_ref = P0 * pow(1.0 + gh / (cp * T0), cp / R);
        rho_ref = (P0 / (R * T0)) * pow(p_ref / P0, 1 / p_gamma);
    } else {
        gh = Scalar(0.0);
        p_ref = P0;
        rho_ref = (P0 / (R * T0));
    }
    applyExplicitBCs(p_ref, true);
    applyExplicitBCs(rho_ref, true);

    // the given state on the interior nodes; ghosts from the boundary conditions
    const Int nb = Mesh::gBCSfield;
    {
        std::FILE* f = std::fopen(argv[2], "rb");
        if (!f) return 3;
        std::vector<double> buf((size_t)nb * 5);
        if (std::fread(buf.data(), sizeof(double), buf.size(), f) != buf.size()) return 4;
        std::fclose(f);
        for (Int i = 0; i < nb; i++) {
            rho[i] = buf[i];
            for (int c = 0; c < 3; c++) U[i][c] = buf[nb + 3 * (size_t)i + c];
            T[i] = buf[4 * (size_t)nb + i];
        }
    }
    applyExplicitBCs(rho, true);
    applyExplicitBCs(U, true);
    applyExplicitBCs(T, true);

    // euler.cpp:181-190 at this state
    T += T0;
    VectorCellField Fc = flxc(rho * U);
    ScalarFacetField F = flx(rho * U);
    ScalarFacetField lambdaMax = (cds(mag(U)) + cds(sqrt(p_gamma * R * T))) / 2;
    ScalarCellField mu;
    if (diffusion) mu = rho * viscosity;
    else mu = Scalar(0);

    // density: euler.cpp:198-203
    ScalarCellField r_rho;
    {
        VectorCellField fq = U * rho;
        ScalarFacetField Fr = flx(U);
        r_rho = divf(fq, false, &Fr, &rho, &lambdaMax);
    }
    // pressure from this state's rho and theta: euler.cpp:211-213
    p = P0 * pow((rho * T * R) / P0, p_gamma);
    applyExplicitBCs(p, true);
    p -= p_ref;
    // momentum: euler.cpp:229-233 (the source enters L, not r)
    VectorCellField r_m;
    {
        TensorCellField fq = mul(Fc, U) + TensorCellField(Constants::I_T) * p - mu * gradf<true>(U, true);
        VectorCellField q = rho * U;
        r_m = divf(fq, false, &F, &q, &lambdaMax);
    }
    // theta: euler.cpp:247-251
    ScalarCellField r_T;
    {
        ScalarCellField imu = mu * iPr;
        VectorCellField fq = Fc * T - imu * gradf<true>(T, true);
        ScalarCellField q = rho * T;
        r_T = divf(fq, false, &F, &q, &lambdaMax);
    }

    std::FILE* o = std::fopen(argv[3], "wb");
    if (!o) return 5;
    std::vector<double> out;
    out.reserve((size_t)nb * 6);
    for (Int i = 0; i < nb; i++) out.push_back(r_rho[i]);
    for (Int i = 0; i < nb; i++)
        for (int c = 0; c < 3; c++) out.push_back(r_m[i][c]);
    for (Int i = 0; i < nb; i++) out.push_back(r_T[i]);
    for (Int i = 0; i < nb; i++) out.push_back(p[i]);
    std::fwrite(out.data(), sizeof(double), out.size(), o);
    std::fclose(o);
    return 0;
}
