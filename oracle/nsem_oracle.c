/*
 * ORACLE — test infrastructure only.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this; it is the checker, never the
 * product path.
 *
 * A plain-C, single-threaded CPU restatement of the reference NebulaSEM
 * algorithms on the hot path, written to follow the reference's own data
 * flow: fields stored on every node including ghost nodes, face values built
 * by gather (cds / rusanov), a cell-by-cell surface integral and a weak-form
 * volume integral, then explicit boundary-condition passes.
 *
 *   om_euler_split_steps  apps/euler/euler.cpp:179-287 (BDF1; diffusion NO or YES):
 *       T += T0; Fc = rho*U; lambda = (cds|U| + cds sqrt(g R T))/2;
 *       rho  <- Solve(divf(U*rho) + ddt)               field.h:3416-3478,3740-3920
 *       p    <- P0 (rho T R / P0)^gamma - p_ref        euler.cpp:211-213
 *       U    <- Solve(divf(Fc (x) U + I p) - src + ddt)  euler.cpp:216-240
 *       T    <- Solve(divf(Fc T) + ddt); T -= T0       euler.cpp:243-286
 *     cds           field.h:2880-2893     rusanov   field.h:2927-2943
 *     div_flux      field.h:3085-3117     divf      field.h:3416-3478
 *     fillBCs       field.h:2730-2776     applyExplicitBCs field.h:2585-2727
 *     sym (SYMMETRY) src/tensor/tensor.h:483-492
 *   om_fv_poisson          apps/potential/potential.cpp:58-64 pressure solve:
 *       divf (FV) field.h:3470-3477, lap (FV) field.h:3571-3615,
 *       applyImplicitBCs field.h:2541-2582, SolveT PCG solve.cpp:259-280,403-417
 *       with the reference's stopping rule sqrt(|sum z^2| / |sum x^2|) (solve.cpp:11-35).
 *   om_euler_hydrostatic   euler.cpp:104-149 background and initial density.
 *   om_euler_diagnostics   euler.cpp:261-283, calc_courant field.cpp:440-448,
 *       reduce_* field.h:953-1006: Courant max/min/avg, mass, energy, volume.
 *   2:1 non-conforming faces: scatter_nc / gather_nc (scatter_non_conforming,
 *       gather_non_conforming, field.h:2018-2248) inside cds, rusanov,
 *       div_flux and grad_flux, with the mortar matrices psiRef / psiCor
 *       (dg.cpp:516-583) and the mortar flags FMC (mesh.cpp:433-444).
 *   om_euler_mol_rhs       the method-of-lines residual L(q) at one state: the
 *       reference's divf(fq, false, &F, &q, &lambdaMax) of the euler.cpp
 *       fluxes (181-251) all formed from that state.  Pinned bit for bit
 *       against the reference's own operators evaluated at mid-RK states of
 *       eight decks (oracle/ref_tendency.cpp; tests/test_oracle_mol_pin.py),
 *       and against the split step by first-order consistency.
 *   om_euler_ssprk3_steps  SSP-RK3 (Shu & Osher 1988) of that residual (no
 *       reference counterpart: the reference's RK schemes reduce to forward
 *       Euler on this path); pinned through the residual above and the
 *       third-order convergence, conservation and rest-state properties of
 *       tests/test_oracle_mol.py.
 *
 * Pinning: om_euler_split_steps and om_fv_poisson are checked against the
 * outputs of the reference build (oracle/ref.mk) committed under
 * tests/golden/, the split step also on grids refined 2:1 by the
 * reference's prepare tool (tests/golden/amr/); see tests/test_oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int32_t npx, npy, npz, np, npf;
    uint32_t nbcs, ncells, nfacets;
    uint64_t nbf, naf;             /* interior nodes, all nodes           */
    const uint32_t *FO, *FN;       /* nfacets*npf                         */
    const double *fN;              /* nfacets*npf*3                       */
    const double *fI;              /* nfacets*npf                         */
    const double *fD;              /* nfacets*npf (FV)                    */
    const double *cV;              /* naf                                 */
    const double *cC;              /* naf*3                               */
    const double *Jinv;            /* naf*9 (XX YY ZZ XY YZ XZ YX ZY ZX)   */
    const double *psi[3], *dpsi[3];
    const uint32_t *allFaces, *fi0, *fi1;
    int32_t nbnd;                  /* boundary patches (name order)       */
    const uint32_t *boff, *bfaces;
    /* 2:1 non-conforming faces; FMC may be null (conforming mesh)        */
    const uint32_t *FMC, *FOC, *FNC;             /* nfacets                */
    const uint32_t *cell_off, *cell_faces, *faceid; /* cell side lists    */
    const double *gFC;                           /* nfacets*3 centres      */
    const double *psiRef[3], *psiCor[3];         /* [2][n*n] per direction */
} om_mesh;

enum { OM_NEUMANN = 0, OM_SYMMETRY = 1, OM_DIRICHLET = 2 };
enum { XX, YY, ZZ, XY, YZ, XZ, YX, ZY, ZX };

typedef struct {
    double dt, P0, T0, cp, cv;
    double g[3];
    int32_t buoyancy;
    /* per patch boundary kinds and values (values: 1 or 3 per patch) */
    const int32_t *bc_rho, *bc_U, *bc_T, *bc_p;
    const double *val_rho, *val_U, *val_T, *val_p;
    /* artificial viscosity (euler block `diffusion YES`): mu = rho viscosity,
     * heat diffusivity mu / Pr (Fluid::viscosity, Fluid::Pr) */
    int32_t diffusion;
    double viscosity, Pr;
    /* method-of-lines expressions: 0 = the reference's, p' = P0 (rho theta
     * R/P0)^gamma - p_ref (euler.cpp:211-213) and the strong gradient of
     * theta = theta' + T0 (euler.cpp:181, 250); 1 = the same quantities free
     * of cancellation: p' = p_ref expm1(gamma log1p(Theta'/(rho_ref T0)))
     * (Theta' = rho theta - rho_ref T0) and the gradient of theta' (the
     * constant T0 has none) */
    int32_t mol_exact;
} om_euler;

static double dot3(const double* p, const double* q) { /* right-nested, tensor.h:117 */
    double r = p[2] * q[2] + 0.0;
    r = p[1] * q[1] + r;
    return p[0] * q[0] + r;
}
static int equal0(double p) { return fabs(p) <= 1e-7; }

/* Built with -DOM_PERTURB_POW the equation of state's pow() is moved by
 * -1, 0 or +1 ulp pseudo-randomly per call: the tests use that build to
 * measure how far a 1-ulp libm difference moves the reference algorithm's
 * outputs.  Built with -DOM_CR_POW the equation of state's and the
 * diagnostics' pow() are the correctly rounded power (x^y in quadruple
 * precision, libquadmath's powq, rounded once to double): the reference's
 * algorithm with a correctly rounded pow, where glibc's pow (below 0.52 ulp)
 * rounds the wrong way for about one argument in a thousand. */
#if defined(OM_CR_POW)
#include <quadmath.h>
static double diag_pow(double a, double b) { return (double)powq((__float128)a, (__float128)b); }
static double eos_pow(double a, double b, unsigned long long i) { (void)i; return diag_pow(a, b); }
#elif defined(OM_PERTURB_POW)
static unsigned long long g_seed = 1;
static double eos_pow(double a, double b, unsigned long long i) {
    double r = pow(a, b);
    g_seed = g_seed * 6364136223846793005ULL + 1442695040888963407ULL + i;
    int k = (int)((g_seed >> 33) % 3) - 1;
    if (k > 0) r = nextafter(r, INFINITY);
    else if (k < 0) r = nextafter(r, -INFINITY);
    return r;
}
static double diag_pow(double a, double b) { return pow(a, b); }
#else
static double eos_pow(double a, double b, unsigned long long i) { (void)i; return pow(a, b); }
static double diag_pow(double a, double b) { return pow(a, b); }
#endif

/* ---------------------------------------------------- boundary passes -- */
static void bc_scalar(const om_mesh* m, const int32_t* kind, const double* val, double* f) {
    for (int32_t b = 0; b < m->nbnd; b++)
        for (uint32_t j = m->boff[b]; j < m->boff[b + 1]; j++) {
            uint32_t fid = m->bfaces[j];
            for (int32_t n = 0; n < m->npf; n++) {
                uint64_t k = (uint64_t)fid * m->npf + n;
                uint32_t c1 = m->FO[k], c2 = m->FN[k];
                if (c2 >= m->naf) continue;
                if (kind[b] == OM_NEUMANN) {
                    double dv[3] = {m->cC[3 * c2] - m->cC[3 * c1], m->cC[3 * c2 + 1] - m->cC[3 * c1 + 1],
                                    m->cC[3 * c2 + 2] - m->cC[3 * c1 + 2]};
                    f[c2] = f[c1] + val[b] * sqrt(dot3(dv, dv));
                } else if (kind[b] == OM_SYMMETRY) {
                    f[c2] = f[c1];
                } else {
                    f[c2] = val[b];
                }
            }
        }
}

static void sym_vec(const double* p, const double* n, double* r) {
    double mn = sqrt(dot3(n, n));
    double en[3] = {n[0] / mn, n[1] / mn, n[2] / mn};
    double A[6];  /* I - en en^T, STensor order XX YY ZZ XY YZ XZ */
    A[0] = 1 - en[0] * en[0];
    A[1] = 1 - en[1] * en[1];
    A[2] = 1 - en[2] * en[2];
    A[3] = 0 - en[0] * en[1];
    A[4] = 0 - en[1] * en[2];
    A[5] = 0 - en[0] * en[2];
    r[0] = A[0] * p[0] + A[3] * p[1] + A[5] * p[2];
    r[1] = A[3] * p[0] + A[1] * p[1] + A[4] * p[2];
    r[2] = A[5] * p[0] + A[4] * p[1] + A[2] * p[2];
    double mr = sqrt(dot3(r, r));
    if (equal0(mr)) return;
    double s = sqrt(dot3(p, p)) / mr;
    r[0] *= s; r[1] *= s; r[2] *= s;
}

static void bc_vector(const om_mesh* m, const int32_t* kind, const double* val, double* f) {
    for (int32_t b = 0; b < m->nbnd; b++)
        for (uint32_t j = m->boff[b]; j < m->boff[b + 1]; j++) {
            uint32_t fid = m->bfaces[j];
            for (int32_t n = 0; n < m->npf; n++) {
                uint64_t k = (uint64_t)fid * m->npf + n;
                uint32_t c1 = m->FO[k], c2 = m->FN[k];
                if (c2 >= m->naf) continue;
                if (kind[b] == OM_NEUMANN) {
                    double dv[3] = {m->cC[3 * c2] - m->cC[3 * c1], m->cC[3 * c2 + 1] - m->cC[3 * c1 + 1],
                                    m->cC[3 * c2 + 2] - m->cC[3 * c1 + 2]};
                    double md = sqrt(dot3(dv, dv));
                    for (int c = 0; c < 3; c++) f[3 * c2 + c] = f[3 * c1 + c] + val[3 * b + c] * md;
                } else if (kind[b] == OM_SYMMETRY) {
                    sym_vec(&f[3 * c1], &m->fN[3 * k], &f[3 * c2]);
                } else {
                    for (int c = 0; c < 3; c++) f[3 * c2 + c] = val[3 * b + c];
                }
            }
        }
}

/* ghost node <- owner node after an integral (fillBCs, field.h:2733-2743) */
static void fill_bcs(const om_mesh* m, double* f, int nc) {
    for (uint32_t i = m->nbcs; i < m->ncells; i++) {
        uint32_t f0 = m->fi0[i];
        for (int32_t n = 0; n < m->npf; n++) {
            uint64_t k = (uint64_t)m->allFaces[f0] * m->npf + n;
            uint32_t c2 = m->FN[k];
            if (c2 < m->naf)
                for (int c = 0; c < nc; c++) f[(uint64_t)c2 * nc + c] = f[(uint64_t)m->FO[k] * nc + c];
        }
    }
}

/* ------------------------------------------------------------------ DG -- */
/* Explicit weak divergence of a flux F with `rows` rows of 3 components
 * (rows=1: vector flux of a scalar equation, rows=3: tensor flux), with
 * Rusanov face flux using q (rows components) and lambda per facet node.
 * r has naf*rows entries.  field.h:3416-3464. */
/* ---- 2:1 non-conforming faces (field.h:2018-2248) ----------------------
 * A mortar face fi (FMC 1 or 2) is one of the faces a coarse cell cn has on
 * one side; its other cell is a refined one.  FMC 2: cn is the face's owner,
 * FMC 1: its neighbour.  The coarse side's values at fi's nodes are the
 * coarse side's nodes interpolated to the half of the side fi covers
 * (psiRef, scatter); a face field is brought back to the coarse side's node
 * positions by psiCor (gather).  The half is found by comparing fi's centre
 * with the side's centre (mean of its faces' centres) along the side's two
 * in-plane node directions. */
typedef struct { uint32_t cn; int fid, h0, h1; } om_mortar;

static uint64_t node_of(const om_mesh* m, uint32_t c, int i, int j, int k) {
    return (uint64_t)c * m->np + (uint64_t)i * m->npy * m->npz + (uint64_t)j * m->npz + k;
}

static int mortar_of(const om_mesh* m, uint32_t fi, om_mortar* mt) {
    const uint32_t fm = m->FMC ? m->FMC[fi] : 0;
    if (fm < 1) return 0;
    const uint32_t cn = fm == 1 ? m->FNC[fi] : m->FOC[fi];
    int fid = -1;
    for (uint32_t j = m->cell_off[cn]; j < m->cell_off[cn + 1]; j++)
        if (m->cell_faces[j] == fi) { fid = (int)m->faceid[j]; break; }
    double ccn[3] = {0, 0, 0};
    int nch = 0;
    for (uint32_t j = m->cell_off[cn]; j < m->cell_off[cn + 1]; j++)
        if ((int)m->faceid[j] == fid) {
            for (int c = 0; c < 3; c++) ccn[c] += m->gFC[3 * (uint64_t)m->cell_faces[j] + c];
            nch++;
        }
    for (int c = 0; c < 3; c++) ccn[c] /= (double)nch;
    const double* cco = &m->gFC[3 * (uint64_t)fi];
    const int NPX = m->npx, NPY = m->npy, NPZ = m->npz;
    uint64_t v0, va, vb;
    if (fid == 0 || fid == 1) {
        const int k = fid == 0 ? 0 : NPZ - 1;
        v0 = node_of(m, cn, 0, 0, k); va = node_of(m, cn, NPX - 1, 0, k); vb = node_of(m, cn, 0, NPY - 1, k);
    } else if (fid == 2 || fid == 3) {
        const int j = fid == 2 ? 0 : NPY - 1;
        v0 = node_of(m, cn, 0, j, 0); va = node_of(m, cn, NPX - 1, j, 0); vb = node_of(m, cn, 0, j, NPZ - 1);
    } else {
        const int i = fid == 4 ? 0 : NPX - 1;
        v0 = node_of(m, cn, i, 0, 0); va = node_of(m, cn, i, NPY - 1, 0); vb = node_of(m, cn, i, 0, NPZ - 1);
    }
    const double *p0 = &m->cC[3 * v0], *pa = &m->cC[3 * va], *pb = &m->cC[3 * vb];
    double dn[3], dO[3], ea[3], eb[3];
    for (int c = 0; c < 3; c++) {
        dn[c] = ccn[c] - p0[c]; dO[c] = cco[c] - p0[c];
        ea[c] = pa[c] - p0[c]; eb[c] = pb[c] - p0[c];
    }
    mt->cn = cn;
    mt->fid = fid;
    mt->h0 = dot3(dn, ea) >= dot3(dO, ea) ? 0 : 1;
    mt->h1 = dot3(dn, eb) >= dot3(dO, eb) ? 0 : 1;
    return (int)fm;
}

/* The side's in-plane directions (d0, d1) and the face-node index of node
 * (a, b) on it, for side id fid. */
static void side_dirs(const om_mesh* m, int fid, int* d0, int* d1) {
    if (fid == 0 || fid == 1) { *d0 = 0; *d1 = 1; }
    else if (fid == 2 || fid == 3) { *d0 = 0; *d1 = 2; }
    else { *d0 = 1; *d1 = 2; }
    (void)m;
}
static int npd(const om_mesh* m, int d) { return d == 0 ? m->npx : d == 1 ? m->npy : m->npz; }
static uint64_t side_node(const om_mesh* m, const om_mortar* mt, int a, int b) {
    const int NPX = m->npx, NPY = m->npy, NPZ = m->npz;
    if (mt->fid == 0 || mt->fid == 1) return node_of(m, mt->cn, a, b, mt->fid == 0 ? 0 : NPZ - 1);
    if (mt->fid == 2 || mt->fid == 3) return node_of(m, mt->cn, a, mt->fid == 2 ? 0 : NPY - 1, b);
    return node_of(m, mt->cn, mt->fid == 4 ? 0 : NPX - 1, a, b);
}

/* Owner- and neighbour-side values of a node field (w doubles per node) at
 * every face node; the coarse side of a mortar face interpolated
 * (scatter_non_conforming, field.h:2134-2248). */
static void scatter_nc(const om_mesh* m, const double* f, int w, double* fO, double* fN) {
    const uint64_t nfn = (uint64_t)m->nfacets * m->npf;
    for (uint64_t k = 0; k < nfn; k++) {
        uint32_t c1 = m->FO[k], c2 = m->FN[k];
        for (int c = 0; c < w; c++) {
            fO[k * w + c] = c1 < m->naf ? f[(uint64_t)c1 * w + c] : 0;
            fN[k * w + c] = c2 < m->naf ? f[(uint64_t)c2 * w + c] : 0;
        }
    }
    if (!m->FMC || m->np <= 1) return;  /* FV: no mortars (NPMAT = 0) */
    for (uint32_t fi = 0; fi < m->nfacets; fi++) {
        om_mortar mt;
        const int fm = mortar_of(m, fi, &mt);
        if (!fm) continue;
        double* side = fm == 1 ? fN : fO;
        int d0, d1;
        side_dirs(m, mt.fid, &d0, &d1);
        const int n0 = npd(m, d0), n1 = npd(m, d1);
        const double* R0 = m->psiRef[d0] + (uint64_t)mt.h0 * n0 * n0;
        const double* R1 = m->psiRef[d1] + (uint64_t)mt.h1 * n1 * n1;
        for (int ao = 0; ao < n0; ao++)
            for (int bo = 0; bo < n1; bo++) {
                const uint64_t out = (uint64_t)fi * m->npf + ao * n1 + bo;
                for (int c = 0; c < w; c++) {
                    double sum = 0.0;
                    for (int an = 0; an < n0; an++)
                        for (int bn = 0; bn < n1; bn++)
                            sum += f[side_node(m, &mt, an, bn) * w + c] * R0[an * n0 + ao] * R1[bn * n1 + bo];
                    side[out * w + c] = sum;
                }
            }
    }
}

/* Owner- and neighbour-side copies of a face field (w doubles per face
 * node); the coarse side of a mortar face projected back to the coarse
 * side's node positions (gather_non_conforming, field.h:2018-2132). */
static void gather_nc(const om_mesh* m, const double* fF, int w, double* fO, double* fN) {
    const uint64_t nfn = (uint64_t)m->nfacets * m->npf;
    memcpy(fO, fF, sizeof(double) * nfn * w);
    memcpy(fN, fF, sizeof(double) * nfn * w);
    if (!m->FMC || m->np <= 1) return;  /* FV: no mortars (NPMAT = 0) */
    for (uint32_t fi = 0; fi < m->nfacets; fi++) {
        om_mortar mt;
        const int fm = mortar_of(m, fi, &mt);
        if (!fm) continue;
        double* side = fm == 1 ? fN : fO;
        int d0, d1;
        side_dirs(m, mt.fid, &d0, &d1);
        const int n0 = npd(m, d0), n1 = npd(m, d1);
        const double* C0 = m->psiCor[d0] + (uint64_t)mt.h0 * n0 * n0;
        const double* C1 = m->psiCor[d1] + (uint64_t)mt.h1 * n1 * n1;
        for (int an = 0; an < n0; an++)
            for (int bn = 0; bn < n1; bn++) {
                const uint64_t out = (uint64_t)fi * m->npf + an * n1 + bn;
                for (int c = 0; c < w; c++) {
                    double sum = 0.0;
                    for (int ao = 0; ao < n0; ao++)
                        for (int bo = 0; bo < n1; bo++)
                            sum += fF[((uint64_t)fi * m->npf + ao * n1 + bo) * w + c] * C0[ao * n0 + an] *
                                   C1[bo * n1 + bn];
                    side[out * w + c] = sum;
                }
            }
    }
}

static void divf_dg(const om_mesh* m, int rows, const double* F, const double* q,
                    const double* lam, double* r) {
    const uint64_t nfn = (uint64_t)m->nfacets * m->npf;
    const int w = rows * 3;
    double* fF = (double*)calloc(nfn * w, sizeof(double));
    double* FOs = (double*)malloc(sizeof(double) * nfn * w);
    double* FNs = (double*)malloc(sizeof(double) * nfn * w);
    double* qO = (double*)malloc(sizeof(double) * nfn * rows);
    double* qN = (double*)malloc(sizeof(double) * nfn * rows);
    scatter_nc(m, F, w, FOs, FNs);
    scatter_nc(m, q, rows, qO, qN);
    /* rusanov = cds(F) - unit(fN) (x) lambda (qN - qO) */
    for (uint64_t k = 0; k < nfn; k++) {
        uint32_t c2 = m->FN[k];
        if (c2 >= m->naf) continue;
        const double fi = m->fI[k];
        const double* nv = &m->fN[3 * k];
        double mn = sqrt(dot3(nv, nv));
        double en[3] = {nv[0] / mn, nv[1] / mn, nv[2] / mn};
        for (int rr = 0; rr < rows; rr++) {
            double dq = lam[k] * (qN[k * rows + rr] - qO[k * rows + rr]);
            for (int c = 0; c < 3; c++) {
                double v = FOs[k * w + rr * 3 + c] * fi + FNs[k * w + rr * 3 + c] * (1 - fi);
                if (rows == 1) v = v - en[c] * dq;
                fF[(k * rows + rr) * 3 + c] = v;
            }
        }
        if (rows == 3) { /* tensor flux: row r gets en_r * (lambda dq_c) in column c */
            for (int rr = 0; rr < 3; rr++)
                for (int c = 0; c < 3; c++) {
                    double dq = lam[k] * (qN[k * 3 + c] - qO[k * 3 + c]);
                    fF[(k * 3 + rr) * 3 + c] -= en[rr] * dq;
                }
        }
    }
    /* surface integral, cell by cell (div_flux<false>), the coarse side of a
     * mortar face taking the gathered flux */
    double* fFO = FOs;
    double* fFN = FNs;
    gather_nc(m, fF, w, fFO, fFN);
    memset(r, 0, sizeof(double) * m->naf * rows);
    for (uint32_t i = 0; i < m->ncells; i++)
        for (uint32_t f = m->fi0[i]; f < m->fi1[i]; f++)
            for (int32_t n = 0; n < m->npf; n++) {
                uint64_t k = (uint64_t)m->allFaces[f] * m->npf + n;
                uint32_t c1 = m->FO[k], c2 = m->FN[k];
                const double* nv = &m->fN[3 * k];
                int own = (c1 >= (uint64_t)i * m->np && c1 < (uint64_t)(i + 1) * m->np);
                if (!own && c2 >= m->naf) continue;
                for (int rr = 0; rr < rows; rr++) {
                    const double* v = own ? &fFO[(k * rows + rr) * 3] : &fFN[(k * rows + rr) * 3];
                    double d = rows == 1 ? dot3(v, nv) : v[0] * nv[0] + v[1] * nv[1] + v[2] * nv[2];
                    if (own) r[(uint64_t)c1 * rows + rr] += d;
                    else r[(uint64_t)c2 * rows + rr] -= d;
                }
            }
    free(fF);
    free(FOs);
    free(FNs);
    free(qO);
    free(qN);
    /* volume integral (weak form) */
    const int NPX = m->npx, NPY = m->npy, NPZ = m->npz;
    for (uint32_t ci = 0; ci < m->nbcs; ci++)
        for (int ii = 0; ii < NPX; ii++)
            for (int jj = 0; jj < NPY; jj++)
                for (int kk = 0; kk < NPZ; kk++) {
                    uint64_t index = (uint64_t)ci * m->np + ii * NPY * NPZ + jj * NPZ + kk;
                    double Jin[9];
                    for (int c = 0; c < 9; c++) Jin[c] = m->Jinv[9 * index + c] * m->cV[index];
                    for (int line = 0; line < 3; line++) {
                        int npts = line == 0 ? NPX : (line == 1 ? NPY : NPZ);
                        for (int t = 0; t < npts; t++) {
                            int im = ii, jm = jj, km = kk;
                            if (line == 0) im = t;
                            else if (line == 1) { if (t == jj) continue; jm = t; }
                            else { if (t == kk) continue; km = t; }
                            uint64_t index1 = (uint64_t)ci * m->np + im * NPY * NPZ + jm * NPZ + km;
                            double d[3];
                            d[0] = m->dpsi[0][ii * NPX + im] * m->psi[1][jj * NPY + jm] * m->psi[2][kk * NPZ + km];
                            d[1] = m->psi[0][ii * NPX + im] * m->dpsi[1][jj * NPY + jm] * m->psi[2][kk * NPZ + km];
                            d[2] = m->psi[0][ii * NPX + im] * m->psi[1][jj * NPY + jm] * m->dpsi[2][kk * NPZ + km];
                            double dj[3] = {Jin[XX] * d[0] + Jin[XY] * d[1] + Jin[XZ] * d[2],
                                            Jin[YX] * d[0] + Jin[YY] * d[1] + Jin[YZ] * d[2],
                                            Jin[ZX] * d[0] + Jin[ZY] * d[1] + Jin[ZZ] * d[2]};
                            for (int rr = 0; rr < rows; rr++) {
                                const double* v = &F[(index * rows + rr) * 3];
                                double s = rows == 1 ? dot3(v, dj) : v[0] * dj[0] + v[1] * dj[1] + v[2] * dj[2];
                                r[index1 * rows + rr] -= s;
                            }
                        }
                    }
                }
    fill_bcs(m, r, rows);
}

/* cds of a node scalar at every facet node (field.h:2880-2893) */
static void cds_scalar(const om_mesh* m, const double* f, double* out) {
    const uint64_t nfn = (uint64_t)m->nfacets * m->npf;
    double* fO = (double*)malloc(sizeof(double) * nfn);
    double* fN = (double*)malloc(sizeof(double) * nfn);
    scatter_nc(m, f, 1, fO, fN);
    for (uint64_t k = 0; k < nfn; k++) {
        uint32_t c2 = m->FN[k];
        if (c2 >= m->naf) { out[k] = 0; continue; }
        out[k] = fO[k] * m->fI[k] + fN[k] * (1 - m->fI[k]);
    }
    free(fO);
    free(fN);
}

/* Strong-form DG gradient of a node field with `rows` components,
 * gradf<true>(p, true) (field.h:3327-3372 with grad_flux<true>,
 * field.h:3050-3082): the surface lift fN (x) (cds(p) - p) summed cell by
 * cell over the cell's faces (every cell, ghosts included), then the local
 * derivative (Jinv cV grad psi_index1(x_index)) (x) p[index1] along the three
 * GLL lines, divided by cV, ghost nodes copied from their owners and then
 * set on the field's NEUMANN and SYMMETRY patches as below.  Output
 * G[(node * 3 + a) * rows + b] = d_a p_b (the reference's mul(Vector, .)). */
static void gradf_strong(const om_mesh* m, int rows, const double* p, const int32_t* kind,
                         const double* val, double* G) {
    const uint64_t nfn = (uint64_t)m->nfacets * m->npf;
    double* fF = (double*)calloc(nfn * rows, sizeof(double));
    double* fO = (double*)malloc(sizeof(double) * nfn * rows);
    double* fNs = (double*)malloc(sizeof(double) * nfn * rows);
    scatter_nc(m, p, rows, fO, fNs);
    for (uint64_t k = 0; k < nfn; k++) {
        uint32_t c2 = m->FN[k];
        if (c2 >= m->naf) continue;
        for (int b = 0; b < rows; b++)
            fF[k * rows + b] = fO[k * rows + b] * m->fI[k] + fNs[k * rows + b] * (1 - m->fI[k]);
    }
    /* grad_flux<true> reads the gathered face values (field.h:3050-3082) */
    gather_nc(m, fF, rows, fO, fNs);
    memset(G, 0, sizeof(double) * m->naf * 3 * rows);
    for (uint32_t i = 0; i < m->ncells; i++)
        for (uint32_t f = m->fi0[i]; f < m->fi1[i]; f++)
            for (int32_t n = 0; n < m->npf; n++) {
                uint64_t k = (uint64_t)m->allFaces[f] * m->npf + n;
                uint32_t c1 = m->FO[k], c2 = m->FN[k];
                const double* nv = &m->fN[3 * k];
                if (c1 >= (uint64_t)i * m->np && c1 < (uint64_t)(i + 1) * m->np) {
                    for (int b = 0; b < rows; b++) {
                        const double dp = fO[k * rows + b] - p[(uint64_t)c1 * rows + b];
                        for (int a = 0; a < 3; a++) G[((uint64_t)c1 * 3 + a) * rows + b] += nv[a] * dp;
                    }
                } else if (c2 < m->naf) {
                    for (int b = 0; b < rows; b++) {
                        const double dp = fNs[k * rows + b] - p[(uint64_t)c2 * rows + b];
                        for (int a = 0; a < 3; a++) G[((uint64_t)c2 * 3 + a) * rows + b] -= nv[a] * dp;
                    }
                }
            }
    free(fF);
    free(fO);
    free(fNs);
    const int NPX = m->npx, NPY = m->npy, NPZ = m->npz;
    for (uint32_t ci = 0; ci < m->nbcs; ci++)
        for (int ii = 0; ii < NPX; ii++)
            for (int jj = 0; jj < NPY; jj++)
                for (int kk = 0; kk < NPZ; kk++) {
                    uint64_t index = (uint64_t)ci * m->np + ii * NPY * NPZ + jj * NPZ + kk;
                    double Jin[9];
                    for (int c = 0; c < 9; c++) Jin[c] = m->Jinv[9 * index + c] * m->cV[index];
                    for (int line = 0; line < 3; line++) {
                        int npts = line == 0 ? NPX : (line == 1 ? NPY : NPZ);
                        for (int t = 0; t < npts; t++) {
                            int im = ii, jm = jj, km = kk;
                            if (line == 0) im = t;
                            else if (line == 1) { if (t == jj) continue; jm = t; }
                            else { if (t == kk) continue; km = t; }
                            uint64_t index1 = (uint64_t)ci * m->np + im * NPY * NPZ + jm * NPZ + km;
                            double d[3];
                            d[0] = m->dpsi[0][ii * NPX + im] * m->psi[1][jj * NPY + jm] * m->psi[2][kk * NPZ + km];
                            d[1] = m->psi[0][ii * NPX + im] * m->dpsi[1][jj * NPY + jm] * m->psi[2][kk * NPZ + km];
                            d[2] = m->psi[0][ii * NPX + im] * m->psi[1][jj * NPY + jm] * m->dpsi[2][kk * NPZ + km];
                            double dj[3] = {Jin[XX] * d[0] + Jin[XY] * d[1] + Jin[XZ] * d[2],
                                            Jin[YX] * d[0] + Jin[YY] * d[1] + Jin[YZ] * d[2],
                                            Jin[ZX] * d[0] + Jin[ZY] * d[1] + Jin[ZZ] * d[2]};
                            for (int a = 0; a < 3; a++)
                                for (int b = 0; b < rows; b++)
                                    G[(index * 3 + a) * rows + b] += dj[a] * p[index1 * rows + b];
                        }
                    }
                }
    for (uint64_t i = 0; i < m->naf * 3 * rows; i++) G[i] = G[i] / m->cV[i / (3 * rows)];
    fill_bcs(m, G, 3 * rows);
    /* fillBCs(r, p.fIndex) (field.h:2745-2768): on the field's own NEUMANN
     * patches the ghost gradient is the patch's BCondition value read as
     * the gradient type, on SYMMETRY patches it is zero.  The value read as
     * a Vector (scalar field) is (value, shape, tvalue), as a Tensor (vector
     * field) it is (value[3], shape, tvalue[3], tshape, zMin) in the order
     * XX YY ZZ XY YZ XZ YX ZY ZX; shape, tvalue, tshape and zMin are 0 for
     * the boundary types handled here, so only value survives, on the
     * diagonal of the tensor. */
    for (int32_t b = 0; b < m->nbnd; b++) {
        if (kind[b] != OM_NEUMANN && kind[b] != OM_SYMMETRY) continue;
        double gv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (kind[b] == OM_NEUMANN) {
            if (rows == 1) gv[0] = val[b];
            else for (int c = 0; c < 3; c++) gv[3 * c + c] = val[3 * b + c];
        }
        for (uint32_t j = m->boff[b]; j < m->boff[b + 1]; j++)
            for (int32_t n = 0; n < m->npf; n++) {
                uint64_t k = (uint64_t)m->bfaces[j] * m->npf + n;
                uint32_t c2 = m->FN[k];
                if (c2 >= m->naf) continue;
                for (int c = 0; c < 3 * rows; c++) G[(uint64_t)c2 * 3 * rows + c] = gv[c];
            }
    }
}

/* Hydrostatic background and initial density from (T', p')
 * (euler.cpp:104-149 with buoyancy, non-spherical). */
void om_euler_hydrostatic(const om_mesh* m, const om_euler* e, const double* T, const double* p,
                          double* p_ref, double* rho_ref, double* rho) {
    const double R = e->cp - e->cv, gam = e->cp / e->cv;
    for (uint64_t i = 0; i < m->naf; i++) {
        double gh = e->buoyancy ? dot3(e->g, &m->cC[3 * i]) : 0.0;
        p_ref[i] = e->buoyancy ? e->P0 * pow(1.0 + gh / (e->cp * e->T0), e->cp / R) : e->P0;
        rho_ref[i] = e->buoyancy ? (e->P0 / (R * e->T0)) * pow(p_ref[i] / e->P0, 1 / gam)
                                 : (e->P0 / (R * e->T0));
        double pt = p[i] + p_ref[i];
        rho[i] = (e->P0 / (R * (T[i] + e->T0))) * pow(pt / e->P0, 1 / gam);
    }
}

/* Per-step diagnostics of apps/euler/euler.cpp:261-283 over the interior
 * nodes, summed in node order as the reference's reduce_* do in a serial
 * build (field.h:953-1006): out = {Courant max, min, avg (calc_courant,
 * field.cpp:440-448: |U| dt / cV^(1/3)), mass sum rho cV, energy
 * sum rho cV (gh + |U|^2/2 + ((p'+p_ref)/P0)^(R/cp) (T'+T0) cv), volume}.
 * T is theta' (the reference adds T0 before the step and removes it after). */
void om_euler_diagnostics(const om_mesh* m, const om_euler* e, const double* p_ref, const double* rho,
                          const double* U, const double* T, const double* p, double* out) {
    const double R = e->cp - e->cv;
    double cmax = -1e30, cmin = 1e30, csum = 0, mass = 0, energy = 0, vol = 0;
    for (uint64_t i = 0; i < m->nbf; i++) {
        const double* u = &U[3 * i];
        const double um = sqrt(dot3(u, u));
        const double c = um * e->dt / diag_pow(m->cV[i], 1.0 / 3);
        if (fabs(c) > cmax) cmax = fabs(c);
        if (fabs(c) < cmin) cmin = fabs(c);
        csum += c;
        mass += rho[i] * m->cV[i];
        const double gh = e->buoyancy ? dot3(e->g, &m->cC[3 * i]) : 0.0;
        const double th = T[i] + e->T0;
        double sf = gh + 0.5 * dot3(u, u) + diag_pow((p[i] + p_ref[i]) / e->P0, R / e->cp) * th * e->cv;
        sf = rho[i] * m->cV[i] * sf;
        energy += sf;
        vol += m->cV[i];
    }
    out[0] = cmax; out[1] = cmin; out[2] = csum / (double)m->nbf;
    out[3] = mass; out[4] = energy; out[5] = vol;
}

/* The reference's split forward-Euler step, nsteps times.  All arrays are
 * node arrays over every node (naf), ghosts included; U and g are xyz per
 * node.  Returns 0, or 5 if a non-finite value appears. */
int om_euler_split_steps(const om_mesh* m, const om_euler* e, const double* p_ref,
                         const double* rho_ref, double* rho, double* U, double* T, double* p,
                         int nsteps) {
    const uint64_t N = m->naf, nfn = (uint64_t)m->nfacets * m->npf;
    const double R = e->cp - e->cv, gam = e->cp / e->cv;
    double* Fc = (double*)malloc(sizeof(double) * N * 3);
    double* lam = (double*)malloc(sizeof(double) * nfn);
    double* t1 = (double*)malloc(sizeof(double) * nfn);
    double* w = (double*)malloc(sizeof(double) * N);
    double* rhof = (double*)malloc(sizeof(double) * N);
    double* fq = (double*)malloc(sizeof(double) * N * 9);
    double* q = (double*)malloc(sizeof(double) * N * 3);
    double* r = (double*)malloc(sizeof(double) * N * 3);
    double* mu = (double*)malloc(sizeof(double) * N);
    double* G = e->diffusion ? (double*)malloc(sizeof(double) * N * 9) : NULL;
    int status = 0;
    for (int step = 0; step < nsteps; step++) {
        for (uint64_t i = 0; i < N; i++) T[i] += e->T0;
        for (uint64_t i = 0; i < N; i++)
            for (int c = 0; c < 3; c++) Fc[3 * i + c] = rho[i] * U[3 * i + c];
        /* lambdaMax = (cds(mag(U)) + cds(sqrt(gamma R T))) / 2 */
        for (uint64_t i = 0; i < N; i++) w[i] = sqrt(dot3(&U[3 * i], &U[3 * i]));
        cds_scalar(m, w, lam);
        for (uint64_t i = 0; i < N; i++) w[i] = sqrt(gam * R * T[i]);
        cds_scalar(m, w, t1);
        for (uint64_t k = 0; k < nfn; k++) lam[k] = (lam[k] + t1[k]) / 2;
        /* artificial viscosity mu = rho viscosity (rho^n) */
        for (uint64_t i = 0; i < N; i++) mu[i] = e->diffusion ? rho[i] * e->viscosity : 0.0;
        memcpy(rhof, rho, sizeof(double) * N);

        /* rho equation */
        for (uint64_t i = 0; i < N; i++)
            for (int c = 0; c < 3; c++) fq[3 * i + c] = U[3 * i + c] * rho[i];
        divf_dg(m, 1, fq, rho, lam, r);
        for (uint64_t i = 0; i < N; i++) {
            double ap = (-1.0 / e->dt) * m->cV[i];
            double su = r[i] + rho[i] * ap;
            rho[i] = su / ap;
        }
        bc_scalar(m, e->bc_rho, e->val_rho, rho);

        /* pressure from the new density and the old temperature */
        for (uint64_t i = 0; i < N; i++) p[i] = e->P0 * eos_pow((rho[i] * T[i] * R) / e->P0, gam, i);
        bc_scalar(m, e->bc_p, e->val_p, p);
        for (uint64_t i = 0; i < N; i++) p[i] -= p_ref[i];

        /* momentum equation: flux Fc (x) U + I p - mu gradf<true>(U),
         * Rusanov on rho^{n+1} U^n */
        if (e->diffusion) gradf_strong(m, 3, U, e->bc_U, e->val_U, G);
        for (uint64_t i = 0; i < N; i++) {
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) {
                    fq[9 * i + 3 * a + b] = Fc[3 * i + a] * U[3 * i + b] + (a == b ? p[i] : 0.0 * p[i]);
                    if (e->diffusion) fq[9 * i + 3 * a + b] = fq[9 * i + 3 * a + b] - mu[i] * G[9 * i + 3 * a + b];
                }
            for (int c = 0; c < 3; c++) q[3 * i + c] = rho[i] * U[3 * i + c];
        }
        divf_dg(m, 3, fq, q, lam, r);
        for (uint64_t i = 0; i < N; i++) {
            double ap0 = (-1.0 / e->dt) * m->cV[i];
            double ap = ap0 * rho[i];
            for (int c = 0; c < 3; c++) {
                double sc = e->buoyancy ? (rho[i] - rho_ref[i]) * e->g[c] : 0.0;
                double su = (r[3 * i + c] - sc * m->cV[i]) + (U[3 * i + c] * rhof[i]) * ap0;
                U[3 * i + c] = su / ap;
            }
        }
        bc_vector(m, e->bc_U, e->val_U, U);

        /* potential temperature equation: flux Fc T - (mu/Pr) gradf<true>(T),
         * Rusanov on rho^{n+1} T^n */
        if (e->diffusion) gradf_strong(m, 1, T, e->bc_T, e->val_T, G);
        for (uint64_t i = 0; i < N; i++) {
            for (int c = 0; c < 3; c++) {
                fq[3 * i + c] = Fc[3 * i + c] * T[i];
                if (e->diffusion) fq[3 * i + c] = fq[3 * i + c] - (mu[i] * (1 / e->Pr)) * G[3 * i + c];
            }
            q[i] = rho[i] * T[i];
        }
        divf_dg(m, 1, fq, q, lam, r);
        for (uint64_t i = 0; i < N; i++) {
            double ap0 = (-1.0 / e->dt) * m->cV[i];
            double su = r[i] + (T[i] * rhof[i]) * ap0;
            T[i] = su / (ap0 * rho[i]);
        }
        bc_scalar(m, e->bc_T, e->val_T, T);
        for (uint64_t i = 0; i < N; i++) T[i] -= e->T0;
        for (uint64_t i = 0; i < m->nbf; i++)
            if (!isfinite(rho[i]) || !isfinite(T[i])) status = 5;
        if (status) break;
    }
    free(Fc); free(lam); free(t1); free(w); free(rhof); free(fq); free(q); free(r); free(mu); free(G);
    return status;
}

/* ------------------------------------------------------------------ FV -- */
/* FV pressure Poisson as the potential app solves it: divU = divf(U);
 * Solve(lap(p, 1, true) == divU) with DIAG-PCG.  bc_p/val_p per patch
 * (NEUMANN or DIRICHLET).  U: naf*3 (ghosts set by its own BCs), p: naf
 * (initial guess, ghosts set).  On return p holds the solution on interior
 * cells; *iters, *res0, *res the reference's log line values.
 * Returns 0 (converged) or 7 (iteration limit). */
int om_fv_poisson(const om_mesh* m, const int32_t* bc_p, const double* val_p, const double* U,
                  double* p, double tol, int32_t max_it, int32_t* iters, double* res0,
                  double* res) {
    const uint64_t N = m->naf, nfn = m->nfacets;
    const uint64_t nb = m->nbf;
    double* divU = (double*)calloc(N, sizeof(double));
    double* ano = (double*)malloc(sizeof(double) * nfn);
    double* ann = (double*)malloc(sizeof(double) * nfn);
    double* ap = (double*)calloc(N, sizeof(double));
    double* Su = (double*)calloc(N, sizeof(double));
    double* rr = (double*)calloc(N, sizeof(double));
    double* z = (double*)calloc(N, sizeof(double));
    double* pp = (double*)calloc(N, sizeof(double));
    double* Ap = (double*)calloc(N, sizeof(double));
    /* divf(U), FV: cds face values, surface sum, ghost fill */
    for (uint32_t i = 0; i < m->ncells; i++)
        for (uint32_t f = m->fi0[i]; f < m->fi1[i]; f++) {
            uint64_t k = m->allFaces[f];
            uint32_t c1 = m->FO[k], c2 = m->FN[k];
            double fF[3];
            for (int c = 0; c < 3; c++) fF[c] = U[3 * c1 + c] * m->fI[k] + U[3 * c2 + c] * (1 - m->fI[k]);
            double d = dot3(fF, &m->fN[3 * k]);
            if (c1 == i) divU[c1] += d;
            else if (c2 < N) divU[c2] -= d;
        }
    fill_bcs(m, divU, 1);
    /* lap(p, 1, true): ano = ann = fD * cds(1) */
    for (uint64_t k = 0; k < nfn; k++) {
        double mu = 1.0 * m->fI[k] + 1.0 * (1 - m->fI[k]);
        ano[k] = m->fD[k] * mu;
        ann[k] = m->fD[k] * mu;
    }
    for (uint32_t i = 0; i < m->ncells; i++)
        for (uint32_t f = m->fi0[i]; f < m->fi1[i]; f++) {
            uint64_t k = m->allFaces[f];
            uint32_t c1 = m->FO[k], c2 = m->FN[k];
            if (c1 == i) ap[c1] += ano[k];
            else if (c2 < N) ap[c2] += ann[k];
        }
    /* lap == divU : Su = 0 - divU (the non-orthogonal correction vanishes for a
     * zero initial guess; this restatement requires p = 0 on entry) */
    for (uint64_t i = 0; i < N; i++) Su[i] = 0.0 - divU[i];
    /* applyImplicitBCs */
    for (int32_t b = 0; b < m->nbnd; b++)
        for (uint32_t j = m->boff[b]; j < m->boff[b + 1]; j++) {
            uint64_t k = m->bfaces[j];
            uint32_t c1 = m->FO[k], c2 = m->FN[k];
            if (c2 >= N) continue;
            if (bc_p[b] == OM_NEUMANN) {
                ap[c1] -= ann[k];
                Su[c1] += ann[k] * (p[c2] - p[c1]);
            } else {
                Su[c1] += ann[k] * p[c2];
            }
            ann[k] = 0;
        }
    (void)val_p;
#define MULT(x, out)                                                          \
    for (uint64_t i = 0; i < nb; i++) {                                        \
        double acc = x[i] * ap[i];                                             \
        for (uint32_t f = m->fi0[i]; f < m->fi1[i]; f++) {                     \
            uint64_t k = m->allFaces[f];                                       \
            uint32_t c1 = m->FO[k], c2 = m->FN[k];                             \
            if (c1 == i) acc -= x[c2] * ann[k];                                \
            else if (c2 < N) acc -= x[c1] * ano[k];                            \
        }                                                                      \
        out[i] = acc;                                                          \
    }
    /* CALC_RESID */
    MULT(p, Ap);
    for (uint64_t i = 0; i < nb; i++) rr[i] = Su[i] - Ap[i];
    for (uint64_t i = 0; i < nb; i++) z[i] = rr[i] * (1 / ap[i]);
    double sz = 0, sx = 0, o_rr = 0, oo_rr;
    for (uint64_t i = 0; i < nb; i++) { sz += z[i] * z[i]; sx += p[i] * p[i]; }
    double r0 = sqrt(sx ? fabs(sz) / fabs(sx) : 0);
    for (uint64_t i = 0; i < nb; i++) o_rr += rr[i] * z[i];
    for (uint64_t i = 0; i < nb; i++) pp[i] = z[i];
    int it = 0, status = 7;
    double rs = r0;
    while (it < max_it) {
        it++;
        MULT(pp, Ap);
        oo_rr = 0;
        for (uint64_t i = 0; i < nb; i++) oo_rr += pp[i] * Ap[i];
        double alpha = oo_rr ? o_rr / oo_rr : 0;
        for (uint64_t i = 0; i < nb; i++) p[i] = p[i] + pp[i] * alpha;
        for (uint64_t i = 0; i < nb; i++) rr[i] = rr[i] + Ap[i] * (-alpha);
        for (uint64_t i = 0; i < nb; i++) z[i] = rr[i] * (1 / ap[i]);
        oo_rr = o_rr;
        o_rr = 0;
        for (uint64_t i = 0; i < nb; i++) o_rr += rr[i] * z[i];
        double beta = oo_rr ? o_rr / oo_rr : 0;
        for (uint64_t i = 0; i < nb; i++) pp[i] = z[i] + pp[i] * beta;
        sz = 0; sx = 0;
        for (uint64_t i = 0; i < nb; i++) { sz += z[i] * z[i]; sx += p[i] * p[i]; }
        rs = sqrt(sx ? fabs(sz) / fabs(sx) : 0);
        if (rs <= tol) { status = 0; break; }
    }
#undef MULT
    *iters = it; *res0 = r0; *res = rs;
    free(divU); free(ano); free(ann); free(ap); free(Su); free(rr); free(z); free(pp); free(Ap);
    return status;
}

/* Velocity correction of the potential app, U -= gradf(p, true)
 * (potential.cpp:66-67) on the interior cells: the FV branch of gradf
 * (field.h:3363-3371) is cds(p) at every face, grad_flux<false>
 * (field.h:3050-3082: + fN cds(p) on the cell's owned faces, - on the others,
 * in the cell's face order), divided by cV.  p has naf entries: its ghosts are
 * first set by p's boundary conditions, as Solve's closing applyExplicitBCs
 * does (solve.cpp:580).  U: naf*3, interior entries updated. */
void om_fv_correct_velocity(const om_mesh* m, const int32_t* bc_p, const double* val_p, double* p, double* U) {
    bc_scalar(m, bc_p, val_p, p);
    for (uint32_t i = 0; i < m->nbf; i++) {
        double g[3] = {0, 0, 0};
        for (uint32_t f = m->fi0[i]; f < m->fi1[i]; f++) {
            uint64_t k = m->allFaces[f];
            uint32_t c1 = m->FO[k], c2 = m->FN[k];
            const double fF = p[c1] * m->fI[k] + p[c2] * (1 - m->fI[k]);
            const double* nv = &m->fN[3 * k];
            for (int a = 0; a < 3; a++) {
                if (c1 == i) g[a] += nv[a] * fF;
                else g[a] -= nv[a] * fF;
            }
        }
        for (int a = 0; a < 3; a++) U[3 * (uint64_t)i + a] = U[3 * (uint64_t)i + a] - g[a] / m->cV[i];
    }
}

/* ---------------------------------------------- method-of-lines (new) -- */
/* No reference counterpart: the reference's explicit schemes are forward
 * Euler on the split update.  This composes the pinned operators above into
 * the method-of-lines right-hand side with every flux at one state
 *   L_rho = -r_rho/cV,  L_m = -r_m/cV + (rho - rho_ref) g,  L_Th = -r_Th/cV
 * (r from divf_dg with F = rho U, rho U (x) U + p' I, rho U theta) and
 * SSP-RK3 (Shu & Osher 1988) on (rho, rho U, rho theta).  With diffusion the
 * momentum and theta fluxes carry -mu grad U and -mu/Pr grad theta' (mu =
 * rho viscosity) at the stage state, from the split step's pinned strong
 * gradient gradf_strong and its ghost rule. */
static void mol_rhs(const om_mesh* m, const om_euler* e, const double* p_ref, const double* rho_ref,
                    double* rho, double* U, double* T, const double* Tp, double* L /* naf*5 */,
                    double* rout /* naf*5 or NULL */) {
    const uint64_t N = m->naf, nfn = (uint64_t)m->nfacets * m->npf;
    const double R = e->cp - e->cv, gam = e->cp / e->cv;
    bc_scalar(m, e->bc_rho, e->val_rho, rho);
    bc_vector(m, e->bc_U, e->val_U, U);
    bc_scalar(m, e->bc_T, e->val_T, T);
    double* lam = (double*)malloc(sizeof(double) * nfn);
    double* t1 = (double*)malloc(sizeof(double) * nfn);
    double* w = (double*)malloc(sizeof(double) * N);
    double* fq = (double*)malloc(sizeof(double) * N * 9);
    double* q = (double*)malloc(sizeof(double) * N * 3);
    double* r = (double*)malloc(sizeof(double) * N * 3);
    double* p = (double*)malloc(sizeof(double) * N);
    /* artificial viscosity at the stage state: mu = rho viscosity, the strong
     * gradients of U and theta' with the split step's ghost rule */
    double* GU = e->diffusion ? (double*)malloc(sizeof(double) * N * 9) : NULL;
    double* GT = e->diffusion ? (double*)malloc(sizeof(double) * N * 3) : NULL;
    if (e->diffusion) {
        gradf_strong(m, 3, U, e->bc_U, e->val_U, GU);
        if (e->mol_exact) {
            gradf_strong(m, 1, T, e->bc_T, e->val_T, GT);
        } else {
            for (uint64_t i = 0; i < N; i++) p[i] = T[i] + e->T0;
            gradf_strong(m, 1, p, e->bc_T, e->val_T, GT);
        }
    }
    for (uint64_t i = 0; i < N; i++) w[i] = sqrt(dot3(&U[3 * i], &U[3 * i]));
    cds_scalar(m, w, lam);
    for (uint64_t i = 0; i < N; i++) w[i] = sqrt(gam * R * (T[i] + e->T0));
    cds_scalar(m, w, t1);
    for (uint64_t k = 0; k < nfn; k++) lam[k] = (lam[k] + t1[k]) / 2;
    if (e->mol_exact && Tp) {
        for (uint64_t i = 0; i < m->nbf; i++) p[i] = p_ref[i] * expm1(gam * log1p(Tp[i] / (rho_ref[i] * e->T0)));
        bc_scalar(m, e->bc_p, e->val_p, p);
    } else {
        for (uint64_t i = 0; i < N; i++) p[i] = e->P0 * pow((rho[i] * (T[i] + e->T0) * R) / e->P0, gam) - p_ref[i];
    }
    /* rho */
    for (uint64_t i = 0; i < N; i++)
        for (int c = 0; c < 3; c++) fq[3 * i + c] = U[3 * i + c] * rho[i];
    divf_dg(m, 1, fq, rho, lam, r);
    for (uint64_t i = 0; i < N; i++) L[5 * i] = -r[i] / m->cV[i];
    if (rout) for (uint64_t i = 0; i < N; i++) rout[5 * i] = r[i];
    /* momentum */
    for (uint64_t i = 0; i < N; i++) {
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                fq[9 * i + 3 * a + b] = rho[i] * U[3 * i + a] * U[3 * i + b] + (a == b ? p[i] : 0.0);
        if (e->diffusion)
            for (int c = 0; c < 9; c++) fq[9 * i + c] = fq[9 * i + c] - (rho[i] * e->viscosity) * GU[9 * i + c];
        for (int c = 0; c < 3; c++) q[3 * i + c] = rho[i] * U[3 * i + c];
    }
    divf_dg(m, 3, fq, q, lam, r);
    for (uint64_t i = 0; i < N; i++)
        for (int c = 0; c < 3; c++)
            L[5 * i + 1 + c] = -r[3 * i + c] / m->cV[i] + (e->buoyancy ? (rho[i] - rho_ref[i]) * e->g[c] : 0.0);
    if (rout) for (uint64_t i = 0; i < N; i++) for (int c = 0; c < 3; c++) rout[5 * i + 1 + c] = r[3 * i + c];
    /* rho theta */
    for (uint64_t i = 0; i < N; i++) {
        for (int c = 0; c < 3; c++) fq[3 * i + c] = rho[i] * U[3 * i + c] * (T[i] + e->T0);
        if (e->diffusion)
            for (int c = 0; c < 3; c++)
                fq[3 * i + c] = fq[3 * i + c] - ((rho[i] * e->viscosity) * (1 / e->Pr)) * GT[3 * i + c];
        q[i] = rho[i] * (T[i] + e->T0);
    }
    divf_dg(m, 1, fq, q, lam, r);
    for (uint64_t i = 0; i < N; i++) L[5 * i + 4] = -r[i] / m->cV[i];
    if (rout) for (uint64_t i = 0; i < N; i++) rout[5 * i + 4] = r[i];
    free(lam); free(t1); free(w); free(fq); free(q); free(r); free(p); free(GU); free(GT);
}

/* SSP-RK3 steps of the method-of-lines form; state in and out as the split
 * steps use it (rho, U, theta', p' on every node).  The steps carry the
 * conserved perturbations over the hydrostatic background, rho' = rho -
 * rho_ref, m = rho U, Theta' = rho theta - rho_ref T0 (the device's state
 * variables), converted from and to the primitive state as the device's
 * set_state / get_state do; p' at the end by the form mol_exact selects. */
int om_euler_ssprk3_steps(const om_mesh* m, const om_euler* e, const double* p_ref,
                          const double* rho_ref, double* rho, double* U, double* T, double* p,
                          int nsteps) {
    const uint64_t N = m->naf, nb = m->nbf;
    const double R = e->cp - e->cv, gam = e->cp / e->cv, T0 = e->T0;
    double* q0 = (double*)malloc(sizeof(double) * N * 5);
    double* qs = (double*)malloc(sizeof(double) * N * 5);
    double* L = (double*)malloc(sizeof(double) * N * 5);
    double* Tp = (double*)malloc(sizeof(double) * N);
    static const double A[3] = {0.0, 0.75, 1.0 / 3.0}, B[3] = {1.0, 0.25, 2.0 / 3.0};
    int status = 0;
    for (uint64_t i = 0; i < nb; i++) {
        qs[5 * i] = rho[i] - rho_ref[i];
        for (int c = 0; c < 3; c++) qs[5 * i + 1 + c] = rho[i] * U[3 * i + c];
        qs[5 * i + 4] = rho[i] * (T[i] + T0) - rho_ref[i] * T0;
    }
    for (int st = 0; st < nsteps; st++) {
        memcpy(q0, qs, sizeof(double) * nb * 5);
        for (int s = 0; s < 3; s++) {
            for (uint64_t i = 0; i < nb; i++) {
                rho[i] = rho_ref[i] + qs[5 * i];
                for (int c = 0; c < 3; c++) U[3 * i + c] = qs[5 * i + 1 + c] / rho[i];
                T[i] = (qs[5 * i + 4] + rho_ref[i] * T0) / rho[i] - T0;
                Tp[i] = qs[5 * i + 4];
            }
            mol_rhs(m, e, p_ref, rho_ref, rho, U, T, Tp, L, NULL);
            for (uint64_t i = 0; i < nb; i++)
                for (int v = 0; v < 5; v++)
                    qs[5 * i + v] = A[s] * q0[5 * i + v] + B[s] * (qs[5 * i + v] + e->dt * L[5 * i + v]);
        }
        for (uint64_t i = 0; i < nb * 5; i++)
            if (!isfinite(qs[i])) status = 5;
        if (status) break;
    }
    for (uint64_t i = 0; i < nb; i++) {
        rho[i] = rho_ref[i] + qs[5 * i];
        for (int c = 0; c < 3; c++) U[3 * i + c] = qs[5 * i + 1 + c] / rho[i];
        T[i] = (qs[5 * i + 4] + rho_ref[i] * T0) / rho[i] - T0;
    }
    bc_scalar(m, e->bc_rho, e->val_rho, rho);
    bc_vector(m, e->bc_U, e->val_U, U);
    bc_scalar(m, e->bc_T, e->val_T, T);
    if (e->mol_exact) {
        for (uint64_t i = 0; i < nb; i++)
            p[i] = p_ref[i] * expm1(gam * log1p(qs[5 * i + 4] / (rho_ref[i] * T0)));
        bc_scalar(m, e->bc_p, e->val_p, p);
    } else {
        for (uint64_t i = 0; i < N; i++) p[i] = e->P0 * pow((rho[i] * (T[i] + T0) * R) / e->P0, gam) - p_ref[i];
    }
    free(q0); free(qs); free(L); free(Tp);
    return status;
}

/* The method-of-lines residual at one state (rho, U, theta' on every node;
 * ghosts are set from the boundary conditions; Tp = Theta' on the interior
 * nodes, read only with mol_exact): L[5 per node] = -r / cV of the density,
 * momentum (+ (rho - rho_ref) g) and rho theta equations, and, when r is
 * not NULL, the Rusanov divergences r themselves (5 per node), which
 * divf(fq, false, &F, &q, &lambdaMax) returns in the reference. */
void om_euler_mol_rhs(const om_mesh* m, const om_euler* e, const double* p_ref, const double* rho_ref,
                      double* rho, double* U, double* T, const double* Tp, double* L, double* r) {
    mol_rhs(m, e, p_ref, rho_ref, rho, U, T, Tp, L, r);
}
