// Discrete adjoint and tangent-linear model of the FP64 method-of-lines
// SSP-RK3 Euler step (no reference counterpart; the reference has no
// derivative code).
//
// Residual level.  L(Q)_n = (vol_n - surf_n) / cV_n + rho'_n g (momentum rows),
// vol = sum_d D_d^T (JcV^T F(Q)) along the GLL lines, surf = the signed
// Rusanov fluxes of the node's faces.  Its vector-Jacobian product
// t = (dL/dQ)^T w is written out structurally:
//   volume   H^d = D_d w^   (w^ = w / cV; the transpose of the weak derivative),
//            Fbar = JcV H, and the pointwise flux Jacobian applied to Fbar;
//   faces    each element takes both copies of a face flux that depend on its
//            own face node (its own residual's and the neighbour's, which is
//            the same function with the opposite sign), so no element writes
//            another element's nodes: coefficient s (w^_own - w^_nbr),
//            s = -1 on the owner side;
//   buoyancy d(rho' g)/d rho'.
// The pointwise Jacobians (flux rows, Rusanov flux with its ghost state) are
// taken by forward-mode dual numbers through the same physics functions the
// forward kernels use (dg_physics.hpp), one seeded direction at a time.  The
// tangent-linear residual is the forward stage kernel instantiated on
// Dual<1>.
//
// Step level (SSP-RK3 in increment form, d1 = dt L(q), d2 = (d1 + dt L(q+d1))/4,
// q+ = q + 2(d2 + dt L(q+d2))/3), backwards, each line one VJP launch with
// the algebra fused into its epilogue:
//   t3 = J(q+d2)^T lam;  ad2 = 2/3 (lam + dt t3)
//   t2 = J(q+d1)^T ad2;  ad1 = 1/4 (ad2 + dt t2)
//   t1 = J(q)^T ad1;     lam <- lam + 2/3 dt t3 + 1/4 dt t2 + dt t1
//                             = 1/3 lam + 3/4 ad2 + ad1 + dt t1
// (the last form needs no accumulator array between the launches)
// with the forward states recomputed from checkpoints stored every
// `checkpoint` steps, one segment at a time (its states and stage increments
// held while the segment is swept backwards).
#pragma clang fp contract(fast)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "dg_context.hpp"
#include "dg_physics.hpp"

namespace nsem {

namespace {

using D1 = Dual<1>;

__device__ __forceinline__ D1 seeded(double v, bool seed) {
    D1 r(v);
    r.d[0] = seed ? 1.0 : 0.0;
    return r;
}

// d p' / d Theta' of p' = p_ref expm1(gamma log1p(Theta' / (rho_ref T0))).
__device__ __forceinline__ double dpp_dtp(const Prim<double>& P, double rref, double pref, double T0, double gam) {
    return gam * (pref + P.pp) / (rref * T0 + P.tp);
}

// The primitives of one node as dual numbers, seeded in conserved
// component k (rho', m_x, m_y, m_z, Theta'), from their values P: the
// derivatives of make_prim written out (d(1/rho), the velocity, theta', p'
// through dpp_dtp, c and |U| with the zero-at-zero convention of r_sqrt).
__device__ __forceinline__ void prim_tangent(const Prim<double>& P, double rref, double pref, double T0, double gam,
                                             double gR, int k, Prim<Dual<1>>& T) {
    const double ir = 1.0 / (rref + P.rp);
    const double drp = k == 0 ? 1.0 : 0.0, dtp = k == 4 ? 1.0 : 0.0;
    const double dm[3] = {k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0, k == 3 ? 1.0 : 0.0};
    const double dir = -ir * ir * drp;
    T.rp.v = P.rp; T.rp.d[0] = drp;
    T.tp.v = P.tp; T.tp.d[0] = dtp;
    double udu = 0.0;
    for (int c = 0; c < 3; c++) {
        T.m[c].v = P.m[c]; T.m[c].d[0] = dm[c];
        T.u[c].v = P.u[c]; T.u[c].d[0] = dm[c] * ir + P.m[c] * dir;
        udu += P.u[c] * T.u[c].d[0];
    }
    const double dthp = (dtp - T0 * drp) * ir + (P.tp - T0 * P.rp) * dir;
    T.thp.v = P.thp; T.thp.d[0] = dthp;
    T.pp.v = P.pp; T.pp.d[0] = dpp_dtp(P, rref, pref, T0, gam) * dtp;
    T.c.v = P.c; T.c.d[0] = P.c > 0.0 ? 0.5 * gR * dthp / P.c : 0.0;
    T.umag.v = P.umag; T.umag.d[0] = P.umag > 0.0 ? udu / P.umag : 0.0;
}

// Cotangent of the conserved state (rho', m, Theta') from cotangents of the
// primitives u, theta and p' (u = m/rho, theta = (Theta' + rho_ref T0)/rho).
__device__ __forceinline__ void prim_to_conserved(const Prim<double>& P, double rref, double pref, double T0,
                                                  double gam, const double ub[3], double thb, double ppb,
                                                  double qb[5]) {
    const double rho = rref + P.rp, ir = 1.0 / rho, th = T0 + P.thp;
    for (int k = 0; k < 3; k++) qb[1 + k] += ub[k] * ir;
    qb[0] -= (ub[0] * P.u[0] + ub[1] * P.u[1] + ub[2] * P.u[2]) * ir + thb * th * ir;
    qb[4] += thb * ir + ppb * dpp_dtp(P, rref, pref, T0, gam);
}

// Vector-Jacobian product of the flux rows F(Q) at one node: Q-cotangent
// (added to qb) of the row cotangents fbar[5][3].
__device__ __forceinline__ void flux_rows_vjp(const Prim<double>& P, double rref, double pref, double T0, double gam,
                                              const double fbar[5][3], double qb[5]) {
    double ub[3] = {0, 0, 0}, thb = 0, ppb = 0;
    for (int c = 0; c < 3; c++) {
        qb[1 + c] += fbar[0][c];                       // F_rho = m
        qb[1 + c] += fbar[4][c] * (T0 + P.thp);        // F_Theta = m theta
        thb += fbar[4][c] * P.m[c];
    }
    for (int r = 0; r < 3; r++) {                      // F_m_r = m_r u + p' e_r
        ppb += fbar[1 + r][r];
        for (int c = 0; c < 3; c++) {
            qb[1 + r] += fbar[1 + r][c] * P.u[c];
            ub[c] += fbar[1 + r][c] * P.m[r];
        }
    }
    prim_to_conserved(P, rref, pref, T0, gam, ub, thb, ppb, qb);
}

// Vector-Jacobian product of the interior Rusanov flux (rusanov(), owner O,
// neighbour N) with respect to the conserved state of one side S, for the
// flux cotangent phi[5]; added to qb.
__device__ __forceinline__ void rusanov_vjp_side(const Prim<double>& O, const Prim<double>& N, const double nv[3],
                                                 double fI, double T0, double gam, double gR, const double phi[5],
                                                 bool side_o, double rrS, double prS, double qb[5]) {
    const double fJ = 1.0 - fI;
    const Prim<double>& S = side_o ? O : N;
    const double fS = side_o ? fI : fJ, sD = side_o ? -1.0 : 1.0;
    const double lam = ((O.umag * fI + N.umag * fJ) + (O.c * fI + N.c * fJ)) * 0.5;
    const double nmag = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    const double en[3] = {nv[0] / nmag, nv[1] / nmag, nv[2] / nmag};
    const double Dm = (N.m[0] - O.m[0]) * nv[0] + (N.m[1] - O.m[1]) * nv[1] + (N.m[2] - O.m[2]) * nv[2];
    const double phe = phi[1] * en[0] + phi[2] * en[1] + phi[3] * en[2];
    const double lbar = -phi[0] * (N.rp - O.rp) * nmag - phe * Dm - phi[4] * (N.tp - O.tp) * nmag;
    const double rho = rrS + S.rp;
    const double mnS = S.m[0] * nv[0] + S.m[1] * nv[1] + S.m[2] * nv[2];
    const double unS = S.u[0] * nv[0] + S.u[1] * nv[1] + S.u[2] * nv[2];
    const double thS = T0 + S.thp;
    const double phm = phi[1] * S.m[0] + phi[2] * S.m[1] + phi[3] * S.m[2];
    const double phn = phi[1] * nv[0] + phi[2] * nv[1] + phi[3] * nv[2];
    qb[0] += -lam * nmag * phi[0] * sD - fS * phm * unS / rho;
    for (int k = 0; k < 3; k++)
        qb[1 + k] += fS * nv[k] * phi[0] + fS * (phi[1 + k] * unS + phm * nv[k] / rho) - phe * lam * sD * nv[k] +
                     fS * phi[4] * thS * nv[k];
    qb[4] += -lam * nmag * phi[4] * sD;
    double thb = fS * phi[4] * mnS;
    const double sb = 0.5 * fS * lbar;               // cotangent of |U_S| and of c_S
    double ub[3] = {0, 0, 0};
    if (S.umag > 0)
        for (int k = 0; k < 3; k++) ub[k] = sb * S.u[k] / S.umag;
    if (S.c > 0) thb += sb * gR / (2.0 * S.c);
    prim_to_conserved(S, rrS, prS, T0, gam, ub, thb, fS * phn, qb);
}

// Boundary face: the derivative of the flux through the ghost state built
// from the owner side with respect to the owner's state component kk,
// contracted with coef, by one forward-mode sweep.
__device__ __forceinline__ double ghost_face_vjp(const double QA[5], const double nvd[3], int kind, const double uv[3],
                                                 double rrA, double prA, double T0, double gam, double gR,
                                                 const double coef[5], int kk) {
    const D1 nv[3] = {D1(nvd[0]), D1(nvd[1]), D1(nvd[2])};
    Prim<D1> A;
    A.rp = seeded(QA[0], kk == 0);
    A.m[0] = seeded(QA[1], kk == 1); A.m[1] = seeded(QA[2], kk == 2); A.m[2] = seeded(QA[3], kk == 3);
    A.tp = seeded(QA[4], kk == 4);
    make_prim(A, D1(rrA), D1(prA), D1(T0), D1(gam), D1(gR));
    Prim<D1> N;
    ghost_state(A, nv, kind, uv, D1(rrA), N);
    D1 flux[5];
    rusanov(A, N, nv, D1(0.0), D1(T0), flux);
    double acc = 0;
    for (int v = 0; v < 5; v++) acc += coef[v] * flux[v].d[0];
    return acc;
}

// DIN: the launch reads a stage increment (compile-time, so the state loads
// carry no null-pointer branches and can be issued together).
// The primitives of element node t from LDS: Q rows, and with PL the
// derived rows too; without PL they are recomputed.
template <bool PL>
__device__ __forceinline__ void load_prim(Prim<double>& P, const double* Ql, const double* Xl, int NP, int t,
                                          double rref, double pref, double T0, double gam, double gR,
                                          const double* ppc) {
    P.rp = Ql[t]; P.m[0] = Ql[NP + t]; P.m[1] = Ql[2 * NP + t]; P.m[2] = Ql[3 * NP + t]; P.tp = Ql[4 * NP + t];
    if constexpr (PL) {
        P.u[0] = Xl[t]; P.u[1] = Xl[NP + t]; P.u[2] = Xl[2 * NP + t]; P.thp = Xl[3 * NP + t];
        P.pp = Xl[4 * NP + t]; P.c = Xl[5 * NP + t]; P.umag = Xl[6 * NP + t];
    } else {
        make_prim(P, rref, pref, T0, gam, gR, ppc);
    }
}

// Whether the VJP kernel for a compile-time element shape keeps the
// primitives (u, theta', p', c, |U|) of its nodes in LDS beside Q and w/cV,
// so the face tasks take their own node's primitives from there instead of
// recomputing them: when a work-group's LDS (vjp_epb elements) still lets
// three work-groups share a CU's 160 KiB.
__host__ __device__ constexpr bool vjp_prims_fit(int nx, int ny, int nz) {
    return nx && ny && nz &&
           ((nx * ny * nz >= 256 ? 1 : 256 / (nx * ny * nz)) *
                (17 * nx * ny * nz + 10 * (nx * ny + nx * nz + ny * nz)) +
            nx * nx + ny * ny + nz * nz) * 8 * 3 <= 160 * 1024;
}

// O32: node byte offsets formed in 32 bits (nb * 8 below 4 GiB).  At most
// 168 VGPRs, three wavefronts per SIMD (the compiler would otherwise take
// up to 176 on the 2D shapes and fall to two).
template <int NXc, int NYc, int NZc, bool DIN, bool O32>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(3))) rk_vjp_kernel(DgDev d, VjpArgs a) {
    const int NX = NXc ? NXc : d.npx, NY = NYc ? NYc : d.npy, NZ = NZc ? NZc : d.npz;
    const int NP = NX * NY * NZ;
    const int NFN = 2 * (NX * NY + NX * NZ + NY * NZ);
    const int epb = a.epb;
    const int le = threadIdx.x / NP, q = threadIdx.x - le * NP;
    const uint32_t e = blockIdx.x * epb + le;
    const bool active = le < epb && e < d.nelem;
    extern __shared__ double lds_v[];
    constexpr bool PL = vjp_prims_fit(NXc, NYc, NZc);
    const int per_elem = (PL ? 17 : 10) * NP + 5 * NFN;
    double* Wl = lds_v + (size_t)le * per_elem;  // w / cV      [5][NP]
    double* Ql = Wl + 5 * NP;                    // Q = q + din [5][NP]
    double* Xl = Ql + 5 * NP;                    // PL: u, theta', p', c, |U| [7][NP]
    double* FL = Xl + (PL ? 7 : 0) * NP;         // face terms  [5][NFN]
    double* Dl = lds_v + (size_t)epb * per_elem;
    const int i = q / (NY * NZ), j = (q / NZ) % NY, k = q % NZ;
    const uint64_t n = (uint64_t)e * NP + q;
    const uint64_t nb = d.nb;
    const double T0 = d.T0, gam = d.gamma, gR = d.gamma * d.R;

    for (int t = threadIdx.x; t < NX * NX + NY * NY + NZ * NZ; t += blockDim.x) {
        const int s1 = NX * NX, s2 = s1 + NY * NY;
        Dl[t] = t < s1 ? d.D[0][t] : t < s2 ? d.D[1][t - s1] : d.D[2][t - s2];
    }
    // phase 1: w / cV and Q = q + din to LDS (w itself kept for the
    // epilogue)
    double wn[5] = {0, 0, 0, 0, 0};
    if (active) {
        const double cvn = ldn<O32>(d.cV, n);
        double Qn[5];
        for (int v = 0; v < 5; v++) {
            wn[v] = ldn<O32>(a.w[v], n);
            Wl[v * NP + q] = wn[v] / cvn;
            Qn[v] = ldn<O32>(a.q[v], n) + (DIN ? ldn<O32>(a.din[v], n) : 0.0);
            Ql[v * NP + q] = Qn[v];
        }
        if constexpr (PL) {
            Prim<double> P;
            P.rp = Qn[0]; P.m[0] = Qn[1]; P.m[1] = Qn[2]; P.m[2] = Qn[3]; P.tp = Qn[4];
            make_prim(P, ldn<O32>(d.rho_ref, n), ldn<O32>(d.p_ref, n), T0, gam, gR, d.ppc);
            const double x[7] = {P.u[0], P.u[1], P.u[2], P.thp, P.pp, P.c, P.umag};
            for (int c = 0; c < 7; c++) Xl[c * NP + q] = x[c];
        }
    }
    __syncthreads();

    // phase 2: the face terms of this element's face nodes
    if (active) {
        for (int t = q; t < NFN; t += NP) {
            int g, fa, fb, fnode;
            face_task(t, NX, NY, NZ, g, fa, fb, fnode);
            const EFace ef = d.eface[(size_t)e * 6 + g];
            double contrib[5] = {0, 0, 0, 0, 0};
            // 2:1 mortar sides: their share comes from mortar_vjp_kernel
            // (madj); ghost sides: from ghost_vjp_kernel (gadj)
            if (((ef.meta >> 4) & 1) && !((ef.meta >> 6) & 1)) {
                const bool own = (ef.meta >> 3) & 1;
                const bool ghost = (ef.meta >> 5) & 1;
                if (!ghost) {
                    // interior face: the hand-written Rusanov VJP
                    const int og = ef.meta & 7;
                    const double wgt = face_weight(d, own ? g : og, fa, fb);
                    const uint64_t na = (uint64_t)e * NP + fnode;
                    const double rrA = ldn<O32>(d.rho_ref, na), prA = ldn<O32>(d.p_ref, na);
                    const double nvd[3] = {d.gFN[3 * (size_t)ef.facet] * wgt, d.gFN[3 * (size_t)ef.facet + 1] * wgt,
                                           d.gFN[3 * (size_t)ef.facet + 2] * wgt};
                    const uint64_t m = face_node(ef.other, og, fa, fb, NX, NY, NZ);
                    const double cvm = ldn<O32>(d.cV, m);
                    Prim<double> Ap, Bp;
                    load_prim<PL>(Ap, Ql, Xl, NP, fnode, rrA, prA, T0, gam, gR, d.ppc);
                    Bp.rp = ldn<O32>(a.q[0], m) + (DIN ? ldn<O32>(a.din[0], m) : 0.0);
                    Bp.m[0] = ldn<O32>(a.q[1], m) + (DIN ? ldn<O32>(a.din[1], m) : 0.0);
                    Bp.m[1] = ldn<O32>(a.q[2], m) + (DIN ? ldn<O32>(a.din[2], m) : 0.0);
                    Bp.m[2] = ldn<O32>(a.q[3], m) + (DIN ? ldn<O32>(a.din[3], m) : 0.0);
                    Bp.tp = ldn<O32>(a.q[4], m) + (DIN ? ldn<O32>(a.din[4], m) : 0.0);
                    make_prim(Bp, ldn<O32>(d.rho_ref, m), ldn<O32>(d.p_ref, m), T0, gam, gR, d.ppc);
                    const double s = own ? -1.0 : 1.0;
                    double coef[5];
                    for (int v = 0; v < 5; v++) coef[v] = s * (Wl[v * NP + fnode] - ldn<O32>(a.w[v], m) / cvm);
                    if (own) rusanov_vjp_side(Ap, Bp, nvd, 0.5, T0, gam, gR, coef, true, rrA, prA, contrib);
                    else rusanov_vjp_side(Bp, Ap, nvd, 0.5, T0, gam, gR, coef, false, rrA, prA, contrib);
                } else {
                    const double* ga = d.gadj + (size_t)d.ghost_of[(size_t)e * 6 + g] * 5 * d.npfm +
                                       fa * (g < 2 ? NY : NZ) + fb;
                    for (int v = 0; v < 5; v++) contrib[v] = ga[v * d.npfm];
                }
            }
            for (int v = 0; v < 5; v++) FL[v * NFN + t] = contrib[v];
        }
    }
    __syncthreads();
    if (!active) return;

    // phase 3: volume term, its flux Jacobian, buoyancy, then the face terms
    // of the faces through this node; the flux-row cotangents are consumed
    // row by row (rows rho, Theta, then momentum, the order in which each
    // accumulator of flux_rows_vjp sums them)
    double qbar[5] = {0, 0, 0, 0, 0};
    {
        double J[9];
        if (d.affine) {
            // the forward stage's per-element metric (JcV = Jinv_e (vol_e W))
            const double* em = d.emet + (size_t)e * 10;
            const double cv64 = em[9] * d.wnode[q];
            for (int c = 0; c < 9; c++) J[c] = em[c] * cv64;
        } else {
            for (int c = 0; c < 9; c++) J[c] = ldn<O32>(d.JcV + (size_t)c * nb, n);
        }
        const double rref = ldn<O32>(d.rho_ref, n), pref = ldn<O32>(d.p_ref, n);
        Prim<double> P;
        load_prim<PL>(P, Ql, Xl, NP, q, rref, pref, T0, gam, gR, d.ppc);
        const double* D0 = Dl;
        const double* Dy = Dl + NX * NX;
        const double* Dz = Dy + NY * NY;
        double ub[3] = {0, 0, 0}, thb = 0, ppb = 0;
#pragma unroll 1
        for (int oi = 0; oi < 5; oi++) {
            const int row = oi == 0 ? 0 : (oi == 1 ? 4 : oi - 1);
            const double* W = Wl + row * NP;
            double h0 = 0, h1 = 0, h2 = 0;
            for (int t = 0; t < NX; t++) h0 += D0[i * NX + t] * W[t * NY * NZ + j * NZ + k];
            for (int t = 0; t < NY; t++) h1 += Dy[j * NY + t] * W[i * NY * NZ + t * NZ + k];
            for (int t = 0; t < NZ; t++) h2 += Dz[k * NZ + t] * W[i * NY * NZ + j * NZ + t];
            const double fb[3] = {h0 * J[0] + h1 * J[3] + h2 * J[5], h0 * J[6] + h1 * J[1] + h2 * J[4],
                                  h0 * J[8] + h1 * J[7] + h2 * J[2]};
            if (row == 0) {                                // F_rho = m
                for (int c = 0; c < 3; c++) qbar[1 + c] += fb[c];
            } else if (row == 4) {                         // F_Theta = m theta
                for (int c = 0; c < 3; c++) {
                    qbar[1 + c] += fb[c] * (T0 + P.thp);
                    thb += fb[c] * P.m[c];
                }
            } else {                                       // F_m_r = m_r u + p' e_r
                // (component r selected, not indexed: a run-time index
                // would put qbar in scratch memory)
                const int r = row - 1;
                const double fr = r == 0 ? fb[0] : (r == 1 ? fb[1] : fb[2]);
                const double mr = r == 0 ? P.m[0] : (r == 1 ? P.m[1] : P.m[2]);
                double qr = r == 0 ? qbar[1] : (r == 1 ? qbar[2] : qbar[3]);
                ppb += fr;
                for (int c = 0; c < 3; c++) {
                    qr += fb[c] * P.u[c];
                    ub[c] += fb[c] * mr;
                }
                if (r == 0) qbar[1] = qr;
                else if (r == 1) qbar[2] = qr;
                else qbar[3] = qr;
            }
        }
        prim_to_conserved(P, rref, pref, T0, gam, ub, thb, ppb, qbar);
    }
    if (d.buoyancy) qbar[0] += wn[1] * d.g[0] + wn[2] * d.g[1] + wn[3] * d.g[2];
    for (int g = 0; g < 6; g++) {
        if (!on_face(g, i, j, k, NX, NY, NZ)) continue;
        if (!((d.eface[(size_t)e * 6 + g].meta >> 4) & 1)) continue;
        int fa, fb;
        face_ab(g, i, j, k, fa, fb);
        const int t = face_task_index(g, fa, fb, NX, NY, NZ);
        for (int v = 0; v < 5; v++) qbar[v] += FL[v * NFN + t];
    }
    if (d.nmort) {
        // the mortar faces' share (mortar_vjp_kernel) sits only on nodes of
        // elements with a 2:1 side; read it there and clear it for the next
        // launch (madj is zero between launches)
        bool mside = false;
        for (int g = 0; g < 6; g++) mside = mside || ((d.eface[(size_t)e * 6 + g].meta >> 6) & 1);
        if (mside)
            for (int v = 0; v < 5; v++) {
                qbar[v] += ldn<O32>(d.madj + (size_t)v * nb, n);
                stv<O32>(d.madj + (size_t)v * nb, n, 0.0);
            }
    }
    for (int v = 0; v < 5; v++) {
        if (a.t[0]) stv<O32>(a.t[v], n, qbar[v]);
        if (a.ad_out[0]) stv<O32>(a.ad_out[v], n, a.c1 * (wn[v] + a.dt * qbar[v]));
        if (a.comb_out[0])
            stv<O32>(a.comb_out[v], n,
                     a.k0 * ldn<O32>(a.y0[v], n) + a.k1 * ldn<O32>(a.y1[v], n) + wn[v] + a.dt * qbar[v]);
    }
}

// Boundary (ghost) sides of the vector-Jacobian product, ahead of the
// element kernel: one work-group per ghost side, five threads per face node
// (one per state component), the flux through the ghost state
// differentiated by a forward-mode sweep (ghost_face_vjp) and contracted
// with the element kernel's coefficient -w/cV, written to gadj for the
// element kernel to add in its face order.  Kept out of the element kernel,
// whose registers the sweeps would otherwise hold.
template <bool DIN>
__global__ void __launch_bounds__(512) ghost_vjp_kernel(DgDev d, VjpArgs a) {
    const int NX = d.npx, NY = d.npy, NZ = d.npz, NP = d.np;
    const int np = d.npfm;
    const int kk = threadIdx.x / np, t = threadIdx.x - kk * np;
    const uint32_t rec = d.ghost_rec[blockIdx.x];
    const uint32_t e = rec / 6;
    const int g = rec - e * 6;
    const int nbf = g < 2 ? NY : NZ, naf = g < 4 ? NX : NY;
    if (kk >= 5 || t >= naf * nbf) return;
    const double T0 = d.T0, gam = d.gamma, gR = d.gamma * d.R;
    const int fa = t / nbf, fb = t - fa * nbf;
    int i, j, k;
    if (g < 2) { i = fa; j = fb; k = (g & 1) ? NZ - 1 : 0; }
    else if (g < 4) { i = fa; j = (g & 1) ? NY - 1 : 0; k = fb; }
    else { i = (g & 1) ? NX - 1 : 0; j = fa; k = fb; }
    const uint64_t na = (uint64_t)e * NP + (uint64_t)(i * NY * NZ + j * NZ + k);
    const EFace ef = d.eface[rec];
    const bool own = (ef.meta >> 3) & 1;
    const int og = ef.meta & 7;
    const double wgt = face_weight(d, own ? g : og, fa, fb);
    const double nvd[3] = {d.gFN[3 * (size_t)ef.facet] * wgt, d.gFN[3 * (size_t)ef.facet + 1] * wgt,
                           d.gFN[3 * (size_t)ef.facet + 2] * wgt};
    const double cvn = d.cV[na];
    double QA[5], coef[5];
    for (int v = 0; v < 5; v++) {
        QA[v] = a.q[v][na] + (DIN ? a.din[v][na] : 0.0);
        coef[v] = -(a.w[v][na] / cvn);
    }
    const BcTable& bt = *d.bc;
    const int pi = ef.meta >> 8;
    const double uv[3] = {bt.value[1][pi], bt.value[2][pi], bt.value[3][pi]};
    d.gadj[((size_t)blockIdx.x * 5 + kk) * np + t] =
        ghost_face_vjp(QA, nvd, bt.kind[1][pi], uv, d.rho_ref[na], d.p_ref[na], T0, gam, gR, coef, kk);
}

// Vector-Jacobian product of the mortar faces' share of the residual (the
// method-of-lines pre-pass of dg_rk.hip: mortar_flux_kernel and
// mortar_gather_kernel).  There the refined node gains sf w_t phi_t and each
// coarse side node t' gains sc w_t' sum_t psiCor(t, t') phi_t (sf = -1 when
// the refined element owns the face, else +1; sc = -sf), phi_t the Rusanov
// flux per unit weight from the refined node's values and the coarse side's
// values interpolated with psiRef.  With w^ = w / cV the cotangent of phi_t
// is  sf w_t w^(refined) + sum_t' psiCor(t, t') sc w_t' w^(coarse t');  phi's
// partial derivatives (with respect to the refined node's state and to the
// interpolated values) are taken by forward-mode dual numbers through the
// same functions as the forward pass; the interpolated values' cotangents go
// back to the coarse nodes through psiRef transposed, and through the node
// values' derivatives to the coarse states.  Contributions are added into
// d.madj [5][nb] (a node can sit on several mortar faces), which the VJP
// kernel adds to its own.  One work-group per mortar face; per face node
// (refined, and in the last phase coarse) kMortarSub threads, thread s of a
// node taking state component k = s and interpolated values j = s, s + 5,
// s + 10, so the 17 (refined) and 5 (coarse) forward-mode sweeps of a node
// run side by side.
constexpr int kMortarSub = 5;
__global__ void __launch_bounds__(512) mortar_vjp_kernel(DgDev d, VjpArgs a) {
    const uint32_t f = blockIdx.x;
    const int np = d.npfm;
    const int sub = threadIdx.x / np, t = threadIdx.x - sub * np;
    const MortarFace mf = d.mort[f];
    const bool act = sub < kMortarSub && t < mf.n0 * mf.n1;
    const int NX = d.npx, NY = d.npy, NZ = d.npz;
    const int ao = t / mf.n1, bo = t - ao * mf.n1;
    const size_t ft = (size_t)f * np + t;
    const uint64_t nb = d.nb;
    const double T0 = d.T0, gam = d.gamma, gR = d.gamma * d.R;
    const double nv[3] = {mf.nf[0], mf.nf[1], mf.nf[2]};
    const D1 nvD[3] = {D1(nv[0]), D1(nv[1]), D1(nv[2])};
    const bool fo = mf.fine_owner != 0;
    const int d0 = mf.side < 4 ? 0 : 1, d1 = mf.side < 2 ? 1 : 2;
    const double* R0 = d.psiRef[d0] + (size_t)mf.h0 * mf.n0 * mf.n0;
    const double* R1 = d.psiRef[d1] + (size_t)mf.h1 * mf.n1 * mf.n1;
    const double* C0 = d.psiCor[d0] + (size_t)mf.h0 * mf.n0 * mf.n0;
    const double* C1 = d.psiCor[d1] + (size_t)mf.h1 * mf.n1 * mf.n1;
    auto Q = [&](uint64_t n, int v) { return a.q[v][n] + (a.din[v] ? a.din[v][n] : 0.0); };
    // a node's primitives (kept in LDS rows PL[12][np]) and its 12 values
    auto values = [&](uint64_t n, double* v, double* PL) {
        Prim<double> P;
        P.rp = Q(n, 0); P.m[0] = Q(n, 1); P.m[1] = Q(n, 2); P.m[2] = Q(n, 3); P.tp = Q(n, 4);
        make_prim(P, d.rho_ref[n], d.p_ref[n], T0, gam, gR, d.ppc);
        mortar_node_values(P, nv, T0, v);
        const double x[12] = {P.rp, P.m[0], P.m[1], P.m[2], P.tp, P.u[0], P.u[1], P.u[2], P.thp, P.pp, P.c, P.umag};
        for (int c = 0; c < 12; c++) PL[c * np + t] = x[c];
    };
    // the node values' derivatives with respect to the node's state k, from
    // the primitives in PL
    auto values_d = [&](uint64_t n, int k, D1* v, const double* PL) {
        Prim<double> P;
        P.rp = PL[t]; P.m[0] = PL[np + t]; P.m[1] = PL[2 * np + t]; P.m[2] = PL[3 * np + t]; P.tp = PL[4 * np + t];
        P.u[0] = PL[5 * np + t]; P.u[1] = PL[6 * np + t]; P.u[2] = PL[7 * np + t]; P.thp = PL[8 * np + t];
        P.pp = PL[9 * np + t]; P.c = PL[10 * np + t]; P.umag = PL[11 * np + t];
        Prim<D1> T;
        prim_tangent(P, d.rho_ref[n], d.p_ref[n], T0, gam, gR, k, T);
        mortar_node_values(T, nvD, D1(T0), v);
    };
    extern __shared__ double mortar_adj_lds[];
    double* Lv = mortar_adj_lds;   // [12][np] coarse node values
    double* Lw = Lv + 12 * np;     // [5][np]  sc w_t' w^(coarse node t')
    double* Lp = Lw + 5 * np;      // [12][np] cotangents of the interpolated values
    double* Lf = Lp + 12 * np;     // [12][np] refined node values
    double* Lc = Lf + 12 * np;     // [12][np] interpolated coarse values; then gbar
    double* Lb = Lc + 12 * np;     // [5][np]  cotangent of phi
    double* Pc = Lb + 5 * np;      // [12][np] primitives of the coarse side nodes
    double* Pf = Pc + 12 * np;     // [12][np] primitives of the refined nodes
    double* LR0 = Pf + 12 * np;    // psiRef and psiCor of the face's half, n0^2 and n1^2 each
    double* LR1 = LR0 + mf.n0 * mf.n0;
    double* LC0 = LR1 + mf.n1 * mf.n1;
    double* LC1 = LC0 + mf.n0 * mf.n0;
    for (int r = threadIdx.x; r < mf.n0 * mf.n0; r += blockDim.x) { LR0[r] = R0[r]; LC0[r] = C0[r]; }
    for (int r = threadIdx.x; r < mf.n1 * mf.n1; r += blockDim.x) { LR1[r] = R1[r]; LC1[r] = C1[r]; }
    const uint64_t nc = face_node(mf.coarse, mf.side, ao, bo, NX, NY, NZ);
    if (act && sub == 0) {
        double v[12];
        values(nc, v, Pc);
        for (int q = 0; q < 12; q++) Lv[q * np + t] = v[q];
        const double sc = fo ? d.mort_w[ft] : -d.mort_w[ft];
        const double icv = 1.0 / d.cV[nc];
        for (int v5 = 0; v5 < 5; v5++) Lw[v5 * np + t] = sc * (a.w[v5][nc] * icv);
    }
    __syncthreads();
    const uint64_t nf = act ? d.mort_fine[ft] : 0;
    if (act) {
        if (sub == 0) {
            double vf[12];
            values(nf, vf, Pf);
            for (int q = 0; q < 12; q++) Lf[q * np + t] = vf[q];
        }
        // interpolated values q = sub, sub + 5, sub + 10
        double vc[3] = {0.0, 0.0, 0.0};
        for (int an = 0; an < mf.n0; an++)
            for (int bn = 0; bn < mf.n1; bn++) {
                const double w = LR0[an * mf.n0 + ao] * LR1[bn * mf.n1 + bo];
                for (int i = 0; i < 3; i++)
                    if (sub + kMortarSub * i < 12) vc[i] += w * Lv[(sub + kMortarSub * i) * np + an * mf.n1 + bn];
            }
        for (int i = 0; i < 3; i++)
            if (sub + kMortarSub * i < 12) Lc[(sub + kMortarSub * i) * np + t] = vc[i];
        // cotangent of phi_t, component sub
        const double sf = fo ? -d.mort_w[ft] : d.mort_w[ft];
        const double icvf = 1.0 / d.cV[nf];
        double phib = sf * (a.w[sub][nf] * icvf);
        for (int an = 0; an < mf.n0; an++)
            for (int bn = 0; bn < mf.n1; bn++) {
                const double c = LC0[ao * mf.n0 + an] * LC1[bo * mf.n1 + bn];
                phib += c * Lw[sub * np + an * mf.n1 + bn];
            }
        Lb[sub * np + t] = phib;
    }
    __syncthreads();
    if (act) {
        const D1 bg(d.mort_bg[ft]);
        double phib[5];
        for (int v5 = 0; v5 < 5; v5++) phib[v5] = Lb[v5 * np + t];
        {
            // the refined node's state component k = sub
            const int k = sub;
            D1 vfD[12], vcD[12], phi[5];
            values_d(nf, k, vfD, Pf);
            for (int q = 0; q < 12; q++) vcD[q] = D1(Lc[q * np + t]);
            mortar_rusanov(vfD, vcD, bg, nvD, fo, D1(T0), phi);
            double acc = 0;
            for (int v5 = 0; v5 < 5; v5++) acc += phib[v5] * phi[v5].d[0];
            atomicAdd(d.madj + (size_t)k * nb + nf, acc);
        }
        // the interpolated values j = sub, sub + 5, sub + 10
#pragma nounroll
        for (int j = sub; j < 12; j += kMortarSub) {
            D1 vfD[12], vcD[12], phi[5];
            for (int q = 0; q < 12; q++) {
                vfD[q] = D1(Lf[q * np + t]);
                vcD[q] = seeded(Lc[q * np + t], q == j);
            }
            mortar_rusanov(vfD, vcD, bg, nvD, fo, D1(T0), phi);
            double acc = 0;
            for (int v5 = 0; v5 < 5; v5++) acc += phib[v5] * phi[v5].d[0];
            Lp[j * np + t] = acc;
        }
    }
    __syncthreads();
    // coarse side node (ao, bo): psiRef transposed, then the node values'
    // derivatives
    if (act) {
        double gb[3] = {0.0, 0.0, 0.0};
        for (int a2 = 0; a2 < mf.n0; a2++)
            for (int b2 = 0; b2 < mf.n1; b2++) {
                const double r = LR0[ao * mf.n0 + a2] * LR1[bo * mf.n1 + b2];
                for (int i = 0; i < 3; i++)
                    if (sub + kMortarSub * i < 12) gb[i] += r * Lp[(sub + kMortarSub * i) * np + a2 * mf.n1 + b2];
            }
        for (int i = 0; i < 3; i++)
            if (sub + kMortarSub * i < 12) Lc[(sub + kMortarSub * i) * np + t] = gb[i];
    }
    __syncthreads();
    if (!act) return;
    {
        const int k = sub;
        D1 v[12];
        values_d(nc, k, v, Pc);
        double acc = 0;
        for (int q = 0; q < 12; q++) acc += Lc[q * np + t] * v[q].d[0];
        atomicAdd(d.madj + (size_t)k * nb + nc, acc);
    }
}

int vjp_epb(int np) { return np >= 256 ? 1 : 256 / np; }

template <int X, int Y, int Z>
hipError_t launch_vjp_t(const DgDev& d, const VjpArgs& a_in, hipStream_t st) {
    VjpArgs a = a_in;
    const int NP = d.np;
    constexpr bool pl = vjp_prims_fit(X, Y, Z);
    // small meshes: fewer elements per work-group, so that every CU gets
    // about four work-groups
    int epb = std::max(1, std::min(vjp_epb(NP), (int)(d.nelem / 1024)));
    while (epb > 1 && rk_vjp_lds_bytes(d.npx, d.npy, d.npz, epb, pl) > 65536) epb--;
    const int threads = (epb * NP + 63) / 64 * 64;
    const unsigned blocks = (d.nelem + epb - 1) / epb;
    const size_t lds = rk_vjp_lds_bytes(d.npx, d.npy, d.npz, epb, pl);
    if (NP > 512 || lds > 65536) return hipErrorInvalidValue;
    a.epb = epb;
    // node byte offsets below 4 GiB (each array is addressed from its own
    // base, JcV and madj from the base of each component's slice)
    const bool o32 = d.nb * sizeof(double) < (1ull << 32);
    if (a.din[0]) {
        if (o32) hipLaunchKernelGGL((rk_vjp_kernel<X, Y, Z, true, true>), dim3(blocks), dim3(threads), lds, st, d, a);
        else hipLaunchKernelGGL((rk_vjp_kernel<X, Y, Z, true, false>), dim3(blocks), dim3(threads), lds, st, d, a);
    } else {
        if (o32) hipLaunchKernelGGL((rk_vjp_kernel<X, Y, Z, false, true>), dim3(blocks), dim3(threads), lds, st, d, a);
        else hipLaunchKernelGGL((rk_vjp_kernel<X, Y, Z, false, false>), dim3(blocks), dim3(threads), lds, st, d, a);
    }
    return hipGetLastError();
}

}  // namespace

size_t rk_vjp_lds_bytes(int nx, int ny, int nz, int epb, bool pl) {
    const int np = nx * ny * nz, nfn = 2 * (nx * ny + nx * nz + ny * nz);
    return ((size_t)epb * ((pl ? 17 : 10) * np + 5 * nfn) + nx * nx + ny * ny + nz * nz) * sizeof(double);
}

hipError_t launch_rk_vjp(const DgDev& d, const VjpArgs& a, hipStream_t st) {
    // madj is zero on entry (cleared when the context is made, and by each
    // element VJP launch after reading it); a failed launch clears it
    auto fail = [&](hipError_t e) {
        if (d.nmort) (void)hipMemsetAsync(d.madj, 0, 5 * d.nb * sizeof(double), st);
        return e;
    };
    if (d.nmort) {
        if (kMortarSub * d.npfm > 512) return hipErrorInvalidValue;
        hipError_t e;
        const int threads = (kMortarSub * d.npfm + 63) / 64 * 64;
        // LDS: 82 rows of npfm values and the four 1-D matrices
        const size_t lds = ((size_t)82 * d.npfm + 2 * (d.npx * d.npx + d.npy * d.npy + d.npz * d.npz)) * sizeof(double);
        hipLaunchKernelGGL(mortar_vjp_kernel, dim3(d.nmort), dim3(threads), lds, st, d, a);
        if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    }
    if (d.nghost) {
        if (5 * d.npfm > 512) return fail(hipErrorInvalidValue);
        const int gthreads = (5 * d.npfm + 63) / 64 * 64;
        if (a.din[0]) hipLaunchKernelGGL(ghost_vjp_kernel<true>, dim3(d.nghost), dim3(gthreads), 0, st, d, a);
        else hipLaunchKernelGGL(ghost_vjp_kernel<false>, dim3(d.nghost), dim3(gthreads), 0, st, d, a);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(e);
    }
    hipError_t e;
    if (d.npx == 8 && d.npy == 8 && d.npz == 8) e = launch_vjp_t<8, 8, 8>(d, a, st);
    else if (d.npx == 6 && d.npy == 6 && d.npz == 6) e = launch_vjp_t<6, 6, 6>(d, a, st);
    else if (d.npx == 6 && d.npy == 1 && d.npz == 6) e = launch_vjp_t<6, 1, 6>(d, a, st);
    else if (d.npx == 5 && d.npy == 1 && d.npz == 5) e = launch_vjp_t<5, 1, 5>(d, a, st);
    else e = launch_vjp_t<0, 0, 0>(d, a, st);
    return e == hipSuccess ? e : fail(e);
}

}  // namespace nsem

using namespace nsem;

namespace {

// Five FP64 node arrays: a view into the handle's workspace.
struct Field5 {
    double* b[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    double* p(int v) const { return b[v]; }
};

// n views of 5 node arrays each, in the handle's workspace (allocated once
// and grown when a call needs more, so repeated calls allocate nothing).
hipError_t workspace(nsem_dg* d, size_t n, std::vector<Field5>& out) {
    const size_t need = n * 5 * d->nb * sizeof(double);
    if (d->adj_ws.bytes < need) {
        if (d->adj_ws.p) {
            hipError_t e = hipStreamSynchronize(d->stream);
            if (e != hipSuccess) return e;
            (void)hipFree(d->adj_ws.p);
            d->adj_ws.p = nullptr;
            d->adj_ws.bytes = 0;
        }
        hipError_t e = d->adj_ws.alloc(need);
        if (e != hipSuccess) return e;
    }
    out.resize(n);
    double* base = d->adj_ws.as<double>();
    for (size_t k = 0; k < n; k++)
        for (int v = 0; v < 5; v++) out[k].b[v] = base + (k * 5 + v) * d->nb;
    return hipSuccess;
}

hipError_t upload5(const Field5& f, const double* host, uint64_t nb, hipStream_t s) {
    for (int v = 0; v < 5; v++) {
        hipError_t e = hipMemcpyAsync(f.p(v), host + (size_t)v * nb, nb * sizeof(double), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
hipError_t download5(double* host, const Field5& f, uint64_t nb, hipStream_t s) {
    for (int v = 0; v < 5; v++) {
        hipError_t e = hipMemcpyAsync(host + (size_t)v * nb, f.p(v), nb * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return e;
    }
    return hipStreamSynchronize(s);
}
hipError_t copy5(const Field5& dst, const double* const src[5], uint64_t nb, hipStream_t s) {
    for (int v = 0; v < 5; v++) {
        hipError_t e = hipMemcpyAsync(dst.p(v), src[v], nb * sizeof(double), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

constexpr double kB[3] = {1.0, 0.25, 2.0 / 3.0};

// One FP64 forward stage from q (+din) into dout (or qout when last).
hipError_t fwd_stage(nsem_dg* d, int stage, const double* const q[5], const double* const din[5],
                     double* const dout[5], double* const qout[5]) {
    RkArgs<double> a{};
    for (int v = 0; v < 5; v++) {
        a.q[v] = q[v];
        a.din[v] = din ? din[v] : nullptr;
        a.dout[v] = dout ? dout[v] : nullptr;
        a.qout[v] = qout ? qout[v] : nullptr;
    }
    a.b = kB[stage];
    a.bdt = kB[stage] * d->dev.dt;
    a.last = stage == 2;
    a.check = stage == 2;
    a.flag = d->flag.as<int>();
    return launch_rk_stage<double, double>(d->dev, a, d->stream);
}

bool all_finite(const double* v, uint64_t n) {
    for (uint64_t i = 0; i < n; i++)
        if (!std::isfinite(v[i])) return false;
    return true;
}

struct AdjointWork {
    Field5 ad1, ad2;
};

void ptrs(const Field5& f, double* out[5]) { for (int v = 0; v < 5; v++) out[v] = f.p(v); }
void cptrs(const Field5& f, const double* out[5]) { for (int v = 0; v < 5; v++) out[v] = f.p(v); }

// lam <- (dq+/dq)^T lam for the step from state q whose stage increments
// d1, d2 were stored by the forward recomputation.
int adjoint_step(nsem_dg* d, const double* const q[5], const double* const d1[5], const double* const d2[5],
                 Field5& lam, AdjointWork& w) {
    double *ad1[5], *ad2[5], *lp[5];
    ptrs(w.ad1, ad1); ptrs(w.ad2, ad2); ptrs(lam, lp);
    const double dt = d->dev.dt;
    auto vjp = [&](const double* const din[5], double* const wv[5], double* const ad_out[5], double c1,
                   bool last) -> hipError_t {
        VjpArgs a{};
        for (int v = 0; v < 5; v++) {
            a.q[v] = q[v]; a.din[v] = din ? din[v] : nullptr; a.w[v] = wv[v];
            a.ad_out[v] = ad_out ? ad_out[v] : nullptr;
            if (last) { a.y0[v] = lp[v]; a.y1[v] = ad2[v]; a.comb_out[v] = lp[v]; }
        }
        a.dt = dt; a.c1 = c1;
        a.k0 = 1.0 / 3.0; a.k1 = 0.75;
        return launch_rk_vjp(d->dev, a, d->stream);
    };
    // t3 = J(q+d2)^T lam: ad2 = 2/3 (lam + dt t3)
    HIP_TRY(vjp(d2, lp, ad2, 2.0 / 3.0, false));
    // t2 = J(q+d1)^T ad2: ad1 = 1/4 (ad2 + dt t2)
    HIP_TRY(vjp(d1, ad2, ad1, 0.25, false));
    // t1 = J(q)^T ad1: lam + 2/3 dt t3 + 1/4 dt t2 + dt t1
    //                = 1/3 lam + 3/4 ad2 + ad1 + dt t1
    HIP_TRY(vjp(nullptr, ad1, nullptr, 0.0, true));
    return NSEM_OK;
}

int check_adjoint_support(const nsem_dg* d) {
    if (int rc = nsem_mol_supported(d)) return rc;
    const DgDev& v = d->dev;
    if (v.np > 512 || rk_vjp_lds_bytes(v.npx, v.npy, v.npz, 1, false) > 65536)
        return nsem_fail(NSEM_EUNSUPPORTED, "element too large for the adjoint kernel");
    return NSEM_OK;
}

}  // namespace

extern "C" {

int nsem_dg_residual_vjp(nsem_dg* d, const double* w, double* t) {
    if (!d || !w || !t) return nsem_fail(NSEM_EINVAL, "null argument");
    int rc = check_adjoint_support(d);
    if (rc) return rc;
    if ((rc = nsem_dg_require_mol(d))) return rc;
    const uint64_t nb = d->nb;
    std::vector<Field5> ws;
    HIP_TRY(workspace(d, 2, ws));
    const Field5 &W = ws[0], &T = ws[1];
    HIP_TRY(upload5(W, w, nb, d->stream));
    VjpArgs a{};
    for (int v = 0; v < 5; v++) {
        a.q[v] = d->mq[d->mcur][v].as<double>(); a.din[v] = nullptr; a.w[v] = W.p(v); a.t[v] = T.p(v);
    }
    // (no epilogue arrays: t only)
    HIP_TRY(launch_rk_vjp(d->dev, a, d->stream));
    HIP_TRY(download5(t, T, nb, d->stream));
    return NSEM_OK;
}

int nsem_dg_residual_tlm(nsem_dg* d, const double* dq, double* L, double* dL) {
    if (!d || !dq || !L || !dL) return nsem_fail(NSEM_EINVAL, "null argument");
    if (int rc0 = nsem_mol_supported(d)) return rc0;
    if (int rc0 = nsem_dg_require_mol(d)) return rc0;
    const uint64_t nb = d->nb;
    std::vector<Field5> ws;
    HIP_TRY(workspace(d, 3, ws));
    const Field5 &TQ = ws[0], &Lout = ws[1], &TL = ws[2];
    HIP_TRY(upload5(TQ, dq, nb, d->stream));
    RkArgs<double> a{};
    for (int v = 0; v < 5; v++) {
        a.q[v] = d->mq[d->mcur][v].as<double>();
        a.dout[v] = Lout.p(v);
        a.tq[v] = TQ.p(v);
        a.tdout[v] = TL.p(v);
    }
    a.b = 1.0;
    a.bdt = 1.0;   // out = L
    a.last = 0;
    a.flag = d->flag.as<int>();
    hipError_t e = launch_rk_stage<Dual<1>, double>(d->dev, a, d->stream);
    if (e == hipErrorInvalidValue) return nsem_fail(NSEM_EUNSUPPORTED, "element too large for the tangent-linear kernel");
    HIP_TRY(e);
    HIP_TRY(download5(L, Lout, nb, d->stream));
    HIP_TRY(download5(dL, TL, nb, d->stream));
    return NSEM_OK;
}

int nsem_dg_rk3_tlm(nsem_dg* d, int32_t nsteps, double* dq) {
    if (!d || !dq || nsteps < 0) return nsem_fail(NSEM_EINVAL, "bad argument");
    if (int rc0 = nsem_mol_supported(d)) return rc0;
    if (int rc0 = nsem_dg_require_mol(d)) return rc0;
    const uint64_t nb = d->nb;
    std::vector<Field5> ws;
    HIP_TRY(workspace(d, 4, ws));
    const Field5 tq[2] = {ws[0], ws[1]}, td[2] = {ws[2], ws[3]};
    HIP_TRY(upload5(tq[0], dq, nb, d->stream));
    int tc = 0;
    for (int32_t s = 0; s < nsteps; s++) {
        DevBuf* qa = d->mq[d->mcur];
        DevBuf* qb = d->mq[1 - d->mcur];
        for (int stage = 0; stage < 3; stage++) {
            RkArgs<double> a{};
            for (int v = 0; v < 5; v++) {
                a.q[v] = qa[v].as<double>();
                a.din[v] = stage == 0 ? nullptr : d->d64[(stage - 1) & 1][v].as<double>();
                a.dout[v] = d->d64[stage & 1][v].as<double>();
                a.qout[v] = qb[v].as<double>();
                a.tq[v] = tq[tc].p(v);
                a.tdin[v] = stage == 0 ? nullptr : td[(stage - 1) & 1].p(v);
                a.tdout[v] = td[stage & 1].p(v);
                a.tqout[v] = tq[1 - tc].p(v);
            }
            a.b = kB[stage];
            a.bdt = kB[stage] * d->dev.dt;
            a.last = stage == 2;
            a.check = stage == 2;
            a.flag = d->flag.as<int>();
            hipError_t e = launch_rk_stage<Dual<1>, double>(d->dev, a, d->stream);
            if (e == hipErrorInvalidValue)
                return nsem_fail(NSEM_EUNSUPPORTED, "element too large for the tangent-linear kernel");
            HIP_TRY(e);
        }
        d->mcur = 1 - d->mcur;
        tc = 1 - tc;
    }
    HIP_TRY(download5(dq, tq[tc], nb, d->stream));
    if (nsteps > 0) {
        d->mol_valid = true;
        d->split_valid = false;
    }
    if (int rc0 = nsem_dg_take_flag(d, "non-finite state in the tangent-linear steps")) return rc0;
    if (!all_finite(dq, 5 * nb)) return nsem_fail(NSEM_ENONFINITE, "non-finite tangent");
    return NSEM_OK;
}

}  // extern "C"

// The adjoint sweep; with `seconds`, the device time from after the upload
// of lam to before the download of the gradient (HIP events on the stream).
static int rk3_adjoint(nsem_dg* d, int32_t nsteps, int32_t checkpoint, double* lam, double* seconds) {
    if (!d || !lam || nsteps < 0 || checkpoint < 1) return nsem_fail(NSEM_EINVAL, "bad argument");
    int rc = check_adjoint_support(d);
    if (rc) return rc;
    if ((rc = nsem_dg_require_mol(d))) return rc;
    const uint64_t nb = d->nb;
    const int nseg = (nsteps + checkpoint - 1) / checkpoint;
    // checkpoints; one segment's states and each step's stage increments
    // d1, d2; the adjoint and its stage work arrays
    std::vector<Field5> ws;
    HIP_TRY(workspace(d, (size_t)nseg + 3 * (size_t)checkpoint + 3, ws));
    const Field5* ck = ws.data();
    const Field5* seg = ck + nseg;
    const Field5* sd1 = seg + checkpoint;
    const Field5* sd2 = sd1 + checkpoint;
    Field5 L = sd2[checkpoint];
    AdjointWork w{sd2[checkpoint + 1], sd2[checkpoint + 2]};
    HIP_TRY(upload5(L, lam, nb, d->stream));
    if (seconds) HIP_TRY(hipEventRecord(d->ev0, d->stream));
    // one segment forward from its first state seg[0], recording each step's
    // state and stage increments; the state after the segment goes to `last`
    // (if given)
    auto record = [&](int s0, int s1, double* const last[5]) -> int {
        for (int s = s0; s < s1; s++) {
            const double* q[5];
            double *d1[5], *d2[5];
            cptrs(seg[s - s0], q); ptrs(sd1[s - s0], d1); ptrs(sd2[s - s0], d2);
            HIP_TRY(fwd_stage(d, 0, q, nullptr, d1, nullptr));
            HIP_TRY(fwd_stage(d, 1, q, d1, d2, nullptr));
            if (s + 1 < s1) {
                double* qn[5];
                ptrs(seg[s - s0 + 1], qn);
                HIP_TRY(fwd_stage(d, 2, q, d2, nullptr, qn));
            } else if (last) {
                HIP_TRY(fwd_stage(d, 2, q, d2, nullptr, last));
            }
        }
        return NSEM_OK;
    };
    // forward sweep: checkpoints at the start of every segment; the last
    // segment is recorded as it runs, so it needs no recomputation
    const int last0 = (nseg - 1) * checkpoint;
    for (int s = 0; s < last0; s++) {
        if (s % checkpoint == 0) {
            const double* src[5];
            for (int v = 0; v < 5; v++) src[v] = d->mq[d->mcur][v].as<double>();
            HIP_TRY(copy5(ck[s / checkpoint], src, nb, d->stream));
        }
        if ((rc = nsem_rk3_fp64_queue(d, 1))) return rc;
    }
    if (nsteps > 0) {
        const double* src[5];
        double* fin[5];
        for (int v = 0; v < 5; v++) {
            src[v] = d->mq[d->mcur][v].as<double>();
            fin[v] = d->mq[1 - d->mcur][v].as<double>();
        }
        HIP_TRY(copy5(seg[0], src, nb, d->stream));
        if ((rc = record(last0, nsteps, fin))) return rc;
        d->mcur = 1 - d->mcur;
        d->mol_valid = true;
        d->split_valid = false;
    }
    // backward sweep, segment by segment (recomputed from its checkpoint
    // except the last, recorded above)
    for (int sg = nseg - 1; sg >= 0; sg--) {
        const int s0 = sg * checkpoint, s1 = std::min(nsteps, s0 + checkpoint);
        if (sg < nseg - 1) {
            const double* src[5];
            cptrs(ck[sg], src);
            HIP_TRY(copy5(seg[0], src, nb, d->stream));
            if ((rc = record(s0, s1, nullptr))) return rc;
        }
        for (int s = s1 - 1; s >= s0; s--) {
            const double *q[5], *d1[5], *d2[5];
            cptrs(seg[s - s0], q); cptrs(sd1[s - s0], d1); cptrs(sd2[s - s0], d2);
            if ((rc = adjoint_step(d, q, d1, d2, L, w))) return rc;
        }
    }
    if (seconds) {
        HIP_TRY(hipEventRecord(d->ev1, d->stream));
        HIP_TRY(hipEventSynchronize(d->ev1));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, d->ev0, d->ev1));
        *seconds = ms * 1e-3;
    }
    HIP_TRY(download5(lam, L, nb, d->stream));
    if ((rc = nsem_dg_take_flag(d, "non-finite state in the adjoint's forward sweep"))) return rc;
    if (!all_finite(lam, 5 * nb)) return nsem_fail(NSEM_ENONFINITE, "non-finite gradient");
    return NSEM_OK;
}

extern "C" {

int nsem_dg_rk3_adjoint(nsem_dg* d, int32_t nsteps, int32_t checkpoint, double* lam) {
    return rk3_adjoint(d, nsteps, checkpoint, lam, nullptr);
}

int nsem_dg_time_rk3_adjoint(nsem_dg* d, int32_t nsteps, int32_t checkpoint, double* lam, double* seconds) {
    if (!seconds) return nsem_fail(NSEM_EINVAL, "null argument");
    return rk3_adjoint(d, nsteps, checkpoint, lam, seconds);
}

}  // extern "C"
