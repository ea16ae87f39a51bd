// The reference's split forward-Euler step of apps/euler/euler.cpp:179-287,
// as two fused element kernels (three with artificial viscosity):
//
//   split_grad: (diffusion YES) strong gradients of U^n and theta^n
//               (gradf<true>(., true)), stored per node
//   split_rho : rho^{n+1} = (r + rho^n ap) / ap,  ap = -cV/dt,
//               r = surface(Rusanov) - volume(weak) of F = U rho
//   split_ut  : p' = P0 (rho^{n+1} theta R / P0)^gamma - p_ref
//               U^{n+1}     from F = (rho^n U) (x) U + p' I - mu grad U,
//                           q = rho^{n+1} U, source (rho^{n+1} - rho_ref) g
//               theta^{n+1} from F = (rho^n U) theta - mu iPr grad theta,
//                           q = rho^{n+1} theta;  mu = rho^n visc (0 without
//                           diffusion)
//
// One work-group holds EPB elements (EPB*NP threads, one per node).  A node
// computes its own flux, writes its contravariant components to LDS, then
// applies the transposed 1-D derivative along its three GLL lines (the
// sum-factorised form of field.h:3443-3464), adds the Rusanov surface terms
// of the faces it lies on (field.h:2927-2943, 3085-3117) reading the
// neighbour's trace straight from global memory, and updates itself
// (field.h:3740-3920 with solve.cpp:563-570).  Ghost traces come from the
// boundary condition of the owner state, never from storage.
#include "dd_math.hpp"
#include "dg_kernels.hpp"
#include "dg_volume.hpp"

namespace nsem {

// ---------------------------------------------------------------------------
// LDS of the density and momentum/theta kernels: per element FR flux rows
// (3 x NP each), then per element Jin (9 x NP), then -- unless they fit in
// the flux rows, which are written only after the face terms are summed --
// the signed face fluxes of the face tasks (SFR values per task), then the
// three 1-D derivative matrices.
__host__ __device__ __forceinline__ int split_nfn(int NX, int NY, int NZ) {
    return 2 * (NX * NY + NX * NZ + NY * NZ);
}
__host__ __device__ __forceinline__ bool split_sf_overlaid(int epb, int NP, int NFN, int FR, int SFR) {
    return NFN * SFR <= FR * NP;
}
__host__ __device__ __forceinline__ size_t split_lds_doubles(int epb, int NP, int NFN, int FR, int SFR) {
    return (size_t)epb * (FR + 9) * NP + (split_sf_overlaid(epb, NP, NFN, FR, SFR) ? 0 : (size_t)epb * NFN * SFR);
}
// Bit g set: node (i, j, k) of element e lies on face g and that face takes
// a flux (the cell's face entry is active).
__device__ __forceinline__ unsigned split_node_faces(const DgDev& d, uint32_t e, int i, int j, int k, int NX, int NY,
                                                     int NZ) {
    unsigned fon = 0;
    for (int g = 0; g < 6; g++)
        if (on_face(g, i, j, k, NX, NY, NZ) && ((d.eface[(size_t)e * 6 + g].meta >> 4) & 1)) fon |= 1u << g;
    return fon;
}
// Stage D0, D1, D2 (NX^2, NY^2, NZ^2 values) in LDS at Dl.
__device__ __forceinline__ void split_stage_d(const DgDev& d, double* Dl, int NX, int NY, int NZ) {
    const int n0 = NX * NX, n1 = NY * NY, n2 = NZ * NZ;
    for (int t = threadIdx.x; t < n0 + n1 + n2; t += blockDim.x)
        Dl[t] = t < n0 ? d.D[0][t] : (t < n0 + n1 ? d.D[1][t - n0] : d.D[2][t - n0 - n1]);
}

// Rusanov flux of the density pass at one face node (field.h:2927-2943 with
// div_flux, field.h:3093-3114), signed as the node's element adds it: + on
// the owner side, - on the neighbour side; 2:1 mortar sides take the signed
// term of the pre-pass.  Own values are re-read from global memory (the
// work-group has just loaded them).
__device__ __forceinline__ bool split_rho_face(const DgDev& d, const SplitState& s, uint32_t e, int g, int a, int b,
                                               int node, int NX, int NY, int NZ, double& out) {
    const EFace ef = d.eface[(size_t)e * 6 + g];
    if (!((ef.meta >> 4) & 1)) return false;
    if ((ef.meta >> 6) & 1) {  // 2:1 mortar side: signed term from the pre-pass
        out = d.msurf[((size_t)(ef.meta >> 8) * d.npfm + a * (g < 2 ? NY : NZ) + b) * 5];
        return true;
    }
    const uint64_t n = (uint64_t)e * (NX * NY * NZ) + node;
    const double rho = s.rho[n], u0 = s.ux[n], u1 = s.uy[n], u2 = s.uz[n];
    const double th = s.T[n] + d.T0;
    const double gRT = d.gamma * d.R;
    const bool own = (ef.meta >> 3) & 1;
    const bool ghost = (ef.meta >> 5) & 1;
    const int og = ef.meta & 7;
    const double wgt = face_weight(d, own ? g : og, a, b);
    const double nx = d.gFN[3 * (size_t)ef.facet] * wgt;
    const double ny = d.gFN[3 * (size_t)ef.facet + 1] * wgt;
    const double nz = d.gFN[3 * (size_t)ef.facet + 2] * wgt;
    const double nv[3] = {nx, ny, nz};
    double orho, ou[3], oth;
    if (!ghost) {
        const uint64_t m = face_node(ef.other, og, a, b, NX, NY, NZ);
        orho = s.rho[m]; ou[0] = s.ux[m]; ou[1] = s.uy[m]; ou[2] = s.uz[m];
        oth = s.T[m] + d.T0;
    } else {
        // scalars: zero-gradient copies; velocity by its own condition
        const double uu[3] = {u0, u1, u2};
        ghost_velocity(*d.bc, ef.meta >> 8, uu, nv, ou);
        orho = rho;
        oth = th;
    }
    const double rO = own ? rho : orho, rN = own ? orho : rho;
    const double uO[3] = {own ? u0 : ou[0], own ? u1 : ou[1], own ? u2 : ou[2]};
    const double uN[3] = {own ? ou[0] : u0, own ? ou[1] : u1, own ? ou[2] : u2};
    const double thO = own ? th : oth, thN = own ? oth : th;
    const double fI = ghost ? 0.0 : 0.5;
    const double magO = sqrt(dot3r(uO[0], uO[1], uO[2], uO[0], uO[1], uO[2]));
    const double magN = sqrt(dot3r(uN[0], uN[1], uN[2], uN[0], uN[1], uN[2]));
    const double cO = sqrt(gRT * thO), cN = sqrt(gRT * thN);
    const double lam = ((magO * fI + magN * (1 - fI)) + (cO * fI + cN * (1 - fI))) / 2;
    const double mn = sqrt(dot3r(nx, ny, nz, nx, ny, nz));
    const double en[3] = {nx / mn, ny / mn, nz / mn};
    const double dq = lam * (rN - rO);
    double fF[3];
    for (int c = 0; c < 3; c++) fF[c] = (uO[c] * rO) * fI + (uN[c] * rN) * (1 - fI) - en[c] * dq;
    const double flux = dot3r(fF[0], fF[1], fF[2], nx, ny, nz);
    out = own ? flux : -flux;
    return true;
}

template <int NXc, int NYc, int NZc>
__global__ void __launch_bounds__(512) split_rho_kernel(DgDev d, SplitState s, double* __restrict__ rho_out,
                                                       double* __restrict__ p_out) {
    const int NX = NXc ? NXc : d.npx, NY = NYc ? NYc : d.npy, NZ = NZc ? NZc : d.npz;
    const int NP = NX * NY * NZ;
    const int NFN = split_nfn(NX, NY, NZ);
    const int epb = blockDim.x / NP;
    const int le = threadIdx.x / NP, q = threadIdx.x - le * NP;
    const uint32_t e = blockIdx.x * epb + le;
    const bool active = le < epb && e < d.nelem;
    extern __shared__ double lds[];
    double* F = lds + (size_t)le * 3 * NP;  // flux U rho (3 rows)
    double* J = lds + (size_t)epb * 3 * NP + (size_t)le * 9 * NP;
    double* SF = split_sf_overlaid(epb, NP, NFN, 3, 1) ? lds : lds + (size_t)epb * 12 * NP;
    double* Dl = lds + split_lds_doubles(epb, NP, NFN, 3, 1);
    split_stage_d(d, Dl, NX, NY, NZ);
    const int i = q / (NY * NZ), j = (q / NZ) % NY, k = q % NZ;
    const uint64_t n = (uint64_t)e * NP + q;

    double rho = 0, u0 = 0, u1 = 0, u2 = 0, th = 0, cvn = 0, pref = 0;
    unsigned fon = 0;
    if (active) {
        rho = s.rho[n]; u0 = s.ux[n]; u1 = s.uy[n]; u2 = s.uz[n];
        th = s.T[n] + d.T0;
        for (int c = 0; c < 9; c++) J[c * NP + q] = d.JcV[(size_t)kJinRow[c] * d.nb + n];
        // read with the state, before the first barrier, so that no global
        // load waits after it: cV, p_ref and the faces through the node
        cvn = d.cV[n];
        pref = d.p_ref[n];
        fon = split_node_faces(d, e, i, j, k, NX, NY, NZ);
    }
    // surface: one thread per face node of the work-group's elements (for
    // N = 7 one face per wavefront, so the face branch is uniform), each
    // node then adds the faces through it in the cell's face order
    // (div_flux, field.h:3093-3114)
    for (int t = threadIdx.x; t < epb * NFN; t += blockDim.x) {
        const int fl = t / NFN, task = t - fl * NFN;
        const uint32_t fe = blockIdx.x * epb + fl;
        if (fe >= d.nelem) break;
        int g, fa, fb, node;
        face_task(task, NX, NY, NZ, g, fa, fb, node);
        double v;
        if (split_rho_face(d, s, fe, g, fa, fb, node, NX, NY, NZ, v)) SF[t] = v;
    }
    __syncthreads();
    double surf = 0;
    if (active)
        for (int g = 0; g < 6; g++) {
            if (!((fon >> g) & 1)) continue;
            int a, b;
            face_ab(g, i, j, k, a, b);
            surf += SF[le * NFN + face_task_index(g, a, b, NX, NY, NZ)];
        }
    __syncthreads();
    if (active) {
        F[0 * NP + q] = u0 * rho;
        F[1 * NP + q] = u1 * rho;
        F[2 * NP + q] = u2 * rho;
    }
    __syncthreads();
    if (!active) return;
    double r[1] = {surf};
    volume_reference_order<1, 1>(r, F, J, NP, i, j, k, NX, NY, NZ, Dl, Dl + NX * NX, Dl + NX * NX + NY * NY);
    const double ap = (-1.0 / d.dt) * cvn;
    const double rho1 = (r[0] + rho * ap) / ap;
    rho_out[n] = rho1;
    // p = P0 (rho T R / P0)^gamma with the new rho, then p - p_ref
    // (euler.cpp:210-212), once per node for the momentum and theta pass
    p_out[n] = d.P0 * dd::pow_dd((rho1 * th * d.R) / d.P0, d.gamma) - pref;
}

// ---------------------------------------------------------------------------
// Strong-form gradients of U and theta at every interior node, as
// gradf<true>(X, true) builds them (field.h:3327-3372 with grad_flux<true>,
// field.h:3050-3082): the surface lift fN (x) (cds(X) - X) over the faces
// through the node, in the cell's face order, added on the owner side and
// subtracted on the neighbour side; then (Jin grad psi_s)(x_t) (x) X_s with
// Jin = Jinv cV of the node itself, over the node's x line (itself included),
// its y line and its z line (itself excluded), in that order; then / cV.
// Ghost traces as in the flux kernels: U by its boundary condition, theta a
// copy.  grad[c][nb]: c = 3a+b for d_a U_b, 9+a for d_a theta.
template <int NXc, int NYc, int NZc>
__global__ void __launch_bounds__(512) split_grad_kernel(DgDev d, SplitState s, double* __restrict__ grad) {
    const int NX = NXc ? NXc : d.npx, NY = NYc ? NYc : d.npy, NZ = NZc ? NZc : d.npz;
    const int NP = NX * NY * NZ;
    const int epb = blockDim.x / NP;
    const int le = threadIdx.x / NP, q = threadIdx.x - le * NP;
    const uint32_t e = blockIdx.x * epb + le;
    const bool active = le < epb && e < d.nelem;
    extern __shared__ double lds[];
    double* P = lds + (size_t)le * 4 * NP;  // U_x, U_y, U_z, theta rows
    const int i = q / (NY * NZ), j = (q / NZ) % NY, k = q % NZ;
    const uint64_t n = (uint64_t)e * NP + q;

    double u[3] = {0, 0, 0}, th = 0;
    if (active) {
        u[0] = s.ux[n]; u[1] = s.uy[n]; u[2] = s.uz[n];
        th = s.T[n] + d.T0;
        P[q] = u[0]; P[NP + q] = u[1]; P[2 * NP + q] = u[2]; P[3 * NP + q] = th;
    }
    __syncthreads();
    if (!active) return;

    double gu[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, gt[3] = {0, 0, 0};
    for (int g = 0; g < 6; g++) {
        if (!on_face(g, i, j, k, NX, NY, NZ)) continue;
        const EFace ef = d.eface[(size_t)e * 6 + g];
        if (!((ef.meta >> 4) & 1)) continue;
        const bool own = (ef.meta >> 3) & 1;
        const bool ghost = (ef.meta >> 5) & 1;
        const int og = ef.meta & 7;
        int a, b;
        face_ab(g, i, j, k, a, b);
        if ((ef.meta >> 6) & 1) {  // 2:1 mortar side: signed lift from the pre-pass
            const double* mg = d.mgrad + ((size_t)(ef.meta >> 8) * d.npfm + a * (g < 2 ? NY : NZ) + b) * 12;
            for (int c = 0; c < 9; c++) gu[c] += mg[c];
            for (int c = 0; c < 3; c++) gt[c] += mg[9 + c];
            continue;
        }
        const double wgt = face_weight(d, own ? g : og, a, b);
        const double nv[3] = {d.gFN[3 * (size_t)ef.facet] * wgt, d.gFN[3 * (size_t)ef.facet + 1] * wgt,
                              d.gFN[3 * (size_t)ef.facet + 2] * wgt};
        double ou[3], oth;
        if (!ghost) {
            const uint64_t m = face_node(ef.other, og, a, b, NX, NY, NZ);
            ou[0] = s.ux[m]; ou[1] = s.uy[m]; ou[2] = s.uz[m];
            oth = s.T[m] + d.T0;
        } else {
            ghost_velocity(*d.bc, ef.meta >> 8, u, nv, ou);
            oth = th;
        }
        const double fI = ghost ? 0.0 : 0.5;
        // cds with the owner's value weighted by fI, minus the node's own
        double du[3];
        for (int c = 0; c < 3; c++)
            du[c] = (own ? u[c] * fI + ou[c] * (1 - fI) : ou[c] * fI + u[c] * (1 - fI)) - u[c];
        const double dth = (own ? th * fI + oth * (1 - fI) : oth * fI + th * (1 - fI)) - th;
        for (int aa = 0; aa < 3; aa++) {
            for (int c = 0; c < 3; c++)
                gu[3 * aa + c] = own ? gu[3 * aa + c] + nv[aa] * du[c] : gu[3 * aa + c] - nv[aa] * du[c];
            gt[aa] = own ? gt[aa] + nv[aa] * dth : gt[aa] - nv[aa] * dth;
        }
    }
    double J[9];
    for (int c = 0; c < 9; c++) J[c] = d.JcV[(size_t)c * d.nb + n];
    auto add = [&](int src, double w0, double w1, double w2) {
        const double p0 = P[src], p1 = P[NP + src], p2 = P[2 * NP + src], p3 = P[3 * NP + src];
        gu[0] += w0 * p0; gu[1] += w0 * p1; gu[2] += w0 * p2;
        gu[3] += w1 * p0; gu[4] += w1 * p1; gu[5] += w1 * p2;
        gu[6] += w2 * p0; gu[7] += w2 * p1; gu[8] += w2 * p2;
        gt[0] += w0 * p3; gt[1] += w1 * p3; gt[2] += w2 * p3;
    };
    // off the node, grad psi_s(x_t) has one non-zero component, so Jin grad
    // psi is one column of Jin (XX0 YY1 ZZ2 XY3 YZ4 XZ5 YX6 ZY7 ZX8) times it
    const int syz = NY * NZ;
    const double* D0 = d.D[0];
    const double* D1 = d.D[1];
    const double* D2 = d.D[2];
    for (int t = 0; t < NX; t++) {
        const int src = t * syz + j * NZ + k;
        if (t == i) {
            const double d0 = D0[i * NX + i], d1 = D1[j * NY + j], d2 = D2[k * NZ + k];
            add(src, J[0] * d0 + J[3] * d1 + J[5] * d2, J[6] * d0 + J[1] * d1 + J[4] * d2,
                J[8] * d0 + J[7] * d1 + J[2] * d2);
        } else {
            const double D = D0[i * NX + t];
            add(src, J[0] * D, J[6] * D, J[8] * D);
        }
    }
    for (int t = 0; t < NY; t++) {
        if (t == j) continue;
        const double D = D1[j * NY + t];
        add(i * syz + t * NZ + k, J[3] * D, J[1] * D, J[7] * D);
    }
    for (int t = 0; t < NZ; t++) {
        if (t == k) continue;
        const double D = D2[k * NZ + t];
        add(i * syz + j * NZ + t, J[5] * D, J[4] * D, J[2] * D);
    }
    const double cv = d.cV[n];
    for (int c = 0; c < 9; c++) grad[(size_t)c * d.nb + n] = gu[c] / cv;
    for (int c = 0; c < 3; c++) grad[(size_t)(9 + c) * d.nb + n] = gt[c] / cv;
}

// Ghost-node gradients (fillBCs(r, X.fIndex), field.h:2745-2768): the
// owner's copy, except on the field's own NEUMANN patches, where it is the
// patch value read as the gradient type (with the default shape, tvalue and
// zMin of 0, the value on the tensor diagonal, or (value, 0, 0) for theta),
// and on its SYMMETRY patches, where it is zero.
__device__ __forceinline__ void ghost_gradient(const BcTable& bt, int pi, const double* own, double* out) {
    const int ku = bt.kind[1][pi], kt = bt.kind[2][pi];
    for (int c = 0; c < 9; c++) out[c] = own[c];
    if (ku == BC_SYMMETRY || ku == BC_NEUMANN)
        for (int c = 0; c < 9; c++) out[c] = 0.0;
    if (ku == BC_NEUMANN) { out[0] = bt.value[1][pi]; out[4] = bt.value[2][pi]; out[8] = bt.value[3][pi]; }
    for (int c = 9; c < 12; c++) out[c] = own[c];
    if (kt == BC_SYMMETRY || kt == BC_NEUMANN)
        for (int c = 9; c < 12; c++) out[c] = 0.0;
    if (kt == BC_NEUMANN) out[9] = bt.value[4][pi];
}

// ---------------------------------------------------------------------------
// Rusanov fluxes of the momentum and theta pass at one face node, signed as
// the node's element adds them (out[0..2] momentum rows, out[3] theta); 2:1
// mortar sides take the signed terms of the pre-pass.
template <bool VISC>
__device__ __forceinline__ bool split_ut_face(const DgDev& d, const SplitState& s, const double* __restrict__ rho1,
                                              const double* __restrict__ p1, uint32_t e, int g, int a, int b,
                                              int node, int NX, int NY, int NZ, double* out) {
    const EFace ef = d.eface[(size_t)e * 6 + g];
    if (!((ef.meta >> 4) & 1)) return false;
    if ((ef.meta >> 6) & 1) {  // 2:1 mortar side: signed terms from the pre-pass
        const double* ms = d.msurf + ((size_t)(ef.meta >> 8) * d.npfm + a * (g < 2 ? NY : NZ) + b) * 5;
        for (int r = 0; r < 4; r++) out[r] = ms[1 + r];
        return true;
    }
    const uint64_t n = (uint64_t)e * (NX * NY * NZ) + node;
    const double rho0 = s.rho[n], rho = rho1[n];
    const double u[3] = {s.ux[n], s.uy[n], s.uz[n]};
    const double th = s.T[n] + d.T0;
    const double pp = p1[n];
    double gown[VISC ? 12 : 1];
    if constexpr (VISC)
        for (int c = 0; c < 12; c++) gown[c] = d.grad[(size_t)c * d.nb + n];
    const double gRT = d.gamma * d.R;
    const bool own = (ef.meta >> 3) & 1;
    const bool ghost = (ef.meta >> 5) & 1;
    const int og = ef.meta & 7;
    const double wgt = face_weight(d, own ? g : og, a, b);
    const double nv[3] = {d.gFN[3 * (size_t)ef.facet] * wgt, d.gFN[3 * (size_t)ef.facet + 1] * wgt,
                          d.gFN[3 * (size_t)ef.facet + 2] * wgt};
    double orho0, orho, ou[3], oth, opp;
    double go[VISC ? 12 : 1];
    if (!ghost) {
        const uint64_t m = face_node(ef.other, og, a, b, NX, NY, NZ);
        orho0 = s.rho[m]; orho = rho1[m];
        ou[0] = s.ux[m]; ou[1] = s.uy[m]; ou[2] = s.uz[m];
        oth = s.T[m] + d.T0;
        opp = p1[m];
        if constexpr (VISC)
            for (int c = 0; c < 12; c++) go[c] = d.grad[(size_t)c * d.nb + m];
    } else {
        ghost_velocity(*d.bc, ef.meta >> 8, u, nv, ou);
        orho0 = rho0;
        orho = rho;
        oth = th;
        opp = pp;
        if constexpr (VISC) ghost_gradient(*d.bc, ef.meta >> 8, gown, go);
    }
    const double r0O = own ? rho0 : orho0, r0N = own ? orho0 : rho0;
    const double rO = own ? rho : orho, rN = own ? orho : rho;
    double uO[3], uN[3];
    for (int c = 0; c < 3; c++) { uO[c] = own ? u[c] : ou[c]; uN[c] = own ? ou[c] : u[c]; }
    const double thO = own ? th : oth, thN = own ? oth : th;
    const double pO = own ? pp : opp, pN = own ? opp : pp;
    const double fI = ghost ? 0.0 : 0.5;
    const double magO = sqrt(dot3r(uO[0], uO[1], uO[2], uO[0], uO[1], uO[2]));
    const double magN = sqrt(dot3r(uN[0], uN[1], uN[2], uN[0], uN[1], uN[2]));
    const double lam = ((magO * fI + magN * (1 - fI)) +
                        (sqrt(gRT * thO) * fI + sqrt(gRT * thN) * (1 - fI))) / 2;
    const double mn = sqrt(dot3r(nv[0], nv[1], nv[2], nv[0], nv[1], nv[2]));
    const double en[3] = {nv[0] / mn, nv[1] / mn, nv[2] / mn};
    const double fcO[3] = {r0O * uO[0], r0O * uO[1], r0O * uO[2]};
    const double fcN[3] = {r0N * uN[0], r0N * uN[1], r0N * uN[2]};
    // viscous parts: mu = rho^n visc on each side, gradients G
    double muO = 0, muN = 0;
    const double* gO = nullptr;
    const double* gN = nullptr;
    if constexpr (VISC) {
        muO = r0O * d.visc;
        muN = r0N * d.visc;
        gO = own ? gown : go;
        gN = own ? go : gown;
    }
    for (int row = 0; row < 3; row++) {
        double fF[3];
        for (int c = 0; c < 3; c++) {
            double FO = fcO[row] * uO[c] + (row == c ? pO : 0.0 * pO);
            double FN = fcN[row] * uN[c] + (row == c ? pN : 0.0 * pN);
            if constexpr (VISC) {
                FO = FO - muO * gO[3 * row + c];
                FN = FN - muN * gN[3 * row + c];
            }
            const double dq = lam * (rN * uN[c] - rO * uO[c]);
            fF[c] = FO * fI + FN * (1 - fI) - en[row] * dq;
        }
        const double flux = fF[0] * nv[0] + fF[1] * nv[1] + fF[2] * nv[2];
        out[row] = own ? flux : -flux;
    }
    {
        const double dq = lam * (rN * thN - rO * thO);
        double fF[3];
        for (int c = 0; c < 3; c++) {
            double FO = fcO[c] * thO, FN = fcN[c] * thN;
            if constexpr (VISC) {
                FO = FO - (muO * d.iPr) * gO[9 + c];
                FN = FN - (muN * d.iPr) * gN[9 + c];
            }
            fF[c] = FO * fI + FN * (1 - fI) - en[c] * dq;
        }
        const double flux = dot3r(fF[0], fF[1], fF[2], nv[0], nv[1], nv[2]);
        out[3] = own ? flux : -flux;
    }
    return true;
}

template <int NXc, int NYc, int NZc, bool VISC>
__global__ void __launch_bounds__(512) split_ut_kernel(DgDev d, SplitState s, const double* __restrict__ rho1,
                                                      SplitOut o) {
    const int NX = NXc ? NXc : d.npx, NY = NYc ? NYc : d.npy, NZ = NZc ? NZc : d.npz;
    const int NP = NX * NY * NZ;
    const int NFN = split_nfn(NX, NY, NZ);
    const int epb = blockDim.x / NP;
    const int le = threadIdx.x / NP, q = threadIdx.x - le * NP;
    const uint32_t e = blockIdx.x * epb + le;
    const bool active = le < epb && e < d.nelem;
    extern __shared__ double lds[];
    double* F = lds + (size_t)le * 6 * NP;  // node states (inviscid) or two flux rows (viscous)
    double* J = lds + (size_t)epb * 6 * NP + (size_t)le * 9 * NP;
    double* SF = split_sf_overlaid(epb, NP, NFN, 6, 4) ? lds : lds + (size_t)epb * 15 * NP;
    double* Dl = lds + split_lds_doubles(epb, NP, NFN, 6, 4);
    split_stage_d(d, Dl, NX, NY, NZ);
    const int i = q / (NY * NZ), j = (q / NZ) % NY, k = q % NZ;
    const uint64_t n = (uint64_t)e * NP + q;

    double rho0 = 0, rho = 0, u[3] = {0, 0, 0}, th = 0, pp = 0, cv = 0, rref = 0;
    unsigned fon = 0;
    // viscous: own flux rows U_x U_y U_z (tensor rows), theta (vector)
    double fr[VISC ? 4 : 1][3];
    if (active) {
        rho0 = s.rho[n]; rho = rho1[n];
        u[0] = s.ux[n]; u[1] = s.uy[n]; u[2] = s.uz[n];
        th = s.T[n] + d.T0;
        pp = o.p[n];
        // read with the state (no global load after the first barrier)
        cv = d.cV[n];
        if (d.buoyancy) rref = d.rho_ref[n];
        fon = split_node_faces(d, e, i, j, k, NX, NY, NZ);
        if constexpr (VISC) {
            const double fc[3] = {rho0 * u[0], rho0 * u[1], rho0 * u[2]};
            for (int row = 0; row < 3; row++)
                for (int c = 0; c < 3; c++) fr[row][c] = fc[row] * u[c] + (row == c ? pp : 0.0 * pp);
            for (int c = 0; c < 3; c++) fr[3][c] = fc[c] * th;
            double gown[12];
            for (int c = 0; c < 12; c++) gown[c] = d.grad[(size_t)c * d.nb + n];
            const double mu = rho0 * d.visc, imu = mu * d.iPr;
            for (int row = 0; row < 3; row++)
                for (int c = 0; c < 3; c++) fr[row][c] = fr[row][c] - mu * gown[3 * row + c];
            for (int c = 0; c < 3; c++) fr[3][c] = fr[3][c] - imu * gown[9 + c];
        }
        for (int c = 0; c < 9; c++) J[c * NP + q] = d.JcV[(size_t)kJinRow[c] * d.nb + n];
    }
    // surface: face tasks (one face per wavefront for N = 7), then each node
    // adds the faces through it in the cell's face order
    for (int t = threadIdx.x; t < epb * NFN; t += blockDim.x) {
        const int fl = t / NFN, task = t - fl * NFN;
        const uint32_t fe = blockIdx.x * epb + fl;
        if (fe >= d.nelem) break;
        int g, fa, fb, node;
        face_task(task, NX, NY, NZ, g, fa, fb, node);
        double v[4];
        if (split_ut_face<VISC>(d, s, rho1, o.p, fe, g, fa, fb, node, NX, NY, NZ, v))
            for (int r = 0; r < 4; r++) SF[4 * t + r] = v[r];
    }
    __syncthreads();
    double res[4] = {0, 0, 0, 0};
    if (active)
        for (int g = 0; g < 6; g++) {
            if (!((fon >> g) & 1)) continue;
            int a, b;
            face_ab(g, i, j, k, a, b);
            const double* sf = SF + 4 * (le * NFN + face_task_index(g, a, b, NX, NY, NZ));
            for (int r = 0; r < 4; r++) res[r] += sf[r];
        }
    __syncthreads();
    const double* D1l = Dl + NX * NX;
    const double* D2l = D1l + NY * NY;
    if constexpr (VISC) {
        // volume terms in two passes of two flux rows through LDS
        if (active)
            for (int c = 0; c < 3; c++) { F[c * NP + q] = fr[0][c]; F[(3 + c) * NP + q] = fr[1][c]; }
        __syncthreads();
        if (active) volume_reference_order<2, 0>(res, F, J, NP, i, j, k, NX, NY, NZ, Dl, D1l, D2l);
        __syncthreads();
        if (active)
            for (int c = 0; c < 3; c++) { F[c * NP + q] = fr[2][c]; F[(3 + c) * NP + q] = fr[3][c]; }
        __syncthreads();
        if (!active) return;
        volume_reference_order<2, 2>(res + 2, F, J, NP, i, j, k, NX, NY, NZ, Dl, D1l, D2l);
    } else {
        // volume terms in one pass: the six state values (rho^n, U, theta,
        // p') of every node in LDS, each source's four flux rows rebuilt
        // from them with the same operations as fr above
        if (active) {
            F[q] = rho0; F[NP + q] = u[0]; F[2 * NP + q] = u[1]; F[3 * NP + q] = u[2];
            F[4 * NP + q] = th; F[5 * NP + q] = pp;
        }
        __syncthreads();
        if (!active) return;
        volume_reference_walk(
            [&](int sn, double w0, double w1, double w2) {
                const double r0 = F[sn], us[3] = {F[NP + sn], F[2 * NP + sn], F[3 * NP + sn]};
                const double ts = F[4 * NP + sn], ps = F[5 * NP + sn];
                const double fc[3] = {r0 * us[0], r0 * us[1], r0 * us[2]};
                for (int row = 0; row < 3; row++) {
                    // tensor flux: row dot left to right
                    const double f0 = fc[row] * us[0] + (row == 0 ? ps : 0.0 * ps);
                    const double f1 = fc[row] * us[1] + (row == 1 ? ps : 0.0 * ps);
                    const double f2 = fc[row] * us[2] + (row == 2 ? ps : 0.0 * ps);
                    res[row] -= f0 * w0 + f1 * w1 + f2 * w2;
                }
                // vector flux: right-nested dot
                res[3] -= dot3r(fc[0] * ts, fc[1] * ts, fc[2] * ts, w0, w1, w2);
            },
            J, NP, i, j, k, NX, NY, NZ, Dl, D1l, D2l);
    }

    const double ap0 = (-1.0 / d.dt) * cv;
    const double ap = ap0 * rho;
    double* outs[3] = {o.ux, o.uy, o.uz};
    double un[3];
    for (int c = 0; c < 3; c++) {
        const double sc = d.buoyancy ? (rho - rref) * d.g[c] : 0.0;
        un[c] = ((res[c] - sc * cv) + (u[c] * rho0) * ap0) / ap;
        outs[c][n] = un[c];
    }
    const double thn = (res[3] + (th * rho0) * ap0) / ap;
    o.T[n] = thn - d.T0;
    // the new rho (this pass's input) and U, theta: any NaN or Inf raises
    // the problem's flag, which the step's caller reads
    if (!(isfinite(rho) && isfinite(un[0]) && isfinite(un[1]) && isfinite(un[2]) && isfinite(thn)))
        *d.flag = 1;
}

// ---------------------------------------------------------------------------
// 2:1 mortar faces of the split step, in the reference's form
// (scatter_non_conforming / gather_non_conforming, field.h:2018-2248, inside
// cds, rusanov and div_flux): the coarse side's node fields -- flux vectors,
// the Rusanov state, |U| and c -- interpolated to each refined face node
// with psiRef, the Rusanov flux vector formed there, and gathered to the
// coarse side's node positions with psiCor before the dot with the face
// normal.  One work-group per mortar face (coarse node fields staged in
// LDS), one thread per face node; then one work-group per split coarse side.
// The terms are stored signed (+ on the owner side, as div_flux adds) in
// msurf: entry 0 for the density pass, 1-3 momentum, 4 theta.  mphi holds
// per face node the flux vectors (density 0-2 in the first pass; momentum
// rows 0-8 and theta 9-11 in the second) and lambda (12).
constexpr int kSplitPhi = 16;

__device__ __forceinline__ void mortar_node_ids(const DgDev& d, const MortarFace& mf, int t, int& ao, int& bo) {
    ao = t / mf.n1;
    bo = t - ao * mf.n1;
    (void)d;
}

// density pass: F = U rho, q = rho; lambda = (cds |U| + cds c) / 2
__global__ void __launch_bounds__(256) split_mortar_rho_kernel(DgDev d, SplitState s) {
    const uint32_t f = blockIdx.x;
    const int t = threadIdx.x;
    const MortarFace mf = d.mort[f];
    const bool act = t < mf.n0 * mf.n1;
    int ao, bo;
    mortar_node_ids(d, mf, t, ao, bo);
    const int np = d.npfm;
    const double gRT = d.gamma * d.R;
    auto vals = [&](uint64_t m, double* v) {
        const double rho = s.rho[m], u0 = s.ux[m], u1 = s.uy[m], u2 = s.uz[m];
        const double th = s.T[m] + d.T0;
        v[0] = u0 * rho; v[1] = u1 * rho; v[2] = u2 * rho;
        v[3] = rho;
        v[4] = sqrt(dot3r(u0, u1, u2, u0, u1, u2));
        v[5] = sqrt(gRT * th);
    };
    extern __shared__ double L[];  // [6][npfm]
    if (act) {
        double v[6];
        vals(face_node(mf.coarse, mf.side, ao, bo, d.npx, d.npy, d.npz), v);
        for (int q = 0; q < 6; q++) L[q * np + t] = v[q];
    }
    __syncthreads();
    if (!act) return;
    const size_t ft = (size_t)f * np + t;
    double vf[6], vc[6] = {0, 0, 0, 0, 0, 0};
    vals(d.mort_fine[ft], vf);
    const int d0 = mf.side < 4 ? 0 : 1, d1 = mf.side < 2 ? 1 : 2;
    const double* R0 = d.psiRef[d0] + (size_t)mf.h0 * mf.n0 * mf.n0;
    const double* R1 = d.psiRef[d1] + (size_t)mf.h1 * mf.n1 * mf.n1;
    for (int q = 0; q < 6; q++) {
        double sum = 0.0;
        for (int an = 0; an < mf.n0; an++)
            for (int bn = 0; bn < mf.n1; bn++)
                sum += L[q * np + an * mf.n1 + bn] * R0[an * mf.n0 + ao] * R1[bn * mf.n1 + bo];
        vc[q] = sum;
    }
    const bool fo = mf.fine_owner;
    const double* vO = fo ? vf : vc;
    const double* vN = fo ? vc : vf;
    const double fI = 0.5;
    const double lam = ((vO[4] * fI + vN[4] * (1 - fI)) + (vO[5] * fI + vN[5] * (1 - fI))) / 2;
    const double wk = d.mort_w[ft];
    const double nv[3] = {mf.nf[0] * wk, mf.nf[1] * wk, mf.nf[2] * wk};
    const double mn = sqrt(dot3r(nv[0], nv[1], nv[2], nv[0], nv[1], nv[2]));
    const double en[3] = {nv[0] / mn, nv[1] / mn, nv[2] / mn};
    const double dq = lam * (vN[3] - vO[3]);
    double fF[3];
    for (int c = 0; c < 3; c++) fF[c] = vO[c] * fI + vN[c] * (1 - fI) - en[c] * dq;
    double* ph = d.mphi + ft * kSplitPhi;
    for (int c = 0; c < 3; c++) ph[c] = fF[c];
    ph[12] = lam;
    const double flux = dot3r(fF[0], fF[1], fF[2], nv[0], nv[1], nv[2]);
    d.msurf[(size_t)d.mort_fms[ft] * 5] = fo ? flux : -flux;
}

// momentum and theta pass: F_U = (rho^n U) (x) U + p' I, q_U = rho^{n+1} U;
// F_theta = (rho^n U) theta, q_theta = rho^{n+1} theta; lambda of the first
// pass
template <bool VISC>
__global__ void __launch_bounds__(256) split_mortar_ut_kernel(DgDev d, SplitState s, const double* rho1,
                                                              const double* p1) {
    const uint32_t f = blockIdx.x;
    const int t = threadIdx.x;
    const MortarFace mf = d.mort[f];
    const bool act = t < mf.n0 * mf.n1;
    int ao, bo;
    mortar_node_ids(d, mf, t, ao, bo);
    const int np = d.npfm;
    const double R = d.R, gam = d.gamma;
    // v: F_U rows (9), F_theta (3), q_U (3), q_theta (1)
    auto vals = [&](uint64_t m, double* v) {
        const double rho0 = s.rho[m], rho = rho1[m];
        const double u[3] = {s.ux[m], s.uy[m], s.uz[m]};
        const double th = s.T[m] + d.T0;
        const double pp = p1[m];
        const double fc[3] = {rho0 * u[0], rho0 * u[1], rho0 * u[2]};
        for (int row = 0; row < 3; row++)
            for (int c = 0; c < 3; c++) v[3 * row + c] = fc[row] * u[c] + (row == c ? pp : 0.0 * pp);
        for (int c = 0; c < 3; c++) v[9 + c] = fc[c] * th;
        if constexpr (VISC) {   // - mu grad U, - mu/Pr grad theta; mu = rho^n visc
            const double mu = rho0 * d.visc, imu = mu * d.iPr;
            for (int c = 0; c < 9; c++) v[c] = v[c] - mu * d.grad[(size_t)c * d.nb + m];
            for (int c = 0; c < 3; c++) v[9 + c] = v[9 + c] - imu * d.grad[(size_t)(9 + c) * d.nb + m];
        }
        for (int c = 0; c < 3; c++) v[12 + c] = rho * u[c];
        v[15] = rho * th;
    };
    extern __shared__ double L[];  // [16][npfm]
    if (act) {
        double v[16];
        vals(face_node(mf.coarse, mf.side, ao, bo, d.npx, d.npy, d.npz), v);
        for (int q = 0; q < 16; q++) L[q * np + t] = v[q];
    }
    __syncthreads();
    if (!act) return;
    const size_t ft = (size_t)f * np + t;
    double vf[16], vc[16];
    vals(d.mort_fine[ft], vf);
    const int d0 = mf.side < 4 ? 0 : 1, d1 = mf.side < 2 ? 1 : 2;
    const double* R0 = d.psiRef[d0] + (size_t)mf.h0 * mf.n0 * mf.n0;
    const double* R1 = d.psiRef[d1] + (size_t)mf.h1 * mf.n1 * mf.n1;
    for (int q = 0; q < 16; q++) {
        double sum = 0.0;
        for (int an = 0; an < mf.n0; an++)
            for (int bn = 0; bn < mf.n1; bn++)
                sum += L[q * np + an * mf.n1 + bn] * R0[an * mf.n0 + ao] * R1[bn * mf.n1 + bo];
        vc[q] = sum;
    }
    const bool fo = mf.fine_owner;
    const double* vO = fo ? vf : vc;
    const double* vN = fo ? vc : vf;
    const double fI = 0.5;
    double* ph = d.mphi + ft * kSplitPhi;
    const double lam = ph[12];
    const double wk = d.mort_w[ft];
    const double nv[3] = {mf.nf[0] * wk, mf.nf[1] * wk, mf.nf[2] * wk};
    const double mn = sqrt(dot3r(nv[0], nv[1], nv[2], nv[0], nv[1], nv[2]));
    const double en[3] = {nv[0] / mn, nv[1] / mn, nv[2] / mn};
    double* ms = d.msurf + (size_t)d.mort_fms[ft] * 5;
    for (int row = 0; row < 3; row++) {
        double fF[3];
        for (int c = 0; c < 3; c++) {
            const double dq = lam * (vN[12 + c] - vO[12 + c]);
            fF[c] = vO[3 * row + c] * fI + vN[3 * row + c] * (1 - fI) - en[row] * dq;
            ph[3 * row + c] = fF[c];
        }
        const double flux = fF[0] * nv[0] + fF[1] * nv[1] + fF[2] * nv[2];
        ms[1 + row] = fo ? flux : -flux;
    }
    {
        const double dq = lam * (vN[15] - vO[15]);
        double fF[3];
        for (int c = 0; c < 3; c++) {
            fF[c] = vO[9 + c] * fI + vN[9 + c] * (1 - fI) - en[c] * dq;
            ph[9 + c] = fF[c];
        }
        const double flux = dot3r(fF[0], fF[1], fF[2], nv[0], nv[1], nv[2]);
        ms[4] = fo ? flux : -flux;
    }
}

// strong-gradient pass (diffusion): cds of U and theta at the refined face
// nodes (scatter), and the lift fN (x) (face value - node value) of the
// refined node, signed as grad_flux<true> adds it (field.h:3050-3082); the
// face values are kept in mphi 12-15 for the gather
__global__ void __launch_bounds__(256) split_mortar_grad_kernel(DgDev d, SplitState s) {
    const uint32_t f = blockIdx.x;
    const int t = threadIdx.x;
    const MortarFace mf = d.mort[f];
    const bool act = t < mf.n0 * mf.n1;
    int ao, bo;
    mortar_node_ids(d, mf, t, ao, bo);
    const int np = d.npfm;
    auto vals = [&](uint64_t m, double* v) {
        v[0] = s.ux[m]; v[1] = s.uy[m]; v[2] = s.uz[m];
        v[3] = s.T[m] + d.T0;
    };
    extern __shared__ double L[];  // [4][npfm]
    if (act) {
        double v[4];
        vals(face_node(mf.coarse, mf.side, ao, bo, d.npx, d.npy, d.npz), v);
        for (int q = 0; q < 4; q++) L[q * np + t] = v[q];
    }
    __syncthreads();
    if (!act) return;
    const size_t ft = (size_t)f * np + t;
    double vf[4], vc[4];
    vals(d.mort_fine[ft], vf);
    const int d0 = mf.side < 4 ? 0 : 1, d1 = mf.side < 2 ? 1 : 2;
    const double* R0 = d.psiRef[d0] + (size_t)mf.h0 * mf.n0 * mf.n0;
    const double* R1 = d.psiRef[d1] + (size_t)mf.h1 * mf.n1 * mf.n1;
    for (int q = 0; q < 4; q++) {
        double sum = 0.0;
        for (int an = 0; an < mf.n0; an++)
            for (int bn = 0; bn < mf.n1; bn++)
                sum += L[q * np + an * mf.n1 + bn] * R0[an * mf.n0 + ao] * R1[bn * mf.n1 + bo];
        vc[q] = sum;
    }
    const bool fo = mf.fine_owner;
    const double* vO = fo ? vf : vc;
    const double* vN = fo ? vc : vf;
    const double fI = 0.5;
    double* ph = d.mphi + ft * kSplitPhi;
    double fF[4];
    for (int q = 0; q < 4; q++) {
        fF[q] = vO[q] * fI + vN[q] * (1 - fI);
        ph[12 + q] = fF[q];
    }
    const double wk = d.mort_w[ft];
    const double nv[3] = {mf.nf[0] * wk, mf.nf[1] * wk, mf.nf[2] * wk};
    double* mg = d.mgrad + (size_t)d.mort_fms[ft] * 12;
    for (int aa = 0; aa < 3; aa++) {
        for (int c = 0; c < 3; c++) {
            const double x = nv[aa] * (fF[c] - vf[c]);
            mg[3 * aa + c] = fo ? x : -x;
        }
        const double x = nv[aa] * (fF[3] - vf[3]);
        mg[9 + aa] = fo ? x : -x;
    }
}

// gather of the gradient pass: per coarse side node, over the side's faces,
// the face values projected with psiCor minus the node's own value, times
// the face normal at that node, signed for the coarse element
__global__ void __launch_bounds__(256) split_mortar_grad_gather_kernel(DgDev d, SplitState s) {
    const uint32_t* cs = d.cslot + (size_t)blockIdx.x * 8;
    const uint32_t elem = cs[0];
    const int side = (int)cs[1], nch = (int)cs[2];
    const uint32_t slot = cs[3];
    const MortarFace m0 = d.mort[cs[4]];
    const int t = threadIdx.x;
    if (t >= m0.n0 * m0.n1) return;
    const int an = t / m0.n1, bn = t - an * m0.n1;
    const int d0 = side < 4 ? 0 : 1, d1 = side < 2 ? 1 : 2;
    const uint64_t nc = face_node(elem, side, an, bn, d.npx, d.npy, d.npz);
    const double xc[4] = {s.ux[nc], s.uy[nc], s.uz[nc], s.T[nc] + d.T0};
    double out[12];
    for (int q = 0; q < 12; q++) out[q] = 0.0;
    for (int ch = 0; ch < nch; ch++) {
        const uint32_t f = cs[4 + ch];
        const MortarFace mf = d.mort[f];
        const double* C0 = d.psiCor[d0] + (size_t)mf.h0 * mf.n0 * mf.n0;
        const double* C1 = d.psiCor[d1] + (size_t)mf.h1 * mf.n1 * mf.n1;
        const double* ph = d.mphi + (size_t)f * d.npfm * kSplitPhi;
        double G[4];
        for (int q = 0; q < 4; q++) {
            double sum = 0.0;
            for (int ao = 0; ao < mf.n0; ao++)
                for (int bo = 0; bo < mf.n1; bo++)
                    sum += ph[(ao * mf.n1 + bo) * kSplitPhi + 12 + q] * C0[ao * mf.n0 + an] * C1[bo * mf.n1 + bn];
            G[q] = sum;
        }
        const double wk = d.mort_w[(size_t)f * d.npfm + t];
        const double nv[3] = {mf.nf[0] * wk, mf.nf[1] * wk, mf.nf[2] * wk};
        const bool coarse_owns = !mf.fine_owner;
        for (int aa = 0; aa < 3; aa++) {
            for (int c = 0; c < 3; c++) {
                const double x = nv[aa] * (G[c] - xc[c]);
                out[3 * aa + c] = coarse_owns ? out[3 * aa + c] + x : out[3 * aa + c] - x;
            }
            const double x = nv[aa] * (G[3] - xc[3]);
            out[9 + aa] = coarse_owns ? out[9 + aa] + x : out[9 + aa] - x;
        }
    }
    double* mg = d.mgrad + ((size_t)slot * d.npfm + t) * 12;
    for (int q = 0; q < 12; q++) mg[q] = out[q];
}

// gather of either pass to the split coarse sides: per coarse side node,
// over the side's faces in order, the flux vectors projected with psiCor and
// dotted with the face's normal at that node; UT selects the second pass
template <bool UT>
__global__ void __launch_bounds__(256) split_mortar_gather_kernel(DgDev d) {
    const uint32_t* cs = d.cslot + (size_t)blockIdx.x * 8;
    const int side = (int)cs[1], nch = (int)cs[2];
    const uint32_t slot = cs[3];
    const MortarFace m0 = d.mort[cs[4]];
    const int t = threadIdx.x;
    if (t >= m0.n0 * m0.n1) return;
    const int an = t / m0.n1, bn = t - an * m0.n1;
    const int d0 = side < 4 ? 0 : 1, d1 = side < 2 ? 1 : 2;
    constexpr int NV = UT ? 12 : 3;
    double out[4] = {0, 0, 0, 0};
    for (int ch = 0; ch < nch; ch++) {
        const uint32_t f = cs[4 + ch];
        const MortarFace mf = d.mort[f];
        const double* C0 = d.psiCor[d0] + (size_t)mf.h0 * mf.n0 * mf.n0;
        const double* C1 = d.psiCor[d1] + (size_t)mf.h1 * mf.n1 * mf.n1;
        const double* ph = d.mphi + (size_t)f * d.npfm * kSplitPhi;
        double G[NV];
        for (int q = 0; q < NV; q++) {
            double sum = 0.0;
            for (int ao = 0; ao < mf.n0; ao++)
                for (int bo = 0; bo < mf.n1; bo++)
                    sum += ph[(ao * mf.n1 + bo) * kSplitPhi + q] * C0[ao * mf.n0 + an] * C1[bo * mf.n1 + bn];
            G[q] = sum;
        }
        const double wk = d.mort_w[(size_t)f * d.npfm + t];
        const double nv[3] = {mf.nf[0] * wk, mf.nf[1] * wk, mf.nf[2] * wk};
        const bool coarse_owns = !mf.fine_owner;
        if constexpr (UT) {
            for (int row = 0; row < 3; row++) {
                const double flux = G[3 * row] * nv[0] + G[3 * row + 1] * nv[1] + G[3 * row + 2] * nv[2];
                out[row] = coarse_owns ? out[row] + flux : out[row] - flux;
            }
            const double flux = dot3r(G[9], G[10], G[11], nv[0], nv[1], nv[2]);
            out[3] = coarse_owns ? out[3] + flux : out[3] - flux;
        } else {
            const double flux = dot3r(G[0], G[1], G[2], nv[0], nv[1], nv[2]);
            out[0] = coarse_owns ? out[0] + flux : out[0] - flux;
        }
    }
    double* ms = d.msurf + ((size_t)slot * d.npfm + t) * 5;
    if constexpr (UT)
        for (int r = 0; r < 4; r++) ms[1 + r] = out[r];
    else
        ms[0] = out[0];
}

template <int X, int Y, int Z>
static hipError_t launch_split_t(const DgDev& d, const SplitState& s, double* rho1, const SplitOut& o,
                                 double* grad, hipStream_t st) {
    const int NP = d.np;
    if (NP > 512) return hipErrorInvalidValue;
    int epb = NP >= 256 ? 1 : 256 / NP;
    int threads = epb * NP;
    threads = (threads + 63) / 64 * 64;
    const unsigned blocks = (d.nelem + epb - 1) / epb;
    const int NFN = split_nfn(d.npx, d.npy, d.npz);
    const int ND = d.npx * d.npx + d.npy * d.npy + d.npz * d.npz;
    size_t lds_a = (split_lds_doubles(epb, NP, NFN, 3, 1) + ND) * sizeof(double);
    size_t lds_b = (split_lds_doubles(epb, NP, NFN, 6, 4) + ND) * sizeof(double);
    if (lds_a > 65536 || lds_b > 65536) return hipErrorInvalidValue;
    size_t lds_g = (size_t)epb * 4 * NP * sizeof(double);
    hipError_t err;
    const int mthreads = (d.npfm + 63) / 64 * 64;
    if (d.nmort && d.npfm > 256) return hipErrorInvalidValue;
    if (d.diffusion) {
        if (!grad || grad != d.grad) return hipErrorInvalidValue;
        if (d.nmort) {
            if (!d.mgrad) return hipErrorInvalidValue;
            hipLaunchKernelGGL(split_mortar_grad_kernel, dim3(d.nmort), dim3(mthreads),
                               (size_t)4 * d.npfm * sizeof(double), st, d, s);
            if (d.ncslot)
                hipLaunchKernelGGL(split_mortar_grad_gather_kernel, dim3(d.ncslot), dim3(mthreads), 0, st, d, s);
            if ((err = hipGetLastError()) != hipSuccess) return err;
        }
        hipLaunchKernelGGL((split_grad_kernel<X, Y, Z>), dim3(blocks), dim3(threads), lds_g, st, d, s, grad);
        if ((err = hipGetLastError()) != hipSuccess) return err;
    }
    if (d.nmort) {
        hipLaunchKernelGGL(split_mortar_rho_kernel, dim3(d.nmort), dim3(mthreads),
                           (size_t)6 * d.npfm * sizeof(double), st, d, s);
        if (d.ncslot) hipLaunchKernelGGL(split_mortar_gather_kernel<false>, dim3(d.ncslot), dim3(mthreads), 0, st, d);
        if ((err = hipGetLastError()) != hipSuccess) return err;
    }
    hipLaunchKernelGGL((split_rho_kernel<X, Y, Z>), dim3(blocks), dim3(threads), lds_a, st, d, s, rho1, o.p);
    if ((err = hipGetLastError()) != hipSuccess) return err;
    if (d.nmort) {
        if (d.diffusion)
            hipLaunchKernelGGL(split_mortar_ut_kernel<true>, dim3(d.nmort), dim3(mthreads),
                               (size_t)16 * d.npfm * sizeof(double), st, d, s, (const double*)rho1,
                               (const double*)o.p);
        else
            hipLaunchKernelGGL(split_mortar_ut_kernel<false>, dim3(d.nmort), dim3(mthreads),
                               (size_t)16 * d.npfm * sizeof(double), st, d, s, (const double*)rho1,
                               (const double*)o.p);
        if (d.ncslot) hipLaunchKernelGGL(split_mortar_gather_kernel<true>, dim3(d.ncslot), dim3(mthreads), 0, st, d);
        if ((err = hipGetLastError()) != hipSuccess) return err;
    }
    if (d.diffusion)
        hipLaunchKernelGGL((split_ut_kernel<X, Y, Z, true>), dim3(blocks), dim3(threads), lds_b, st, d, s, rho1, o);
    else
        hipLaunchKernelGGL((split_ut_kernel<X, Y, Z, false>), dim3(blocks), dim3(threads), lds_b, st, d, s, rho1, o);
    return hipGetLastError();
}

hipError_t launch_split_step(const DgDev& d, const SplitState& s, double* rho1, const SplitOut& o,
                             double* grad, hipStream_t st) {
    if (d.npx == 8 && d.npy == 8 && d.npz == 8) return launch_split_t<8, 8, 8>(d, s, rho1, o, grad, st);
    if (d.npx == 5 && d.npy == 1 && d.npz == 5) return launch_split_t<5, 1, 5>(d, s, rho1, o, grad, st);
    if (d.npx == 4 && d.npy == 4 && d.npz == 4) return launch_split_t<4, 4, 4>(d, s, rho1, o, grad, st);
    return launch_split_t<0, 0, 0>(d, s, rho1, o, grad, st);
}

}  // namespace nsem
