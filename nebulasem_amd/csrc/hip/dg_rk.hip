// Method-of-lines dGSEM Euler residual fused with an SSP-RK3 stage
// (Shu-Osher form, written as increments over the step's base state):
//
//   d1 = dt L(q)            d2 = (d1 + dt L(q + d1)) / 4
//   d3 = 2 (d2 + dt L(q + d2)) / 3                  q^{n+1} = q + d3
//
// State: conserved perturbations over the FP64 hydrostatic background,
//   rho' = rho - rho_ref,  m = rho U,  Theta' = rho theta - rho_ref T0.
// Spatial operator: the reference's (divf + rusanov + div_flux + src,
// field.h:2927-2943, 3085-3117, 3416-3478, 3711-3723) with all fluxes
// evaluated at one state: F_rho = m, F_m = m (x) U + p' I, F_Theta = m theta,
// lambda = (avg|U| + avg sqrt(gamma R theta)) / 2, normal-only momentum
// dissipation, buoyancy (rho - rho_ref) g.  p' = p_ref expm1(gamma
// log1p(Theta'/(rho_ref T0))) is the exact p(rho theta) - p_ref without the
// cancellation, so the flux never touches the background's hydrostatic
// balance.
//
// Template parameters: R the arithmetic type of the residual (double, or
// float for the FP32-perturbation form), S the storage type of the stage
// increments.  The base state and the final accumulation q + d3 are FP64 in
// both forms; the background is FP64 and converted once per node.
//
// One work-group per element (NP threads, NP = 512 at N = 7): contravariant
// fluxes of the 5 equations staged in LDS (15 x NP values), transposed 1-D
// derivative along the three GLL lines, Rusanov surface terms from the
// neighbour traces (global, L2/MALL resident), update fused.  Work-groups are
// mapped to elements XCD-contiguously so face neighbours share an L2.
// The method-of-lines path has no bit-level counterpart in the reference, so
// products and sums are fused into FMAs here, in the headers' inline physics
// too (the split reference-order kernels keep the build-wide
// -ffp-contract=off).
#pragma clang fp contract(fast)

#include "dg_kernels.hpp"
#include "dg_physics.hpp"

namespace nsem {

namespace {

// Stores of the stage output bypass the caches' retention (non-temporal):
// they are not read again in this launch, and the L2 keeps the neighbour
// traces instead.
template <bool O32, typename T>
__device__ __forceinline__ void stn(T* base, uint64_t i, T v) {
    if constexpr (O32)
        __builtin_nontemporal_store(v, reinterpret_cast<T*>(reinterpret_cast<char*>(base) + byte_off(i, sizeof(T))));
    else
        __builtin_nontemporal_store(v, base + i);
}

// Increment of variable v at node n (DIN: the stage has one; a compile-time
// choice, so the loads carry no null-pointer branches between them).
template <bool O32, bool DIN, typename S>
__device__ __forceinline__ double ld_delta(const S* d, uint64_t i) {
    if constexpr (DIN) return (double)ldn<O32>(d, i);
    else return 0.0;
}

// State q + d_in of variable v at node n in the residual type (with its
// tangent for the tangent-linear instantiation).
template <typename R, typename S, bool O32, bool DIN>
__device__ __forceinline__ R load_state(const RkArgs<S>& a, int v, uint64_t n) {
    const double x = ldn<O32>(a.q[v], n) + ld_delta<O32, DIN>(a.din[v], n);
    if constexpr (is_dual<R>::value) {
        R r(x);
        r.d[0] = ldn<O32>(a.tq[v], n) + ld_delta<O32, DIN>(a.tdin[v], n);
        return r;
    } else {
        return (R)x;
    }
}

}  // namespace

// Everything one face-node task reads from global memory, gathered before
// the volume term so that its latency overlaps the element's own loads:
// the face entry, the weighted facet normal and the neighbour's state (on a
// boundary or absent face the owner's own face node is read instead, so the
// loads need no branch; only its background density is used there).
template <typename R>
struct FaceTask {
    uint32_t meta;
    int g, fa, fb, fnode;
    uint64_t node;    // the node the state was read from (neighbour, or own on a boundary)
    R nv[3];
    R rp, m[3], tp, rref, pref;
};

template <typename R, typename S, bool O32, bool DIN>
__device__ __forceinline__ void prepare_face_task(const DgDev& d, const RkArgs<S>& a, uint32_t e, int t,
                                                  int NX, int NY, int NZ, FaceTask<R>& F) {
    face_task(t, NX, NY, NZ, F.g, F.fa, F.fb, F.fnode);
    // one 12-byte load of the face entry (facet, other, meta)
    const uint3 efv = *reinterpret_cast<const uint3*>(d.eface + (size_t)e * 6 + F.g);
    const EFace ef{efv.x, efv.y, efv.z};
    F.meta = ef.meta;
    const bool present = (ef.meta >> 4) & 1;
    const bool interior = present && !((ef.meta >> 5) & 1);
    const bool own = (ef.meta >> 3) & 1;
    const int og = ef.meta & 7;
    // facet-node weight w_a w_b / 4 in the owner's face layout (face_weight)
    const int wg = own ? F.g : og;
    const double* wa = wg < 4 ? d.w[0] : d.w[1];
    const double* wb = wg < 2 ? d.w[1] : d.w[2];
    const double wgt = wa[F.fa] * wb[F.fb] / 4;
    const size_t f3 = 3 * (size_t)(present ? ef.facet : 0);
    F.nv[0] = (R)(d.gFN[f3] * wgt);
    F.nv[1] = (R)(d.gFN[f3 + 1] * wgt);
    F.nv[2] = (R)(d.gFN[f3 + 2] * wgt);
    const uint64_t m = interior ? face_node(ef.other, og, F.fa, F.fb, NX, NY, NZ)
                                : (uint64_t)e * (NX * NY * NZ) + F.fnode;
    F.node = m;
    F.rp = load_state<R, S, O32, DIN>(a, 0, m);
    F.m[0] = load_state<R, S, O32, DIN>(a, 1, m);
    F.m[1] = load_state<R, S, O32, DIN>(a, 2, m);
    F.m[2] = load_state<R, S, O32, DIN>(a, 3, m);
    F.tp = load_state<R, S, O32, DIN>(a, 4, m);
    F.rref = (R)ldn<O32>(d.rho_ref, m);
    F.pref = (R)ldn<O32>(d.p_ref, m);
}

// Ghost-node gradients of the viscous terms, as the split step forms them
// (fillBCs(r, X.fIndex), field.h:2745-2768): the owner's copy, zero on the
// field's SYMMETRY patches, the patch value on the tensor diagonal (theta:
// (value, 0, 0)) on its NEUMANN patches.
template <typename R>
__device__ __forceinline__ void ghost_gradient_r(const BcTable& bt, int pi, const R* own, R* out) {
    const int ku = bt.kind[1][pi], kt = bt.kind[2][pi];
    const bool zu = ku == BC_SYMMETRY || ku == BC_NEUMANN, zt = kt == BC_SYMMETRY || kt == BC_NEUMANN;
    for (int c = 0; c < 9; c++) out[c] = zu ? R(0) : own[c];
    if (ku == BC_NEUMANN) { out[0] = (R)bt.value[1][pi]; out[4] = (R)bt.value[2][pi]; out[8] = (R)bt.value[3][pi]; }
    for (int c = 9; c < 12; c++) out[c] = zt ? R(0) : own[c];
    if (kt == BC_NEUMANN) out[9] = (R)bt.value[4][pi];
}

// Normal components of the viscous fluxes mu grad U (rows 0-2) and
// mu/Pr grad theta' (row 3) through the area-weighted normal nv.
template <typename R>
__device__ __forceinline__ void viscous_normal(const R* g, R mu, R iPr, const R* nv, R* v) {
    for (int r = 0; r < 3; r++) v[r] = mu * (g[3 * r] * nv[0] + g[3 * r + 1] * nv[1] + g[3 * r + 2] * nv[2]);
    v[3] = (mu * iPr) * (g[9] * nv[0] + g[10] * nv[1] + g[11] * nv[2]);
}

// Primitive rows kept in LDS for the face tasks (the owner side of each face
// node): rho', m (3), Theta', U (3), theta - T0, p'.  c and |U| are
// recomputed from them with the same operations as make_prim.  The rows are
// padded by one slot per 8 nodes: a face of constant k reads every 8th node,
// which would otherwise hit the same few LDS banks.
constexpr int kPrimRows = 10;

// Work-group of one element (NP threads, one per node; several elements per
// group when NP < 256).  Phases, each closed by a barrier:
//   1. own state -> primitives -> contravariant fluxes (15 rows) in LDS;
//      face tasks issue their neighbour loads
//   2. volume term along the three GLL lines -> registers
//   3. own primitives -> LDS (over the flux rows)
//   4. one thread per face node (6 faces x (N+1)^2 = 384 tasks at N = 7,
//      one face per wavefront): Rusanov flux -> LDS
//   5. each node subtracts/adds the fluxes of its faces in face-group order,
//      then the stage update.
// AFF: the elements are affine, so the metric is formed from one Jinv and
// cell volume per element (JcV = Jinv_e (vol_e W), cV = vol_e W with W the
// node's LGL weight product) instead of being read per node.
// VISC: artificial viscosity: the flux rows carry -mu grad U and -mu/Pr grad
// theta' (mu = rho visc) from a.grad, on both sides of every face.
// MORT: the mesh has 2:1 mortar sides, whose surface terms come from the
// stage's mortar pre-pass (a separate instantiation, so that conforming
// meshes run the kernel without the test).
template <typename R, typename S, int NXc, int NYc, int NZc, bool O32, bool DIN, bool AFF, bool VISC,
          bool MORT = false>
__device__ __forceinline__ void rk_stage_body(const DgDev& d, const RkArgs<S>& a) {
    const int NX = NXc ? NXc : d.npx, NY = NYc ? NYc : d.npy, NZ = NZc ? NZc : d.npz;
    const int NP = NX * NY * NZ;
    const int NFN = 2 * (NX * NY + NX * NZ + NY * NZ);
    // a compile-time shape of more than 128 nodes runs one element per
    // work-group (the launcher guarantees epb = 1), so the element index is
    // uniform across the work-group and lives in scalar registers
    constexpr bool ONE = NXc * NYc * NZc > 128;
    const int epb = ONE ? 1 : a.epb;
    const int le = ONE ? 0 : threadIdx.x / NP, q = threadIdx.x - le * NP;
    // XCD-contiguous element mapping: work-group b runs on XCD b % 8
    const uint32_t per_xcd = (gridDim.x + 7) / 8;
    const uint32_t bmap = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    const uint32_t e = bmap * epb + le;
    const bool active = (ONE ? (int)threadIdx.x < NP : le < epb) && e < d.nelem;
    extern __shared__ unsigned char lds_raw[];
    const int per_elem = rk_lds_per_element(NX, NY, NZ);
    R* G = reinterpret_cast<R*>(lds_raw) + (size_t)le * per_elem;
    const int NPP = rk_prim_stride(NP);
    R* FL = G + kPrimRows * NPP;                     // face fluxes [5][NFN]
    R* Dl = reinterpret_cast<R*>(lds_raw) + (size_t)epb * per_elem;  // D0, D1, D2
    const int i = q / (NY * NZ), j = (q / NZ) % NY, k = q % NZ;
    const uint64_t n = (uint64_t)e * NP + q;
    const uint64_t nb = d.nb;
    const R T0 = (R)d.T0, gam = (R)d.gamma, gR = (R)(d.gamma * d.R);

    for (int t = threadIdx.x; t < NX * NX + NY * NY + NZ * NZ; t += blockDim.x) {
        const int s1 = NX * NX, s2 = s1 + NY * NY;
        Dl[t] = (R)(t < s1 ? d.D[0][t] : t < s2 ? d.D[1][t - s1] : d.D[2][t - s2]);
    }


    Prim<R> P{};
    // the part of the stage output known before the residual:
    // b din (+ q on the last stage); the output is pre + b dt L(q + din)
    double pre[5] = {0, 0, 0, 0, 0};
    double tpre[5] = {0, 0, 0, 0, 0};  // its tangent (tangent-linear instantiation)
    R rref = 0, pref = 0, cvn = 1;
    FaceTask<R> ft{};   // global reads of this thread's first face task
    constexpr int NG = VISC ? 12 : 1;
    R gown[NG];          // this node's gradients (VISC)
    R gA[NG], gB[NG];    // the face task's own-side and far-side gradients (VISC)
    R rrefA = 0;         // the face task's own-side background density (VISC)
    const R* grd = static_cast<const R*>(a.grad);
    if (active) {
        // all of this phase's global reads first (own node, then the face
        // task), straight-line, so they are in flight together
        double base[5], din[5];
        for (int v = 0; v < 5; v++) {
            base[v] = ldn<O32>(a.q[v], n);
            din[v] = ld_delta<O32, DIN>(a.din[v], n);
        }
        rref = (R)ldn<O32>(d.rho_ref, n);
        pref = (R)ldn<O32>(d.p_ref, n);
        R J[9];
        if constexpr (AFF) {
            const double* em = d.emet + (size_t)e * 10;
            const double cv64 = em[9] * d.wnode[q];
            cvn = (R)cv64;
            for (int c = 0; c < 9; c++) J[c] = (R)(em[c] * cv64);
        } else {
            if constexpr (sizeof(R) == 4) cvn = ldn<O32>(d.cV32, n);
            else cvn = (R)ldn<O32>(d.cV, n);
            if constexpr (sizeof(R) == 4)
                for (int c = 0; c < 9; c++) J[c] = ldn<O32>(d.JcV32 + (size_t)c * nb, n);
            else
                for (int c = 0; c < 9; c++) J[c] = (R)ldn<O32>(d.JcV + (size_t)c * nb, n);
        }
        if (q < NFN) prepare_face_task<R, S, O32, DIN>(d, a, e, q, NX, NY, NZ, ft);
        if constexpr (VISC) {
            for (int c = 0; c < 12; c++) gown[c] = ldn<O32>(grd + (size_t)c * nb, n);
            if (q < NFN) {
                const uint64_t na = (uint64_t)e * NP + ft.fnode;
                for (int c = 0; c < 12; c++) {
                    gA[c] = ldn<O32>(grd + (size_t)c * nb, na);
                    gB[c] = ldn<O32>(grd + (size_t)c * nb, ft.node);
                }
                rrefA = (R)ldn<O32>(d.rho_ref, na);
            }
        }
        for (int v = 0; v < 5; v++) {
            pre[v] = a.last ? base[v] + a.b * din[v] : a.b * din[v];
            if constexpr (is_dual<R>::value) {
                const double tb = ldn<O32>(a.tq[v], n), td = ld_delta<O32, DIN>(a.tdin[v], n);
                tpre[v] = a.last ? tb + a.b * td : a.b * td;
            }
        }
        if constexpr (is_dual<R>::value) {
            P.rp = load_state<R, S, O32, DIN>(a, 0, n);
            P.m[0] = load_state<R, S, O32, DIN>(a, 1, n);
            P.m[1] = load_state<R, S, O32, DIN>(a, 2, n);
            P.m[2] = load_state<R, S, O32, DIN>(a, 3, n);
            P.tp = load_state<R, S, O32, DIN>(a, 4, n);
        } else {
            P.rp = (R)(base[0] + din[0]);
            P.m[0] = (R)(base[1] + din[1]); P.m[1] = (R)(base[2] + din[2]); P.m[2] = (R)(base[3] + din[3]);
            P.tp = (R)(base[4] + din[4]);
        }
        make_prim(P, rref, pref, T0, gam, gR, d.ppc);
        R mu = 0, imu = 0;
        if constexpr (VISC) {
            mu = (rref + P.rp) * (R)d.visc;
            imu = mu * (R)d.iPr;
        }
        // rows: rho, m_x, m_y, m_z, Theta
        for (int row = 0; row < 5; row++) {
            R f0, f1, f2, g0, g1, g2;
            flux_row(P, row, T0, f0, f1, f2);
            if constexpr (VISC) {
                if (row >= 1 && row <= 3) {
                    const int r3 = 3 * (row - 1);
                    f0 = f0 - mu * gown[r3]; f1 = f1 - mu * gown[r3 + 1]; f2 = f2 - mu * gown[r3 + 2];
                } else if (row == 4) {
                    f0 = f0 - imu * gown[9]; f1 = f1 - imu * gown[10]; f2 = f2 - imu * gown[11];
                }
            }
            contravariant(J, f0, f1, f2, g0, g1, g2);
            G[(row * 3 + 0) * NP + q] = g0;
            G[(row * 3 + 1) * NP + q] = g1;
            G[(row * 3 + 2) * NP + q] = g2;
        }
    }
    __syncthreads();

    R res[5] = {0, 0, 0, 0, 0};  // volume - surface (= -r of the reference)
    if (active) {
        const R* D0 = Dl;
        const R* D1 = Dl + NX * NX;
        const R* D2 = D1 + NY * NY;
        for (int row = 0; row < 5; row++) {
            const R* G0 = G + (row * 3 + 0) * NP;
            const R* G1 = G + (row * 3 + 1) * NP;
            const R* G2 = G + (row * 3 + 2) * NP;
            R v = 0;
            for (int t = 0; t < NX; t++) v += D0[t * NX + i] * G0[t * NY * NZ + j * NZ + k];
            for (int t = 0; t < NY; t++) v += D1[t * NY + j] * G1[i * NY * NZ + t * NZ + k];
            for (int t = 0; t < NZ; t++) v += D2[t * NZ + k] * G2[i * NY * NZ + j * NZ + t];
            res[row] = v;
        }
    }
    __syncthreads();
    if (active) {
        const R pr[kPrimRows] = {P.rp, P.m[0], P.m[1], P.m[2], P.tp, P.u[0], P.u[1], P.u[2], P.thp, P.pp};
        for (int r = 0; r < kPrimRows; r++) G[r * NPP + rk_prim_slot(q)] = pr[r];
    }
    __syncthreads();

    if (active) {
        for (int t = q; t < NFN; t += NP) {
            if (t != q) {
                prepare_face_task<R, S, O32, DIN>(d, a, e, t, NX, NY, NZ, ft);
                if constexpr (VISC) {
                    const uint64_t na = (uint64_t)e * NP + ft.fnode;
                    for (int c = 0; c < 12; c++) {
                        gA[c] = ldn<O32>(grd + (size_t)c * nb, na);
                        gB[c] = ldn<O32>(grd + (size_t)c * nb, ft.node);
                    }
                    rrefA = (R)ldn<O32>(d.rho_ref, na);
                }
            }
            const FaceTask<R>& F = ft;
            if (!((F.meta >> 4) & 1)) continue;
            if constexpr (MORT)
                if ((F.meta >> 6) & 1) continue;  // mortar side: surface term from the pre-pass
            const bool own = (F.meta >> 3) & 1;
            const bool ghost = (F.meta >> 5) & 1;
            const int fnode = F.fnode;
            const R* nv = F.nv;
            Prim<R> A{};
            const R* Pf = G + rk_prim_slot(fnode);
            A.rp = Pf[0 * NPP];
            A.m[0] = Pf[1 * NPP]; A.m[1] = Pf[2 * NPP]; A.m[2] = Pf[3 * NPP];
            A.tp = Pf[4 * NPP];
            A.u[0] = Pf[5 * NPP]; A.u[1] = Pf[6 * NPP]; A.u[2] = Pf[7 * NPP];
            A.thp = Pf[8 * NPP];
            A.pp = Pf[9 * NPP];
            A.c = r_sqrt(gR * (T0 + A.thp));
            A.umag = r_sqrt(A.u[0] * A.u[0] + A.u[1] * A.u[1] + A.u[2] * A.u[2]);
            Prim<R> Q{};
            if (!ghost) {
                Q.rp = F.rp; Q.m[0] = F.m[0]; Q.m[1] = F.m[1]; Q.m[2] = F.m[2]; Q.tp = F.tp;
                make_prim(Q, F.rref, F.pref, T0, gam, gR, d.ppc);
            } else {
                const BcTable& bt = *d.bc;
                const int pi = F.meta >> 8;
                const double uv[3] = {bt.value[1][pi], bt.value[2][pi], bt.value[3][pi]};
                ghost_state(A, nv, bt.kind[1][pi], uv, F.rref, Q);
            }
            R flux[5];
            if (own) rusanov(A, Q, nv, ghost ? R(0) : R(0.5), T0, flux);
            else rusanov(Q, A, nv, ghost ? R(0) : R(0.5), T0, flux);
            if constexpr (VISC) {
                // central average of -mu grad . n; the far side of a boundary
                // face takes the ghost gradient and the own density
                R gG[12];
                const R* gFar = gB;
                if (ghost) {
                    ghost_gradient_r(*d.bc, (int)(F.meta >> 8), gA, gG);
                    gFar = gG;
                }
                const R muA = (rrefA + A.rp) * (R)d.visc;
                const R muB = ghost ? muA : (F.rref + F.rp) * (R)d.visc;
                R vA[4], vB[4];
                viscous_normal(gA, muA, (R)d.iPr, nv, vA);
                viscous_normal(gFar, muB, (R)d.iPr, nv, vB);
                const R fI = ghost ? R(0) : R(0.5), fJ = R(1) - fI;
                const R* vO = own ? vA : vB;
                const R* vN = own ? vB : vA;
                for (int r = 0; r < 4; r++) flux[1 + r] = flux[1 + r] - (vO[r] * fI + vN[r] * fJ);
            }
            for (int v = 0; v < 5; v++) FL[v * NFN + t] = flux[v];
        }
    }
    __syncthreads();
    if (!active) return;

    for (int g = 0; g < 6; g++) {
        if (!on_face(g, i, j, k, NX, NY, NZ)) continue;
        const uint32_t meta = d.eface[(size_t)e * 6 + g].meta;
        if (!((meta >> 4) & 1)) continue;
        int fa, fb;
        face_ab(g, i, j, k, fa, fb);
        if constexpr (MORT) {
            if ((meta >> 6) & 1) {
                // 2:1 mortar side: its surface term, sign included, from the
                // stage's mortar pre-pass
                const size_t mi = ((size_t)(meta >> 8) * d.npfm + fa * (g < 2 ? NY : NZ) + fb) * 5;
                for (int v = 0; v < 5; v++) {
                    R x = (R)d.msurf[mi + v];
                    if constexpr (is_dual<R>::value) x.d[0] = d.msurf_t[mi + v];
                    res[v] += x;
                }
                continue;
            }
        }
        const int t = face_task_index(g, fa, fb, NX, NY, NZ);
        if ((meta >> 3) & 1)
            for (int v = 0; v < 5; v++) res[v] -= FL[v * NFN + t];
        else
            for (int v = 0; v < 5; v++) res[v] += FL[v * NFN + t];
    }
    // L(q) = (volume - surface)/cV + buoyancy; increments and base
    const R icv = r_div(R(1), cvn);
    R Lq[5];
    for (int v = 0; v < 5; v++) Lq[v] = res[v] * icv;
    if (d.buoyancy) {
        Lq[1] += P.rp * (R)d.g[0];
        Lq[2] += P.rp * (R)d.g[1];
        Lq[3] += P.rp * (R)d.g[2];
    }
    for (int v = 0; v < 5; v++) {
        const double out = pre[v] + a.bdt * value_of(Lq[v]);
        if (a.last) stn<O32>(a.qout[v], n, out);
        else stn<O32>(a.dout[v], n, (S)out);
        if constexpr (is_dual<R>::value) {
            const double tout = tpre[v] + a.bdt * Lq[v].d[0];
            if (a.last) stn<O32>(a.tqout[v], n, tout);
            else stn<O32>(a.tdout[v], n, tout);
        }
    }
    if (a.check) {
        bool bad = false;
        for (int v = 0; v < 5; v++) bad |= !isfinite(value_of(Lq[v]));
        if (bad) *a.flag = 1;
    }
}

// ---------------------------------------------------------------------------
// Strong gradients of U and theta' at the stage state Q = q + din of the
// method-of-lines form, as the split step's gradf<true> forms them (surface
// lift fN (x) (cds(X) - X) over the faces through the node in face-group
// order, then (Jin grad psi_s)(x_t) (x) X_s over the node's x line and its y
// and z lines without the node, then / cV), in the residual type R.  Ghost
// traces: U by its boundary condition (ghost_state), theta' a copy.
template <typename S>
struct MolGradIn {
    const double* q[5];
    const S* din[5];
};

template <typename R, typename S, int NXc, int NYc, int NZc, bool DIN, bool AFF>
__global__ void __launch_bounds__(512) mol_grad_kernel(DgDev d, MolGradIn<S> in, R* __restrict__ grad) {
    const int NX = NXc ? NXc : d.npx, NY = NYc ? NYc : d.npy, NZ = NZc ? NZc : d.npz;
    const int NP = NX * NY * NZ;
    const int epb = blockDim.x / NP;
    const int le = threadIdx.x / NP, q = threadIdx.x - le * NP;
    const uint32_t e = blockIdx.x * epb + le;
    const bool active = le < epb && e < d.nelem;
    extern __shared__ unsigned char lds_raw[];
    R* Pl = reinterpret_cast<R*>(lds_raw) + (size_t)le * 4 * NP;   // U_x U_y U_z theta'
    const int i = q / (NY * NZ), j = (q / NZ) % NY, k = q % NZ;
    const uint64_t n = (uint64_t)e * NP + q;
    const R T0 = (R)d.T0;
    auto prims = [&](uint64_t m, R* u, R& th) {
        R Qv[5];
        for (int v = 0; v < 5; v++) Qv[v] = (R)(in.q[v][m] + (DIN ? (double)in.din[v][m] : 0.0));
        const R rho = (R)d.rho_ref[m] + Qv[0];
        const R ir = r_div(R(1), rho);
        u[0] = Qv[1] * ir; u[1] = Qv[2] * ir; u[2] = Qv[3] * ir;
        th = (Qv[4] - T0 * Qv[0]) * ir;
    };
    R u[3] = {0, 0, 0}, th = 0;
    if (active) {
        prims(n, u, th);
        Pl[q] = u[0]; Pl[NP + q] = u[1]; Pl[2 * NP + q] = u[2]; Pl[3 * NP + q] = th;
    }
    __syncthreads();
    if (!active) return;

    R gu[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, gt[3] = {0, 0, 0};
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
            for (int c = 0; c < 9; c++) gu[c] += (R)mg[c];
            for (int c = 0; c < 3; c++) gt[c] += (R)mg[9 + c];
            continue;
        }
        const double wgt = face_weight(d, own ? g : og, a, b);
        const R nv[3] = {(R)(d.gFN[3 * (size_t)ef.facet] * wgt), (R)(d.gFN[3 * (size_t)ef.facet + 1] * wgt),
                         (R)(d.gFN[3 * (size_t)ef.facet + 2] * wgt)};
        R ou[3], oth;
        if (!ghost) {
            prims(face_node(ef.other, og, a, b, NX, NY, NZ), ou, oth);
        } else {
            const BcTable& bt = *d.bc;
            const int pi = ef.meta >> 8;
            const int kind = bt.kind[1][pi];
            if (kind == BC_SYMMETRY) sym_r(u, nv, ou);
            else if (kind == BC_DIRICHLET) { ou[0] = (R)bt.value[1][pi]; ou[1] = (R)bt.value[2][pi]; ou[2] = (R)bt.value[3][pi]; }
            else { ou[0] = u[0]; ou[1] = u[1]; ou[2] = u[2]; }
            oth = th;
        }
        const R fI = ghost ? R(0) : R(0.5), fJ = R(1) - fI;
        R du[3];
        for (int c = 0; c < 3; c++) du[c] = (own ? u[c] * fI + ou[c] * fJ : ou[c] * fI + u[c] * fJ) - u[c];
        const R dth = (own ? th * fI + oth * fJ : oth * fI + th * fJ) - th;
        const R sg = own ? R(1) : R(-1);
        for (int aa = 0; aa < 3; aa++) {
            for (int c = 0; c < 3; c++) gu[3 * aa + c] += sg * (nv[aa] * du[c]);
            gt[aa] += sg * (nv[aa] * dth);
        }
    }
    R J[9], cv;
    if constexpr (AFF) {
        const double* em = d.emet + (size_t)e * 10;
        const double cv64 = em[9] * d.wnode[q];
        cv = (R)cv64;
        for (int c = 0; c < 9; c++) J[c] = (R)(em[c] * cv64);
    } else {
        cv = (R)d.cV[n];
        for (int c = 0; c < 9; c++) J[c] = (R)d.JcV[(size_t)c * d.nb + n];
    }
    auto add = [&](int src, R w0, R w1, R w2) {
        const R p0 = Pl[src], p1 = Pl[NP + src], p2 = Pl[2 * NP + src], p3 = Pl[3 * NP + src];
        gu[0] += w0 * p0; gu[1] += w0 * p1; gu[2] += w0 * p2;
        gu[3] += w1 * p0; gu[4] += w1 * p1; gu[5] += w1 * p2;
        gu[6] += w2 * p0; gu[7] += w2 * p1; gu[8] += w2 * p2;
        gt[0] += w0 * p3; gt[1] += w1 * p3; gt[2] += w2 * p3;
    };
    const int syz = NY * NZ;
    for (int t = 0; t < NX; t++) {
        const int src = t * syz + j * NZ + k;
        if (t == i) {
            const R d0 = (R)d.D[0][i * NX + i], d1 = (R)d.D[1][j * NY + j], d2 = (R)d.D[2][k * NZ + k];
            add(src, J[0] * d0 + J[3] * d1 + J[5] * d2, J[6] * d0 + J[1] * d1 + J[4] * d2,
                J[8] * d0 + J[7] * d1 + J[2] * d2);
        } else {
            const R D = (R)d.D[0][i * NX + t];
            add(src, J[0] * D, J[6] * D, J[8] * D);
        }
    }
    for (int t = 0; t < NY; t++) {
        if (t == j) continue;
        const R D = (R)d.D[1][j * NY + t];
        add(i * syz + t * NZ + k, J[3] * D, J[1] * D, J[7] * D);
    }
    for (int t = 0; t < NZ; t++) {
        if (t == k) continue;
        const R D = (R)d.D[2][k * NZ + t];
        add(i * syz + j * NZ + t, J[5] * D, J[4] * D, J[2] * D);
    }
    const R icv = r_div(R(1), cv);
    for (int c = 0; c < 9; c++) grad[(size_t)c * d.nb + n] = gu[c] * icv;
    for (int c = 0; c < 3; c++) grad[(size_t)(9 + c) * d.nb + n] = gt[c] * icv;
}

// 2:1 mortar faces of the method-of-lines gradient (as the split step's:
// cds of U and theta' at the refined face nodes with the coarse side
// interpolated by psiRef, the refined node's lift fN (x) (face - node), the
// face values gathered by psiCor for the coarse side's lift), in R, stored
// in mgrad as double; mphi 12-15 hold the face values.
template <typename R, typename S, bool DIN>
__global__ void __launch_bounds__(256) mol_mortar_grad_kernel(DgDev d, MolGradIn<S> in) {
    const uint32_t f = blockIdx.x;
    const int t = threadIdx.x;
    const MortarFace mf = d.mort[f];
    const bool act = t < mf.n0 * mf.n1;
    const int ao = t / mf.n1, bo = t - ao * mf.n1;
    const int np = d.npfm;
    const R T0 = (R)d.T0;
    auto vals = [&](uint64_t m, R* v) {
        R Qv[5];
        for (int q = 0; q < 5; q++) Qv[q] = (R)(in.q[q][m] + (DIN ? (double)in.din[q][m] : 0.0));
        const R ir = r_div(R(1), (R)d.rho_ref[m] + Qv[0]);
        v[0] = Qv[1] * ir; v[1] = Qv[2] * ir; v[2] = Qv[3] * ir;
        v[3] = (Qv[4] - T0 * Qv[0]) * ir;
    };
    extern __shared__ unsigned char mortar_grad_lds[];
    R* L = reinterpret_cast<R*>(mortar_grad_lds);   // [4][npfm]
    if (act) {
        R v[4];
        vals(face_node(mf.coarse, mf.side, ao, bo, d.npx, d.npy, d.npz), v);
        for (int q = 0; q < 4; q++) L[q * np + t] = v[q];
    }
    __syncthreads();
    if (!act) return;
    const size_t ft = (size_t)f * np + t;
    R vf[4], vc[4] = {0, 0, 0, 0};
    vals(d.mort_fine[ft], vf);
    const int d0 = mf.side < 4 ? 0 : 1, d1 = mf.side < 2 ? 1 : 2;
    const double* R0 = d.psiRef[d0] + (size_t)mf.h0 * mf.n0 * mf.n0;
    const double* R1 = d.psiRef[d1] + (size_t)mf.h1 * mf.n1 * mf.n1;
    for (int an = 0; an < mf.n0; an++)
        for (int bn = 0; bn < mf.n1; bn++) {
            const R w = (R)(R0[an * mf.n0 + ao] * R1[bn * mf.n1 + bo]);
            for (int q = 0; q < 4; q++) vc[q] += w * L[q * np + an * mf.n1 + bn];
        }
    R fF[4];
    for (int q = 0; q < 4; q++) {
        fF[q] = (vf[q] + vc[q]) * R(0.5);
        d.mphi[ft * 16 + 12 + q] = (double)fF[q];
    }
    const double wk = d.mort_w[ft];
    const R nv[3] = {(R)(mf.nf[0] * wk), (R)(mf.nf[1] * wk), (R)(mf.nf[2] * wk)};
    const R sg = mf.fine_owner ? R(1) : R(-1);
    double* mg = d.mgrad + (size_t)d.mort_fms[ft] * 12;
    for (int aa = 0; aa < 3; aa++) {
        for (int c = 0; c < 3; c++) mg[3 * aa + c] = (double)(sg * (nv[aa] * (fF[c] - vf[c])));
        mg[9 + aa] = (double)(sg * (nv[aa] * (fF[3] - vf[3])));
    }
}

template <typename R, typename S, bool DIN>
__global__ void __launch_bounds__(256) mol_mortar_grad_gather_kernel(DgDev d, MolGradIn<S> in) {
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
    const R T0 = (R)d.T0;
    R Qv[5];
    for (int q = 0; q < 5; q++) Qv[q] = (R)(in.q[q][nc] + (DIN ? (double)in.din[q][nc] : 0.0));
    const R ir = r_div(R(1), (R)d.rho_ref[nc] + Qv[0]);
    const R xc[4] = {Qv[1] * ir, Qv[2] * ir, Qv[3] * ir, (Qv[4] - T0 * Qv[0]) * ir};
    R out[12];
    for (int q = 0; q < 12; q++) out[q] = R(0);
    for (int ch = 0; ch < nch; ch++) {
        const uint32_t f = cs[4 + ch];
        const MortarFace mf = d.mort[f];
        const double* C0 = d.psiCor[d0] + (size_t)mf.h0 * mf.n0 * mf.n0;
        const double* C1 = d.psiCor[d1] + (size_t)mf.h1 * mf.n1 * mf.n1;
        const double* ph = d.mphi + (size_t)f * d.npfm * 16;
        R G[4] = {0, 0, 0, 0};
        for (int ao = 0; ao < mf.n0; ao++)
            for (int bo = 0; bo < mf.n1; bo++) {
                const R w = (R)(C0[ao * mf.n0 + an] * C1[bo * mf.n1 + bn]);
                for (int q = 0; q < 4; q++) G[q] += w * (R)ph[(ao * mf.n1 + bo) * 16 + 12 + q];
            }
        const double wk = d.mort_w[(size_t)f * d.npfm + t];
        const R nv[3] = {(R)(mf.nf[0] * wk), (R)(mf.nf[1] * wk), (R)(mf.nf[2] * wk)};
        const R sg = mf.fine_owner ? R(-1) : R(1);
        for (int aa = 0; aa < 3; aa++) {
            for (int c = 0; c < 3; c++) out[3 * aa + c] += sg * (nv[aa] * (G[c] - xc[c]));
            out[9 + aa] += sg * (nv[aa] * (G[3] - xc[3]));
        }
    }
    double* mg = d.mgrad + ((size_t)slot * d.npfm + t) * 12;
    for (int q = 0; q < 12; q++) mg[q] = (double)out[q];
}

template <typename R, typename S, int X, int Y, int Z>
static hipError_t launch_mol_gradient_t(const DgDev& d, const MolGradIn<S>& in, R* grad, hipStream_t st) {
    const int NP = d.np;
    if (NP > 512) return hipErrorInvalidValue;
    const int epb = NP >= 256 ? 1 : 256 / NP;
    const int threads = (epb * NP + 63) / 64 * 64;
    const unsigned blocks = (d.nelem + epb - 1) / epb;
    const size_t lds = (size_t)epb * 4 * NP * sizeof(R);
    const bool din = in.din[0] != nullptr;
    if (d.affine) {
        if (din) hipLaunchKernelGGL((mol_grad_kernel<R, S, X, Y, Z, true, true>), dim3(blocks), dim3(threads), lds, st, d, in, grad);
        else hipLaunchKernelGGL((mol_grad_kernel<R, S, X, Y, Z, false, true>), dim3(blocks), dim3(threads), lds, st, d, in, grad);
    } else {
        if (din) hipLaunchKernelGGL((mol_grad_kernel<R, S, X, Y, Z, true, false>), dim3(blocks), dim3(threads), lds, st, d, in, grad);
        else hipLaunchKernelGGL((mol_grad_kernel<R, S, X, Y, Z, false, false>), dim3(blocks), dim3(threads), lds, st, d, in, grad);
    }
    return hipGetLastError();
}

template <typename R, typename S>
hipError_t launch_mol_gradient(const DgDev& d, const double* const q[5], const S* const din[5], R* grad,
                               hipStream_t st) {
    if (d.nmort) {
        if (!d.mgrad || d.npfm > 256) return hipErrorInvalidValue;
        MolGradIn<S> mi{};
        for (int v = 0; v < 5; v++) { mi.q[v] = q[v]; mi.din[v] = din[v]; }
        const int threads = (d.npfm + 63) / 64 * 64;
        const size_t lds = (size_t)4 * d.npfm * sizeof(R);
        if (din[0]) {
            hipLaunchKernelGGL((mol_mortar_grad_kernel<R, S, true>), dim3(d.nmort), dim3(threads), lds, st, d, mi);
            if (d.ncslot)
                hipLaunchKernelGGL((mol_mortar_grad_gather_kernel<R, S, true>), dim3(d.ncslot), dim3(threads), 0, st,
                                   d, mi);
        } else {
            hipLaunchKernelGGL((mol_mortar_grad_kernel<R, S, false>), dim3(d.nmort), dim3(threads), lds, st, d, mi);
            if (d.ncslot)
                hipLaunchKernelGGL((mol_mortar_grad_gather_kernel<R, S, false>), dim3(d.ncslot), dim3(threads), 0,
                                   st, d, mi);
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    MolGradIn<S> in{};
    for (int v = 0; v < 5; v++) { in.q[v] = q[v]; in.din[v] = din[v]; }
    if (d.npx == 8 && d.npy == 8 && d.npz == 8) return launch_mol_gradient_t<R, S, 8, 8, 8>(d, in, grad, st);
    if (d.npx == 5 && d.npy == 1 && d.npz == 5) return launch_mol_gradient_t<R, S, 5, 1, 5>(d, in, grad, st);
    if (d.npx == 4 && d.npy == 4 && d.npz == 4) return launch_mol_gradient_t<R, S, 4, 4, 4>(d, in, grad, st);
    return launch_mol_gradient_t<R, S, 0, 0, 0>(d, in, grad, st);
}

template hipError_t launch_mol_gradient<double, double>(const DgDev&, const double* const[5], const double* const[5],
                                                        double*, hipStream_t);
template hipError_t launch_mol_gradient<float, float>(const DgDev&, const double* const[5], const float* const[5],
                                                      float*, hipStream_t);

template <typename R, typename S, int X, int Y, int Z, bool DIN, bool AFF, bool VISC, bool MORT = false>
__global__ void __launch_bounds__(512) rk_stage_kernel(DgDev d, RkArgs<S> a) {
    rk_stage_body<R, S, X, Y, Z, false, DIN, AFF, VISC, MORT>(d, a);
}

// Compile-time element shape (X x Y x Z nodes), node offsets in 32 bits,
// registers capped at W waves per SIMD.  Measured on the 64x64x32 N=7
// workload, FP32 runs fastest at 6 (three work-groups per CU) and FP64 at 4
// (two work-groups per CU).  Shapes of up to 128 nodes hold several elements
// per work-group, whose element index is not uniform and costs registers, so
// they take one wave less (FP64 at 4^3 nodes two less: at 3 waves it spills
// 56 bytes per lane and ran 8% slower than at 2); the viscous variant's extra
// gradient state takes two less.  FP32 at 6^3 nodes (N = 5) spills a few
// dwords at 6 and runs 5-6% faster at 5.  Measured and not taken: FP32 4^3
// at 4 (2% slower than 5), FP64 5^3 at 2 (18% slower than 3).
template <typename R, int X, int Y, int Z, bool VISC>
constexpr int rk_waves_per_simd() {
    constexpr int np = X * Y * Z;
    if constexpr (sizeof(R) == 4) return VISC ? 4 : (np > 128 ? (np == 216 ? 5 : 6) : 5);
    return VISC ? 2 : (np > 128 ? 4 : np > 64 ? 3 : 2);
}

template <typename R, typename S, int X, int Y, int Z, int W, bool DIN, bool AFF, bool VISC, bool MORT = false>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(W)))
rk_stage_kernel_fx(DgDev d, RkArgs<S> a) {
    rk_stage_body<R, S, X, Y, Z, true, DIN, AFF, VISC, MORT>(d, a);
}

template <typename R, typename S, int X, int Y, int Z, bool DIN, bool AFF, bool VISC, bool MORT = false>
static void launch_rk_kernel(const DgDev& d, const RkArgs<S>& a, unsigned blocks, int threads, size_t lds,
                             hipStream_t st) {
    constexpr int W = rk_waves_per_simd<R, X, Y, Z, VISC>();
    if constexpr (X > 0 && !is_dual<R>::value) {
        if (d.nb * sizeof(double) < (1ull << 32)) {
            hipLaunchKernelGGL((rk_stage_kernel_fx<R, S, X, Y, Z, W, DIN, AFF, VISC, MORT>), dim3(blocks),
                               dim3(threads), lds, st, d, a);
            return;
        }
    }
    hipLaunchKernelGGL((rk_stage_kernel<R, S, X, Y, Z, DIN, AFF, VISC, MORT>), dim3(blocks), dim3(threads), lds, st,
                       d, a);
}

template <typename R, typename S, int X, int Y, int Z, bool DIN>
static void launch_rk_metric(const DgDev& d, const RkArgs<S>& a, unsigned blocks, int threads, size_t lds,
                             hipStream_t st) {
    // the tangent-linear instantiation has no viscous variant (the
    // derivative entry points refuse diffusion)
    {
        // 2:1 non-conforming meshes (inviscid)
        if (d.nmort) {
            if constexpr (!is_dual<R>::value) {
                if (a.grad) {
                    if (d.affine) launch_rk_kernel<R, S, X, Y, Z, DIN, true, true, true>(d, a, blocks, threads, lds, st);
                    else launch_rk_kernel<R, S, X, Y, Z, DIN, false, true, true>(d, a, blocks, threads, lds, st);
                    return;
                }
            }
            if (d.affine) launch_rk_kernel<R, S, X, Y, Z, DIN, true, false, true>(d, a, blocks, threads, lds, st);
            else launch_rk_kernel<R, S, X, Y, Z, DIN, false, false, true>(d, a, blocks, threads, lds, st);
            return;
        }
    }
    if constexpr (!is_dual<R>::value) {
        if (a.grad) {
            if (d.affine) launch_rk_kernel<R, S, X, Y, Z, DIN, true, true>(d, a, blocks, threads, lds, st);
            else launch_rk_kernel<R, S, X, Y, Z, DIN, false, true>(d, a, blocks, threads, lds, st);
            return;
        }
    }
    if (d.affine) launch_rk_kernel<R, S, X, Y, Z, DIN, true, false>(d, a, blocks, threads, lds, st);
    else launch_rk_kernel<R, S, X, Y, Z, DIN, false, false>(d, a, blocks, threads, lds, st);
}

template <typename R, typename S, int X, int Y, int Z>
static hipError_t launch_rk_t(const DgDev& d, const RkArgs<S>& a_in, hipStream_t st) {
    RkArgs<S> a = a_in;
    const int NP = d.np;
    if (NP > 512) return hipErrorInvalidValue;
    // elements per work-group: fill 256 threads, as far as the LDS allows
    int epb = NP >= 256 ? 1 : 256 / NP;
    while (epb > 1 && rk_lds_bytes(d.npx, d.npy, d.npz, epb, sizeof(R)) > 65536) epb--;
    int threads = (epb * NP + 63) / 64 * 64;
    unsigned blocks = (d.nelem + epb - 1) / epb;
    blocks = (blocks + 7) / 8 * 8;  // whole XCD chunks
    const size_t lds = rk_lds_bytes(d.npx, d.npy, d.npz, epb, sizeof(R));
    if (lds > 65536) return hipErrorInvalidValue;
    if (X * Y * Z > 128 && epb != 1) return hipErrorInvalidValue;  // the body assumes one element
    a.epb = epb;
    if (a.din[0]) launch_rk_metric<R, S, X, Y, Z, true>(d, a, blocks, threads, lds, st);
    else launch_rk_metric<R, S, X, Y, Z, false>(d, a, blocks, threads, lds, st);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// 2:1 non-conforming faces of the method-of-lines stage (field.h:2018-2248,
// 2927-2943, 3085-3117).  On a mortar face the coarse side's node values
// reach the refined face's nodes through psiRef (scatter): the normal
// fluxes, the state, |U| and c are interpolated, not recomputed from an
// interpolated state.  The Rusanov flux at the refined face's nodes then
// goes to the refined element as it is and to the coarse element's side
// nodes through psiCor (gather).  Both contributions are linear in the face
// normal, so the flux is formed per unit face-node weight (phi) and
// weighted where it is used.
//
// mortar_flux_kernel: one work-group per mortar face, one thread per face
// node: the coarse side's node values once per node into LDS, then per
// refined face node their interpolation, phi, and the refined side's
// surface term.  On dual numbers (tangent-linear stages) the tangents go to
// mphi_t / msurf_t.
template <typename R, typename S, bool DIN>
__global__ void __launch_bounds__(256) mortar_flux_kernel(DgDev d, RkArgs<S> a) {
    const uint32_t f = blockIdx.x;
    const int t = threadIdx.x;
    const MortarFace mf = d.mort[f];
    const bool act = t < mf.n0 * mf.n1;
    const int NX = d.npx, NY = d.npy, NZ = d.npz;
    const int ao = t / mf.n1, bo = t - ao * mf.n1;
    const size_t ft = (size_t)f * d.npfm + t;
    const R T0 = (R)d.T0, gam = (R)d.gamma, gR = (R)(d.gamma * d.R);
    const R nv[3] = {(R)mf.nf[0], (R)mf.nf[1], (R)mf.nf[2]};
    auto load_prim = [&](uint64_t n, Prim<R>& P) {
        P.rp = load_state<R, S, false, DIN>(a, 0, n);
        P.m[0] = load_state<R, S, false, DIN>(a, 1, n);
        P.m[1] = load_state<R, S, false, DIN>(a, 2, n);
        P.m[2] = load_state<R, S, false, DIN>(a, 3, n);
        P.tp = load_state<R, S, false, DIN>(a, 4, n);
        make_prim(P, (R)d.rho_ref[n], (R)d.p_ref[n], T0, gam, gR, d.ppc);
    };
    extern __shared__ unsigned char mortar_lds[];
    R* L = reinterpret_cast<R*>(mortar_lds);   // [12][npfm]
    const int np = d.npfm;
    // artificial viscosity: the normal fluxes carry -mu grad U . n and
    // -mu/Pr grad theta' . n (mu = rho visc), as the face flux rows do
    const R* grd = static_cast<const R*>(a.grad);
    auto viscous = [&](uint64_t n, const Prim<R>& P, R* v) {
        if constexpr (!is_dual<R>::value) {
            if (!grd) return;
            R gg[12], vis[4];
            for (int c = 0; c < 12; c++) gg[c] = grd[(size_t)c * d.nb + n];
            const R mu = ((R)d.rho_ref[n] + P.rp) * (R)d.visc;
            viscous_normal(gg, mu, (R)d.iPr, nv, vis);
            for (int r = 0; r < 4; r++) v[1 + r] = v[1 + r] - vis[r];
        }
    };
    if (act) {
        Prim<R> P{};
        const uint64_t ncn = face_node(mf.coarse, mf.side, ao, bo, NX, NY, NZ);
        load_prim(ncn, P);
        R v[12];
        mortar_node_values(P, nv, T0, v);
        viscous(ncn, P, v);
        for (int q = 0; q < 12; q++) L[q * np + t] = v[q];
    }
    __syncthreads();
    if (!act) return;
    Prim<R> Pf{};
    load_prim(d.mort_fine[ft], Pf);
    R vf[12], vc[12];
    mortar_node_values(Pf, nv, T0, vf);
    viscous(d.mort_fine[ft], Pf, vf);
    for (int q = 0; q < 12; q++) vc[q] = R(0);
    const int d0 = mf.side < 4 ? 0 : 1, d1 = mf.side < 2 ? 1 : 2;
    const double* R0 = d.psiRef[d0] + (size_t)mf.h0 * mf.n0 * mf.n0;
    const double* R1 = d.psiRef[d1] + (size_t)mf.h1 * mf.n1 * mf.n1;
    for (int an = 0; an < mf.n0; an++)
        for (int bn = 0; bn < mf.n1; bn++) {
            const R w = (R)(R0[an * mf.n0 + ao] * R1[bn * mf.n1 + bo]);
            const int s2 = an * mf.n1 + bn;
            for (int q = 0; q < 12; q++) vc[q] += w * L[q * np + s2];
        }
    R phi[5];
    mortar_rusanov(vf, vc, (R)d.mort_bg[ft], nv, mf.fine_owner != 0, T0, phi);
    // the refined element: -w phi when it owns the face, +w phi otherwise
    const double wgt = d.mort_w[ft], sw = mf.fine_owner ? -wgt : wgt;
    double* ms = d.msurf + (size_t)d.mort_fms[ft] * 5;
    for (int v = 0; v < 5; v++) {
        d.mphi[ft * 5 + v] = value_of(phi[v]);
        ms[v] = sw * value_of(phi[v]);
        if constexpr (is_dual<R>::value) {
            d.mphi_t[ft * 5 + v] = phi[v].d[0];
            d.msurf_t[(size_t)d.mort_fms[ft] * 5 + v] = sw * phi[v].d[0];
        }
    }
}

// mortar_gather_kernel: one work-group per split coarse side, one thread per
// side node: the sum over the side's faces of w(node) psiCor(phi), with the
// coarse element's sign.
__global__ void __launch_bounds__(256) mortar_gather_kernel(DgDev d, const double* mphi, double* msurf) {
    const uint32_t* cs = d.cslot + (size_t)blockIdx.x * 8;
    const int side = (int)cs[1], nch = (int)cs[2];
    const uint32_t slot = cs[3];
    const MortarFace m0 = d.mort[cs[4]];
    const int t = threadIdx.x;
    if (t >= m0.n0 * m0.n1) return;
    const int an = t / m0.n1, bn = t - an * m0.n1;
    const int d0 = side < 4 ? 0 : 1, d1 = side < 2 ? 1 : 2;
    double sum[5] = {0, 0, 0, 0, 0};
    for (int ch = 0; ch < nch; ch++) {
        const uint32_t f = cs[4 + ch];
        const MortarFace mf = d.mort[f];
        const double* C0 = d.psiCor[d0] + (size_t)mf.h0 * mf.n0 * mf.n0;
        const double* C1 = d.psiCor[d1] + (size_t)mf.h1 * mf.n1 * mf.n1;
        const double* ph = mphi + (size_t)f * d.npfm * 5;
        double g[5] = {0, 0, 0, 0, 0};
        for (int ao = 0; ao < mf.n0; ao++)
            for (int bo = 0; bo < mf.n1; bo++) {
                const double w = C0[ao * mf.n0 + an] * C1[bo * mf.n1 + bn];
                for (int v = 0; v < 5; v++) g[v] += w * ph[(ao * mf.n1 + bo) * 5 + v];
            }
        // the weight of this node on the refined face (the reference's
        // fN[k] at the same face-node index)
        const double wk = d.mort_w[(size_t)f * d.npfm + t];
        const double sgn = mf.fine_owner ? wk : -wk;   // the coarse element owns the face: subtract
        for (int v = 0; v < 5; v++) sum[v] += sgn * g[v];
    }
    double* ms = msurf + ((size_t)slot * d.npfm + t) * 5;
    for (int v = 0; v < 5; v++) ms[v] = sum[v];
}

template <typename R, typename S>
hipError_t launch_rk_shape(const DgDev& d, const RkArgs<S>& a, hipStream_t st);

template <typename R, typename S>
static hipError_t launch_mortar(const DgDev& d, const RkArgs<S>& a, hipStream_t st) {
    if (d.npfm > 256) return hipErrorInvalidValue;
    const int threads = (d.npfm + 63) / 64 * 64;
    const size_t lds = (size_t)12 * d.npfm * sizeof(R);
    if (a.din[0])
        hipLaunchKernelGGL((mortar_flux_kernel<R, S, true>), dim3(d.nmort), dim3(threads), lds, st, d, a);
    else
        hipLaunchKernelGGL((mortar_flux_kernel<R, S, false>), dim3(d.nmort), dim3(threads), lds, st, d, a);
    if (d.ncslot) {
        hipLaunchKernelGGL(mortar_gather_kernel, dim3(d.ncslot), dim3(threads), 0, st, d, (const double*)d.mphi,
                           d.msurf);
        if constexpr (is_dual<R>::value)
            hipLaunchKernelGGL(mortar_gather_kernel, dim3(d.ncslot), dim3(threads), 0, st, d,
                               (const double*)d.mphi_t, d.msurf_t);
    }
    return hipGetLastError();
}

template <typename R, typename S>
hipError_t launch_rk_stage(const DgDev& d, const RkArgs<S>& a, hipStream_t st) {
    if (d.nmort) {
        const hipError_t e = launch_mortar<R, S>(d, a, st);
        if (e != hipSuccess) return e;
    }
    return launch_rk_shape<R, S>(d, a, st);
}

template <typename R, typename S>
hipError_t launch_rk_shape(const DgDev& d, const RkArgs<S>& a, hipStream_t st) {
    if (d.npx == 8 && d.npy == 8 && d.npz == 8) return launch_rk_t<R, S, 8, 8, 8>(d, a, st);
    if (d.npx == 7 && d.npy == 7 && d.npz == 7) return launch_rk_t<R, S, 7, 7, 7>(d, a, st);
    if (d.npx == 6 && d.npy == 6 && d.npz == 6) return launch_rk_t<R, S, 6, 6, 6>(d, a, st);
    if (d.npx == 5 && d.npy == 5 && d.npz == 5) return launch_rk_t<R, S, 5, 5, 5>(d, a, st);
    if (d.npx == 4 && d.npy == 4 && d.npz == 4) return launch_rk_t<R, S, 4, 4, 4>(d, a, st);
    if (d.npx == 5 && d.npy == 1 && d.npz == 5) return launch_rk_t<R, S, 5, 1, 5>(d, a, st);
    return launch_rk_t<R, S, 0, 0, 0>(d, a, st);
}

template hipError_t launch_rk_stage<double, double>(const DgDev&, const RkArgs<double>&, hipStream_t);
template hipError_t launch_rk_stage<float, float>(const DgDev&, const RkArgs<float>&, hipStream_t);
template hipError_t launch_rk_stage<Dual<1>, double>(const DgDev&, const RkArgs<double>&, hipStream_t);

}  // namespace nsem
