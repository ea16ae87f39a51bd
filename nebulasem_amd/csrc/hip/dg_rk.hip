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
#include <cstdlib>

#include "dg_kernels.hpp"

// The method-of-lines path has no bit-level counterpart in the reference, so
// products and sums are fused into FMAs here (the split reference-order
// kernels keep the build-wide -ffp-contract=off).
#pragma clang fp contract(fast)

namespace nsem {

namespace {

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
__device__ __forceinline__ R r_expm1(R x);
template <>
__device__ __forceinline__ double r_expm1<double>(double x) { return expm1(x); }
template <>
__device__ __forceinline__ float r_expm1<float>(float x) { return expm1f(x); }
template <typename R>
__device__ __forceinline__ R r_log1p(R x);
template <>
__device__ __forceinline__ double r_log1p<double>(double x) { return log1p(x); }
template <>
__device__ __forceinline__ float r_log1p<float>(float x) { return log1pf(x); }
template <typename R>
__device__ __forceinline__ R r_sqrt(R x);
template <>
__device__ __forceinline__ double r_sqrt<double>(double x) { return sqrt(x); }
template <>
__device__ __forceinline__ float r_sqrt<float>(float x) { return sqrtf(x); }

template <typename R>
__device__ __forceinline__ void make_prim(Prim<R>& P, R rho_ref, R p_ref, R T0, R gam, R gR) {
    const R rho = rho_ref + P.rp;
    const R ir = R(1) / rho;
    P.u[0] = P.m[0] * ir; P.u[1] = P.m[1] * ir; P.u[2] = P.m[2] * ir;
    P.thp = (P.tp - T0 * P.rp) * ir;
    P.pp = p_ref * r_expm1<R>(gam * r_log1p<R>(P.tp / (rho_ref * T0)));
    P.c = r_sqrt<R>(gR * (T0 + P.thp));
    P.umag = r_sqrt<R>(P.u[0] * P.u[0] + P.u[1] * P.u[1] + P.u[2] * P.u[2]);
}

// SYMMETRY ghost of a vector in arithmetic type R.
template <typename R>
__device__ __forceinline__ void sym_r(const R* p, const R* n, R* r) {
    const R mn = r_sqrt<R>(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const R e0 = n[0] / mn, e1 = n[1] / mn, e2 = n[2] / mn;
    const R dn = e0 * p[0] + e1 * p[1] + e2 * p[2];
    r[0] = p[0] - e0 * dn; r[1] = p[1] - e1 * dn; r[2] = p[2] - e2 * dn;
    const R mr = r_sqrt<R>(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    if (mr <= R(1e-7)) return;
    const R s = r_sqrt<R>(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) / mr;
    r[0] *= s; r[1] *= s; r[2] *= s;
}

template <typename S>
__device__ __forceinline__ double ld_delta(const S* d, uint64_t i) { return d ? (double)d[i] : 0.0; }

}  // namespace

// Face-node task t (0 .. NFN-1) of an element: face group g, face-local
// (a, b) and the element-local node.  Groups 0,1 (k = 0, NZ-1) have NX*NY
// nodes, 2,3 (j) NX*NZ, 4,5 (i) NY*NZ, enumerated group by group in the
// (a, b) order of face_ab.
__device__ __forceinline__ void face_task(int t, int NX, int NY, int NZ, int& g, int& fa, int& fb,
                                          int& node) {
    const int FK = NX * NY, FJ = NX * NZ, FI = NY * NZ;
    int i, j, k, l;
    if (t < 2 * FK) {
        g = t >= FK; l = t - g * FK;
        fa = l / NY; fb = l - fa * NY;
        i = fa; j = fb; k = g ? NZ - 1 : 0;
    } else if (t < 2 * (FK + FJ)) {
        l = t - 2 * FK;
        const int h = l >= FJ;
        g = 2 + h; l -= h * FJ;
        fa = l / NZ; fb = l - fa * NZ;
        i = fa; j = h ? NY - 1 : 0; k = fb;
    } else {
        l = t - 2 * (FK + FJ);
        const int h = l >= FI;
        g = 4 + h; l -= h * FI;
        fa = l / NZ; fb = l - fa * NZ;
        i = h ? NX - 1 : 0; j = fa; k = fb;
    }
    node = i * NY * NZ + j * NZ + k;
}

// Inverse of face_task for a node on face group g.
__device__ __forceinline__ int face_task_index(int g, int fa, int fb, int NX, int NY, int NZ) {
    const int FK = NX * NY, FJ = NX * NZ, FI = NY * NZ;
    switch (g) {
        case 0: return fa * NY + fb;
        case 1: return FK + fa * NY + fb;
        case 2: return 2 * FK + fa * NZ + fb;
        case 3: return 2 * FK + FJ + fa * NZ + fb;
        case 4: return 2 * (FK + FJ) + fa * NZ + fb;
        default: return 2 * (FK + FJ) + FI + fa * NZ + fb;
    }
}

// Everything one face-node task reads from global memory, gathered before
// the volume term so that its latency overlaps the element's own loads:
// the face entry, the weighted facet normal and the neighbour's state (or,
// on a boundary face, the owner node's background density).
template <typename R>
struct FaceTask {
    uint32_t meta;
    int g, fa, fb, fnode;
    R nv[3];
    R rp, m[3], tp, rref, pref;
};

template <typename R, typename S>
__device__ __forceinline__ void prepare_face_task(const DgDev& d, const RkArgs<S>& a, uint32_t e, int t,
                                                  int NX, int NY, int NZ, FaceTask<R>& F) {
    face_task(t, NX, NY, NZ, F.g, F.fa, F.fb, F.fnode);
    const EFace ef = d.eface[(size_t)e * 6 + F.g];
    F.meta = ef.meta;
    if (!((ef.meta >> 4) & 1)) return;
    const bool own = (ef.meta >> 3) & 1;
    const int og = ef.meta & 7;
    const double wgt = face_weight(d, own ? F.g : og, F.fa, F.fb);
    F.nv[0] = (R)(d.gFN[3 * (size_t)ef.facet] * wgt);
    F.nv[1] = (R)(d.gFN[3 * (size_t)ef.facet + 1] * wgt);
    F.nv[2] = (R)(d.gFN[3 * (size_t)ef.facet + 2] * wgt);
    if (!((ef.meta >> 5) & 1)) {
        const uint64_t m = face_node(ef.other, og, F.fa, F.fb, NX, NY, NZ);
        F.rp = (R)(a.q[0][m] + ld_delta(a.din[0], m));
        F.m[0] = (R)(a.q[1][m] + ld_delta(a.din[1], m));
        F.m[1] = (R)(a.q[2][m] + ld_delta(a.din[2], m));
        F.m[2] = (R)(a.q[3][m] + ld_delta(a.din[3], m));
        F.tp = (R)(a.q[4][m] + ld_delta(a.din[4], m));
        F.rref = (R)d.rho_ref[m];
        F.pref = (R)d.p_ref[m];
    } else {
        F.rref = (R)d.rho_ref[(uint64_t)e * (NX * NY * NZ) + F.fnode];
        F.pref = 0;
    }
}

// Primitive rows kept in LDS for the face tasks (the owner side of each face
// node): rho', m (3), Theta', U (3), theta - T0, p'.  c and |U| are
// recomputed from them with the same operations as make_prim.
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
template <typename R, typename S, int NXc, int NYc, int NZc>
__device__ __forceinline__ void rk_stage_body(const DgDev& d, const RkArgs<S>& a) {
    const int NX = NXc ? NXc : d.npx, NY = NYc ? NYc : d.npy, NZ = NZc ? NZc : d.npz;
    const int NP = NX * NY * NZ;
    const int NFN = 2 * (NX * NY + NX * NZ + NY * NZ);
    const int epb = blockDim.x / NP;
    const int le = threadIdx.x / NP, q = threadIdx.x - le * NP;
    // XCD-contiguous element mapping: work-group b runs on XCD b % 8
    const uint32_t per_xcd = (gridDim.x + 7) / 8;
    const uint32_t bmap = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    const uint32_t e = bmap * epb + le;
    const bool active = le < epb && e < d.nelem;
    extern __shared__ unsigned char lds_raw[];
    const int per_elem = rk_lds_per_element(NX, NY, NZ);
    R* G = reinterpret_cast<R*>(lds_raw) + (size_t)le * per_elem;
    R* FL = G + kPrimRows * NP;                      // face fluxes [5][NFN]
    R* Dl = reinterpret_cast<R*>(lds_raw) + (size_t)epb * per_elem;  // D0, D1, D2
    const int i = q / (NY * NZ), j = (q / NZ) % NY, k = q % NZ;
    const uint64_t n = (uint64_t)e * NP + q;
    const uint64_t nb = d.nb;
    const R T0 = (R)d.T0, gam = (R)d.gamma, gR = (R)(d.gamma * d.R);

    for (int t = threadIdx.x; t < NX * NX + NY * NY + NZ * NZ; t += blockDim.x) {
        const int s1 = NX * NX, s2 = s1 + NY * NY;
        Dl[t] = (R)(t < s1 ? d.D[0][t] : t < s2 ? d.D[1][t - s1] : d.D[2][t - s2]);
    }

    // global reads of this thread's first face task
    FaceTask<R> ft{};
    if (active && q < NFN) prepare_face_task<R, S>(d, a, e, q, NX, NY, NZ, ft);

    Prim<R> P{};
    // the part of the stage output known before the residual:
    // b din (+ q on the last stage); the output is pre + b dt L(q + din)
    double pre[5] = {0, 0, 0, 0, 0};
    R rref = 0, pref = 0, cvn = 1;
    if (active) {
        double base[5], din[5];
        for (int v = 0; v < 5; v++) {
            base[v] = a.q[v][n];
            din[v] = a.din[v] ? (double)a.din[v][n] : 0.0;
            pre[v] = a.last ? base[v] + a.b * din[v] : a.b * din[v];
        }
        rref = (R)d.rho_ref[n];
        pref = (R)d.p_ref[n];
        if constexpr (sizeof(R) == 4) cvn = d.cV32[n];
        else cvn = (R)d.cV[n];
        P.rp = (R)(base[0] + din[0]);
        P.m[0] = (R)(base[1] + din[1]); P.m[1] = (R)(base[2] + din[2]); P.m[2] = (R)(base[3] + din[3]);
        P.tp = (R)(base[4] + din[4]);
        make_prim(P, rref, pref, T0, gam, gR);
        R J[9];
        if constexpr (sizeof(R) == 4)
            for (int c = 0; c < 9; c++) J[c] = d.JcV32[(size_t)c * nb + n];
        else
            for (int c = 0; c < 9; c++) J[c] = (R)d.JcV[(size_t)c * nb + n];
        // rows: rho, m_x, m_y, m_z, Theta
        for (int row = 0; row < 5; row++) {
            R f0, f1, f2;
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
            G[(row * 3 + 0) * NP + q] = f0 * J[0] + f1 * J[6] + f2 * J[8];
            G[(row * 3 + 1) * NP + q] = f0 * J[3] + f1 * J[1] + f2 * J[7];
            G[(row * 3 + 2) * NP + q] = f0 * J[5] + f1 * J[4] + f2 * J[2];
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
        for (int r = 0; r < kPrimRows; r++) G[r * NP + q] = pr[r];
    }
    __syncthreads();

    if (active) {
        for (int t = q; t < NFN; t += NP) {
            if (t != q) prepare_face_task<R, S>(d, a, e, t, NX, NY, NZ, ft);
            const FaceTask<R>& F = ft;
            if (!((F.meta >> 4) & 1)) continue;
            const bool own = (F.meta >> 3) & 1;
            const bool ghost = (F.meta >> 5) & 1;
            const int fnode = F.fnode;
            const R* nv = F.nv;
            Prim<R> A{};
            A.rp = G[0 * NP + fnode];
            A.m[0] = G[1 * NP + fnode]; A.m[1] = G[2 * NP + fnode]; A.m[2] = G[3 * NP + fnode];
            A.tp = G[4 * NP + fnode];
            A.u[0] = G[5 * NP + fnode]; A.u[1] = G[6 * NP + fnode]; A.u[2] = G[7 * NP + fnode];
            A.thp = G[8 * NP + fnode];
            A.pp = G[9 * NP + fnode];
            A.c = r_sqrt<R>(gR * (T0 + A.thp));
            A.umag = r_sqrt<R>(A.u[0] * A.u[0] + A.u[1] * A.u[1] + A.u[2] * A.u[2]);
            Prim<R> Q{};
            if (!ghost) {
                Q.rp = F.rp; Q.m[0] = F.m[0]; Q.m[1] = F.m[1]; Q.m[2] = F.m[2]; Q.tp = F.tp;
                make_prim(Q, F.rref, F.pref, T0, gam, gR);
            } else {
                Q = A;
                const BcTable& bt = *d.bc;
                const int pi = F.meta >> 8;
                const int kind = bt.kind[1][pi];
                if (kind == BC_SYMMETRY) {
                    sym_r(A.u, nv, Q.u);
                    sym_r(A.m, nv, Q.m);
                } else if (kind == BC_DIRICHLET) {
                    Q.u[0] = (R)bt.value[1][pi]; Q.u[1] = (R)bt.value[2][pi]; Q.u[2] = (R)bt.value[3][pi];
                    const R rho = F.rref + A.rp;
                    Q.m[0] = rho * Q.u[0]; Q.m[1] = rho * Q.u[1]; Q.m[2] = rho * Q.u[2];
                }
                Q.umag = r_sqrt<R>(Q.u[0] * Q.u[0] + Q.u[1] * Q.u[1] + Q.u[2] * Q.u[2]);
            }
            const Prim<R>& O = own ? A : Q;
            const Prim<R>& N = own ? Q : A;
            const R fI = ghost ? R(0) : R(0.5);
            const R fJ = R(1) - fI;
            const R lam = ((O.umag * fI + N.umag * fJ) + (O.c * fI + N.c * fJ)) * R(0.5);
            const R mn = r_sqrt<R>(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
            const R en[3] = {nv[0] / mn, nv[1] / mn, nv[2] / mn};
            R flux[5];
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
            {
                const R mnO = O.m[0] * nv[0] + O.m[1] * nv[1] + O.m[2] * nv[2];
                const R mnN = N.m[0] * nv[0] + N.m[1] * nv[1] + N.m[2] * nv[2];
                flux[4] = (mnO * T0 + mnO * O.thp) * fI + (mnN * T0 + mnN * N.thp) * fJ - lam * (N.tp - O.tp) * mn;
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
        const int t = face_task_index(g, fa, fb, NX, NY, NZ);
        if ((meta >> 3) & 1)
            for (int v = 0; v < 5; v++) res[v] -= FL[v * NFN + t];
        else
            for (int v = 0; v < 5; v++) res[v] += FL[v * NFN + t];
    }
    // L(q) = (volume - surface)/cV + buoyancy; increments and base
    const R icv = R(1) / cvn;
    R Lq[5];
    for (int v = 0; v < 5; v++) Lq[v] = res[v] * icv;
    if (d.buoyancy) {
        Lq[1] += P.rp * (R)d.g[0];
        Lq[2] += P.rp * (R)d.g[1];
        Lq[3] += P.rp * (R)d.g[2];
    }
    const double dt = d.dt;
    for (int v = 0; v < 5; v++) {
        const double out = pre[v] + a.b * dt * (double)Lq[v];
        if (a.last) a.qout[v][n] = out;
        else a.dout[v][n] = (S)out;
    }
    if (a.check) {
        bool bad = false;
        for (int v = 0; v < 5; v++) bad |= !isfinite((double)Lq[v]);
        if (bad) *a.flag = 1;
    }
}

template <typename R, typename S, int X, int Y, int Z>
__global__ void __launch_bounds__(512) rk_stage_kernel(DgDev d, RkArgs<S> a) {
    rk_stage_body<R, S, X, Y, Z>(d, a);
}

// N = 7 (512 nodes, one element per work-group): registers capped so that two
// work-groups (16 waves) share a CU in both precisions.
template <typename R, typename S, int W>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(W)))
rk_stage_kernel_n7(DgDev d, RkArgs<S> a) {
    rk_stage_body<R, S, 8, 8, 8>(d, a);
}

template <typename R, typename S, int X, int Y, int Z>
static hipError_t launch_rk_t(const DgDev& d, const RkArgs<S>& a, hipStream_t st) {
    const int NP = d.np;
    if (NP > 512) return hipErrorInvalidValue;
    int epb = NP >= 256 ? 1 : 256 / NP;
    int threads = (epb * NP + 63) / 64 * 64;
    unsigned blocks = (d.nelem + epb - 1) / epb;
    blocks = (blocks + 7) / 8 * 8;  // whole XCD chunks
    const size_t lds = rk_lds_bytes(d.npx, d.npy, d.npz, epb, sizeof(R));
    if (lds > 65536) return hipErrorInvalidValue;
    if constexpr (X == 8 && Y == 8 && Z == 8) {
        static const int w = [] {
            const char* s = getenv("NSEM_RK_WAVES");
            return s ? atoi(s) : 0;
        }();
        // measured on the 64x64x32 N=7 workload: FP32 fastest at 6 waves/SIMD
        // (three work-groups per CU), FP64 at 4 (two work-groups per CU)
        if (w == 8)
            hipLaunchKernelGGL((rk_stage_kernel_n7<R, S, 8>), dim3(blocks), dim3(threads), lds, st, d, a);
        else if (w == 6 || (w == 0 && sizeof(R) == 4))
            hipLaunchKernelGGL((rk_stage_kernel_n7<R, S, 6>), dim3(blocks), dim3(threads), lds, st, d, a);
        else if (w == 5)
            hipLaunchKernelGGL((rk_stage_kernel_n7<R, S, 5>), dim3(blocks), dim3(threads), lds, st, d, a);
        else
            hipLaunchKernelGGL((rk_stage_kernel_n7<R, S, 4>), dim3(blocks), dim3(threads), lds, st, d, a);
    } else
        hipLaunchKernelGGL((rk_stage_kernel<R, S, X, Y, Z>), dim3(blocks), dim3(threads), lds, st, d, a);
    return hipGetLastError();
}

template <typename R, typename S>
hipError_t launch_rk_stage(const DgDev& d, const RkArgs<S>& a, hipStream_t st) {
    if (d.npx == 8 && d.npy == 8 && d.npz == 8) return launch_rk_t<R, S, 8, 8, 8>(d, a, st);
    if (d.npx == 5 && d.npy == 1 && d.npz == 5) return launch_rk_t<R, S, 5, 1, 5>(d, a, st);
    if (d.npx == 6 && d.npy == 6 && d.npz == 6) return launch_rk_t<R, S, 6, 6, 6>(d, a, st);
    if (d.npx == 4 && d.npy == 4 && d.npz == 4) return launch_rk_t<R, S, 4, 4, 4>(d, a, st);
    return launch_rk_t<R, S, 0, 0, 0>(d, a, st);
}

template hipError_t launch_rk_stage<double, double>(const DgDev&, const RkArgs<double>&, hipStream_t);
template hipError_t launch_rk_stage<float, float>(const DgDev&, const RkArgs<float>&, hipStream_t);

}  // namespace nsem
