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
// q+ = q + 2(d2 + dt L(q+d2))/3), backwards:
//   t3 = J(q+d2)^T lam;  ad2 = 2/3 (lam + dt t3);  acc = lam + 2/3 dt t3
//   t2 = J(q+d1)^T ad2;  ad1 = 1/4 (ad2 + dt t2);  acc += 1/4 dt t2
//   t1 = J(q)^T ad1;                               lam = acc + dt t1
// with the forward states recomputed from checkpoints stored every
// `checkpoint` steps (one segment of states held at a time).
#pragma clang fp contract(fast)

#include <hip/hip_runtime.h>

#include <algorithm>
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

template <int NXc, int NYc, int NZc>
__global__ void __launch_bounds__(512) rk_vjp_kernel(DgDev d, VjpArgs a) {
    const int NX = NXc ? NXc : d.npx, NY = NYc ? NYc : d.npy, NZ = NZc ? NZc : d.npz;
    const int NP = NX * NY * NZ;
    const int NFN = 2 * (NX * NY + NX * NZ + NY * NZ);
    const int epb = a.epb;
    const int le = threadIdx.x / NP, q = threadIdx.x - le * NP;
    const uint32_t e = blockIdx.x * epb + le;
    const bool active = le < epb && e < d.nelem;
    extern __shared__ double lds_v[];
    const int per_elem = 10 * NP + 5 * NFN;
    double* Wl = lds_v + (size_t)le * per_elem;  // w / cV      [5][NP]
    double* Ql = Wl + 5 * NP;                    // Q = q + din [5][NP]
    double* FL = Ql + 5 * NP;                    // face terms  [5][NFN]
    double* Dl = lds_v + (size_t)epb * per_elem;
    const int i = q / (NY * NZ), j = (q / NZ) % NY, k = q % NZ;
    const uint64_t n = (uint64_t)e * NP + q;
    const uint64_t nb = d.nb;
    const double T0 = d.T0, gam = d.gamma, gR = d.gamma * d.R;

    for (int t = threadIdx.x; t < NX * NX + NY * NY + NZ * NZ; t += blockDim.x) {
        const int s1 = NX * NX, s2 = s1 + NY * NY;
        Dl[t] = t < s1 ? d.D[0][t] : t < s2 ? d.D[1][t - s1] : d.D[2][t - s2];
    }
    double Qn[5] = {0, 0, 0, 0, 0}, wn[5] = {0, 0, 0, 0, 0}, J[9];
    double rref = 0, pref = 0;
    if (active) {
        const double cvn = d.cV[n];
        for (int v = 0; v < 5; v++) {
            Qn[v] = a.q[v][n] + (a.din[v] ? a.din[v][n] : 0.0);
            wn[v] = a.w[v][n];
            Wl[v * NP + q] = wn[v] / cvn;
            Ql[v * NP + q] = Qn[v];
        }
        rref = d.rho_ref[n];
        pref = d.p_ref[n];
        for (int c = 0; c < 9; c++) J[c] = d.JcV[(size_t)c * nb + n];
    }
    __syncthreads();

    double qbar[5] = {0, 0, 0, 0, 0};
    if (active) {
        const double* D0 = Dl;
        const double* Dy = Dl + NX * NX;
        const double* Dz = Dy + NY * NY;
        double fbar[5][3];
        for (int row = 0; row < 5; row++) {
            const double* W = Wl + row * NP;
            double h0 = 0, h1 = 0, h2 = 0;
            for (int t = 0; t < NX; t++) h0 += D0[i * NX + t] * W[t * NY * NZ + j * NZ + k];
            for (int t = 0; t < NY; t++) h1 += Dy[j * NY + t] * W[i * NY * NZ + t * NZ + k];
            for (int t = 0; t < NZ; t++) h2 += Dz[k * NZ + t] * W[i * NY * NZ + j * NZ + t];
            fbar[row][0] = h0 * J[0] + h1 * J[3] + h2 * J[5];
            fbar[row][1] = h0 * J[6] + h1 * J[1] + h2 * J[4];
            fbar[row][2] = h0 * J[8] + h1 * J[7] + h2 * J[2];
        }
#pragma nounroll
        for (int kk = 0; kk < 5; kk++) {
            Prim<D1> P;
            P.rp = seeded(Qn[0], kk == 0);
            P.m[0] = seeded(Qn[1], kk == 1); P.m[1] = seeded(Qn[2], kk == 2); P.m[2] = seeded(Qn[3], kk == 3);
            P.tp = seeded(Qn[4], kk == 4);
            make_prim(P, D1(rref), D1(pref), D1(T0), D1(gam), D1(gR));
            double acc = 0;
            for (int row = 0; row < 5; row++) {
                D1 f0, f1, f2;
                flux_row(P, row, D1(T0), f0, f1, f2);
                acc += fbar[row][0] * f0.d[0] + fbar[row][1] * f1.d[0] + fbar[row][2] * f2.d[0];
            }
            qbar[kk] += acc;
        }
        if (d.buoyancy) qbar[0] += wn[1] * d.g[0] + wn[2] * d.g[1] + wn[3] * d.g[2];

        for (int t = q; t < NFN; t += NP) {
            int g, fa, fb, fnode;
            face_task(t, NX, NY, NZ, g, fa, fb, fnode);
            const EFace ef = d.eface[(size_t)e * 6 + g];
            double contrib[5] = {0, 0, 0, 0, 0};
            if ((ef.meta >> 4) & 1) {
                const bool own = (ef.meta >> 3) & 1;
                const bool ghost = (ef.meta >> 5) & 1;
                const int og = ef.meta & 7;
                const double wgt = face_weight(d, own ? g : og, fa, fb);
                const D1 nv[3] = {D1(d.gFN[3 * (size_t)ef.facet] * wgt), D1(d.gFN[3 * (size_t)ef.facet + 1] * wgt),
                                  D1(d.gFN[3 * (size_t)ef.facet + 2] * wgt)};
                const uint64_t na = (uint64_t)e * NP + fnode;
                const double rrA = d.rho_ref[na], prA = d.p_ref[na];
                double QA[5], coef[5];
                for (int v = 0; v < 5; v++) QA[v] = Ql[v * NP + fnode];
                Prim<D1> B;
                int kind = 0;
                double uv[3] = {0, 0, 0};
                if (!ghost) {
                    const uint64_t m = face_node(ef.other, og, fa, fb, NX, NY, NZ);
                    const double cvm = d.cV[m];
                    B.rp = D1(a.q[0][m] + (a.din[0] ? a.din[0][m] : 0.0));
                    B.m[0] = D1(a.q[1][m] + (a.din[1] ? a.din[1][m] : 0.0));
                    B.m[1] = D1(a.q[2][m] + (a.din[2] ? a.din[2][m] : 0.0));
                    B.m[2] = D1(a.q[3][m] + (a.din[3] ? a.din[3][m] : 0.0));
                    B.tp = D1(a.q[4][m] + (a.din[4] ? a.din[4][m] : 0.0));
                    make_prim(B, D1(d.rho_ref[m]), D1(d.p_ref[m]), D1(T0), D1(gam), D1(gR));
                    const double s = own ? -1.0 : 1.0;
                    for (int v = 0; v < 5; v++) coef[v] = s * (Wl[v * NP + fnode] - a.w[v][m] / cvm);
                } else {
                    const BcTable& bt = *d.bc;
                    const int pi = ef.meta >> 8;
                    kind = bt.kind[1][pi];
                    uv[0] = bt.value[1][pi]; uv[1] = bt.value[2][pi]; uv[2] = bt.value[3][pi];
                    for (int v = 0; v < 5; v++) coef[v] = -Wl[v * NP + fnode];
                }
#pragma nounroll
                for (int kk = 0; kk < 5; kk++) {
                    Prim<D1> A;
                    A.rp = seeded(QA[0], kk == 0);
                    A.m[0] = seeded(QA[1], kk == 1); A.m[1] = seeded(QA[2], kk == 2); A.m[2] = seeded(QA[3], kk == 3);
                    A.tp = seeded(QA[4], kk == 4);
                    make_prim(A, D1(rrA), D1(prA), D1(T0), D1(gam), D1(gR));
                    Prim<D1> N;
                    if (!ghost) N = B;
                    else ghost_state(A, nv, kind, uv, D1(rrA), N);
                    D1 flux[5];
                    const D1 fI = ghost ? D1(0.0) : D1(0.5);
                    if (own) rusanov(A, N, nv, fI, D1(T0), flux);
                    else rusanov(N, A, nv, fI, D1(T0), flux);
                    double acc = 0;
                    for (int v = 0; v < 5; v++) acc += coef[v] * flux[v].d[0];
                    contrib[kk] = acc;
                }
            }
            for (int v = 0; v < 5; v++) FL[v * NFN + t] = contrib[v];
        }
    }
    __syncthreads();
    if (!active) return;
    for (int g = 0; g < 6; g++) {
        if (!on_face(g, i, j, k, NX, NY, NZ)) continue;
        if (!((d.eface[(size_t)e * 6 + g].meta >> 4) & 1)) continue;
        int fa, fb;
        face_ab(g, i, j, k, fa, fb);
        const int t = face_task_index(g, fa, fb, NX, NY, NZ);
        for (int v = 0; v < 5; v++) qbar[v] += FL[v * NFN + t];
    }
    for (int v = 0; v < 5; v++) a.t[v][n] = qbar[v];
}

__global__ void __launch_bounds__(256) adj_combine_kernel(uint64_t nb, AdjCombineArgs a) {
    const uint64_t n = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (n >= nb) return;
    for (int v = 0; v < 5; v++) {
        const double t = a.t[v][n];
        if (a.ad_out[v]) a.ad_out[v][n] = a.c1 * (a.ad_in[v][n] + a.dt * t);
        a.acc_out[v][n] = a.acc_in[v][n] + a.c2 * a.dt * t;
    }
}

int vjp_epb(int np) { return np >= 256 ? 1 : 256 / np; }

template <int X, int Y, int Z>
hipError_t launch_vjp_t(const DgDev& d, const VjpArgs& a_in, hipStream_t st) {
    VjpArgs a = a_in;
    const int NP = d.np;
    int epb = vjp_epb(NP);
    while (epb > 1 && rk_vjp_lds_bytes(d.npx, d.npy, d.npz, epb) > 65536) epb--;
    const int threads = (epb * NP + 63) / 64 * 64;
    const unsigned blocks = (d.nelem + epb - 1) / epb;
    const size_t lds = rk_vjp_lds_bytes(d.npx, d.npy, d.npz, epb);
    if (NP > 512 || lds > 65536) return hipErrorInvalidValue;
    a.epb = epb;
    hipLaunchKernelGGL((rk_vjp_kernel<X, Y, Z>), dim3(blocks), dim3(threads), lds, st, d, a);
    return hipGetLastError();
}

}  // namespace

size_t rk_vjp_lds_bytes(int nx, int ny, int nz, int epb) {
    const int np = nx * ny * nz, nfn = 2 * (nx * ny + nx * nz + ny * nz);
    return ((size_t)epb * (10 * np + 5 * nfn) + nx * nx + ny * ny + nz * nz) * sizeof(double);
}

hipError_t launch_rk_vjp(const DgDev& d, const VjpArgs& a, hipStream_t st) {
    if (d.npx == 8 && d.npy == 8 && d.npz == 8) return launch_vjp_t<8, 8, 8>(d, a, st);
    if (d.npx == 6 && d.npy == 1 && d.npz == 6) return launch_vjp_t<6, 1, 6>(d, a, st);
    if (d.npx == 5 && d.npy == 1 && d.npz == 5) return launch_vjp_t<5, 1, 5>(d, a, st);
    return launch_vjp_t<0, 0, 0>(d, a, st);
}

hipError_t launch_adj_combine(uint64_t nb, const AdjCombineArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(adj_combine_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, nb, a);
    return hipGetLastError();
}

}  // namespace nsem

using namespace nsem;

namespace {

// Five FP64 node arrays.
struct Field5 {
    DevBuf b[5];
    hipError_t alloc(uint64_t nb) {
        for (int v = 0; v < 5; v++) {
            hipError_t e = b[v].alloc(nb * sizeof(double));
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    double* p(int v) const { return b[v].as<double>(); }
};

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
    a.flag = d->flag.as<int>();
    return launch_rk_stage<double, double>(d->dev, a, d->stream);
}

struct AdjointWork {
    Field5 d1, d2, t, ad1, ad2, acc;
    hipError_t alloc(uint64_t nb) {
        hipError_t e;
        if ((e = d1.alloc(nb)) != hipSuccess || (e = d2.alloc(nb)) != hipSuccess || (e = t.alloc(nb)) != hipSuccess ||
            (e = ad1.alloc(nb)) != hipSuccess || (e = ad2.alloc(nb)) != hipSuccess || (e = acc.alloc(nb)) != hipSuccess)
            return e;
        return hipSuccess;
    }
};

void ptrs(const Field5& f, double* out[5]) { for (int v = 0; v < 5; v++) out[v] = f.p(v); }
void cptrs(const Field5& f, const double* out[5]) { for (int v = 0; v < 5; v++) out[v] = f.p(v); }

// lam <- (dq+/dq)^T lam for the step from state q.
int adjoint_step(nsem_dg* d, const double* const q[5], Field5& lam, AdjointWork& w) {
    double *d1[5], *d2[5], *t[5], *ad1[5], *ad2[5], *acc[5], *lp[5];
    ptrs(w.d1, d1); ptrs(w.d2, d2); ptrs(w.t, t); ptrs(w.ad1, ad1); ptrs(w.ad2, ad2); ptrs(w.acc, acc);
    ptrs(lam, lp);
    const uint64_t nb = d->nb;
    const double dt = d->dev.dt;
    HIP_TRY(fwd_stage(d, 0, q, nullptr, d1, nullptr));
    HIP_TRY(fwd_stage(d, 1, q, d1, d2, nullptr));
    auto vjp = [&](const double* const din[5], double* const wv[5]) -> hipError_t {
        VjpArgs a{};
        for (int v = 0; v < 5; v++) {
            a.q[v] = q[v]; a.din[v] = din ? din[v] : nullptr; a.w[v] = wv[v]; a.t[v] = t[v];
        }
        return launch_rk_vjp(d->dev, a, d->stream);
    };
    auto combine = [&](double* const ad_in[5], double* const ad_out[5], double* const acc_in[5],
                       double* const acc_out[5], double c1, double c2) -> hipError_t {
        AdjCombineArgs c{};
        for (int v = 0; v < 5; v++) {
            c.t[v] = t[v];
            c.ad_in[v] = ad_in ? ad_in[v] : nullptr;
            c.ad_out[v] = ad_out ? ad_out[v] : nullptr;
            c.acc_in[v] = acc_in[v];
            c.acc_out[v] = acc_out[v];
        }
        c.dt = dt; c.c1 = c1; c.c2 = c2;
        return launch_adj_combine(nb, c, d->stream);
    };
    HIP_TRY(vjp(d2, lp));
    HIP_TRY(combine(lp, ad2, lp, acc, 2.0 / 3.0, 2.0 / 3.0));
    HIP_TRY(vjp(d1, ad2));
    HIP_TRY(combine(ad2, ad1, acc, acc, 0.25, 0.25));
    HIP_TRY(vjp(nullptr, ad1));
    HIP_TRY(combine(nullptr, nullptr, acc, lp, 0.0, 1.0));
    return NSEM_OK;
}

int check_adjoint_support(const nsem_dg* d) {
    const DgDev& v = d->dev;
    if (v.np > 512 || rk_vjp_lds_bytes(v.npx, v.npy, v.npz, 1) > 65536)
        return nsem_fail(NSEM_EUNSUPPORTED, "element too large for the adjoint kernel");
    return NSEM_OK;
}

}  // namespace

extern "C" {

int nsem_dg_residual_vjp(nsem_dg* d, const double* w, double* t) {
    if (!d || !w || !t) return nsem_fail(NSEM_EINVAL, "null argument");
    int rc = check_adjoint_support(d);
    if (rc) return rc;
    const uint64_t nb = d->nb;
    Field5 W, T;
    HIP_TRY(W.alloc(nb));
    HIP_TRY(T.alloc(nb));
    HIP_TRY(upload5(W, w, nb, d->stream));
    VjpArgs a{};
    for (int v = 0; v < 5; v++) {
        a.q[v] = d->mq[d->mcur][v].as<double>(); a.din[v] = nullptr; a.w[v] = W.p(v); a.t[v] = T.p(v);
    }
    HIP_TRY(launch_rk_vjp(d->dev, a, d->stream));
    HIP_TRY(download5(t, T, nb, d->stream));
    return NSEM_OK;
}

int nsem_dg_residual_tlm(nsem_dg* d, const double* dq, double* L, double* dL) {
    if (!d || !dq || !L || !dL) return nsem_fail(NSEM_EINVAL, "null argument");
    const uint64_t nb = d->nb;
    Field5 TQ, Lout, TL;
    HIP_TRY(TQ.alloc(nb));
    HIP_TRY(Lout.alloc(nb));
    HIP_TRY(TL.alloc(nb));
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
    const uint64_t nb = d->nb;
    Field5 tq[2], td[2];
    for (int k = 0; k < 2; k++) { HIP_TRY(tq[k].alloc(nb)); HIP_TRY(td[k].alloc(nb)); }
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
    d->mol_current = true;
    return NSEM_OK;
}

int nsem_dg_rk3_adjoint(nsem_dg* d, int32_t nsteps, int32_t checkpoint, double* lam) {
    if (!d || !lam || nsteps < 0 || checkpoint < 1) return nsem_fail(NSEM_EINVAL, "bad argument");
    int rc = check_adjoint_support(d);
    if (rc) return rc;
    const uint64_t nb = d->nb;
    const int nseg = (nsteps + checkpoint - 1) / checkpoint;
    std::vector<Field5> ck(nseg);
    for (auto& f : ck) HIP_TRY(f.alloc(nb));
    std::vector<Field5> seg(checkpoint);
    for (auto& f : seg) HIP_TRY(f.alloc(nb));
    Field5 L;
    AdjointWork w;
    HIP_TRY(L.alloc(nb));
    HIP_TRY(w.alloc(nb));
    HIP_TRY(upload5(L, lam, nb, d->stream));
    // forward sweep, checkpoints at the start of every segment
    for (int s = 0; s < nsteps; s++) {
        if (s % checkpoint == 0) {
            const double* src[5];
            for (int v = 0; v < 5; v++) src[v] = d->mq[d->mcur][v].as<double>();
            HIP_TRY(copy5(ck[s / checkpoint], src, nb, d->stream));
        }
        if ((rc = nsem_dg_rk3_steps(d, 1, 64))) return rc;
    }
    // backward sweep: recompute each segment's states from its checkpoint
    for (int sg = nseg - 1; sg >= 0; sg--) {
        const int s0 = sg * checkpoint, s1 = std::min(nsteps, s0 + checkpoint);
        const double* src[5];
        cptrs(ck[sg], src);
        HIP_TRY(copy5(seg[0], src, nb, d->stream));
        for (int s = s0; s + 1 < s1; s++) {
            const double* q[5];
            double *d1[5], *d2[5], *qn[5];
            cptrs(seg[s - s0], q); ptrs(w.d1, d1); ptrs(w.d2, d2); ptrs(seg[s - s0 + 1], qn);
            HIP_TRY(fwd_stage(d, 0, q, nullptr, d1, nullptr));
            HIP_TRY(fwd_stage(d, 1, q, d1, d2, nullptr));
            HIP_TRY(fwd_stage(d, 2, q, d2, nullptr, qn));
        }
        for (int s = s1 - 1; s >= s0; s--) {
            const double* q[5];
            cptrs(seg[s - s0], q);
            if ((rc = adjoint_step(d, q, L, w))) return rc;
        }
    }
    HIP_TRY(download5(lam, L, nb, d->stream));
    return NSEM_OK;
}

}  // extern "C"
