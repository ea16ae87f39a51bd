// FV pressure Poisson on the device: divergence right-hand side, the stored
// Laplacian operator, and Jacobi-preconditioned CG with device-resident
// scalars (no host round trip per iteration).
//
// Reference path: lap (field.h:3571-3615) + applyImplicitBCs (2541-2582) ->
// SolveT DIAG-PCG (solve.cpp:259-280, 403-417), mul (field.h:2369-2420),
// getResidual (solve.cpp:11-35), divf FV (field.h:3470-3477).
//
// Operator storage (interior cells only, boundary faces folded):
//   diag ap[n], and per cell up to six off-diagonal couplings.  Structured
//   block grids (the mesh generator's numbering c = (i*ny + j)*nz + k) keep
//   only the coupling to the +x, +y, +z neighbour, cpl[3][n] (the matrix is
//   symmetric, so the -x coupling of c is cpl[0][c - sx]); other meshes keep
//   six (neighbour, coupling) slots.
//
// One CG iteration = three grid-stride kernels:
//   spmv_dot : Ap = A p, partial p.Ap  -> last block: alpha
//   update   : x += alpha p, r -= alpha Ap, z = r / ap,
//              partials r.z z.z x.x r.r -> last block: beta, residuals, stop flag
//   pupdate  : p = z + beta p
// Partial sums go to per-block slots; a one-block kernel reduces them in a
// fixed order right after (a kernel boundary, so every block's slot is
// visible across XCDs), so results are run-to-run identical.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fv_kernels.hpp"

namespace nsem {

namespace {

constexpr int kThreads = 256;
constexpr int kMarch = 4;   // x-planes per pass of the structured SpMV march

template <typename R>
__device__ __forceinline__ R wave_sum(R v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Block sum of NV values into slots[blockIdx.x*NV + j].
template <int NV>
__device__ void block_partials(double (&v)[NV], double* slots) {
    __shared__ double red[kThreads / 64][NV];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int j = 0; j < NV; j++) {
        double s = wave_sum(v[j]);
        if (lane == 0) red[wid][j] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int j = 0; j < NV; j++) {
            double s = 0;
            for (int w = 0; w < kThreads / 64; w++) s += red[w][j];
            slots[(size_t)blockIdx.x * NV + j] = s;
        }
}

// One block: fixed-order sum of nblocks slot rows.
template <int NV>
__device__ void reduce_slots(const double* slots, unsigned nblocks, double (&out)[NV]) {
    __shared__ double red[kThreads][NV];
    double acc[NV];
    for (int j = 0; j < NV; j++) acc[j] = 0;
    for (unsigned b = threadIdx.x; b < nblocks; b += kThreads)
        for (int j = 0; j < NV; j++) acc[j] += slots[(size_t)b * NV + j];
    for (int j = 0; j < NV; j++) red[threadIdx.x][j] = acc[j];
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int j = 0; j < NV; j++) red[threadIdx.x][j] += red[threadIdx.x + s][j];
        __syncthreads();
    }
    for (int j = 0; j < NV; j++) out[j] = red[0][j];
}

__device__ __forceinline__ double sdiv(double p, double q) { return q != 0 ? p / q : 0; }

template <typename R>
__device__ __forceinline__ R apply_row(const FvOp<R>& A, const R* __restrict__ x, uint64_t n) {
    R acc = x[n] * A.diag[n];
    if (A.structured) {
        const int64_t s[3] = {A.sx, A.sy, A.sz};
        const uint32_t nn = (uint32_t)n, nyz = (uint32_t)(A.ny * A.nz);
        const int ii = nn / nyz;
        const uint32_t rem = nn - ii * nyz;
        const int jj = rem / A.nz, kk = rem - jj * A.nz;
        const int lo[3] = {ii > 0, jj > 0, kk > 0};
        const int hi[3] = {ii + 1 < A.nx, jj + 1 < A.ny, kk + 1 < A.nz};
        for (int d = 0; d < 3; d++) {
            if (hi[d]) acc -= x[n + s[d]] * A.cpl[(size_t)d * A.n + n];
            if (lo[d]) acc -= x[n - s[d]] * A.cpl[(size_t)d * A.n + n - s[d]];
        }
    } else {
        for (int sl = 0; sl < 6; sl++) {
            const uint32_t m = A.nbr[(size_t)sl * A.n + n];
            if (m != 0xffffffffu) acc -= x[m] * A.off[(size_t)sl * A.n + n];
        }
    }
    return acc;
}

template <typename R>
__global__ void __launch_bounds__(kThreads) spmv_dot_kernel(FvOp<R> A, const R* __restrict__ p,
                                                           R* __restrict__ Ap, const CgScalars* sc,
                                                           double* slots) {
    if (sc->done) return;
    double v[1] = {0};
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t n = (uint64_t)blockIdx.x * kThreads + threadIdx.x; n < A.n; n += stride) {
        R y = apply_row(A, p, n);
        Ap[n] = y;
        v[0] += (double)p[n] * (double)y;
    }
    block_partials<1>(v, slots);
}

// Structured grids: Ap = A p marching along x.  Each thread owns one (j, k)
// column of a chunk of x-planes and keeps p and the x-coupling of the
// previous, current and next plane in registers, so every array is read from
// HBM once; the y and z neighbours are the adjacent threads' loads (L2).
// Work-groups are mapped XCD-contiguously (work-group b on XCD b % 8 takes
// logical group (b % 8) G/8 + b/8), so the column blocks of one x-chunk, and
// their y neighbours, share an L2.  Same row arithmetic as apply_row.
struct MarchGrid {
    uint32_t ncolblk, nchunk, chunk, groups;   // groups: multiple of 8
};

inline MarchGrid march_grid(int nx, int ny, int nz) {
    MarchGrid m{};
    const uint64_t ncol = (uint64_t)ny * nz;
    m.ncolblk = (uint32_t)((ncol + kThreads - 1) / kThreads);
    uint32_t nchunk = kCgSlots / std::max<uint32_t>(m.ncolblk, 1);
    nchunk = std::max<uint32_t>(1, std::min<uint32_t>(nchunk, (uint32_t)nx));
    m.chunk = (nx + nchunk - 1) / nchunk;
    m.nchunk = (nx + m.chunk - 1) / m.chunk;
    m.groups = (m.nchunk * m.ncolblk + 7) / 8 * 8;
    return m;
}

// UNI: the couplings are the three constants A.cu (bitwise equal to every
// stored coupling), so only p, ap and Ap move.
template <typename R, bool UNI>
__global__ void __launch_bounds__(kThreads) spmv_dot_march_kernel(FvOp<R> A, MarchGrid mg, const R* __restrict__ p,
                                                                 R* __restrict__ Ap, const CgScalars* sc,
                                                                 double* slots) {
    if (sc->done) return;
    double v[1] = {0};
    const uint32_t lb = (blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8;
    if (lb < mg.nchunk * mg.ncolblk) {
        const uint32_t c = lb / mg.ncolblk, cb = lb - c * mg.ncolblk;
        const uint64_t sx = (uint64_t)A.ny * A.nz;
        const uint64_t col = (uint64_t)cb * kThreads + threadIdx.x;
        if (col < sx) {
            const int j = (int)(col / A.nz), k = (int)(col - (uint64_t)j * A.nz);
            const int i0 = c * mg.chunk, i1 = min(A.nx, (int)(i0 + mg.chunk));
            const R* c0 = A.cpl;
            const R* c1 = A.cpl + A.n;
            const R* c2 = A.cpl + 2 * A.n;
            uint64_t n = (uint64_t)i0 * sx + col;
            R pm = i0 > 0 ? p[n - sx] : R(0), c0m = UNI ? A.cu[0] : (i0 > 0 ? c0[n - sx] : R(0));
            R pc = p[n];
            const bool hyp = j + 1 < A.ny, hym = j > 0, hzp = k + 1 < A.nz, hzm = k > 0;
            // kMarch planes per pass: every load of the pass is issued before
            // the first row is formed, so a wave keeps kMarch planes of reads
            // in flight (one plane at a time leaves the march latency-bound);
            // each row is then formed exactly as one plane at a time forms it
            for (int i = i0; i < i1; i += kMarch, n += (uint64_t)kMarch * sx) {
                R pn[kMarch], dg[kMarch], yp[kMarch], ym[kMarch], zp[kMarch], zm[kMarch];
                R cx[kMarch], cyp[kMarch], cym[kMarch], czp[kMarch], czm[kMarch];
#pragma unroll
                for (int u = 0; u < kMarch; u++) {
                    const uint64_t m = n + (uint64_t)u * sx;
                    const bool in = i + u < i1;
                    pn[u] = (in && i + u + 1 < A.nx) ? p[m + sx] : R(0);
                    dg[u] = in ? A.diag[m] : R(0);
                    yp[u] = (in && hyp) ? p[m + A.nz] : R(0);
                    ym[u] = (in && hym) ? p[m - A.nz] : R(0);
                    zp[u] = (in && hzp) ? p[m + 1] : R(0);
                    zm[u] = (in && hzm) ? p[m - 1] : R(0);
                    if constexpr (UNI) {
                        cx[u] = A.cu[0]; cyp[u] = cym[u] = A.cu[1]; czp[u] = czm[u] = A.cu[2];
                    } else {
                        cx[u] = in ? c0[m] : R(0);
                        cyp[u] = (in && hyp) ? c1[m] : R(0);
                        cym[u] = (in && hym) ? c1[m - A.nz] : R(0);
                        czp[u] = (in && hzp) ? c2[m] : R(0);
                        czm[u] = (in && hzm) ? c2[m - 1] : R(0);
                    }
                }
#pragma unroll
                for (int u = 0; u < kMarch; u++) {
                    if (i + u >= i1) break;
                    R acc = pc * dg[u];
                    if (i + u + 1 < A.nx) acc -= pn[u] * cx[u];
                    if (i + u > 0) acc -= pm * c0m;
                    if (hyp) acc -= yp[u] * cyp[u];
                    if (hym) acc -= ym[u] * cym[u];
                    if (hzp) acc -= zp[u] * czp[u];
                    if (hzm) acc -= zm[u] * czm[u];
                    Ap[n + (uint64_t)u * sx] = acc;
                    v[0] += (double)pc * (double)acc;
                    pm = pc; pc = pn[u]; c0m = cx[u];
                }
            }
        }
    }
    block_partials<1>(v, slots);
}

__global__ void __launch_bounds__(kThreads) alpha_kernel(CgScalars* sc, const double* slots, unsigned nblocks) {
    if (sc->done) return;
    double tot[1];
    reduce_slots<1>(slots, nblocks, tot);
    if (threadIdx.x == 0) {
        sc->pAp = tot[0];
        sc->alpha = sdiv(sc->rz, tot[0]);
        if (!(tot[0] > 0) && sc->rz != 0 && !sc->keep_going) sc->breakdown = 1;
    }
}

template <typename R>
__global__ void __launch_bounds__(kThreads) update_kernel(uint64_t n_cells, R* __restrict__ x, R* __restrict__ r,
                                                         R* __restrict__ z, const R* __restrict__ p,
                                                         const R* __restrict__ Ap, const R* __restrict__ idiag,
                                                         const CgScalars* sc, double* slots) {
    if (sc->done) return;
    const R alpha = (R)sc->alpha;
    double v[4] = {0, 0, 0, 0};
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t n = (uint64_t)blockIdx.x * kThreads + threadIdx.x; n < n_cells; n += stride) {
        R xn = x[n] + p[n] * alpha;
        R rn = r[n] + Ap[n] * (-alpha);
        R zn = rn * idiag[n];
        x[n] = xn;
        r[n] = rn;
        z[n] = zn;
        v[0] += (double)rn * (double)zn;
        v[1] += (double)zn * (double)zn;
        v[2] += (double)xn * (double)xn;
        v[3] += (double)rn * (double)rn;
    }
    block_partials<4>(v, slots);
}

__global__ void __launch_bounds__(kThreads) beta_kernel(CgScalars* sc, const double* slots, unsigned nblocks) {
    if (sc->done) return;
    double tot[4];
    reduce_slots<4>(slots, nblocks, tot);
    if (threadIdx.x == 0) {
        const double oo = sc->rz;
        sc->rz = tot[0];
        sc->beta = sdiv(tot[0], oo);
        sc->iters += 1;
        sc->res_ref = sqrt(sdiv(fabs(tot[1]), fabs(tot[2])));
        sc->res_true = sc->bnorm > 0 ? sqrt(tot[3]) / sc->bnorm : 0;
        const double crit = sc->mode == 0 ? sc->res_ref : sc->res_true;
        if (crit <= sc->tol || sc->iters >= sc->max_it || sc->breakdown) sc->done = 1;
    }
}

template <typename R>
__global__ void __launch_bounds__(kThreads) pupdate_kernel(uint64_t n_cells, R* __restrict__ p, const R* __restrict__ z,
                                                          const CgScalars* sc) {
    if (sc->done) return;
    const R beta = (R)sc->beta;
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t n = (uint64_t)blockIdx.x * kThreads + threadIdx.x; n < n_cells; n += stride)
        p[n] = z[n] + p[n] * beta;
}

// r = b - A x; z = r / ap; p = z; partials r.z z.z x.x r.r
template <typename R>
__global__ void __launch_bounds__(kThreads) init_kernel(FvOp<R> A, const R* __restrict__ b, const R* __restrict__ x,
                                                       R* __restrict__ r, R* __restrict__ z, R* __restrict__ p,
                                                       double* slots) {
    double v[4] = {0, 0, 0, 0};
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t n = (uint64_t)blockIdx.x * kThreads + threadIdx.x; n < A.n; n += stride) {
        R rn = b[n] - apply_row(A, x, n);
        R zn = rn * A.idiag[n];
        r[n] = rn;
        z[n] = zn;
        p[n] = zn;
        v[0] += (double)rn * (double)zn;
        v[1] += (double)zn * (double)zn;
        v[2] += (double)x[n] * (double)x[n];
        v[3] += (double)rn * (double)rn;
    }
    block_partials<4>(v, slots);
}

__global__ void __launch_bounds__(kThreads) init_scalars_kernel(CgScalars* sc, const double* slots, unsigned nblocks) {
    double tot[4];
    reduce_slots<4>(slots, nblocks, tot);
    if (threadIdx.x == 0) {
        sc->rz = tot[0];
        sc->res_ref = sqrt(sdiv(fabs(tot[1]), fabs(tot[2])));
        sc->res0_ref = sc->res_ref;
        sc->rnorm0 = sqrt(tot[3]);
        sc->bnorm = sqrt(sc->bb);
        sc->res_true = sc->bnorm > 0 ? sqrt(tot[3]) / sc->bnorm : 0;
        sc->iters = 0;
        sc->breakdown = 0;
        const double crit = sc->mode == 0 ? sc->res_ref : sc->res_true;
        sc->done = (sc->mode == 1 && crit <= sc->tol) || sc->max_it <= 0;
    }
}

// sum of squares -> per-block slots (reduce with sum1_kernel)
template <typename R>
__global__ void __launch_bounds__(kThreads) norm2_kernel(uint64_t n_cells, const R* __restrict__ v_, double* slots) {
    double v[1] = {0};
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t n = (uint64_t)blockIdx.x * kThreads + threadIdx.x; n < n_cells; n += stride)
        v[0] += (double)v_[n] * (double)v_[n];
    block_partials<1>(v, slots);
}

__global__ void __launch_bounds__(kThreads) sum1_kernel(const double* slots, unsigned nblocks, double* out) {
    double tot[1];
    reduce_slots<1>(slots, nblocks, tot);
    if (threadIdx.x == 0) *out = tot[0];
}

template <typename R>
__global__ void __launch_bounds__(kThreads) apply_kernel(FvOp<R> A, const R* __restrict__ x, R* __restrict__ y) {
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t n = (uint64_t)blockIdx.x * kThreads + threadIdx.x; n < A.n; n += stride) y[n] = apply_row(A, x, n);
}

// y = b - A x in double (refinement residual), also |y|^2 -> *out
__global__ void __launch_bounds__(kThreads) residual64_kernel(FvOp<double> A, const double* __restrict__ b,
                                                             const double* __restrict__ x, double* __restrict__ y,
                                                             float* __restrict__ y32, double* slots) {
    double v[1] = {0};
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t n = (uint64_t)blockIdx.x * kThreads + threadIdx.x; n < A.n; n += stride) {
        double rn = b[n] - apply_row(A, x, n);
        y[n] = rn;
        if (y32) y32[n] = (float)rn;
        v[0] += rn * rn;
    }
    block_partials<1>(v, slots);
}

// x64 += (double) e32
__global__ void __launch_bounds__(kThreads) correct_kernel(uint64_t n_cells, double* __restrict__ x,
                                                          const float* __restrict__ e) {
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t n = (uint64_t)blockIdx.x * kThreads + threadIdx.x; n < n_cells; n += stride)
        x[n] += (double)e[n];
}

// Su = -divf(U) over interior cells (field.h:3470-3477 with div_flux<false>),
// ghost velocities from the boundary condition.
__global__ void __launch_bounds__(kThreads) divu_kernel(FvDivInput in, double* __restrict__ su) {
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t c = (uint64_t)blockIdx.x * kThreads + threadIdx.x; c < in.ncells; c += stride) {
        double acc = 0;
        for (uint32_t f = in.fi0[c]; f < in.fi1[c]; f++) {
            const uint32_t k = in.allFaces[f];
            const uint32_t c1 = in.FO[k], c2 = in.FN[k];
            const double* nv = &in.fN[3 * (size_t)k];
            const double fi = in.fI[k];
            double u2[3];
            if (c2 < in.ncells) {
                u2[0] = in.U[3 * (size_t)c2]; u2[1] = in.U[3 * (size_t)c2 + 1]; u2[2] = in.U[3 * (size_t)c2 + 2];
            } else {
                const int pi = in.patch_of_ghost[c2 - in.ncells];
                const double* u1 = &in.U[3 * (size_t)c1];
                if (in.bc_kind[pi] == 2) {
                    u2[0] = in.bc_value[3 * pi]; u2[1] = in.bc_value[3 * pi + 1]; u2[2] = in.bc_value[3 * pi + 2];
                } else if (in.bc_kind[pi] == 1) {
                    double mn = sqrt(nv[2] * nv[2] + 0.0 + nv[1] * nv[1] + nv[0] * nv[0]);
                    (void)mn;
                    double e0, e1, e2;
                    {
                        double m2 = nv[2] * nv[2] + 0.0;
                        m2 = nv[1] * nv[1] + m2;
                        m2 = nv[0] * nv[0] + m2;
                        const double mm = sqrt(m2);
                        e0 = nv[0] / mm; e1 = nv[1] / mm; e2 = nv[2] / mm;
                    }
                    const double A0 = 1 - e0 * e0, A1 = 1 - e1 * e1, A2 = 1 - e2 * e2;
                    const double A3 = 0 - e0 * e1, A4 = 0 - e1 * e2, A5 = 0 - e0 * e2;
                    double r0 = A0 * u1[0] + A3 * u1[1] + A5 * u1[2];
                    double r1 = A3 * u1[0] + A1 * u1[1] + A4 * u1[2];
                    double r2 = A5 * u1[0] + A4 * u1[1] + A2 * u1[2];
                    double mr = r2 * r2 + 0.0;
                    mr = r1 * r1 + mr;
                    mr = sqrt(r0 * r0 + mr);
                    if (!(fabs(mr) <= 1e-7)) {
                        double pm = u1[2] * u1[2] + 0.0;
                        pm = u1[1] * u1[1] + pm;
                        pm = sqrt(u1[0] * u1[0] + pm);
                        const double s = pm / mr;
                        r0 *= s; r1 *= s; r2 *= s;
                    }
                    u2[0] = r0; u2[1] = r1; u2[2] = r2;
                } else {
                    u2[0] = u1[0]; u2[1] = u1[1]; u2[2] = u1[2];
                }
            }
            const double* u1 = &in.U[3 * (size_t)c1];
            const double f0 = u1[0] * fi + u2[0] * (1 - fi);
            const double f1 = u1[1] * fi + u2[1] * (1 - fi);
            const double f2 = u1[2] * fi + u2[2] * (1 - fi);
            double dd = f2 * nv[2] + 0.0;
            dd = f1 * nv[1] + dd;
            dd = f0 * nv[0] + dd;
            if (c1 == c) acc += dd;
            else acc -= dd;
        }
        su[c] = in.su_bc[c] - acc;
    }
}

// U -= gradf(p, true) on the interior cells (potential.cpp:66-67; the FV
// branch of gradf, field.h:3363-3371, with grad_flux<false>,
// field.h:3050-3082): cds(p) at each face of the cell, + fN cds on the faces
// it owns and - on the others, in the cell's face order, then / cV.  Ghost p
// from the pressure boundary condition as applyExplicitBCs sets it after the
// solve: DIRICHLET the patch value, NEUMANN the owner plus value |dv|
// (goff holds the value, or value |dv|, per ghost cell).
__global__ void __launch_bounds__(kThreads) gradcorr_kernel(FvDivInput in, const double* __restrict__ p,
                                                           const int32_t* __restrict__ gkind,
                                                           const double* __restrict__ goff,
                                                           const double* __restrict__ cV, uint64_t nghost,
                                                           double* __restrict__ U) {
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t c = (uint64_t)blockIdx.x * kThreads + threadIdx.x; c < in.ncells; c += stride) {
        double g0 = 0, g1 = 0, g2 = 0;
        for (uint32_t f = in.fi0[c]; f < in.fi1[c]; f++) {
            const uint32_t k = in.allFaces[f];
            const uint32_t c1 = in.FO[k], c2 = in.FN[k];
            const double* nv = &in.fN[3 * (size_t)k];
            const double fi = in.fI[k];
            const double p1 = p[c1];
            double p2;
            if (c2 < in.ncells) p2 = p[c2];
            else {
                const uint64_t gi = c2 - in.ncells;
                p2 = gi >= nghost ? p1 : (gkind[gi] == 0 ? p1 + goff[gi] : goff[gi]);
            }
            const double fF = p1 * fi + p2 * (1 - fi);
            if (c1 == c) { g0 += nv[0] * fF; g1 += nv[1] * fF; g2 += nv[2] * fF; }
            else { g0 -= nv[0] * fF; g1 -= nv[1] * fF; g2 -= nv[2] * fF; }
        }
        const double v = cV[c];
        U[3 * c] = U[3 * c] - g0 / v;
        U[3 * c + 1] = U[3 * c + 1] - g1 / v;
        U[3 * c + 2] = U[3 * c + 2] - g2 / v;
    }
}

template <typename R>
__global__ void __launch_bounds__(kThreads) convert_kernel(uint64_t n, const double* __restrict__ a,
                                                          R* __restrict__ b) {
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) b[i] = (R)a[i];
}

unsigned grid_for(uint64_t n) {
    uint64_t b = (n + kThreads - 1) / kThreads;
    if (b > (uint64_t)kCgSlots) b = kCgSlots;
    if (b < 1) b = 1;
    return (unsigned)b;
}

}  // namespace

template <typename R>
hipError_t fv_cg_init(const FvOp<R>& A, const R* b, const R* x, R* r, R* z, R* p, CgScalars* sc,
                      double* slots, hipStream_t s) {
    const unsigned g = grid_for(A.n);
    hipLaunchKernelGGL(norm2_kernel<R>, dim3(g), dim3(kThreads), 0, s, A.n, b, slots);
    hipLaunchKernelGGL(sum1_kernel, dim3(1), dim3(kThreads), 0, s, slots, g, &sc->bb);
    hipLaunchKernelGGL(init_kernel<R>, dim3(g), dim3(kThreads), 0, s, A, b, x, r, z, p, slots);
    hipLaunchKernelGGL(init_scalars_kernel, dim3(1), dim3(kThreads), 0, s, sc, slots, g);
    return hipGetLastError();
}

template <typename R>
hipError_t fv_cg_iterations(const FvOp<R>& A, int count, R* x, R* r, R* z, R* p, R* Ap, CgScalars* sc,
                            double* slots, hipStream_t s) {
    const unsigned g = grid_for(A.n);
    // p = z + beta p sums nothing, so in FP64 it takes one cell per thread
    // (the widest grid, the most reads in flight; 256^3: 89 -> 70 us); in
    // FP32 the slot-bounded grid measured faster (33 us against 42 us)
    const unsigned gp = sizeof(R) == 8 ? (unsigned)std::min<uint64_t>((A.n + kThreads - 1) / kThreads, 1u << 30) : g;
    const MarchGrid mg = A.structured ? march_grid(A.nx, A.ny, A.nz) : MarchGrid{};
    const bool march = A.structured && mg.groups <= (unsigned)kCgSlots;
    for (int it = 0; it < count; it++) {
        if (march) {
            if (A.uniform)
                hipLaunchKernelGGL((spmv_dot_march_kernel<R, true>), dim3(mg.groups), dim3(kThreads), 0, s, A, mg, p,
                                   Ap, sc, slots);
            else
                hipLaunchKernelGGL((spmv_dot_march_kernel<R, false>), dim3(mg.groups), dim3(kThreads), 0, s, A, mg,
                                   p, Ap, sc, slots);
            hipLaunchKernelGGL(alpha_kernel, dim3(1), dim3(kThreads), 0, s, sc, slots, mg.groups);
        } else {
            hipLaunchKernelGGL(spmv_dot_kernel<R>, dim3(g), dim3(kThreads), 0, s, A, p, Ap, sc, slots);
            hipLaunchKernelGGL(alpha_kernel, dim3(1), dim3(kThreads), 0, s, sc, slots, g);
        }
        hipLaunchKernelGGL(update_kernel<R>, dim3(g), dim3(kThreads), 0, s, A.n, x, r, z, p, Ap, A.idiag, sc, slots);
        hipLaunchKernelGGL(beta_kernel, dim3(1), dim3(kThreads), 0, s, sc, slots, g);
        hipLaunchKernelGGL(pupdate_kernel<R>, dim3(gp), dim3(kThreads), 0, s, A.n, p, z, sc);
    }
    return hipGetLastError();
}

template <typename R>
hipError_t fv_norm2(uint64_t n, const R* v, double* out, double* slots, hipStream_t s) {
    const unsigned g = grid_for(n);
    hipLaunchKernelGGL(norm2_kernel<R>, dim3(g), dim3(kThreads), 0, s, n, v, slots);
    hipLaunchKernelGGL(sum1_kernel, dim3(1), dim3(kThreads), 0, s, slots, g, out);
    return hipGetLastError();
}

hipError_t fv_residual64(const FvOp<double>& A, const double* b, const double* x, double* y, float* y32,
                         double* out, double* slots, hipStream_t s) {
    const unsigned g = grid_for(A.n);
    hipLaunchKernelGGL(residual64_kernel, dim3(g), dim3(kThreads), 0, s, A, b, x, y, y32, slots);
    hipLaunchKernelGGL(sum1_kernel, dim3(1), dim3(kThreads), 0, s, slots, g, out);
    return hipGetLastError();
}

hipError_t fv_correct(uint64_t n, double* x, const float* e, hipStream_t s) {
    hipLaunchKernelGGL(correct_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, n, x, e);
    return hipGetLastError();
}

hipError_t fv_gradcorr(const FvDivInput& in, const double* p, const int32_t* gkind, const double* goff,
                       const double* cV, uint64_t nghost, double* U, hipStream_t s) {
    hipLaunchKernelGGL(gradcorr_kernel, dim3(grid_for(in.ncells)), dim3(kThreads), 0, s, in, p, gkind, goff, cV,
                       nghost, U);
    return hipGetLastError();
}

hipError_t fv_divu(const FvDivInput& in, double* su, hipStream_t s) {
    hipLaunchKernelGGL(divu_kernel, dim3(grid_for(in.ncells)), dim3(kThreads), 0, s, in, su);
    return hipGetLastError();
}

template <typename R>
hipError_t fv_convert(uint64_t n, const double* a, R* b, hipStream_t s) {
    hipLaunchKernelGGL(convert_kernel<R>, dim3(grid_for(n)), dim3(kThreads), 0, s, n, a, b);
    return hipGetLastError();
}

template <typename R>
hipError_t fv_apply(const FvOp<R>& A, const R* x, R* y, hipStream_t s) {
    hipLaunchKernelGGL(apply_kernel<R>, dim3(grid_for(A.n)), dim3(kThreads), 0, s, A, x, y);
    return hipGetLastError();
}

template hipError_t fv_cg_init<double>(const FvOp<double>&, const double*, const double*, double*, double*,
                                       double*, CgScalars*, double*, hipStream_t);
template hipError_t fv_cg_init<float>(const FvOp<float>&, const float*, const float*, float*, float*,
                                      float*, CgScalars*, double*, hipStream_t);
template hipError_t fv_cg_iterations<double>(const FvOp<double>&, int, double*, double*, double*, double*,
                                             double*, CgScalars*, double*, hipStream_t);
template hipError_t fv_cg_iterations<float>(const FvOp<float>&, int, float*, float*, float*, float*, float*,
                                            CgScalars*, double*, hipStream_t);
template hipError_t fv_norm2<double>(uint64_t, const double*, double*, double*, hipStream_t);
template hipError_t fv_norm2<float>(uint64_t, const float*, double*, double*, hipStream_t);
template hipError_t fv_convert<float>(uint64_t, const double*, float*, hipStream_t);
template hipError_t fv_apply<double>(const FvOp<double>&, const double*, double*, hipStream_t);
template hipError_t fv_apply<float>(const FvOp<float>&, const float*, float*, hipStream_t);

}  // namespace nsem
