// Device kernels of the field API's operators, in the reference's order of
// accumulation (src/field/field.h): face interpolation and the Rusanov
// flux per face node, the surface integrals per node over the face nodes of
// its own cell in the order div_flux visits them, the DG volume terms along
// the GLL lines (the split step's reference-order routine), boundary passes
// per patch, the FV laplacian's coefficients and non-orthogonal
// correction, reductions, and Solve.
#include <cstdio>

#include "dg_volume.hpp"
#include "field.h"
#include "field_mesh.h"
#include "nebulasem_amd.h"
#include "solve.h"

namespace nsem_field {
namespace {

constexpr int kThreads = 256;
inline unsigned blocks_for(size_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

template <int NC> struct TypeOf;
template <> struct TypeOf<1> { using T = Scalar; };
template <> struct TypeOf<3> { using T = Vector; };
template <> struct TypeOf<6> { using T = STensor; };
template <> struct TypeOf<9> { using T = Tensor; };

__device__ inline Vector normal(const MeshTables& m, uint64_t k) {
    return Vector(m.fN[3 * k], m.fN[3 * k + 1], m.fN[3 * k + 2]);
}
struct DevTables {
    const uint32_t *FO, *FN;
    const double *fN, *fI, *fD, *cC, *cV, *JcV, *D0, *D1, *D2;
    const uint32_t* ent_off;
    const uint64_t* ent;
    uint64_t nbf, nall, nfn;
    int npx, npy, npz, np;
};
DevTables dev() {
    const MeshTables& m = mesh_tables();
    DevTables t{};
    t.FO = m.FO; t.FN = m.FN; t.fN = m.fN; t.fI = m.fI; t.fD = m.fD; t.cC = m.cC; t.cV = m.cV; t.JcV = m.JcV;
    t.D0 = m.D[0]; t.D1 = m.D[1]; t.D2 = m.D[2];
    t.ent_off = m.ent_off; t.ent = m.ent;
    t.nbf = m.nbf; t.nall = m.nall; t.nfn = (uint64_t)m.nfacets * m.npf;
    t.npx = m.npx; t.npy = m.npy; t.npz = m.npz; t.np = m.np;
    return t;
}
__device__ inline Vector fnormal(const DevTables& t, uint64_t k) {
    return Vector(t.fN[3 * k], t.fN[3 * k + 1], t.fN[3 * k + 2]);
}

// -------------------------------------------------------------- kernels --
template <class T>
__global__ void cds_kernel(DevTables t, const T* cell, T* facet) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= t.nfn) return;
    const uint32_t c1 = t.FO[k], c2 = t.FN[k];
    const T O = c1 < t.nall ? cell[c1] : T(0);
    const T N = c2 < t.nall ? cell[c2] : T(0);
    const Scalar fI = t.fI[k];
    facet[k] = (O * fI) + (N * (1 - fI));
}

template <class T1, class T2>
__global__ void rusanov_kernel(DevTables t, const T1* F, const T2* q, const Scalar* lam, T1* fF) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= t.nfn) return;
    const uint32_t c1 = t.FO[k], c2 = t.FN[k];
    const T1 FOv = c1 < t.nall ? F[c1] : T1(0);
    const T1 FNv = c2 < t.nall ? F[c2] : T1(0);
    const T2 qO = c1 < t.nall ? q[c1] : T2(0);
    const T2 qN = c2 < t.nall ? q[c2] : T2(0);
    const Scalar fI = t.fI[k];
    const T1 c = (FOv * fI) + (FNv * (1 - fI));
    fF[k] = c - mul(unit(fnormal(t, k)), lam[k] * (qN - qO));
}

// r[t] = +-dot(fF[k], fN[k]) over the node's entries (weak div_flux)
template <class T1, class R>
__global__ void div_flux_kernel(DevTables t, const T1* fF, R* r) {
    const uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= t.nbf) return;
    R acc(0);
    for (uint32_t e = t.ent_off[n]; e < t.ent_off[n + 1]; e++) {
        const uint64_t v = t.ent[e];
        const uint64_t k = v >> 1;
        if (v & 1) acc -= dot(fF[k], fnormal(t, k));
        else acc += dot(fF[k], fnormal(t, k));
    }
    r[n] = acc;
}

// the weak DG volume term of one flux row (rows of a tensor flux are the
// rows of dot(Tensor, Vector); a vector flux is one row, dotted from the
// right), subtracted in the reference's order
__global__ void __launch_bounds__(512) divf_volume_kernel(DevTables t, const double* F, int ncF, int row,
                                                          double* r, int ncR) {
    extern __shared__ double lds[];
    const int NP = t.np, NX = t.npx, NY = t.npy, NZ = t.npz;
    const int q = threadIdx.x;
    const uint64_t e = blockIdx.x;
    const uint64_t n = e * NP + q;
    double* Fr = lds;            // 3 x NP: the row's components
    double* J = lds + 3 * NP;    // 9 x NP
    // component (row, c) of the flux value: a vector flux is its own row;
    // tensor storage XX YY ZZ XY YZ XZ YX ZY ZX
    static constexpr int tmap[3][3] = {{0, 3, 5}, {6, 1, 4}, {8, 7, 2}};
    if (q < NP) {
        for (int c = 0; c < 3; c++) Fr[c * NP + q] = F[n * ncF + (ncF == 3 ? c : tmap[row][c])];
        for (int c = 0; c < 9; c++) J[c * NP + q] = t.JcV[9 * n + nsem::kJinRow[c]];
    }
    __syncthreads();
    if (q >= NP) return;
    const int i = q / (NY * NZ), j = (q / NZ) % NY, k = q % NZ;
    double res[1] = {r[n * ncR + row]};
    if (ncF == 3)
        nsem::volume_reference_order<1, 1>(res, Fr, J, NP, i, j, k, NX, NY, NZ, t.D0, t.D1, t.D2);
    else
        nsem::volume_reference_order<1, 0>(res, Fr, J, NP, i, j, k, NX, NY, NZ, t.D0, t.D1, t.D2);
    r[n * ncR + row] = res[0];
}

// r[t] = +-fN (x) (fF - p[t]) (strong) or +-fN (x) fF over the node's entries
template <class T, class R>
__global__ void grad_flux_kernel(DevTables t, const T* p, const T* fF, int strong, R* r) {
    const uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= t.nbf) return;
    R acc(0);
    const T pn = p[n];
    for (uint32_t e = t.ent_off[n]; e < t.ent_off[n + 1]; e++) {
        const uint64_t v = t.ent[e];
        const uint64_t k = v >> 1;
        const T d = strong ? T(fF[k] - pn) : fF[k];
        if (v & 1) acc -= mul(fnormal(t, k), d);
        else acc += mul(fnormal(t, k), d);
    }
    r[n] = acc;
}

// strong DG gradient volume term (GRADD, field.h:3340-3354): at target
// (ii,jj,kk), sources on the x line (all), the y line and the z line
// (without the target), each mul(Jin grad psi, p[source]) with Jin the
// target's Jinv cV
template <class T, class R>
__global__ void __launch_bounds__(512) grad_volume_kernel(DevTables t, const T* p, R* r) {
    extern __shared__ double lds[];
    T* P = reinterpret_cast<T*>(lds);
    const int NP = t.np, NX = t.npx, NY = t.npy, NZ = t.npz;
    const int q = threadIdx.x;
    const uint64_t n = (uint64_t)blockIdx.x * NP + q;
    if (q < NP) P[q] = p[n];
    __syncthreads();
    if (q >= NP) return;
    const int ii = q / (NY * NZ), jj = (q / NZ) % NY, kk = q % NZ;
    Tensor Jin;
    for (int c = 0; c < 9; c++) Jin[c] = t.JcV[9 * n + c];
    R acc = r[n];
    const int syz = NY * NZ;
    auto dpsi = [&](int im, int jm, int km) {
        // DPSI_ with the identity basis: products with psi of 1 or 0
        const double p0 = (ii == im) ? 1.0 : 0.0, p1 = (jj == jm) ? 1.0 : 0.0, p2 = (kk == km) ? 1.0 : 0.0;
        const double d0 = t.D0[ii * NX + im] * p1 * p2;
        const double d1 = p0 * t.D1[jj * NY + jm] * p2;
        const double d2 = p0 * p1 * t.D2[kk * NZ + km];
        return dot(Jin, Vector(d0, d1, d2));
    };
    for (int im = 0; im < NX; im++) acc += mul(dpsi(im, jj, kk), P[im * syz + jj * NZ + kk]);
    for (int jm = 0; jm < NY; jm++)
        if (jm != jj) acc += mul(dpsi(ii, jm, kk), P[ii * syz + jm * NZ + kk]);
    for (int km = 0; km < NZ; km++)
        if (km != kk) acc += mul(dpsi(ii, jj, km), P[ii * syz + jj * NZ + km]);
    r[n] = acc;
}

template <class T>
__global__ void sum_facet_kernel(DevTables t, const T* fF, T* r) {
    const uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= t.nall) return;
    T acc(0);
    for (uint32_t e = t.ent_off[n]; e < t.ent_off[n + 1]; e++) {
        const uint64_t v = t.ent[e];
        if (v & 1) acc -= fF[v >> 1];
        else acc += fF[v >> 1];
    }
    r[n] = acc;
}

// r -= (or +=) x[other node] a over the faces of the node's cell: a = c0[k]
// where the node owns face node k, c1[k] where it is the neighbour
template <class T>
__global__ void offdiag_kernel(DevTables t, const T* x, const Scalar* c0, const Scalar* c1, int add, T* r) {
    const uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= t.nbf) return;
    T acc = r[n];
    for (uint32_t e = t.ent_off[n]; e < t.ent_off[n + 1]; e++) {
        const uint64_t v = t.ent[e];
        const uint64_t k = v >> 1;
        if (add) {
            if (v & 1) acc += x[t.FO[k]] * c1[k];
            else acc += x[t.FN[k]] * c0[k];
        } else {
            if (v & 1) acc -= x[t.FO[k]] * c1[k];
            else acc -= x[t.FN[k]] * c0[k];
        }
    }
    r[n] = acc;
}

__global__ void lap_face_kernel(DevTables t, const Scalar* mu, int use_fD, Scalar* ano, Scalar* ann) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= t.nfn) return;
    const Scalar a = use_fD ? t.fD[k] * mu[k] : mu[k];
    ano[k] = a;
    ann[k] = a;
}
// ap += the couplings of the faces of each node's cell, ghost cells
// included (the reference's loops run over gNCells)
__global__ void diag_sum_kernel(DevTables t, const Scalar* ano, const Scalar* ann, Scalar* ap) {
    const uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= t.nall) return;
    Scalar acc = ap[n];
    for (uint32_t e = t.ent_off[n]; e < t.ent_off[n + 1]; e++) {
        const uint64_t v = t.ent[e];
        acc += (v & 1) ? ann[v >> 1] : ano[v >> 1];
    }
    ap[n] = acc;
}

// FV implicit convection coefficients (div_flux_implicit, field.h:3147-3189):
// blending factor gamma per face (CDS 1, BLENDED the blend factor, HYBRID
// from the face Peclet comparison, otherwise upwind), then
// ano = G (-F fI) + (1 - G)(-max(F, 0)), ann = G (F (1 - fI)) + (1 - G)(-max(-F, 0))
__global__ void conv_face_kernel(DevTables t, const Scalar* flux, const Scalar* mu, int scheme, Scalar blend,
                                 Scalar* ano, Scalar* ann) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= t.nfn) return;
    const Scalar F = flux[k], fI = t.fI[k];
    Scalar G = 0;
    if (scheme == Controls::CDS) G = 1;
    else if (scheme == Controls::BLENDED) G = blend;
    else if (scheme == Controls::HYBRID && mu && t.FN[k] < t.nall) {
        const Scalar D = t.fD[k] * mu[k];
        if (F < 0) G = (F * fI + D >= 0) ? 1 : 0;
        else G = (F * (1 - fI) - D > 0) ? 0 : 1;
    }
    const Scalar mF = F >= 0 ? F : Scalar(0);
    const Scalar nF = -F >= 0 ? -F : Scalar(0);
    ano[k] = ((G) * (-F * (fI)) + (1 - G) * (-mF));
    ann[k] = ((G) * (F * (1 - fI)) + (1 - G) * (-nF));
}

template <class T, class G>
__global__ void nonortho_kernel(DevTables t, const G* g, const T* x, const Scalar* ano, T* r) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= t.nfn) return;
    const uint32_t c1 = t.FO[k], c2 = t.FN[k];
    if (c1 >= t.nall || c2 >= t.nall) {
        r[k] = T(0);
        return;
    }
    const Vector dv = Vector(t.cC[3 * c2], t.cC[3 * c2 + 1], t.cC[3 * c2 + 2]) -
                      Vector(t.cC[3 * c1], t.cC[3 * c1 + 1], t.cC[3 * c1 + 2]);
    const Vector K = fnormal(t, k) - t.fD[k] * dv;
    T v = dot(g[k], K);
    const T res = ano[k] * (x[c2] - x[c1]);
    if (mag(v) > Scalar(0.5) * mag(res)) v = Scalar(0.5) * res;
    r[k] = v;
}

// boundary face nodes of one patch
struct PatchArgs {
    const uint32_t* k;
    const uint32_t* slot;
    uint32_t n;
};

template <class T>
__global__ void explicit_bc_kernel(DevTables t, PatchArgs pa, int kind, T value, Scalar shape, T tvalue, T* f,
                                   T* fixed, int capture) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pa.n) return;
    const uint64_t k = pa.k[i];
    const uint32_t c1 = t.FO[k], c2 = t.FN[k];
    if (kind == BC_NEUMANN) {
        const Vector dv = Vector(t.cC[3 * c2], t.cC[3 * c2 + 1], t.cC[3 * c2 + 2]) -
                          Vector(t.cC[3 * c1], t.cC[3 * c1 + 1], t.cC[3 * c1 + 2]);
        f[c2] = f[c1] + value * mag(dv);
    } else if (kind == BC_ROBIN) {
        // field.h:2677-2680
        const Vector dv = Vector(t.cC[3 * c2], t.cC[3 * c2 + 1], t.cC[3 * c2 + 2]) -
                          Vector(t.cC[3 * c1], t.cC[3 * c1 + 1], t.cC[3 * c1 + 2]);
        f[c2] = shape * value + (1 - shape) * (f[c1] + tvalue * mag(dv));
    } else if (kind == BC_SYMMETRY) {
        f[c2] = sym(f[c1], fnormal(t, k));
    } else if (kind == BC_DIRICHLET) {
        f[c2] = value;
    } else if (kind == BC_CALC_DIRICHLET) {
        if (capture) {
            fixed[pa.slot[i]] = f[c1];
            f[c2] = f[c1];
        } else {
            f[c2] = fixed[pa.slot[i]];
        }
    }
}

template <class T>
__global__ void ghost_copy_kernel(const uint32_t* go, const uint32_t* gn, uint32_t n, T* f) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[gn[i]] = f[go[i]];
}
template <class T>
__global__ void ghost_value_kernel(DevTables t, PatchArgs pa, T value, T* f) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < pa.n) f[t.FN[pa.k[i]]] = value;
}

// applyImplicitBCs on one patch's face nodes, sequentially in the
// reference's order (a node can appear on several faces of a patch)
template <class T>
__global__ void implicit_bc_kernel(DevTables t, PatchArgs pa, int kind, Scalar shape, Scalar* ap, Scalar* ann, T* Su,
                                   const T* cF) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (uint32_t i = 0; i < pa.n; i++) {
        const uint64_t k = pa.k[i];
        const uint32_t c1 = t.FO[k], c2 = t.FN[k];
        if (kind == BC_NEUMANN) {
            ap[c1] -= ann[k];
            Su[c1] += ann[k] * (cF[c2] - cF[c1]);
        } else if (kind == BC_ROBIN) {
            // field.h:2571-2573
            ap[c1] -= (1 - shape) * ann[k];
            Su[c1] += ann[k] * (cF[c2] - (1 - shape) * cF[c1]);
        } else {
            Su[c1] += ann[k] * cF[c2];
        }
        ann[k] = 0;
    }
}

// deterministic reductions over the interior nodes: per-block partials,
// then one block in a fixed order
constexpr int kRedBlocks = 512;
template <int NC, int KIND>
__global__ void reduce_partial_kernel(const double* d, uint64_t n, double* part) {
    __shared__ double sh[kThreads][NC];
    double acc[NC];
    for (int c = 0; c < NC; c++) acc[c] = KIND == 0 ? 0.0 : (KIND == 1 ? -1e300 : 1e300);
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads) {
        if (KIND == 0) {
            for (int c = 0; c < NC; c++) acc[c] += d[i * NC + c];
        } else {
            double m2 = 0;
            if (NC == 1) m2 = fabs(d[i]);
            else {
                double s = 0;
                for (int c = NC - 1; c >= 0; c--) s = d[i * NC + c] * d[i * NC + c] + s;
                m2 = sqrt(s);
            }
            acc[0] = KIND == 1 ? fmax(acc[0], m2) : fmin(acc[0], m2);
        }
    }
    for (int c = 0; c < NC; c++) sh[threadIdx.x][c] = acc[c];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int s = 1; s < kThreads; s++)
            for (int c = 0; c < NC; c++)
                acc[c] = KIND == 0 ? acc[c] + sh[s][c] : (KIND == 1 ? fmax(acc[c], sh[s][c]) : fmin(acc[c], sh[s][c]));
        for (int c = 0; c < NC; c++) part[blockIdx.x * NC + c] = acc[c];
    }
}
template <int NC, int KIND>
__global__ void reduce_final_kernel(const double* part, int nb, double* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double acc[NC];
    for (int c = 0; c < NC; c++) acc[c] = part[c];
    for (int b = 1; b < nb; b++)
        for (int c = 0; c < NC; c++)
            acc[c] = KIND == 0 ? acc[c] + part[b * NC + c] : (KIND == 1 ? fmax(acc[c], part[b * NC + c])
                                                                          : fmin(acc[c], part[b * NC + c]));
    for (int c = 0; c < NC; c++) out[c] = acc[c];
}

template <int NC>
void reduce_nc(const Scalar* d, int kind, Scalar* out, uint64_t n) {
    double* part = static_cast<double*>(dev_alloc(kRedBlocks * NC * sizeof(double)));
    double* res = static_cast<double*>(dev_alloc(NC * sizeof(double)));
    if (kind == 0) {
        hipLaunchKernelGGL((reduce_partial_kernel<NC, 0>), dim3(kRedBlocks), dim3(kThreads), 0, stream(), d, n, part);
        hipLaunchKernelGGL((reduce_final_kernel<NC, 0>), dim3(1), dim3(64), 0, stream(), part, kRedBlocks, res);
    } else if (kind == 1) {
        hipLaunchKernelGGL((reduce_partial_kernel<NC, 1>), dim3(kRedBlocks), dim3(kThreads), 0, stream(), d, n, part);
        hipLaunchKernelGGL((reduce_final_kernel<1, 1>), dim3(1), dim3(64), 0, stream(), part, kRedBlocks, res);
    } else {
        hipLaunchKernelGGL((reduce_partial_kernel<NC, 2>), dim3(kRedBlocks), dim3(kThreads), 0, stream(), d, n, part);
        hipLaunchKernelGGL((reduce_final_kernel<1, 2>), dim3(1), dim3(64), 0, stream(), part, kRedBlocks, res);
    }
    check(hipGetLastError(), "reduction");
    download(out, res, (kind == 0 ? NC : 1) * sizeof(double));
    dev_free(part, kRedBlocks * NC * sizeof(double));
    dev_free(res, NC * sizeof(double));
}

PatchArgs patch(uint32_t b) {
    const MeshTables& m = mesh_tables();
    const uint32_t o = m.patch_off_host[b];
    return PatchArgs{m.patch_k + o, m.patch_slot + o, m.patch_off_host[b + 1] - o};
}

// device copies of the CALC_DIRICHLET values, per field and patch
std::map<std::pair<Int, uint32_t>, std::pair<void*, size_t>>& fixed_store() {
    static auto* s = new std::map<std::pair<Int, uint32_t>, std::pair<void*, size_t>>();
    return *s;
}

template <class T>
void explicit_bcs_t(T* f, Int fIndex) {
    // fields without a name share index 0 (the hash of ""), and conditions
    // copied to such a field apply to all of them, as in the reference
    if (!has_bcs(fIndex)) return;
    constexpr int NC = (int)(sizeof(T) / sizeof(Scalar));
    FieldBCs& b = bcs_of(fIndex, NC);
    const MeshTables& m = mesh_tables();
    for (uint32_t p = 0; p < m.npatch; p++) {
        const int kind = b.kind[p];
        if (kind == BC_NONE) continue;
        const PatchArgs pa = patch(p);
        if (!pa.n) continue;
        T value, tvalue;
        for (int c = 0; c < NC; c++) {
            reinterpret_cast<Scalar*>(&value)[c] = b.value[p * NC + c];
            reinterpret_cast<Scalar*>(&tvalue)[c] = b.tvalue[p * NC + c];
        }
        T* fixed = nullptr;
        int capture = 0;
        if (kind == BC_CALC_DIRICHLET) {
            auto& slot = fixed_store()[{fIndex, p}];
            if (b.fixed[p].empty()) {
                if (!slot.first) slot = {dev_alloc(pa.n * sizeof(T)), pa.n * sizeof(T)};
                b.fixed[p].assign(1, 1.0);  // captured from now on
                capture = 1;
            }
            fixed = static_cast<T*>(slot.first);
        }
        hipLaunchKernelGGL((explicit_bc_kernel<T>), dim3(blocks_for(pa.n)), dim3(kThreads), 0, stream(), dev(), pa, kind,
                           value, b.shape[p], tvalue, f, fixed, capture);
    }
    check(hipGetLastError(), "boundary conditions");
}

template <class T>
void fill_bcs_t(T* f, Int bind, int nc_src) {
    const MeshTables& m = mesh_tables();
    if (m.nghostnodes)
        hipLaunchKernelGGL((ghost_copy_kernel<T>), dim3(blocks_for(m.nghostnodes)), dim3(kThreads), 0, stream(),
                           m.ghost_owner, m.ghost_node, m.nghostnodes, f);
    if (bind && has_bcs(bind)) {
        FieldBCs& b = bcs_of(bind, nc_src);
        for (uint32_t p = 0; p < m.npatch; p++) {
            const int kind = b.kind[p];
            if (kind != BC_NEUMANN && kind != BC_SYMMETRY) continue;
            const PatchArgs pa = patch(p);
            if (!pa.n) continue;
            // the patch value read as the gradient's type: the value in its
            // first components, the condition's other parameters (zero) after
            T v(0);
            if (kind == BC_NEUMANN)
                for (int c = 0; c < nc_src; c++) reinterpret_cast<Scalar*>(&v)[c] = b.value[p * nc_src + c];
            hipLaunchKernelGGL((ghost_value_kernel<T>), dim3(blocks_for(pa.n)), dim3(kThreads), 0, stream(), dev(), pa, v, f);
        }
    }
    check(hipGetLastError(), "fillBCs");
}

}  // namespace

// ------------------------------------------------------------ entries ----
#define NSEM_DISPATCH(nc, CALL)                                                 \
    switch (nc) {                                                               \
        case 1: { using T = Scalar; CALL; break; }                              \
        case 3: { using T = Vector; CALL; break; }                              \
        case 6: { using T = STensor; CALL; break; }                             \
        case 9: { using T = Tensor; CALL; break; }                              \
        default: fail("unsupported field type");                                \
    }

void cds(const Scalar* cell, int nc, Scalar* facet) {
    const DevTables t = dev();
    NSEM_DISPATCH(nc, hipLaunchKernelGGL((cds_kernel<T>), dim3(blocks_for(t.nfn)), dim3(kThreads), 0, stream(), t,
                                         reinterpret_cast<const T*>(cell), reinterpret_cast<T*>(facet)));
    check(hipGetLastError(), "cds");
}
void rusanov(const Scalar* F, int ncF, const Scalar* q, int ncq, const Scalar* lam, Scalar* fF) {
    const DevTables t = dev();
    if (ncF == 3 && ncq == 1)
        hipLaunchKernelGGL((rusanov_kernel<Vector, Scalar>), dim3(blocks_for(t.nfn)), dim3(kThreads), 0, stream(), t,
                           reinterpret_cast<const Vector*>(F), q, lam, reinterpret_cast<Vector*>(fF));
    else if (ncF == 9 && ncq == 3)
        hipLaunchKernelGGL((rusanov_kernel<Tensor, Vector>), dim3(blocks_for(t.nfn)), dim3(kThreads), 0, stream(), t,
                           reinterpret_cast<const Tensor*>(F), reinterpret_cast<const Vector*>(q), lam,
                           reinterpret_cast<Tensor*>(fF));
    else
        fail("rusanov: vector flux with a scalar state or tensor flux with a vector state");
    check(hipGetLastError(), "rusanov");
}
void div_flux(const Scalar* fF, int ncF, Scalar* r) {
    const DevTables t = dev();
    if (ncF == 3)
        hipLaunchKernelGGL((div_flux_kernel<Vector, Scalar>), dim3(blocks_for(t.nbf)), dim3(kThreads), 0, stream(), t,
                           reinterpret_cast<const Vector*>(fF), r);
    else if (ncF == 9)
        hipLaunchKernelGGL((div_flux_kernel<Tensor, Vector>), dim3(blocks_for(t.nbf)), dim3(kThreads), 0, stream(), t,
                           reinterpret_cast<const Tensor*>(fF), reinterpret_cast<Vector*>(r));
    else
        fail("divf of a scalar or symmetric-tensor flux is not built");
    check(hipGetLastError(), "div_flux");
}
void divf_volume(const Scalar* F, int ncF, Scalar* r) {
    const MeshTables& m = mesh_tables();
    if (m.np > 512) fail("divf: at most 512 nodes per element");
    const int threads = (int)((m.np + 63) / 64 * 64);
    const size_t lds = (size_t)12 * m.np * sizeof(double);
    const int rows = ncF == 3 ? 1 : 3, ncR = ncF == 3 ? 1 : 3;
    for (int row = 0; row < rows; row++)
        hipLaunchKernelGGL(divf_volume_kernel, dim3(m.nbcs), dim3(threads), lds, stream(), dev(), F, ncF, row, r, ncR);
    check(hipGetLastError(), "divf volume");
}
void grad_flux(const Scalar* p, int nc, const Scalar* fF, bool strong, Scalar* r) {
    const DevTables t = dev();
    if (nc == 1)
        hipLaunchKernelGGL((grad_flux_kernel<Scalar, Vector>), dim3(blocks_for(t.nbf)), dim3(kThreads), 0, stream(), t,
                           p, fF, strong ? 1 : 0, reinterpret_cast<Vector*>(r));
    else if (nc == 3)
        hipLaunchKernelGGL((grad_flux_kernel<Vector, Tensor>), dim3(blocks_for(t.nbf)), dim3(kThreads), 0, stream(), t,
                           reinterpret_cast<const Vector*>(p), reinterpret_cast<const Vector*>(fF), strong ? 1 : 0,
                           reinterpret_cast<Tensor*>(r));
    else
        fail("gradf of a tensor field is not built");
    check(hipGetLastError(), "grad_flux");
}
void grad_volume(const Scalar* p, int nc, Scalar* r) {
    const MeshTables& m = mesh_tables();
    if (m.np > 512) fail("gradf: at most 512 nodes per element");
    const int threads = (int)((m.np + 63) / 64 * 64);
    if (nc == 1)
        hipLaunchKernelGGL((grad_volume_kernel<Scalar, Vector>), dim3(m.nbcs), dim3(threads), m.np * sizeof(Scalar),
                           stream(), dev(), p, reinterpret_cast<Vector*>(r));
    else
        hipLaunchKernelGGL((grad_volume_kernel<Vector, Tensor>), dim3(m.nbcs), dim3(threads), m.np * sizeof(Vector),
                           stream(), dev(), reinterpret_cast<const Vector*>(p), reinterpret_cast<Tensor*>(r));
    check(hipGetLastError(), "grad volume");
}
void sum_facet(const Scalar* fF, int nc, Scalar* r) {
    const DevTables t = dev();
    NSEM_DISPATCH(nc, hipLaunchKernelGGL((sum_facet_kernel<T>), dim3(blocks_for(t.nall)), dim3(kThreads), 0, stream(), t,
                                         reinterpret_cast<const T*>(fF), reinterpret_cast<T*>(r)));
    check(hipGetLastError(), "sum");
}
void offdiag(const Scalar* x, int nc, const Scalar* c0, const Scalar* c1, bool add, Scalar* r) {
    const DevTables t = dev();
    NSEM_DISPATCH(nc, hipLaunchKernelGGL((offdiag_kernel<T>), dim3(blocks_for(t.nbf)), dim3(kThreads), 0, stream(), t,
                                         reinterpret_cast<const T*>(x), c0, c1, add ? 1 : 0, reinterpret_cast<T*>(r)));
    check(hipGetLastError(), "matrix off-diagonal product");
}
void lap_coeffs(const Scalar* mu, bool use_fD, Scalar* ano, Scalar* ann, Scalar* ap) {
    const DevTables t = dev();
    hipLaunchKernelGGL(lap_face_kernel, dim3(blocks_for(t.nfn)), dim3(kThreads), 0, stream(), t, mu, use_fD ? 1 : 0, ano, ann);
    hipLaunchKernelGGL(diag_sum_kernel, dim3(blocks_for(t.nall)), dim3(kThreads), 0, stream(), t, ano, ann, ap);
    check(hipGetLastError(), "lap");
}
void conv_coeffs(const Scalar* flux, const Scalar* mu_facet, int scheme, Scalar blend, Scalar* ano, Scalar* ann,
                 Scalar* ap) {
    const DevTables t = dev();
    hipLaunchKernelGGL(conv_face_kernel, dim3(blocks_for(t.nfn)), dim3(kThreads), 0, stream(), t, flux, mu_facet, scheme,
                       blend, ano, ann);
    hipLaunchKernelGGL(diag_sum_kernel, dim3(blocks_for(t.nall)), dim3(kThreads), 0, stream(), t, ano, ann, ap);
    check(hipGetLastError(), "convection");
}
void nonortho_flux(const Scalar* g, const Scalar* x, const Scalar* ano, Scalar* r, int nc) {
    const DevTables t = dev();
    if (nc == 1)
        hipLaunchKernelGGL((nonortho_kernel<Scalar, Vector>), dim3(blocks_for(t.nfn)), dim3(kThreads), 0, stream(), t,
                           reinterpret_cast<const Vector*>(g), x, ano, r);
    else if (nc == 3)
        hipLaunchKernelGGL((nonortho_kernel<Vector, Tensor>), dim3(blocks_for(t.nfn)), dim3(kThreads), 0, stream(), t,
                           reinterpret_cast<const Tensor*>(g), reinterpret_cast<const Vector*>(x), ano,
                           reinterpret_cast<Vector*>(r));
    else
        fail("lap: the non-orthogonal correction of a tensor field is not built");
    check(hipGetLastError(), "non-orthogonal correction");
}
void implicit_bcs(Scalar* ap, Scalar* ann, Scalar* Su, const Scalar* cF, int nc, Int fIndex) {
    if (!has_bcs(fIndex)) return;
    const MeshTables& m = mesh_tables();
    FieldBCs& b = bcs_of(fIndex, nc);
    for (uint32_t p = 0; p < m.npatch; p++) {
        if (b.kind[p] == BC_NONE) continue;
        const PatchArgs pa = patch(p);
        if (!pa.n) continue;
        NSEM_DISPATCH(nc, hipLaunchKernelGGL((implicit_bc_kernel<T>), dim3(1), dim3(64), 0, stream(), dev(), pa, b.kind[p],
                                             b.shape[p], ap, ann, reinterpret_cast<T*>(Su),
                                             reinterpret_cast<const T*>(cF)));
    }
    check(hipGetLastError(), "applyImplicitBCs");
}
void explicit_bcs(Scalar* d, int nc, Int fIndex) {
    NSEM_DISPATCH(nc, explicit_bcs_t<T>(reinterpret_cast<T*>(d), fIndex));
}
void fill_bcs(Scalar* d, int nc, Int bind, int nc_src) {
    NSEM_DISPATCH(nc, fill_bcs_t<T>(reinterpret_cast<T*>(d), bind, nc_src));
}
void reduce(const Scalar* d, int nc, int kind, Scalar* out, size_t n) {
    switch (nc) {
        case 1: reduce_nc<1>(d, kind, out, n); break;
        case 3: reduce_nc<3>(d, kind, out, n); break;
        case 6: reduce_nc<6>(d, kind, out, n); break;
        case 9: reduce_nc<9>(d, kind, out, n); break;
        default: fail("reduction of an unsupported type");
    }
}

void solve_fv_pcg(const Scalar* ap, const Scalar* ano, const Scalar* ann, const Scalar* Su, Scalar* x, Int) {
    const MeshTables& m = mesh_tables();
    if (Controls::Solver != Controls::PCG || Controls::Preconditioner != Controls::DIAG)
        fail("Solve: the FV system is solved by the DIAG-preconditioned conjugate gradient; set method PCG and "
             "preconditioner DIAG");
    std::vector<double> hap(m.ncells), hano((size_t)m.nfacets), hann((size_t)m.nfacets), hsu(m.nbcs), hx(m.nbcs);
    download(hap.data(), ap, m.ncells * sizeof(double));
    download(hano.data(), ano, m.nfacets * sizeof(double));
    download(hann.data(), ann, m.nfacets * sizeof(double));
    download(hsu.data(), Su, m.nbcs * sizeof(double));
    download(hx.data(), x, m.nbcs * sizeof(double));
    nsem_fv* fv = nullptr;
    if (nsem_fv_create_coeffs(mesh_handle(), hap.data(), hano.data(), hann.data(), &fv))
        fail(std::string("Solve: ") + nsem_last_error());
    if (nsem_fv_set_rhs(fv, hsu.data())) fail(std::string("Solve: ") + nsem_last_error());
    nsem_fv_result res{};
    const int rc = nsem_fv_solve(fv, hx.data(), 0 | NSEM_FV_KEEP_GOING, Controls::tolerance,
                                 (int32_t)Controls::max_iterations, &res);
    nsem_fv_destroy(fv);
    if (rc != NSEM_OK && rc != NSEM_ENOTCONVERGED) fail(std::string("Solve: ") + nsem_last_error());
    upload(x, hx.data(), m.nbcs * sizeof(double));
    if (MP::printOn) {
        MP::printH("SYMM-FULL-DIAG-PCG :");
        MP::print("Iterations %d Initial Residual %.5e Final Residual %.5e\n", res.iterations, res.res0_ref, res.res_ref);
    }
}

}  // namespace nsem_field

// ----------------------------------------------------------------- Solve --
template <class T1, class T2, class T3>
static void solve_impl(const MeshMatrix<T1, T2, T3>& cA) {
    MeshMatrix<T1, T2, T3>& A = const_cast<MeshMatrix<T1, T2, T3>&>(cA);
    if (!A.cF) nsem_field::fail("Solve: the matrix has no solution field");
    constexpr int NC = nsem_field::NC<T1>;
    // applyImplicitBCs (field.h:2541-2582)
    nsem_field::implicit_bcs(A.ap.dptr_w(), A.ann.dptr_w(), nsem_field::S(A.Su.dptr_w()),
                             nsem_field::S(A.cF->dptr()), NC, A.cF->fIndex);
    using namespace Controls;
    if (A.flags & A.DIAGONAL) {
        // SolveTexplicit (solve.cpp:563-570)
        *A.cF = A.Su / A.ap;
        if (MP::printOn) {
            MP::printH("SYMM-DIAG-DIAG-JAC:");
            MP::print("Iterations %d Initial Residual %.5e Final Residual %.5e\n", 1, 0.0, 0.0);
        }
    } else if (NC == 1 && Solver == PCG && Preconditioner == DIAG && (A.flags & A.SYMMETRIC)) {
        // the pressure-equation case: the library's conjugate gradient
        nsem_field::solve_fv_pcg(A.ap.dptr(), A.ano.dptr(), A.ann.dptr(), nsem_field::S(A.Su.dptr()),
                                 nsem_field::S(A.cF->dptr_w()), A.cF->fIndex);
    } else {
        nsem_field::solve_general(A.ap.dptr(), A.ano.dptr(), A.ann.dptr(), nsem_field::S(A.Su.dptr()),
                                  nsem_field::S(A.cF->dptr_w()), NC, (int)A.flags);
    }
    applyExplicitBCs(*A.cF, true);
}

void Solve(const MeshMatrix<Scalar>& A) { solve_impl(A); }
void Solve(const MeshMatrix<Vector>& A) { solve_impl(A); }
void Solve(const MeshMatrix<STensor>& A) { solve_impl(A); }
void Solve(const MeshMatrix<Tensor>& A) { solve_impl(A); }
