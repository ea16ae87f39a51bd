// FV pressure Poisson problem on the device and its C-ABI.
//
// Assembly (once per matrix, host): lap(p, mu, true) on the FV mesh
// (field.h:3571-3615: ano = ann = fD cds(mu), ap = sum of the couplings of a
// cell) and the folding of boundary conditions (applyImplicitBCs,
// field.h:2541-2582: DIRICHLET adds ann*value to the source, NEUMANN removes
// the face from the diagonal).  The right-hand side -divf(U)
// (potential.cpp:59-64) and the CG iterations run on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../mesh_access.hpp"
#include "../status.hpp"
#include "fv_kernels.hpp"
#include "nebulasem_amd.h"

using namespace nsem;

#define HIP_TRY(x)                                                                          \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess)                                                               \
            return nsem_fail(NSEM_EDEVICE, std::string(#x ": ") + hipGetErrorString(e_));  \
    } while (0)

namespace {

struct Buf {
    void* p = nullptr;
    ~Buf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t n) { return hipMalloc(&p, n ? n : 8); }
    template <typename T> T* as() const { return (T*)p; }
};

template <typename T>
hipError_t put(Buf& b, const std::vector<T>& v, hipStream_t s) {
    hipError_t e = b.alloc(v.size() * sizeof(T));
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
}

inline double dot3(const V3& p, const V3& q) {
    double r = p[2] * q[2] + 0.0;
    r = p[1] * q[1] + r;
    return p[0] * q[0] + r;
}

}  // namespace

struct nsem_fv {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t n = 0;
    FvOp<double> A64;
    FvOp<float> A32;
    Buf diag, idiag, cpl, nbr, off, diag32, idiag32, cpl32, off32;
    Buf su_bc, su;                       // boundary part of the source, full source
    // divergence inputs
    Buf fi0, fi1, allFaces, FO, FN, fN, fI, patch_of_ghost, ubc_kind, ubc_value, U;
    // velocity correction: pressure ghost rule per ghost cell, cell volumes
    Buf pg_kind, pg_off, cV;
    uint64_t nghost = 0;
    FvDivInput div{};
    // CG work
    Buf x, r, z, p, Ap, x32, r32, z32, p32, Ap32, e32, b32, y64, slots, sc, tmp;
    std::vector<int32_t> patch_kind;
    ~nsem_fv() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

static int fv_build(nsem_fv& d, const Geometry& g, size_t npatch, const std::vector<double>& ap,
                    const std::vector<double>& ano, const std::vector<double>& ann, const std::vector<double>& subc,
                    const std::vector<int32_t>& patch_of_ghost, const std::vector<int32_t>& pg_kind,
                    const std::vector<double>& pg_off);

extern "C" {

int nsem_fv_create(const nsem_mesh* mesh, const double* mu, const int32_t* bc_kind, const double* bc_value,
                   nsem_fv** out) {
    if (!mesh || !bc_kind || !bc_value || !out) return nsem_fail(NSEM_EINVAL, "null argument");
    const Geometry& g = nsem_mesh_geometry(mesh);
    const LoadedMesh& lm = nsem_mesh_loaded(mesh);
    const auto& names = nsem_mesh_boundary_names(mesh);
    if (g.np != 1) return nsem_fail(NSEM_EINVAL, "FV path needs a mesh loaded with order 0");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return nsem_fail(NSEM_EDEVICE, "no HIP device available");
    for (size_t b = 0; b < names.size(); b++)
        if (bc_kind[b] != NSEM_BC_NEUMANN && bc_kind[b] != NSEM_BC_DIRICHLET)
            return nsem_fail(NSEM_EUNSUPPORTED, "pressure boundary must be NEUMANN or DIRICHLET: " + names[b]);
    auto d = std::make_unique<nsem_fv>();
    const uint64_t n = g.nbcs, nc = g.ncells, nf = g.nfacets;
    d->n = n;
    HIP_TRY(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&d->ev0));
    HIP_TRY(hipEventCreate(&d->ev1));
    hipStream_t s = d->stream;

    // couplings ano = ann = fD * cds(mu)
    std::vector<double> ano(nf), ann(nf);
    for (uint64_t k = 0; k < nf; k++) {
        const Int c1 = g.FO[k], c2 = g.FN[k];
        const double m1 = mu ? mu[c1] : 1.0, m2 = mu ? mu[c2] : 1.0;
        const double mf = m1 * g.fI[k] + m2 * (1 - g.fI[k]);
        ano[k] = g.fD[k] * mf;
        ann[k] = g.fD[k] * mf;
    }
    std::vector<double> ap(nc, 0.0), subc(n, 0.0);
    for (uint64_t i = 0; i < nc; i++)
        for (Int f = g.faceIndices0[i]; f < g.faceIndices1[i]; f++) {
            const Int k = g.allFaces[f];
            const Int c1 = g.FO[k], c2 = g.FN[k];
            if (c1 == i) ap[c1] += ano[k];
            else if (c2 < nc) ap[c2] += ann[k];
        }
    d->nghost = nc - n;
    std::vector<int32_t> patch_of_ghost(nc - n, 0), pg_kind(nc - n, 0);
    std::vector<double> pg_off(nc - n, 0.0);
    {
        size_t b = 0;
        for (const auto& kv : lm.boundaries) {
            for (Int k : kv.second) {
                const Int c1 = g.FO[k], c2 = g.FN[k];
                if (c2 >= nc) continue;
                patch_of_ghost[c2 - n] = (int32_t)b;
                if (bc_kind[b] == NSEM_BC_NEUMANN) {
                    const V3 dv(g.cC[c2][0] - g.cC[c1][0], g.cC[c2][1] - g.cC[c1][1], g.cC[c2][2] - g.cC[c1][2]);
                    ap[c1] -= ann[k];
                    subc[c1] += ann[k] * (bc_value[b] * std::sqrt(dot3(dv, dv)));
                    pg_kind[c2 - n] = 0;
                    pg_off[c2 - n] = bc_value[b] * std::sqrt(dot3(dv, dv));
                } else {
                    subc[c1] += ann[k] * bc_value[b];
                    pg_kind[c2 - n] = 2;
                    pg_off[c2 - n] = bc_value[b];
                }
                ann[k] = 0;
            }
            b++;
        }
    }
    int rc = fv_build(*d, g, names.size(), ap, ano, ann, subc, patch_of_ghost, pg_kind, pg_off);
    if (rc) return rc;
    *out = d.release();
    return NSEM_OK;
}

}  // extern "C"

// The operator layout (structured block or six slots per cell), the FP32
// copies, the divergence and correction inputs and the CG work vectors,
// from assembled coefficients: ap per cell, ano / ann per face (boundary
// faces already folded into ap and the source), subc the boundary part of
// the source.
static int fv_build(nsem_fv& dref, const Geometry& g, size_t npatch, const std::vector<double>& ap,
                    const std::vector<double>& ano, const std::vector<double>& ann, const std::vector<double>& subc,
                    const std::vector<int32_t>& patch_of_ghost, const std::vector<int32_t>& pg_kind,
                    const std::vector<double>& pg_off) {
    nsem_fv* d = &dref;
    const uint64_t n = g.nbcs, nf = g.nfacets;
    hipStream_t s = d->stream;
    std::vector<double> dg(ap.begin(), ap.begin() + n), idg(n);
    for (uint64_t i = 0; i < n; i++) idg[i] = 1 / dg[i];

    // structured block numbering?  c = (i*ny + j)*nz + k
    bool structured = true;
    int64_t nz = 0, ny = 0, nx = 0;
    {
        // z extent: the run of cells linked to their successor from cell 0
        nz = 1; ny = 1; nx = 1;
        {
            uint64_t c = 0;
            while (c + 1 < n) {
                bool adj = false;
                for (Int f = g.faceIndices0[c]; f < g.faceIndices1[c]; f++) {
                    const Int k = g.allFaces[f];
                    const Int o = g.FO[k] == (Int)c ? g.FN[k] : g.FO[k];
                    if (o == (Int)(c + 1)) adj = true;
                }
                if (!adj) break;
                c++;
            }
            nz = (int64_t)c + 1;
        }
        // y extent: neighbour of 0 at +nz repeated
        auto linked = [&](int64_t a, int64_t b) {
            if (a < 0 || b < 0 || (uint64_t)a >= n || (uint64_t)b >= n) return false;
            for (Int f = g.faceIndices0[a]; f < g.faceIndices1[a]; f++) {
                const Int k = g.allFaces[f];
                const Int o = g.FO[k] == (Int)a ? g.FN[k] : g.FO[k];
                if (o == (Int)b) return true;
            }
            return false;
        };
        ny = 1;
        while (linked((ny - 1) * nz, ny * nz)) ny++;
        if (n % (uint64_t)(ny * nz) != 0) structured = false;
        else nx = n / (ny * nz);
        if (nx * ny * nz != (int64_t)n || nx > (1 << 20) || ny > (1 << 20) || nz > (1 << 20)) structured = false;
    }
    std::vector<double> cpl;
    std::vector<uint32_t> nbr;
    std::vector<double> off;
    if (structured) {
        cpl.assign(3 * n, 0.0);
        const int64_t st[3] = {ny * nz, nz, 1};
        for (uint64_t k = 0; k < nf && structured; k++) {
            const Int c1 = g.FO[k], c2 = g.FN[k];
            if (c2 >= n) continue;
            const int64_t lo = std::min<int64_t>(c1, c2), hi = std::max<int64_t>(c1, c2);
            const int64_t dlt = hi - lo;
            int dir = -1;
            const int64_t i0 = lo / (ny * nz), j0 = (lo / nz) % ny, k0 = lo % nz;
            if (dlt == st[0] && i0 + 1 < nx) dir = 0;
            else if (dlt == st[1] && j0 + 1 < ny) dir = 1;
            else if (dlt == st[2] && k0 + 1 < nz) dir = 2;
            if (dir < 0 || ano[k] != ann[k]) { structured = false; break; }
            cpl[(size_t)dir * n + lo] = ano[k];
        }
        // every interior adjacency of the block must be present
        if (structured) {
            uint64_t want = (uint64_t)(nx - 1) * ny * nz + (uint64_t)nx * (ny - 1) * nz + (uint64_t)nx * ny * (nz - 1);
            uint64_t have = 0;
            for (uint64_t k = 0; k < nf; k++) if (g.FN[k] < n) have++;
            if (want != have) structured = false;
        }
    }
    if (!structured) {
        nbr.assign(6 * n, 0xffffffffu);
        off.assign(6 * n, 0.0);
        for (uint64_t i = 0; i < n; i++) {
            int sl = 0;
            for (Int f = g.faceIndices0[i]; f < g.faceIndices1[i]; f++) {
                const Int k = g.allFaces[f];
                const Int c1 = g.FO[k], c2 = g.FN[k];
                Int o;
                double a;
                if (c1 == i) { o = c2; a = ann[k]; }
                else { o = c1; a = ano[k]; }
                if (o >= n || a == 0) continue;
                if (sl >= 6) return nsem_fail(NSEM_EUNSUPPORTED, "more than six neighbours per cell");
                nbr[(size_t)sl * n + i] = o;
                off[(size_t)sl * n + i] = a;
                sl++;
            }
        }
    }
    hipError_t e;
    std::vector<float> dg32(dg.begin(), dg.end()), idg32(n);
    for (uint64_t i = 0; i < n; i++) idg32[i] = 1.0f / dg32[i];
    if ((e = put(d->diag, dg, s)) || (e = put(d->idiag, idg, s)) || (e = put(d->su_bc, subc, s)) ||
        (e = put(d->diag32, dg32, s)) || (e = put(d->idiag32, idg32, s)))
        return nsem_fail(NSEM_EDEVICE, std::string("upload: ") + hipGetErrorString(e));
    FvOp<double>& A = d->A64;
    FvOp<float>& B = d->A32;
    A.n = B.n = n;
    A.diag = d->diag.as<double>(); A.idiag = d->idiag.as<double>();
    B.diag = d->diag32.as<float>(); B.idiag = d->idiag32.as<float>();
    if (structured) {
        std::vector<float> c32(cpl.begin(), cpl.end());
        if ((e = put(d->cpl, cpl, s)) || (e = put(d->cpl32, c32, s)))
            return nsem_fail(NSEM_EDEVICE, std::string("upload: ") + hipGetErrorString(e));
        A.structured = B.structured = 1;
        A.nx = B.nx = (int)nx; A.ny = B.ny = (int)ny; A.nz = B.nz = (int)nz;
        A.sx = B.sx = ny * nz; A.sy = B.sy = nz; A.sz = B.sz = 1;
        A.cpl = d->cpl.as<double>(); B.cpl = d->cpl32.as<float>();
        // uniform couplings: every interior adjacency of a direction carries
        // the same value, bit for bit (a uniform grid, constant mu)
        bool uni = true;
        double cu[3] = {0, 0, 0};
        bool seen[3] = {false, false, false};
        const int64_t st[3] = {ny * nz, nz, 1};
        for (uint64_t k = 0; k < nf && uni; k++) {
            const Int c1 = g.FO[k], c2 = g.FN[k];
            if (c2 >= n) continue;
            const int64_t lo = std::min<int64_t>(c1, c2), dlt = std::max<int64_t>(c1, c2) - lo;
            const int dir = dlt == st[0] ? 0 : dlt == st[1] ? 1 : 2;
            const double v = cpl[(size_t)dir * n + lo];
            if (!seen[dir]) { cu[dir] = v; seen[dir] = true; }
            else if (std::memcmp(&v, &cu[dir], sizeof(double)) != 0) uni = false;
        }
        if (uni) {
            A.uniform = B.uniform = 1;
            for (int q = 0; q < 3; q++) { A.cu[q] = cu[q]; B.cu[q] = (float)cu[q]; }
        }
    } else {
        std::vector<float> o32(off.begin(), off.end());
        if ((e = put(d->nbr, nbr, s)) || (e = put(d->off, off, s)) || (e = put(d->off32, o32, s)))
            return nsem_fail(NSEM_EDEVICE, std::string("upload: ") + hipGetErrorString(e));
        A.nbr = B.nbr = d->nbr.as<uint32_t>();
        A.off = d->off.as<double>(); B.off = d->off32.as<float>();
    }
    // divergence inputs
    std::vector<double> fnv(3 * nf);
    for (uint64_t k = 0; k < nf; k++) for (int c = 0; c < 3; c++) fnv[3 * k + c] = g.fN[k][c];
    if ((e = put(d->fi0, g.faceIndices0, s)) || (e = put(d->fi1, g.faceIndices1, s)) ||
        (e = put(d->allFaces, g.allFaces, s)) || (e = put(d->FO, g.FO, s)) || (e = put(d->FN, g.FN, s)) ||
        (e = put(d->fN, fnv, s)) || (e = put(d->fI, g.fI, s)) || (e = put(d->patch_of_ghost, patch_of_ghost, s)) ||
        (e = put(d->pg_kind, pg_kind, s)) || (e = put(d->pg_off, pg_off, s)) ||
        (e = put(d->cV, std::vector<double>(g.cV.begin(), g.cV.begin() + n), s)))
        return nsem_fail(NSEM_EDEVICE, std::string("upload: ") + hipGetErrorString(e));
    HIP_TRY(d->ubc_kind.alloc(npatch * sizeof(int32_t) + 8));
    HIP_TRY(d->ubc_value.alloc(npatch * 3 * sizeof(double) + 8));
    HIP_TRY(d->U.alloc(3 * n * sizeof(double)));
    FvDivInput& dv = d->div;
    dv.ncells = n;
    dv.fi0 = d->fi0.as<uint32_t>(); dv.fi1 = d->fi1.as<uint32_t>(); dv.allFaces = d->allFaces.as<uint32_t>();
    dv.FO = d->FO.as<uint32_t>(); dv.FN = d->FN.as<uint32_t>();
    dv.fN = d->fN.as<double>(); dv.fI = d->fI.as<double>(); dv.U = d->U.as<double>();
    dv.patch_of_ghost = d->patch_of_ghost.as<int32_t>();
    dv.bc_kind = d->ubc_kind.as<int32_t>(); dv.bc_value = d->ubc_value.as<double>();
    dv.su_bc = d->su_bc.as<double>();
    // work vectors
    for (Buf* b : {&d->su, &d->x, &d->r, &d->z, &d->p, &d->Ap, &d->y64}) HIP_TRY(b->alloc(n * sizeof(double)));
    for (Buf* b : {&d->x32, &d->r32, &d->z32, &d->p32, &d->Ap32, &d->e32, &d->b32}) HIP_TRY(b->alloc(n * sizeof(float)));
    HIP_TRY(d->slots.alloc(kCgSlots * 4 * sizeof(double)));
    HIP_TRY(d->sc.alloc(sizeof(CgScalars)));
    HIP_TRY(d->tmp.alloc(8 * sizeof(double)));
    HIP_TRY(hipMemcpyAsync(d->su.p, subc.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return NSEM_OK;
}

extern "C" {

int nsem_fv_create_coeffs(const nsem_mesh* mesh, const double* ap, const double* ano, const double* ann,
                          nsem_fv** out) {
    if (!mesh || !ap || !ano || !ann || !out) return nsem_fail(NSEM_EINVAL, "null argument");
    const Geometry& g = nsem_mesh_geometry(mesh);
    if (g.np != 1) return nsem_fail(NSEM_EINVAL, "FV path needs a mesh loaded with order 0");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return nsem_fail(NSEM_EDEVICE, "no HIP device available");
    const uint64_t n = g.nbcs, nc = g.ncells, nf = g.nfacets;
    for (uint64_t i = 0; i < n; i++)
        if (!(ap[i] != 0) || !std::isfinite(ap[i])) return nsem_fail(NSEM_EINVAL, "zero or non-finite diagonal");
    auto d = std::make_unique<nsem_fv>();
    d->n = n;
    HIP_TRY(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&d->ev0));
    HIP_TRY(hipEventCreate(&d->ev1));
    d->nghost = nc - n;
    std::vector<double> vap(ap, ap + nc), vano(ano, ano + nf), vann(ann, ann + nf), subc(n, 0.0);
    // couplings to ghost cells: the caller has folded the boundary
    // conditions into the diagonal and the source
    for (uint64_t k = 0; k < nf; k++)
        if (g.FN[k] >= n) vann[k] = 0;
    std::vector<int32_t> pog(nc - n, 0), pgk(nc - n, 0);
    std::vector<double> pgo(nc - n, 0.0);
    int rc = fv_build(*d, g, nsem_mesh_boundary_names(mesh).size(), vap, vano, vann, subc, pog, pgk, pgo);
    if (rc) return rc;
    *out = d.release();
    return NSEM_OK;
}

void nsem_fv_destroy(nsem_fv* d) { delete d; }

int nsem_fv_info(const nsem_fv* d, uint64_t* ncells, int32_t* structured) {
    if (!d || !ncells || !structured) return nsem_fail(NSEM_EINVAL, "null argument");
    *ncells = d->n;
    // 0 general, 1 structured block numbering, 2 structured with uniform couplings
    *structured = d->A64.structured ? (d->A64.uniform ? 2 : 1) : 0;
    return NSEM_OK;
}

int nsem_fv_set_rhs_divergence(nsem_fv* d, const double* U, const int32_t* bc_kind, const double* bc_value,
                               int32_t nboundaries) {
    if (!d || !U || !bc_kind || !bc_value || nboundaries < 0) return nsem_fail(NSEM_EINVAL, "bad argument");
    hipStream_t s = d->stream;
    HIP_TRY(hipMemcpyAsync(d->U.p, U, 3 * d->n * sizeof(double), hipMemcpyHostToDevice, s));
    if (nboundaries) {
        HIP_TRY(hipMemcpyAsync(d->ubc_kind.p, bc_kind, nboundaries * sizeof(int32_t), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d->ubc_value.p, bc_value, 3 * nboundaries * sizeof(double), hipMemcpyHostToDevice, s));
    }
    HIP_TRY(fv_divu(d->div, d->su.as<double>(), s));
    HIP_TRY(hipStreamSynchronize(s));
    return NSEM_OK;
}

int nsem_fv_set_rhs(nsem_fv* d, const double* su) {
    if (!d || !su) return nsem_fail(NSEM_EINVAL, "null argument");
    HIP_TRY(hipMemcpyAsync(d->su.p, su, d->n * sizeof(double), hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    return NSEM_OK;
}

int nsem_fv_get_rhs(nsem_fv* d, double* su) {
    if (!d || !su) return nsem_fail(NSEM_EINVAL, "null argument");
    HIP_TRY(hipMemcpyAsync(su, d->su.p, d->n * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    return NSEM_OK;
}

int nsem_fv_correct_velocity(nsem_fv* d, const double* p, double* U) {
    if (!d || !p || !U) return nsem_fail(NSEM_EINVAL, "null argument");
    hipStream_t s = d->stream;
    HIP_TRY(hipMemcpyAsync(d->y64.p, p, d->n * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d->U.p, U, 3 * d->n * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(fv_gradcorr(d->div, d->y64.as<double>(), d->pg_kind.as<int32_t>(), d->pg_off.as<double>(),
                        d->cV.as<double>(), d->nghost, d->U.as<double>(), s));
    HIP_TRY(hipMemcpyAsync(U, d->U.p, 3 * d->n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return NSEM_OK;
}

int nsem_fv_apply(nsem_fv* d, const double* x, double* y) {
    if (!d || !x || !y) return nsem_fail(NSEM_EINVAL, "null argument");
    hipStream_t s = d->stream;
    HIP_TRY(hipMemcpyAsync(d->x.p, x, d->n * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(fv_apply(d->A64, d->x.as<double>(), d->y64.as<double>(), s));
    HIP_TRY(hipMemcpyAsync(y, d->y64.p, d->n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return NSEM_OK;
}

}  // extern "C"

namespace {

template <typename R>
int run_cg(nsem_fv* d, const FvOp<R>& A, const R* b, R* x, R* r, R* z, R* p, R* Ap, int mode, double tol,
           int max_it, CgScalars& hs) {
    hipStream_t s = d->stream;
    CgScalars* sc = d->sc.as<CgScalars>();
    CgScalars init{};
    init.tol = tol;
    init.max_it = max_it;
    init.mode = mode & 3;
    init.keep_going = (mode & NSEM_FV_KEEP_GOING) ? 1 : 0;
    HIP_TRY(hipMemcpyAsync(sc, &init, sizeof(init), hipMemcpyHostToDevice, s));
    HIP_TRY(fv_cg_init<R>(A, b, x, r, z, p, sc, d->slots.as<double>(), s));
    int launched = 0;
    int batch = 8;
    while (true) {
        HIP_TRY(fv_cg_iterations<R>(A, batch, x, r, z, p, Ap, sc, d->slots.as<double>(), s));
        launched += batch;
        HIP_TRY(hipMemcpyAsync(&hs, sc, sizeof(hs), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (hs.done || launched > max_it + 8) break;
        if (batch < 64) batch *= 2;
    }
    return NSEM_OK;
}

}  // namespace

extern "C" int nsem_fv_solve(nsem_fv* d, double* x_io, int32_t mode, double tol, int32_t max_it,
                             nsem_fv_result* res) {
    if (!d || !x_io || !res) return nsem_fail(NSEM_EINVAL, "null argument");
    const int32_t keep_going = mode & NSEM_FV_KEEP_GOING;
    mode &= ~NSEM_FV_KEEP_GOING;
    if (mode < 0 || mode > 2 || !(tol > 0) || max_it < 1) return nsem_fail(NSEM_EINVAL, "bad solver settings");
    if (keep_going && mode != 0) return nsem_fail(NSEM_EINVAL, "NSEM_FV_KEEP_GOING applies to mode 0 only");
    std::memset(res, 0, sizeof(*res));
    hipStream_t s = d->stream;
    const uint64_t n = d->n;
    HIP_TRY(hipMemcpyAsync(d->x.p, x_io, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipEventRecord(d->ev0, s));
    CgScalars hs{};
    int rc = NSEM_OK;
    if (mode <= 1) {
        rc = run_cg<double>(d, d->A64, d->su.as<double>(), d->x.as<double>(), d->r.as<double>(), d->z.as<double>(),
                            d->p.as<double>(), d->Ap.as<double>(), mode | keep_going, tol, max_it, hs);
        if (rc) return rc;
        res->iterations = hs.iters;
        res->res0_ref = hs.res0_ref;
        res->res_ref = hs.res_ref;
        res->outer_iterations = 1;
    } else {
        // FP32 inner CG, FP64 residual and correction (iterative refinement)
        double* tmp = d->tmp.as<double>();
        double bb = 0;
        HIP_TRY(fv_norm2<double>(n, d->su.as<double>(), tmp, d->slots.as<double>(), s));
        HIP_TRY(hipMemcpyAsync(&bb, tmp, sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        const double bn = std::sqrt(bb);
        int total = 0, outer = 0;
        double rel = 0;
        for (; outer < 30; outer++) {
            double rr = 0;
            HIP_TRY(fv_residual64(d->A64, d->su.as<double>(), d->x.as<double>(), d->y64.as<double>(),
                                  d->b32.as<float>(), tmp, d->slots.as<double>(), s));
            HIP_TRY(hipMemcpyAsync(&rr, tmp, sizeof(double), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            rel = bn > 0 ? std::sqrt(rr) / bn : 0;
            if (outer == 0) res->res0_ref = rel;
            if (rel <= tol || total >= max_it) break;
            HIP_TRY(hipMemsetAsync(d->e32.p, 0, n * sizeof(float), s));
            const double inner_tol = std::max(1e-5, 0.5 * tol / std::max(rel, 1e-300));
            rc = run_cg<float>(d, d->A32, d->b32.as<float>(), d->e32.as<float>(), d->r32.as<float>(),
                               d->z32.as<float>(), d->p32.as<float>(), d->Ap32.as<float>(), 1,
                               std::min(inner_tol, 1e-2), max_it - total, hs);
            if (rc) return rc;
            total += hs.iters;
            HIP_TRY(fv_correct(n, d->x.as<double>(), d->e32.as<float>(), s));
        }
        res->iterations = total;
        res->outer_iterations = outer;
        res->res_ref = 0;
        hs.res_true = rel;
        hs.done = rel <= tol;
    }
    HIP_TRY(hipEventRecord(d->ev1, s));
    HIP_TRY(hipEventSynchronize(d->ev1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, d->ev0, d->ev1));
    res->seconds = ms * 1e-3;
    // true residual of the returned solution
    double rr = 0, bb = 0;
    double* tmp2 = d->tmp.as<double>();
    HIP_TRY(fv_residual64(d->A64, d->su.as<double>(), d->x.as<double>(), d->y64.as<double>(), nullptr,
                          tmp2, d->slots.as<double>(), s));
    HIP_TRY(fv_norm2<double>(n, d->su.as<double>(), tmp2 + 1, d->slots.as<double>(), s));
    HIP_TRY(hipMemcpyAsync(&rr, tmp2, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&bb, tmp2 + 1, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(x_io, d->x.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    res->res_true = bb > 0 ? std::sqrt(rr / bb) : 0;
    res->breakdown = hs.breakdown;
    (void)0;
    if (hs.breakdown) return nsem_fail(NSEM_EBREAKDOWN, "PCG breakdown (p.Ap <= 0)");
    const double crit = mode == 0 ? res->res_ref : res->res_true;
    if (!(crit <= tol)) return nsem_fail(NSEM_ENOTCONVERGED, "iteration limit reached");
    return NSEM_OK;
}
