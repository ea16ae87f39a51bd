// DG Euler problem on the device: set-up from a loaded mesh, state transfer,
// and the C-ABI entry points that run steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../host/parallel.hpp"
#include "../mesh_access.hpp"
#include "../status.hpp"
#include "dg_common.hpp"
#include "dg_kernels.hpp"
#include "nebulasem_amd.h"

using namespace nsem;

#include "dg_context.hpp"

namespace {

template <typename T>
hipError_t upload(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
    hipError_t e = b.alloc(v.size() * sizeof(T));
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
}

}  // namespace

extern "C" {

int nsem_dg_create(const nsem_mesh* mesh, const nsem_euler_params* prm, const int32_t* bc_kind,
                   const double* bc_value, nsem_dg** out) {
    if (!mesh || !prm || !bc_kind || !bc_value || !out) return nsem_fail(NSEM_EINVAL, "null argument");
    const Geometry& g = nsem_mesh_geometry(mesh);
    const LoadedMesh& lm = nsem_mesh_loaded(mesh);
    const std::vector<std::string>& names = nsem_mesh_boundary_names(mesh);
    if (g.np < 2) return nsem_fail(NSEM_EINVAL, "DG path needs polynomial order >= 1");
    if (g.np > 1024) return nsem_fail(NSEM_EUNSUPPORTED, "at most 1024 nodes per element");
    if (names.size() > (size_t)kMaxPatches) return nsem_fail(NSEM_EUNSUPPORTED, "too many boundary patches");
    {
        const int epb = g.np >= 256 ? 1 : 256 / (int)g.np;
        if (15 * (size_t)g.np * sizeof(double) > 65536 || g.np > 512 ||
            rk_lds_bytes(g.npx, g.npy, g.npz, epb, sizeof(double)) > 65536)
            return nsem_fail(NSEM_EUNSUPPORTED, "element too large for the LDS staging (order <= 7 in 3D)");
    }
    if (!(prm->dt > 0)) return nsem_fail(NSEM_EINVAL, "dt must be positive");
    if (prm->diffusion && !(prm->viscosity >= 0 && prm->Pr > 0))
        return nsem_fail(NSEM_EINVAL, "diffusion needs viscosity >= 0 and Pr > 0");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return nsem_fail(NSEM_EDEVICE, "no HIP device available");

    auto d = std::make_unique<nsem_dg>();
    d->prm = *prm;
    const uint64_t nb = g.nbcs_field;
    d->nb = nb;
    HIP_TRY(hipGetDevice(&d->device));
    HIP_TRY(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&d->ev0));
    HIP_TRY(hipEventCreate(&d->ev1));
    hipStream_t s = d->stream;

    // boundary table: scalars must be zero-gradient on this path
    BcTable bt;
    std::memset(&bt, 0, sizeof(bt));
    for (size_t b = 0; b < names.size(); b++) {
        for (int f = 0; f < 4; f++) {
            int k = bc_kind[4 * b + f];
            if (k < 0) return nsem_fail(NSEM_EINVAL, "no boundary condition for patch " + names[b]);
            if (k > BC_DIRICHLET)
                return nsem_fail(NSEM_EUNSUPPORTED, "boundary kind " + std::to_string(k) + " on patch " + names[b] +
                                                        " is not supported by the DG path");
            if (f != 1 && k == BC_DIRICHLET)
                return nsem_fail(NSEM_EUNSUPPORTED, "DIRICHLET on rho/T/p is not supported by the DG path");
            bt.kind[f][b] = k;
        }
        for (int v = 0; v < 6; v++) bt.value[v][b] = bc_value[6 * b + v];
    }
    // facet -> patch
    std::vector<int32_t> patch_of(g.nfacets, -1);
    {
        int b = 0;
        for (const auto& kv : lm.boundaries) {
            for (Int f : kv.second) patch_of[f] = b;
            b++;
        }
    }
    // element face table; a side holding a 2:1 mortar face (FMC 1 or 2) --
    // the refined element's face, or a coarse element's split side -- gets
    // a slot in the mortar surface-term buffer
    std::vector<EFace> ef((size_t)g.nbcs * 6);
    std::vector<int64_t> side_slot((size_t)g.nbcs * 6, -1);
    uint32_t nslots = 0;
    for (Int ci = 0; ci < g.nbcs; ci++) {
        const IntVec& c = lm.cells[ci];
        for (size_t mm = 0; mm < c.size(); mm++) {
            Int fi = c[mm];
            Int grp = lm.faceid[ci][mm];
            if (grp > 5) return nsem_fail(NSEM_EUNSUPPORTED, "bad face group");
            EFace& e = ef[(size_t)ci * 6 + grp];
            if (lm.FMC[fi] > 0) {
                if (side_slot[(size_t)ci * 6 + grp] < 0) {
                    if (nslots >= (1u << 24)) return nsem_fail(NSEM_EUNSUPPORTED, "too many mortar sides");
                    side_slot[(size_t)ci * 6 + grp] = nslots;
                    e.facet = fi;
                    e.other = ci;
                    e.meta = 16u | 32u | 64u | (nslots << 8);
                    nslots++;
                }
                continue;
            }
            if ((e.meta >> 4) & 1) return nsem_fail(NSEM_EUNSUPPORTED, "a cell side with several conforming faces");
            bool own = lm.FOC[fi] == ci;
            Int other = own ? lm.FNC[fi] : lm.FOC[fi];
            Int og = grp ^ 1;
            bool ghost = other >= g.nbcs;
            if (!ghost) {
                const IntVec& cn = lm.cells[other];
                bool found = false;
                for (size_t t = 0; t < cn.size(); t++)
                    if (cn[t] == fi) { og = lm.faceid[other][t]; found = true; break; }
                if (!found) return nsem_fail(NSEM_EINVAL, "inconsistent face connectivity");
                // both sides must index the face with the same extents
                auto ext = [&](Int gg, Int& na, Int& nb2) {
                    if (gg < 2) { na = g.npx; nb2 = g.npy; }
                    else if (gg < 4) { na = g.npx; nb2 = g.npz; }
                    else { na = g.npy; nb2 = g.npz; }
                };
                Int a0, b0, a1, b1;
                ext(grp, a0, b0);
                ext(og, a1, b1);
                if (a0 != a1 || b0 != b1)
                    return nsem_fail(NSEM_EUNSUPPORTED, "face extents differ across an interior face");
            } else {
                if (!own) return nsem_fail(NSEM_EINVAL, "boundary face not owned by an interior cell");
                if (patch_of[fi] < 0) return nsem_fail(NSEM_EINVAL, "boundary face outside every patch");
            }
            e.facet = fi;
            e.other = other;
            e.meta = og | (own ? 8u : 0u) | 16u | (ghost ? 32u : 0u) |
                     ((uint32_t)(ghost ? patch_of[fi] : 0) << 8);
        }
    }
    // 2:1 mortar faces (field.h:2157-2183): the coarse element and its side,
    // the half of that side the face covers (its centre against the side's
    // centre along the side's node directions), the refined side's node and
    // surface-term entry per face node
    const int NPF = g.npf;
    const int npfm = std::max(g.npx * g.npy, std::max(g.npx * g.npz, g.npy * g.npz));
    std::vector<MortarFace> mfaces;
    std::vector<Int> mface_of;
    std::vector<uint64_t> mfine;
    std::vector<uint32_t> mfms;
    std::vector<double> mw;
    for (Int fi = 0; fi < g.nfacets; fi++) {
        const Int fm = lm.FMC[fi];
        if (fm < 1) continue;
        if (g.np <= 1) return nsem_fail(NSEM_EUNSUPPORTED, "mortar faces on a finite-volume mesh");
        MortarFace mf{};
        const Int cn = fm == 1 ? lm.FNC[fi] : lm.FOC[fi];
        const Int cf = fm == 1 ? lm.FOC[fi] : lm.FNC[fi];
        if (cn >= g.nbcs || cf >= g.nbcs) return nsem_fail(NSEM_EINVAL, "mortar face on the boundary");
        auto side_of = [&](Int cell, Int face) -> int {
            const IntVec& c = lm.cells[cell];
            for (size_t t = 0; t < c.size(); t++) if (c[t] == face) return (int)lm.faceid[cell][t];
            return -1;
        };
        const int fid = side_of(cn, fi), gf = side_of(cf, fi);
        if (fid < 0 || gf < 0) return nsem_fail(NSEM_EINVAL, "inconsistent mortar connectivity");
        double ccn[3] = {0, 0, 0};
        int nch = 0;
        {
            const IntVec& c = lm.cells[cn];
            for (size_t t = 0; t < c.size(); t++)
                if ((int)lm.faceid[cn][t] == fid) {
                    for (int q = 0; q < 3; q++) ccn[q] += lm.fc[c[t]][q];
                    nch++;
                }
            for (int q = 0; q < 3; q++) ccn[q] /= (double)nch;
        }
        const int n0 = fid < 2 ? g.npx : fid < 4 ? g.npx : g.npy;
        const int n1 = fid < 2 ? g.npy : g.npz;
        auto node = [&](Int cell, int side, int a, int b) {
            const int v = (side & 1) ? (side < 2 ? g.npz - 1 : side < 4 ? g.npy - 1 : g.npx - 1) : 0;
            int i, j, k;
            if (side < 2) { i = a; j = b; k = v; }
            else if (side < 4) { i = a; j = v; k = b; }
            else { i = v; j = a; k = b; }
            return (uint64_t)cell * g.np + (uint64_t)(i * g.npy * g.npz + j * g.npz + k);
        };
        auto dotr = [](const double* p, const double* q) {
            double r = p[2] * q[2] + 0.0;
            r = p[1] * q[1] + r;
            return p[0] * q[0] + r;
        };
        {
            const V3& p0 = g.cC[node(cn, fid, 0, 0)];
            const V3& pa = g.cC[node(cn, fid, n0 - 1, 0)];
            const V3& pb = g.cC[node(cn, fid, 0, n1 - 1)];
            double dn[3], dO[3], ea[3], eb[3];
            for (int q = 0; q < 3; q++) {
                dn[q] = ccn[q] - p0[q];
                dO[q] = lm.fc[fi][q] - p0[q];
                ea[q] = pa[q] - p0[q];
                eb[q] = pb[q] - p0[q];
            }
            mf.h0 = dotr(dn, ea) >= dotr(dO, ea) ? 0 : 1;
            mf.h1 = dotr(dn, eb) >= dotr(dO, eb) ? 0 : 1;
        }
        mf.coarse = (uint32_t)cn;
        mf.side = fid;
        mf.fine_owner = fm == 1 ? 1 : 0;
        mf.n0 = n0;
        mf.n1 = n1;
        for (int q = 0; q < 3; q++) mf.nf[q] = lm.fn[fi][q];
        const int64_t fslot = side_slot[(size_t)cf * 6 + gf];
        if (fslot < 0) return nsem_fail(NSEM_EINVAL, "mortar face without a slot");
        const int dA = fid < 4 ? 0 : 1, dB = fid < 2 ? 1 : 2;
        const int nbx = gf < 2 ? g.npy : g.npz;
        for (int t = 0; t < npfm; t++) {
            uint64_t fnode = (uint64_t)cf * g.np;
            uint32_t fms = 0;
            double wk = 0;
            if (t < n0 * n1) {
                const uint64_t k = (uint64_t)fi * NPF + t;
                fnode = fm == 2 ? g.FN[k] : g.FO[k];
                if (fnode / g.np != (uint64_t)cf) return nsem_fail(NSEM_EINVAL, "mortar face node outside its element");
                const int q = (int)(fnode - (uint64_t)cf * g.np);
                const int i = q / (g.npy * g.npz), j = (q / g.npz) % g.npy, kk = q % g.npz;
                int fa, fb;
                if (gf < 2) { fa = i; fb = j; }
                else if (gf < 4) { fa = i; fb = kk; }
                else { fa = j; fb = kk; }
                fms = (uint32_t)(fslot * npfm + fa * nbx + fb);
                const int a = t / n1, b = t - a * n1;
                wk = g.basis[dA].w[a] * g.basis[dB].w[b] / 4;
            }
            mfine.push_back(fnode);
            mfms.push_back(fms);
            mw.push_back(wk);
        }
        mfaces.push_back(mf);
        mface_of.push_back(fi);
    }
    // split coarse sides: element, side, children, slot, child faces
    std::vector<uint32_t> cslot;
    {
        std::vector<std::vector<uint32_t>> kids((size_t)g.nbcs * 6);
        for (size_t f = 0; f < mfaces.size(); f++)
            kids[(size_t)mfaces[f].coarse * 6 + mfaces[f].side].push_back((uint32_t)f);
        for (size_t cs = 0; cs < kids.size(); cs++) {
            if (kids[cs].empty()) continue;
            if (kids[cs].size() > 4) return nsem_fail(NSEM_EUNSUPPORTED, "a split side with more than 4 faces");
            const int64_t sl = side_slot[cs];
            if (sl < 0) return nsem_fail(NSEM_EINVAL, "split side without a slot");
            uint32_t rec[8] = {(uint32_t)(cs / 6), (uint32_t)(cs % 6), (uint32_t)kids[cs].size(), (uint32_t)sl,
                               0, 0, 0, 0};
            for (size_t q = 0; q < kids[cs].size(); q++) rec[4 + q] = kids[cs][q];
            cslot.insert(cslot.end(), rec, rec + 8);
        }
    }
    // metric, volumes, background (euler.cpp:104-131)
    std::vector<double> JcV(9 * nb), cV(nb), pref(nb), rref(nb), ghv(nb, 0.0);
    std::vector<float> JcV32(9 * nb), cV32(nb);
    const double R = prm->cp - prm->cv, gam = prm->cp / prm->cv;
    parallel_for(nb, [&](uint64_t nbeg, uint64_t nend) {
    for (uint64_t n = nbeg; n < nend; n++) {
        for (int c = 0; c < 9; c++) {
            JcV[c * nb + n] = g.Jinv[n][c] * g.cV[n];
            JcV32[c * nb + n] = (float)JcV[c * nb + n];
        }
        cV[n] = g.cV[n];
        cV32[n] = (float)cV[n];
        const V3& x = g.cC[n];
        if (prm->buoyancy) {
            double gh = prm->gravity[2] * x[2] + 0.0;
            gh = prm->gravity[1] * x[1] + gh;
            gh = prm->gravity[0] * x[0] + gh;
            ghv[n] = gh;
            pref[n] = prm->P0 * std::pow(1.0 + gh / (prm->cp * prm->T0), prm->cp / R);
            rref[n] = (prm->P0 / (R * prm->T0)) * std::pow(pref[n] / prm->P0, 1 / gam);
        } else {
            pref[n] = prm->P0;
            rref[n] = prm->P0 / (R * prm->T0);
        }
    }
    });
    // per-element metric of affine elements: Jinv_e from the element's first
    // node, accepted when every node's Jinv agrees to 1e-10 of its largest
    // component.  An affine map has a constant Jacobian, and the per-node
    // values then differ by the round-off of their computation only (about
    // 1e-12 at 10 km / 156 m elements); trapezoidal, graded or curved
    // elements differ by 1e-2 and more.  cV is the cell volume times
    // w_i w_j w_k / 8 on every element (field.cpp:171-384, dg.cpp:175-325).
    const int NPX = g.npx, NPY = g.npy, NPZ = g.npz, NPn = g.np;
    std::vector<double> wnode(NPn);
    for (int i = 0; i < NPX; i++)
        for (int j = 0; j < NPY; j++)
            for (int k = 0; k < NPZ; k++)
                wnode[(i * NPY + j) * NPZ + k] = g.basis[0].w[i] * g.basis[1].w[j] * g.basis[2].w[k] / 8;
    std::vector<double> emet((size_t)g.nbcs * 10);
    std::vector<uint8_t> aff(g.nbcs, 1);
    parallel_for(g.nbcs, [&](uint64_t cb, uint64_t ce) {
    for (uint64_t c = cb; c < ce; c++) {
        const uint64_t n0 = c * NPn;
        double scale = 0;
        for (int q = 0; q < 9; q++) {
            emet[c * 10 + q] = g.Jinv[n0][q];
            scale = std::max(scale, std::fabs(g.Jinv[n0][q]));
        }
        emet[c * 10 + 9] = lm.cv[c];
        for (int t = 0; t < NPn && aff[c]; t++) {
            for (int q = 0; q < 9; q++)
                if (!(std::fabs(g.Jinv[n0 + t][q] - g.Jinv[n0][q]) <= 1e-10 * scale)) aff[c] = 0;
            if (g.cV[n0 + t] != lm.cv[c] * wnode[t]) aff[c] = 0;
        }
    }
    });
    bool all_affine = true;
    for (uint8_t a : aff) all_affine = all_affine && a;
    // background density interpolated from the coarse side minus the refined
    // side's, per mortar face node
    std::vector<double> mbg(mfine.size(), 0.0);
    for (size_t f = 0; f < mfaces.size(); f++) {
        const MortarFace& mf = mfaces[f];
        const int d0 = mf.side < 4 ? 0 : 1, d1 = mf.side < 2 ? 1 : 2;
        const double* R0 = g.basis[d0].ref[mf.h0].data();
        const double* R1 = g.basis[d1].ref[mf.h1].data();
        for (int t = 0; t < mf.n0 * mf.n1; t++) {
            const int ao = t / mf.n1, bo = t - ao * mf.n1;
            double sum = 0;
            for (int an = 0; an < mf.n0; an++)
                for (int bn = 0; bn < mf.n1; bn++) {
                    const int v = (mf.side & 1) ? (mf.side < 2 ? NPZ - 1 : mf.side < 4 ? NPY - 1 : NPX - 1) : 0;
                    int i, j, k;
                    if (mf.side < 2) { i = an; j = bn; k = v; }
                    else if (mf.side < 4) { i = an; j = v; k = bn; }
                    else { i = v; j = an; k = bn; }
                    const uint64_t nc = (uint64_t)mf.coarse * NPn + (uint64_t)(i * NPY * NPZ + j * NPZ + k);
                    sum += R0[an * mf.n0 + ao] * R1[bn * mf.n1 + bo] * rref[nc];
                }
            mbg[f * npfm + t] = sum - rref[mfine[f * npfm + t]];
        }
    }
    // ghost sides (boundary faces that are not 2:1 mortar sides)
    std::vector<uint32_t> ghost_rec, ghost_of((size_t)g.nbcs * 6, ~0u);
    for (size_t r = 0; r < ef.size(); r++) {
        const uint32_t m = ef[r].meta;
        if (((m >> 4) & 1) && ((m >> 5) & 1) && !((m >> 6) & 1)) {
            ghost_of[r] = (uint32_t)ghost_rec.size();
            ghost_rec.push_back((uint32_t)r);
        }
    }
    std::vector<double> gfn(3 * (size_t)g.nfacets);
    for (Int f = 0; f < g.nfacets; f++)
        for (int c = 0; c < 3; c++) gfn[3 * (size_t)f + c] = lm.fn[f][c];
    hipError_t e;
    if ((e = upload(d->JcV, JcV, s)) != hipSuccess || (e = upload(d->cV, cV, s)) != hipSuccess ||
        (e = upload(d->JcV32, JcV32, s)) != hipSuccess || (e = upload(d->cV32, cV32, s)) != hipSuccess ||
        (e = upload(d->p_ref, pref, s)) != hipSuccess || (e = upload(d->rho_ref, rref, s)) != hipSuccess ||
        (e = upload(d->eface, ef, s)) != hipSuccess || (e = upload(d->gFN, gfn, s)) != hipSuccess ||
        (e = upload(d->gh, ghv, s)) != hipSuccess || (e = upload(d->emet, emet, s)) != hipSuccess ||
        (e = upload(d->wnode, wnode, s)) != hipSuccess || (e = upload(d->ghost_rec, ghost_rec, s)) != hipSuccess ||
        (e = upload(d->ghost_of, ghost_of, s)) != hipSuccess)
        return nsem_fail(NSEM_EDEVICE, std::string("upload: ") + hipGetErrorString(e));
    for (int dir = 0; dir < 3; dir++) {
        if ((e = upload(d->D[dir], g.basis[dir].dpsi, s)) != hipSuccess ||
            (e = upload(d->w[dir], g.basis[dir].w, s)) != hipSuccess)
            return nsem_fail(NSEM_EDEVICE, std::string("upload: ") + hipGetErrorString(e));
    }
    if (!mfaces.empty()) {
        std::vector<double> zero(1, 0.0);
        if ((e = upload(d->mort, mfaces, s)) != hipSuccess || (e = upload(d->mort_fine, mfine, s)) != hipSuccess ||
            (e = upload(d->mort_fms, mfms, s)) != hipSuccess || (e = upload(d->mort_w, mw, s)) != hipSuccess ||
            (e = upload(d->mort_bg, mbg, s)) != hipSuccess ||
            (e = upload(d->cslot, cslot.empty() ? std::vector<uint32_t>(8, 0) : cslot, s)) != hipSuccess)
            return nsem_fail(NSEM_EDEVICE, std::string("upload: ") + hipGetErrorString(e));
        for (int dir = 0; dir < 3; dir++) {
            std::vector<double> rf = g.basis[dir].ref[0], cr = g.basis[dir].cor[0];
            rf.insert(rf.end(), g.basis[dir].ref[1].begin(), g.basis[dir].ref[1].end());
            cr.insert(cr.end(), g.basis[dir].cor[1].begin(), g.basis[dir].cor[1].end());
            if ((e = upload(d->psiRef[dir], rf, s)) != hipSuccess || (e = upload(d->psiCor[dir], cr, s)) != hipSuccess)
                return nsem_fail(NSEM_EDEVICE, std::string("upload: ") + hipGetErrorString(e));
        }
        // per face node: 5 values of the method-of-lines pre-pass, 16 of the
        // split step's (flux vectors and lambda)
        HIP_TRY(d->mphi.alloc(mfaces.size() * npfm * 16 * sizeof(double)));
        HIP_TRY(d->msurf.alloc((size_t)nslots * npfm * 5 * sizeof(double)));
        HIP_TRY(hipMemsetAsync(d->msurf.p, 0, (size_t)nslots * npfm * 5 * sizeof(double), s));
        HIP_TRY(d->mphi_t.alloc(mfaces.size() * npfm * 5 * sizeof(double)));
        HIP_TRY(d->msurf_t.alloc((size_t)nslots * npfm * 5 * sizeof(double)));
        HIP_TRY(hipMemsetAsync(d->msurf_t.p, 0, (size_t)nslots * npfm * 5 * sizeof(double), s));
        HIP_TRY(d->madj.alloc(5 * nb * sizeof(double)));
        HIP_TRY(hipMemsetAsync(d->madj.p, 0, 5 * nb * sizeof(double), s));
        if (prm->diffusion) HIP_TRY(d->mgrad.alloc((size_t)nslots * npfm * 12 * sizeof(double)));
    }
    HIP_TRY(d->bc.alloc(sizeof(BcTable)));
    HIP_TRY(hipMemcpyAsync(d->bc.p, &bt, sizeof(BcTable), hipMemcpyHostToDevice, s));
    for (int a = 0; a < 2; a++)
        for (int v = 0; v < 6; v++) {
            HIP_TRY(d->st[a][v].alloc(nb * sizeof(double)));
            HIP_TRY(hipMemsetAsync(d->st[a][v].p, 0, nb * sizeof(double), s));
        }
    for (int a = 0; a < 2; a++)
        for (int v = 0; v < 5; v++) {
            HIP_TRY(d->mq[a][v].alloc(nb * sizeof(double)));
            HIP_TRY(d->d64[a][v].alloc(nb * sizeof(double)));
            HIP_TRY(d->d32[a][v].alloc(nb * sizeof(float)));
        }
    if (prm->diffusion) HIP_TRY(d->grad.alloc(12 * nb * sizeof(double)));
    HIP_TRY(d->diag_slots.alloc(kDiagBlocks * 6 * sizeof(double)));
    HIP_TRY(d->diag_out.alloc(6 * sizeof(double)));
    HIP_TRY(d->flag.alloc(sizeof(int)));
    HIP_TRY(hipMemsetAsync(d->flag.p, 0, sizeof(int), s));
    HIP_TRY(d->gadj.alloc(ghost_rec.size() * 5 * npfm * sizeof(double)));
    HIP_TRY(hipStreamSynchronize(s));
    d->host_p_ref = pref;
    d->host_rho_ref = rref;

    DgDev& dv = d->dev;
    dv.npx = g.npx; dv.npy = g.npy; dv.npz = g.npz; dv.np = g.np;
    dv.nelem = g.nbcs; dv.nb = nb;
    dv.dt = prm->dt; dv.P0 = prm->P0; dv.T0 = prm->T0; dv.R = R; dv.gamma = gam; dv.cp = prm->cp;
    dv.ppc[0] = gam;
    for (int k = 1; k < 8; k++) dv.ppc[k] = dv.ppc[k - 1] * (gam - k) / (k + 1);
    for (int c = 0; c < 3; c++) dv.g[c] = prm->gravity[c];
    dv.buoyancy = prm->buoyancy;
    dv.JcV = d->JcV.as<double>(); dv.cV = d->cV.as<double>();
    dv.JcV32 = d->JcV32.as<float>(); dv.cV32 = d->cV32.as<float>();
    dv.p_ref = d->p_ref.as<double>(); dv.rho_ref = d->rho_ref.as<double>();
    dv.eface = d->eface.as<EFace>(); dv.gFN = d->gFN.as<double>();
    for (int dir = 0; dir < 3; dir++) { dv.D[dir] = d->D[dir].as<double>(); dv.w[dir] = d->w[dir].as<double>(); }
    dv.bc = d->bc.as<BcTable>();
    dv.diffusion = prm->diffusion ? 1 : 0;
    dv.visc = prm->viscosity;
    dv.iPr = 1 / prm->Pr;
    dv.grad = d->grad.as<double>();
    dv.affine = all_affine ? 1 : 0;
    dv.emet = d->emet.as<double>();
    dv.wnode = d->wnode.as<double>();
    dv.nmort = (int)mfaces.size();
    dv.npfm = npfm;
    dv.ncslot = (int)(cslot.size() / 8);
    dv.mort = d->mort.as<MortarFace>();
    dv.mort_fine = d->mort_fine.as<uint64_t>();
    dv.mort_fms = d->mort_fms.as<uint32_t>();
    dv.mort_w = d->mort_w.as<double>();
    dv.mort_bg = d->mort_bg.as<double>();
    for (int dir = 0; dir < 3; dir++) {
        dv.psiRef[dir] = d->psiRef[dir].as<double>();
        dv.psiCor[dir] = d->psiCor[dir].as<double>();
    }
    dv.cslot = d->cslot.as<uint32_t>();
    dv.mphi = d->mphi.as<double>();
    dv.msurf = d->msurf.as<double>();
    dv.mphi_t = d->mphi_t.as<double>();
    dv.msurf_t = d->msurf_t.as<double>();
    dv.madj = d->madj.as<double>();
    dv.mgrad = d->mgrad.as<double>();
    dv.flag = d->flag.as<int>();
    dv.nghost = (int)ghost_rec.size();
    dv.ghost_rec = d->ghost_rec.as<uint32_t>();
    dv.ghost_of = d->ghost_of.as<uint32_t>();
    dv.gadj = d->gadj.as<double>();
    *out = d.release();
    return NSEM_OK;
}

void nsem_dg_destroy(nsem_dg* d) { delete d; }

int nsem_dg_set_state(nsem_dg* d, const double* rho, const double* U, const double* T, const double* p) {
    if (!d || !rho || !U || !T || !p) return nsem_fail(NSEM_EINVAL, "null argument");
    const uint64_t nb = d->nb;
    std::vector<double> ux(nb), uy(nb), uz(nb);
    for (uint64_t n = 0; n < nb; n++) { ux[n] = U[3 * n]; uy[n] = U[3 * n + 1]; uz[n] = U[3 * n + 2]; }
    const double* src[6] = {rho, ux.data(), uy.data(), uz.data(), T, p};
    for (int v = 0; v < 6; v++)
        HIP_TRY(hipMemcpyAsync(d->st[d->cur][v].p, src[v], nb * sizeof(double), hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    // conserved perturbations for the method-of-lines path
    const double T0 = d->prm.T0;
    std::vector<double> q[5];
    for (int v = 0; v < 5; v++) q[v].resize(nb);
    for (uint64_t n = 0; n < nb; n++) {
        const double rr = d->host_rho_ref[n];
        q[0][n] = rho[n] - rr;
        q[1][n] = rho[n] * ux[n];
        q[2][n] = rho[n] * uy[n];
        q[3][n] = rho[n] * uz[n];
        q[4][n] = rho[n] * (T[n] + T0) - rr * T0;
    }
    for (int v = 0; v < 5; v++)
        HIP_TRY(hipMemcpyAsync(d->mq[d->mcur][v].p, q[v].data(), nb * sizeof(double), hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipMemsetAsync(d->flag.p, 0, sizeof(int), d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    d->split_valid = d->mol_valid = true;
    return NSEM_OK;
}

int nsem_dg_set_conserved(nsem_dg* d, const double* rho_p, const double* m, const double* theta_p) {
    if (!d || !rho_p || !m || !theta_p) return nsem_fail(NSEM_EINVAL, "null argument");
    const uint64_t nb = d->nb;
    std::vector<double> mx(nb), my(nb), mz(nb);
    for (uint64_t n = 0; n < nb; n++) { mx[n] = m[3 * n]; my[n] = m[3 * n + 1]; mz[n] = m[3 * n + 2]; }
    const double* src[5] = {rho_p, mx.data(), my.data(), mz.data(), theta_p};
    for (int v = 0; v < 5; v++)
        HIP_TRY(hipMemcpyAsync(d->mq[d->mcur][v].p, src[v], nb * sizeof(double), hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipMemsetAsync(d->flag.p, 0, sizeof(int), d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    d->mol_valid = true;
    d->split_valid = false;
    return NSEM_OK;
}

int nsem_dg_get_conserved(nsem_dg* d, double* rho_p, double* m, double* theta_p) {
    if (!d || !rho_p || !m || !theta_p) return nsem_fail(NSEM_EINVAL, "null argument");
    if (int rc = nsem_dg_require_mol(d)) return rc;
    const uint64_t nb = d->nb;
    std::vector<double> mx(nb), my(nb), mz(nb);
    double* dst[5] = {rho_p, mx.data(), my.data(), mz.data(), theta_p};
    for (int v = 0; v < 5; v++)
        HIP_TRY(hipMemcpyAsync(dst[v], d->mq[d->mcur][v].p, nb * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    for (uint64_t n = 0; n < nb; n++) { m[3 * n] = mx[n]; m[3 * n + 1] = my[n]; m[3 * n + 2] = mz[n]; }
    return NSEM_OK;
}

int nsem_dg_get_state(nsem_dg* d, double* rho, double* U, double* T, double* p) {
    if (!d || !rho || !U || !T || !p) return nsem_fail(NSEM_EINVAL, "null argument");
    const uint64_t nb = d->nb;
    if (!d->split_valid) {
        std::vector<double> rp(nb), m(3 * nb), tp(nb);
        int rc = nsem_dg_get_conserved(d, rp.data(), m.data(), tp.data());
        if (rc) return rc;
        const double T0 = d->prm.T0, gam = d->prm.cp / d->prm.cv;
        for (uint64_t n = 0; n < nb; n++) {
            const double rr = d->host_rho_ref[n];
            rho[n] = rr + rp[n];
            for (int c = 0; c < 3; c++) U[3 * n + c] = m[3 * n + c] / rho[n];
            T[n] = (tp[n] + rr * T0) / rho[n] - T0;
            p[n] = d->host_p_ref[n] * std::expm1(gam * std::log1p(tp[n] / (rr * T0)));
        }
        return NSEM_OK;
    }
    std::vector<double> ux(nb), uy(nb), uz(nb);
    double* dst[6] = {rho, ux.data(), uy.data(), uz.data(), T, p};
    for (int v = 0; v < 6; v++)
        HIP_TRY(hipMemcpyAsync(dst[v], d->st[d->cur][v].p, nb * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    for (uint64_t n = 0; n < nb; n++) { U[3 * n] = ux[n]; U[3 * n + 1] = uy[n]; U[3 * n + 2] = uz[n]; }
    return NSEM_OK;
}

int nsem_dg_diagnostics(nsem_dg* d, double* out) {
    if (!d || !out) return nsem_fail(NSEM_EINVAL, "null argument");
    const DiagState ds = nsem_dg_diag_state(d);
    HIP_TRY(launch_diagnostics(d->dev, ds, d->gh.as<double>(), d->diag_slots.as<double>(),
                               d->diag_out.as<double>(), d->stream));
    HIP_TRY(hipMemcpyAsync(out, d->diag_out.p, 6 * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    return NSEM_OK;
}

int nsem_dg_get_background(const nsem_dg* d, double* p_ref, double* rho_ref) {
    if (!d || !p_ref || !rho_ref) return nsem_fail(NSEM_EINVAL, "null argument");
    std::memcpy(p_ref, d->host_p_ref.data(), d->nb * sizeof(double));
    std::memcpy(rho_ref, d->host_rho_ref.data(), d->nb * sizeof(double));
    return NSEM_OK;
}

int nsem_euler_initial_density(const nsem_dg* d, const double* T, const double* p, double* rho) {
    if (!d || !T || !p || !rho) return nsem_fail(NSEM_EINVAL, "null argument");
    const nsem_euler_params& e = d->prm;
    const double R = e.cp - e.cv, gam = e.cp / e.cv;
    for (uint64_t n = 0; n < d->nb; n++) {
        const double pt = p[n] + d->host_p_ref[n];
        rho[n] = (e.P0 / (R * (T[n] + e.T0))) * std::pow(pt / e.P0, 1 / gam);
    }
    return NSEM_OK;
}

}  // extern "C"

// nsteps split steps queued on the handle's stream (no wait, flag unread).
static int split_steps_queue(nsem_dg* d, int32_t nsteps) {
    if (int rc = nsem_dg_require_split(d)) return rc;
    for (int32_t s = 0; s < nsteps; s++) {
        DevBuf* A = d->st[d->cur];
        DevBuf* B = d->st[1 - d->cur];
        SplitState in{A[0].as<double>(), A[1].as<double>(), A[2].as<double>(), A[3].as<double>(), A[4].as<double>()};
        SplitOut out{B[1].as<double>(), B[2].as<double>(), B[3].as<double>(), B[4].as<double>(), B[5].as<double>()};
        if (d->profile) HIP_TRY(hipEventRecord(d->ev0, d->stream));
        HIP_TRY(launch_split_step(d->dev, in, B[0].as<double>(), out, d->grad.as<double>(), d->stream));
        if (d->profile) {
            HIP_TRY(hipEventRecord(d->ev1, d->stream));
            HIP_TRY(hipEventSynchronize(d->ev1));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, d->ev0, d->ev1));
            d->prof_ms[0] += ms;
            d->prof_n[0]++;
        }
        d->cur = 1 - d->cur;
    }
    if (nsteps > 0) {
        d->split_valid = true;
        d->mol_valid = false;
    }
    return NSEM_OK;
}

extern "C" {

int nsem_dg_split_steps(nsem_dg* d, int32_t nsteps) {
    if (!d || nsteps < 0) return nsem_fail(NSEM_EINVAL, "bad argument");
    if (int rc = split_steps_queue(d, nsteps)) return rc;
    return nsem_dg_take_flag(d, "non-finite state in the split step");
}

}  // extern "C"

namespace {

template <typename R, typename S>
int rk3_steps_t(nsem_dg* d, int32_t nsteps, DevBuf (&dl)[2][5]) {
    static const double kB[3] = {1.0, 0.25, 2.0 / 3.0};
    for (int32_t st = 0; st < nsteps; st++) {
        DevBuf* qa = d->mq[d->mcur];
        DevBuf* qb = d->mq[1 - d->mcur];
        for (int stage = 0; stage < 3; stage++) {
            RkArgs<S> a{};
            for (int v = 0; v < 5; v++) {
                a.q[v] = qa[v].as<double>();
                a.din[v] = stage == 0 ? nullptr : dl[(stage - 1) & 1][v].as<S>();
                a.dout[v] = dl[stage & 1][v].as<S>();
                a.qout[v] = qb[v].as<double>();
            }
            a.b = kB[stage];
            a.bdt = kB[stage] * d->dev.dt;
            a.last = stage == 2;
            a.check = (stage == 2);
            a.flag = d->flag.as<int>();
            if (d->dev.diffusion) {
                // the stage state's gradients for the viscous fluxes
                R* grad = d->grad.as<R>();
                HIP_TRY((launch_mol_gradient<R, S>(d->dev, a.q, a.din, grad, d->stream)));
                a.grad = grad;
            }
            HIP_TRY((launch_rk_stage<R, S>(d->dev, a, d->stream)));
        }
        d->mcur = 1 - d->mcur;
    }
    if (nsteps > 0) {
        d->mol_valid = true;
        d->split_valid = false;
    }
    return NSEM_OK;
}

}  // namespace

int nsem_rk3_fp64_queue(nsem_dg* d, int32_t nsteps) {
    if (int rc = nsem_dg_require_mol(d)) return rc;
    return rk3_steps_t<double, double>(d, nsteps, d->d64);
}

int nsem_dg_require_mol(nsem_dg* d) {
    if (d->mol_valid) return NSEM_OK;
    DevBuf* A = d->st[d->cur];
    SplitState s{A[0].as<double>(), A[1].as<double>(), A[2].as<double>(), A[3].as<double>(), A[4].as<double>()};
    ConsOut q{};
    for (int v = 0; v < 5; v++) q.q[v] = d->mq[d->mcur][v].as<double>();
    HIP_TRY(launch_prim_to_cons(d->dev, s, q, d->stream));
    d->mol_valid = true;
    return NSEM_OK;
}

int nsem_dg_require_split(nsem_dg* d) {
    if (d->split_valid) return NSEM_OK;
    DevBuf* A = d->st[d->cur];
    ConsIn q{};
    for (int v = 0; v < 5; v++) q.q[v] = d->mq[d->mcur][v].as<double>();
    SplitOut o{A[1].as<double>(), A[2].as<double>(), A[3].as<double>(), A[4].as<double>(), A[5].as<double>()};
    HIP_TRY(launch_cons_to_prim(d->dev, q, A[0].as<double>(), o, d->stream));
    d->split_valid = true;
    return NSEM_OK;
}

int nsem_dg_take_flag(nsem_dg* d, const char* what) {
    int flag = 0;
    HIP_TRY(hipMemcpyAsync(&flag, d->flag.p, sizeof(int), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    if (!flag) return NSEM_OK;
    HIP_TRY(hipMemsetAsync(d->flag.p, 0, sizeof(int), d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    return nsem_fail(NSEM_ENONFINITE, what);
}

int nsem_mol_supported(const nsem_dg* d) {
    if (d->prm.diffusion)
        return nsem_fail(NSEM_EUNSUPPORTED, "the derivative entry points do not carry artificial viscosity");
    return NSEM_OK;
}



extern "C" {

int nsem_dg_rk3_steps(nsem_dg* d, int32_t nsteps, int32_t precision) {
    if (!d || nsteps < 0) return nsem_fail(NSEM_EINVAL, "bad argument");
    if (precision != 64 && precision != 32) return nsem_fail(NSEM_EINVAL, "precision must be 64 or 32");
    int rc = nsem_dg_require_mol(d);
    if (rc) return rc;
    if (precision == 64) rc = rk3_steps_t<double, double>(d, nsteps, d->d64);
    else rc = rk3_steps_t<float, float>(d, nsteps, d->d32);
    if (rc) return rc;
    return nsem_dg_take_flag(d, "non-finite residual in the RK stages");
}

int nsem_dg_time_rk3(nsem_dg* d, int32_t nsteps, int32_t precision, double* seconds) {
    if (!d || !seconds) return nsem_fail(NSEM_EINVAL, "null argument");
    if (precision != 64 && precision != 32) return nsem_fail(NSEM_EINVAL, "precision must be 64 or 32");
    int rc = nsem_dg_require_mol(d);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(d->ev0, d->stream));
    if (precision == 64) rc = rk3_steps_t<double, double>(d, nsteps, d->d64);
    else rc = rk3_steps_t<float, float>(d, nsteps, d->d32);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(d->ev1, d->stream));
    HIP_TRY(hipEventSynchronize(d->ev1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, d->ev0, d->ev1));
    *seconds = ms * 1e-3;
    return nsem_dg_take_flag(d, "non-finite residual in the RK stages");
}

int nsem_dg_time_split(nsem_dg* d, int32_t nsteps, double* seconds) {
    if (!d || !seconds || nsteps < 0) return nsem_fail(NSEM_EINVAL, "bad argument");
    if (int rc = nsem_dg_require_split(d)) return rc;
    HIP_TRY(hipEventRecord(d->ev0, d->stream));
    int rc = split_steps_queue(d, nsteps);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(d->ev1, d->stream));
    HIP_TRY(hipEventSynchronize(d->ev1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, d->ev0, d->ev1));
    *seconds = ms * 1e-3;
    return nsem_dg_take_flag(d, "non-finite state in the split step");
}

int nsem_dg_affine(const nsem_dg* d, int32_t* all_affine) {
    if (!d || !all_affine) return nsem_fail(NSEM_EINVAL, "null argument");
    *all_affine = d->dev.affine;
    return NSEM_OK;
}

int nsem_dg_synchronize(nsem_dg* d) {
    if (!d) return nsem_fail(NSEM_EINVAL, "null argument");
    HIP_TRY(hipStreamSynchronize(d->stream));
    return NSEM_OK;
}

}  // extern "C"
