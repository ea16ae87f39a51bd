// C-ABI: mesh generation, loading and geometry export (host side).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

#include "host/mesh.hpp"
#include "nebulasem_amd.h"
#include "mesh_access.hpp"
#include "status.hpp"

struct nsem_mesh {
    nsem::Grid grid;
    nsem::LoadedMesh lm;
    nsem::Geometry geo;
    std::vector<std::string> bnames;
};

using namespace nsem;

namespace {

int finish_mesh(std::unique_ptr<nsem_mesh>& m, const int32_t order[3], int32_t nonortho,
                nsem_mesh** out) {
    std::string err;
    if (!load_mesh(m->grid, m->lm, &err)) return nsem_fail(NSEM_EUNSUPPORTED, err);
    int ord[3] = {order[0], order[1], order[2]};
    for (int d = 0; d < 3; d++)
        if (ord[d] < 0 || ord[d] > 15) return nsem_fail(NSEM_EINVAL, "polynomial order out of range");
    if (!build_geometry(m->lm, ord, nonortho, m->geo, &err)) return nsem_fail(NSEM_EUNSUPPORTED, err);
    for (auto& kv : m->lm.boundaries) m->bnames.push_back(kv.first);
    *out = m.release();
    return NSEM_OK;
}

template <typename T>
int copy_out(const std::vector<T>& v, int32_t kind_expected, int32_t kind, void* dst,
             uint64_t count, uint64_t per) {
    if (kind != kind_expected) return nsem_fail(NSEM_EINVAL, "array kind mismatch");
    if (count != v.size() * per) return nsem_fail(NSEM_EINVAL, "array size mismatch");
    if (count) std::memcpy(dst, v.data(), count * (kind == 0 ? 4 : 8));
    return NSEM_OK;
}

// Flattened view of one exported array.
struct ArrayView {
    const void* data = nullptr;
    uint64_t count = 0;
    int32_t kind = 0;  // 0 = u32, 1 = f64
    std::vector<uint32_t> u;
    std::vector<double> d;
};

bool view_array(const nsem_mesh* m, const std::string& name, ArrayView& a) {
    const Geometry& g = m->geo;
    const LoadedMesh& lm = m->lm;
    auto v3 = [&](const std::vector<V3>& src) {
        a.d.resize(src.size() * 3);
        for (size_t i = 0; i < src.size(); i++)
            for (int c = 0; c < 3; c++) a.d[3 * i + c] = src[i][c];
        a.kind = 1;
    };
    auto csr = [&](const std::vector<IntVec>& lists, bool offsets) {
        a.kind = 0;
        if (offsets) {
            a.u.push_back(0);
            for (auto& l : lists) a.u.push_back(a.u.back() + l.size());
        } else {
            for (auto& l : lists) a.u.insert(a.u.end(), l.begin(), l.end());
        }
    };
    if (name == "vertices") v3(lm.vertices);
    else if (name == "facet_off") csr(lm.facets, true);
    else if (name == "facet_vtx") csr(lm.facets, false);
    else if (name == "cell_off") csr(lm.cells, true);
    else if (name == "cell_faces") csr(lm.cells, false);
    else if (name == "faceid_off") csr(lm.faceid, true);
    else if (name == "faceid") csr(lm.faceid, false);
    else if (name == "FOC") { a.u = lm.FOC; a.kind = 0; }
    else if (name == "FNC") { a.u = lm.FNC; a.kind = 0; }
    else if (name == "FMC") { a.u = lm.FMC; a.kind = 0; }
    else if (name == "has_boundary") { a.u.assign(lm.has_boundary.begin(), lm.has_boundary.end()); a.kind = 0; }
    else if (name == "boff") {
        a.kind = 0; a.u.push_back(0);
        for (auto& kv : lm.boundaries) a.u.push_back(a.u.back() + kv.second.size());
    } else if (name == "bfaces") {
        a.kind = 0;
        for (auto& kv : lm.boundaries) a.u.insert(a.u.end(), kv.second.begin(), kv.second.end());
    }
    else if (name == "gFC") v3(lm.fc);
    else if (name == "gFN") v3(lm.fn);
    else if (name == "gCC") v3(lm.cc);
    else if (name == "gCV") { a.d = lm.cv; a.kind = 1; }
    else if (name == "FO") { a.u = g.FO; a.kind = 0; }
    else if (name == "FN") { a.u = g.FN; a.kind = 0; }
    else if (name == "fN") v3(g.fN);
    else if (name == "fC") v3(g.fC);
    else if (name == "fI") { a.d = g.fI; a.kind = 1; }
    else if (name == "fD") { a.d = g.fD; a.kind = 1; }
    else if (name == "cC") v3(g.cC);
    else if (name == "cV") { a.d = g.cV; a.kind = 1; }
    else if (name == "Jinv") {
        a.kind = 1; a.d.resize(g.Jinv.size() * 9);
        for (size_t i = 0; i < g.Jinv.size(); i++)
            for (int c = 0; c < 9; c++) a.d[9 * i + c] = g.Jinv[i][c];
    }
    else if (name == "allFaces") { a.u = g.allFaces; a.kind = 0; }
    else if (name == "faceIndices0") { a.u = g.faceIndices0; a.kind = 0; }
    else if (name == "faceIndices1") { a.u = g.faceIndices1; a.kind = 0; }
    else if (name.size() == 4 && (name.compare(0, 3, "xgl") == 0 || name.compare(0, 3, "wgl") == 0) &&
             name[3] >= '0' && name[3] <= '2') {
        const Basis1D& b = g.basis[name[3] - '0'];
        a.d = name[0] == 'x' ? b.x : b.w; a.kind = 1;
    } else if ((name.size() == 4 && name.compare(0, 3, "psi") == 0) ||
               (name.size() == 5 && name.compare(0, 4, "dpsi") == 0)) {
        int d = name.back() - '0';
        if (d < 0 || d > 2) return false;
        const Basis1D& b = g.basis[d];
        a.d = name[0] == 'd' ? b.dpsi : b.psi; a.kind = 1;
    } else if (name.size() == 7 && (name.compare(0, 6, "psiRef") == 0 || name.compare(0, 6, "psiCor") == 0)) {
        // both halves, lower then upper, n*n each
        int d = name.back() - '0';
        if (d < 0 || d > 2) return false;
        const Basis1D& b = g.basis[d];
        const std::vector<double>* h = name[3] == 'R' ? b.ref : b.cor;
        a.d = h[0];
        a.d.insert(a.d.end(), h[1].begin(), h[1].end());
        a.kind = 1;
    } else return false;
    a.count = a.kind == 0 ? a.u.size() : a.d.size();
    a.data = a.kind == 0 ? (const void*)a.u.data() : (const void*)a.d.data();
    return true;
}

}  // namespace

extern "C" {

int nsem_grid_generate(const char* spec_path, const char* grid_path) {
    if (!spec_path || !grid_path) return nsem_fail(NSEM_EINVAL, "null path");
    HexBlock blk;
    std::vector<PatchSpec> patches;
    std::string err;
    if (!read_block_spec(spec_path, blk, patches, &err)) return nsem_fail(NSEM_EUNSUPPORTED, err);
    Grid g;
    if (!generate_block_grid(blk, patches, g, &err)) return nsem_fail(NSEM_EUNSUPPORTED, err);
    if (!write_grid(grid_path, g, &err)) return nsem_fail(NSEM_EIO, err);
    return NSEM_OK;
}

int nsem_mesh_load(const char* grid_path, const int32_t order[3], int32_t nonortho, nsem_mesh** out) {
    if (!grid_path || !order || !out) return nsem_fail(NSEM_EINVAL, "null argument");
    auto m = std::make_unique<nsem_mesh>();
    std::string err;
    if (!read_grid(grid_path, m->grid, &err)) return nsem_fail(NSEM_EIO, err);
    return finish_mesh(m, order, nonortho, out);
}

int nsem_mesh_from_spec(const char* spec_path, const int32_t order[3], int32_t nonortho,
                        nsem_mesh** out) {
    if (!spec_path || !order || !out) return nsem_fail(NSEM_EINVAL, "null argument");
    auto m = std::make_unique<nsem_mesh>();
    HexBlock blk;
    std::vector<PatchSpec> patches;
    std::string err;
    if (!read_block_spec(spec_path, blk, patches, &err)) return nsem_fail(NSEM_EUNSUPPORTED, err);
    if (!generate_block_grid(blk, patches, m->grid, &err)) return nsem_fail(NSEM_EUNSUPPORTED, err);
    return finish_mesh(m, order, nonortho, out);
}

int nsem_mesh_box(const double corners[24], const uint32_t ndiv[3], const int32_t spacing[12],
                  const double ratio[12], int32_t npatch, const char* const* names,
                  const int32_t* npoints, const double* points, const int32_t order[3],
                  int32_t nonortho, nsem_mesh** out) {
    if (!corners || !ndiv || !order || !out) return nsem_fail(NSEM_EINVAL, "null argument");
    if (npatch < 0 || (npatch && (!names || !npoints || !points)))
        return nsem_fail(NSEM_EINVAL, "bad patch arguments");
    HexBlock blk;
    for (int i = 0; i < 8; i++) blk.corners[i] = V3(corners[3 * i], corners[3 * i + 1], corners[3 * i + 2]);
    for (int d = 0; d < 3; d++) {
        if (ndiv[d] < 1) return nsem_fail(NSEM_EINVAL, "division counts must be >= 1");
        blk.n[d] = ndiv[d];
    }
    for (int i = 0; i < 12; i++) {
        blk.type[i] = spacing ? spacing[i] : LINEAR;
        blk.s[i] = ratio ? ratio[i] : 1.0;
    }
    std::vector<PatchSpec> patches;
    const double* p = points;
    for (int32_t i = 0; i < npatch; i++) {
        PatchSpec ps;
        ps.name = names[i];
        for (int32_t k = 0; k < npoints[i]; k++, p += 3) ps.points.push_back(V3(p[0], p[1], p[2]));
        patches.push_back(ps);
    }
    auto m = std::make_unique<nsem_mesh>();
    std::string err;
    auto t0 = std::chrono::steady_clock::now();
    if (!generate_block_grid(blk, patches, m->grid, &err)) return nsem_fail(NSEM_EUNSUPPORTED, err);
    if (std::getenv("NSEM_TIMING"))
        std::fprintf(stderr, "[nsem] %-28s %8.3f s\n", "generate_block_grid",
                     std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    return finish_mesh(m, order, nonortho, out);
}

int nsem_mesh_write_grid(const nsem_mesh* m, const char* grid_path) {
    if (!m || !grid_path) return nsem_fail(NSEM_EINVAL, "null argument");
    std::string err;
    if (!write_grid(grid_path, m->grid, &err)) return nsem_fail(NSEM_EIO, err);
    return NSEM_OK;
}

int nsem_mesh_get_info(const nsem_mesh* m, nsem_mesh_info* info) {
    if (!m || !info) return nsem_fail(NSEM_EINVAL, "null argument");
    const Geometry& g = m->geo;
    info->npx = g.npx; info->npy = g.npy; info->npz = g.npz; info->np = g.np; info->npf = g.npf;
    info->nbcs = g.nbcs; info->ncells = g.ncells; info->nfacets = g.nfacets;
    info->nvertices = m->lm.vertices.size();
    info->nbcs_field = g.nbcs_field; info->nall_field = g.nall_field;
    info->nboundaries = m->bnames.size();
    return NSEM_OK;
}

int nsem_mesh_array_size(const nsem_mesh* m, const char* name, uint64_t* count, int32_t* kind) {
    if (!m || !name || !count || !kind) return nsem_fail(NSEM_EINVAL, "null argument");
    ArrayView a;
    if (!view_array(m, name, a)) return nsem_fail(NSEM_EINVAL, std::string("unknown array ") + name);
    *count = a.count;
    *kind = a.kind;
    return NSEM_OK;
}

int nsem_mesh_get_array(const nsem_mesh* m, const char* name, void* dst, uint64_t count) {
    if (!m || !name || (!dst && count)) return nsem_fail(NSEM_EINVAL, "null argument");
    ArrayView a;
    if (!view_array(m, name, a)) return nsem_fail(NSEM_EINVAL, std::string("unknown array ") + name);
    if (count != a.count) return nsem_fail(NSEM_EINVAL, "array size mismatch");
    if (count) std::memcpy(dst, a.data, count * (a.kind == 0 ? 4 : 8));
    return NSEM_OK;
}

int nsem_mesh_boundary_name(const nsem_mesh* m, uint32_t i, char* buf, uint64_t len) {
    if (!m || !buf || !len) return nsem_fail(NSEM_EINVAL, "null argument");
    if (i >= m->bnames.size()) return nsem_fail(NSEM_EINVAL, "boundary index out of range");
    std::strncpy(buf, m->bnames[i].c_str(), len - 1);
    buf[len - 1] = 0;
    return NSEM_OK;
}

void nsem_mesh_destroy(nsem_mesh* m) { delete m; }

}  // extern "C"

const nsem::Geometry& nsem_mesh_geometry(const nsem_mesh* m) { return m->geo; }
const nsem::LoadedMesh& nsem_mesh_loaded(const nsem_mesh* m) { return m->lm; }
const std::vector<std::string>& nsem_mesh_boundary_names(const nsem_mesh* m) { return m->bnames; }
