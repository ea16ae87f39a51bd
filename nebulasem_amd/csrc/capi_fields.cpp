// C-ABI: reading the reference's field files (initial values + BCs).
#include <cmath>
#include <cstring>
#include <string>

#include "host/fields.hpp"
#include "host/mesh.hpp"
#include "mesh_access.hpp"
#include "nebulasem_amd.h"
#include "status.hpp"

namespace {

double dot_r(const double* p, const double* q) {  // right-nested, tensor.h:117
    double r = p[2] * q[2] + 0.0;
    r = p[1] * q[1] + r;
    return p[0] * q[0] + r;
}

// The tangential part of p on a face of normal n, rescaled to |p|
// (tensor.h:486-494).
void reflect(const double* p, const double* n, double* r) {
    const double mn = std::sqrt(dot_r(n, n));
    const double en[3] = {n[0] / mn, n[1] / mn, n[2] / mn};
    const double A[6] = {1 - en[0] * en[0], 1 - en[1] * en[1], 1 - en[2] * en[2],
                         0 - en[0] * en[1], 0 - en[1] * en[2], 0 - en[0] * en[2]};
    r[0] = A[0] * p[0] + A[3] * p[1] + A[5] * p[2];
    r[1] = A[3] * p[0] + A[1] * p[1] + A[4] * p[2];
    r[2] = A[5] * p[0] + A[4] * p[1] + A[2] * p[2];
    const double mr = std::sqrt(dot_r(r, r));
    if (std::fabs(mr) <= 1e-7) return;
    const double s = std::sqrt(dot_r(p, p)) / mr;
    for (int c = 0; c < 3; c++) r[c] *= s;
}

// Ghost nodes from the boundary conditions just read, as a field read does
// (MeshField::read_, field.h:1557-1564, then applyExplicitBCs 2585-2727):
// NEUMANN: owner + value |x_ghost - x_owner|; ROBIN: shape value + (1 -
// shape)(owner + tvalue |x_ghost - x_owner|) (field.h:2677-2680); SYMMETRY:
// a vector reflected about the face, a scalar copied; DIRICHLET: the value.
void apply_explicit_bcs(const nsem::Geometry& g, const nsem::LoadedMesh& lm, int ncomp,
                        const std::vector<nsem::BoundaryCondition>& bcs, double* v) {
    for (const auto& bc : bcs) {
        auto it = lm.boundaries.find(bc.patch);
        if (it == lm.boundaries.end()) continue;
        for (nsem::Int face : it->second)
            for (nsem::Int n = 0; n < g.npf; n++) {
                const uint64_t k = uint64_t(face) * g.npf + n;
                const uint64_t c1 = g.FO[k], c2 = g.FN[k];
                if (c2 >= g.nall_field) continue;
                double* dst = v + c2 * ncomp;
                const double* src = v + c1 * ncomp;
                if (bc.type == "NEUMANN") {
                    double dv[3];
                    for (int c = 0; c < 3; c++) dv[c] = g.cC[c2][c] - g.cC[c1][c];
                    const double md = std::sqrt(dot_r(dv, dv));
                    for (int c = 0; c < ncomp; c++) dst[c] = src[c] + bc.value[c] * md;
                } else if (bc.type == "ROBIN") {
                    double dv[3];
                    for (int c = 0; c < 3; c++) dv[c] = g.cC[c2][c] - g.cC[c1][c];
                    const double md = std::sqrt(dot_r(dv, dv));
                    for (int c = 0; c < ncomp; c++)
                        dst[c] = bc.shape * bc.value[c] + (1 - bc.shape) * (src[c] + bc.tvalue[c] * md);
                } else if (bc.type == "SYMMETRY") {
                    if (ncomp == 3) {
                        const double nv[3] = {g.fN[k][0], g.fN[k][1], g.fN[k][2]};
                        reflect(src, nv, dst);
                    } else {
                        for (int c = 0; c < ncomp; c++) dst[c] = src[c];
                    }
                } else if (bc.type == "DIRICHLET") {
                    for (int c = 0; c < ncomp; c++) dst[c] = bc.value[c];
                }
            }
    }
}

int field_read(const nsem_mesh* m, const char* path, int32_t ncomp, const double gravity[3], double* values,
               int32_t* bc_kind, double* bc_value, double* bc_shape, double* bc_tvalue) {
    if (!m || !path || !values || !bc_kind || !bc_value) return nsem_fail(NSEM_EINVAL, "null argument");
    if (ncomp != 1 && ncomp != 3) return nsem_fail(NSEM_EINVAL, "ncomp must be 1 or 3");
    const bool robin = bc_shape && bc_tvalue;
    const nsem::Geometry& g = nsem_mesh_geometry(m);
    std::vector<double> cc(g.nall_field * 3);
    for (uint64_t i = 0; i < g.nall_field; i++)
        for (int c = 0; c < 3; c++) cc[3 * i + c] = g.cC[i][c];
    nsem::FieldContext ctx;
    ctx.nall = g.nall_field;
    ctx.nbcs = g.nbcs_field;
    ctx.cC = cc.data();
    if (gravity)
        for (int c = 0; c < 3; c++) ctx.gravity[c] = gravity[c];
    nsem::FieldData fd;
    std::string err;
    if (!nsem::read_field_file(path, ctx, ncomp, fd, &err)) return nsem_fail(NSEM_EIO, err);
    const std::vector<std::string>& names = nsem_mesh_boundary_names(m);
    for (size_t b = 0; b < names.size(); b++) {
        bc_kind[b] = -1;
        for (int c = 0; c < ncomp; c++) bc_value[b * ncomp + c] = 0;
        if (robin) {
            bc_shape[b] = 0;
            for (int c = 0; c < ncomp; c++) bc_tvalue[b * ncomp + c] = 0;
        }
        for (const auto& bc : fd.bcs) {
            if (bc.patch != names[b]) continue;
            if (bc.type == "NEUMANN") bc_kind[b] = NSEM_BC_NEUMANN;
            else if (bc.type == "SYMMETRY") bc_kind[b] = NSEM_BC_SYMMETRY;
            else if (bc.type == "DIRICHLET") bc_kind[b] = NSEM_BC_DIRICHLET;
            else if (bc.type == "ROBIN" && robin) bc_kind[b] = NSEM_BC_ROBIN;
            else return nsem_fail(NSEM_EUNSUPPORTED, "boundary type " + bc.type + " is not supported");
            for (int c = 0; c < ncomp; c++) bc_value[b * ncomp + c] = bc.value[c];
            if (robin) {
                bc_shape[b] = bc.shape;
                for (int c = 0; c < ncomp; c++) bc_tvalue[b * ncomp + c] = bc.tvalue[c];
            }
        }
    }
    apply_explicit_bcs(g, nsem_mesh_loaded(m), ncomp, fd.bcs, fd.values.data());
    std::memcpy(values, fd.values.data(), sizeof(double) * fd.values.size());
    return NSEM_OK;
}

int field_write(const nsem_mesh* m, const char* path, int32_t ncomp, const double* values, const int32_t* bc_kind,
                const double* bc_value, const double* bc_shape, const double* bc_tvalue) {
    if (!m || !path || !values || !bc_kind || !bc_value) return nsem_fail(NSEM_EINVAL, "null argument");
    if (ncomp != 1 && ncomp != 3) return nsem_fail(NSEM_EINVAL, "ncomp must be 1 or 3");
    const bool robin = bc_shape && bc_tvalue;
    const nsem::Geometry& g = nsem_mesh_geometry(m);
    const std::vector<std::string>& names = nsem_mesh_boundary_names(m);
    std::vector<nsem::BoundaryCondition> bcs;
    for (size_t b = 0; b < names.size(); b++) {
        if (bc_kind[b] < 0) continue;
        nsem::BoundaryCondition bc;
        bc.patch = names[b];
        if (bc_kind[b] == NSEM_BC_NEUMANN) bc.type = "NEUMANN";
        else if (bc_kind[b] == NSEM_BC_SYMMETRY) bc.type = "SYMMETRY";
        else if (bc_kind[b] == NSEM_BC_DIRICHLET) bc.type = "DIRICHLET";
        else if (bc_kind[b] == NSEM_BC_ROBIN && robin) bc.type = "ROBIN";
        else return nsem_fail(NSEM_EINVAL, "unknown boundary kind for " + names[b]);
        bc.value.assign(bc_value + b * ncomp, bc_value + (b + 1) * ncomp);
        if (robin) {
            bc.shape = bc_shape[b];
            bc.tvalue.assign(bc_tvalue + b * ncomp, bc_tvalue + (b + 1) * ncomp);
        }
        bcs.push_back(bc);
    }
    std::string err;
    if (!nsem::write_field_file(path, ncomp, g.nbcs_field, values, bcs, &err)) return nsem_fail(NSEM_EIO, err);
    return NSEM_OK;
}

}  // namespace

extern "C" int nsem_field_read(const nsem_mesh* m, const char* path, int32_t ncomp, const double gravity[3],
                               double* values, int32_t* bc_kind, double* bc_value) {
    return field_read(m, path, ncomp, gravity, values, bc_kind, bc_value, nullptr, nullptr);
}
extern "C" int nsem_field_read_bc(const nsem_mesh* m, const char* path, int32_t ncomp, const double gravity[3],
                                  double* values, int32_t* bc_kind, double* bc_value, double* bc_shape,
                                  double* bc_tvalue) {
    if (!bc_shape || !bc_tvalue) return nsem_fail(NSEM_EINVAL, "null argument");
    return field_read(m, path, ncomp, gravity, values, bc_kind, bc_value, bc_shape, bc_tvalue);
}

extern "C" int nsem_field_write(const nsem_mesh* m, const char* path, int32_t ncomp, const double* values,
                                const int32_t* bc_kind, const double* bc_value) {
    return field_write(m, path, ncomp, values, bc_kind, bc_value, nullptr, nullptr);
}
extern "C" int nsem_field_write_bc(const nsem_mesh* m, const char* path, int32_t ncomp, const double* values,
                                   const int32_t* bc_kind, const double* bc_value, const double* bc_shape,
                                   const double* bc_tvalue) {
    if (!bc_shape || !bc_tvalue) return nsem_fail(NSEM_EINVAL, "null argument");
    return field_write(m, path, ncomp, values, bc_kind, bc_value, bc_shape, bc_tvalue);
}

extern "C" int nsem_amr_tree_read(const char* path, uint32_t* nodes, uint64_t capacity, uint64_t* n_out) {
    if (!path || !n_out) return nsem_fail(NSEM_EINVAL, "null argument");
    std::vector<uint32_t> v;
    std::string err;
    if (!nsem::read_amr_tree(path, v, &err)) return nsem_fail(NSEM_EIO, err);
    *n_out = v.size() / 4;
    if (!nodes) return NSEM_OK;
    if (capacity < v.size() / 4) return nsem_fail(NSEM_EINVAL, "node buffer too small");
    std::memcpy(nodes, v.data(), v.size() * sizeof(uint32_t));
    return NSEM_OK;
}

extern "C" int nsem_amr_tree_write(const char* path, const uint32_t* nodes, uint64_t n) {
    if (!path || (!nodes && n)) return nsem_fail(NSEM_EINVAL, "null argument");
    std::string err;
    if (!nsem::write_amr_tree(path, nodes, n, &err)) return nsem_fail(NSEM_EIO, err);
    return NSEM_OK;
}
