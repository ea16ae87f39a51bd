// Device tables of the loaded mesh that the field API's operator kernels
// read (internal to the field library).
#ifndef NSEM_FIELD_MESH_TABLES_H
#define NSEM_FIELD_MESH_TABLES_H

#include <string>
#include <vector>

#include "field.h"
#include "nebulasem_amd.h"

namespace nsem_field {
struct MeshTables {
    uint32_t npx = 1, npy = 1, npz = 1, np = 1, npf = 1;
    uint32_t nbcs = 0, ncells = 0, nfacets = 0, nvertices = 0, npatch = 0, nghostnodes = 0;
    uint64_t nbf = 0, nall = 0;
    // per face node
    const uint32_t* FO = nullptr;
    const uint32_t* FN = nullptr;
    const double* fN = nullptr;   // 3 per face node (area normal x face-node weight)
    const double* fI = nullptr;
    const double* fD = nullptr;
    // per node (all nodes)
    const double* cC = nullptr;   // 3 per node
    const double* cV = nullptr;
    const double* JcV = nullptr;  // 9 per node: Jinv cV (DG)
    const double* D[3] = {nullptr, nullptr, nullptr};  // dpsi per direction (DG)
    // node t (interior and ghost): entries ent[ent_off[t] .. ent_off[t+1]) =
    // (k << 1) | side, side 0 when t is the face node's owner node FO[k], 1
    // when it is FN[k], in the order the reference's cell loops (div_flux,
    // mul, the matrix diagonals) visit the faces of t's own cell
    const uint32_t* ent_off = nullptr;
    const uint64_t* ent = nullptr;
    // host copies for the level schedules of the sequential solvers
    std::vector<uint32_t> ent_off_h, FO_h, FN_h;
    std::vector<uint64_t> ent_h;
    std::vector<uint8_t> has_boundary_h;  // per interior cell: a face on a boundary patch
    uint64_t generation = 0;              // bumped by every mesh load
    // ghost nodes and their owner nodes
    const uint32_t* ghost_owner = nullptr;
    const uint32_t* ghost_node = nullptr;
    // boundary face nodes per patch (patch_off_host[b] .. [b+1]): face node k
    // and its slot within the patch (face j * npf + n)
    std::vector<uint32_t> patch_off_host;
    const uint32_t* patch_k = nullptr;
    const uint32_t* patch_slot = nullptr;
    std::vector<std::string> patch_names;
};
const MeshTables& mesh_tables();
nsem_mesh* mesh_handle();
}  // namespace nsem_field

#endif
