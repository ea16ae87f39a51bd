// Field-file reader: initial values and boundary conditions of one field.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace nsem {

struct BoundaryCondition {
    std::string patch;
    std::string type = "NEUMANN";  // upper-case BC name
    std::vector<double> value;     // ncomp entries
    double shape = 0;
    std::vector<double> tvalue;    // ncomp entries (ROBIN)
};

struct FieldData {
    std::vector<double> values;  // nall * ncomp, node-major (AoS per node)
    std::vector<BoundaryCondition> bcs;
};

struct FieldContext {
    uint64_t nall = 0;        // all nodes (interior + ghost)
    uint64_t nbcs = 0;        // interior nodes
    const double* cC = nullptr;  // nall*3 node coordinates
    double gravity[3] = {0, 0, 0};
};

bool read_field_file(const std::string& path, const FieldContext& ctx, int ncomp, FieldData& out,
                     std::string* err);

// Writer of the same formats (MeshField::write_, field.h:1588-1663, with
// BCondition's operator<<, field.h:187-216): ".bin" paths get the binary
// token stream, others the text form with 12 significant digits.  values:
// n interior entries of ncomp; one boundary block per condition in bcs.
bool write_field_file(const std::string& path, int ncomp, uint64_t n, const double* values,
                      const std::vector<BoundaryCondition>& bcs, std::string* err);

// The refinement tree amrTree_<step>.bin / .txt (mesh.h:49-72 written as a
// std::vector by util.h:129-148): a count, then per node id, nchildren,
// cid (first child's node index), level.  nodes: 4 u32 per node.
bool read_amr_tree(const std::string& path, std::vector<uint32_t>& nodes, std::string* err);
bool write_amr_tree(const std::string& path, const uint32_t* nodes, uint64_t n, std::string* err);

}  // namespace nsem
