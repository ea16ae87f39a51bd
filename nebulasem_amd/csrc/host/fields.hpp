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

}  // namespace nsem
