// Host-side mesh and DG geometry for the MI355X dGSEM / FV path.
//
// This is the layer the reference calls L1 (src/mesh) plus the geometry half
// of L2 (src/field/dg.cpp, src/field/field.cpp:95-384), restated from scratch.
// Everything here runs once per mesh load on the host; its integer outputs
// (cell/face numbering, face-node maps FO/FN, face groups) must equal the
// reference's exactly because the device kernels and every parity check index
// by them.  Floating-point geometry follows the reference's operation order so
// it matches to the last bit on the same libm.
#pragma once
#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace nsem {

using Int = uint32_t;
using IntVec = std::vector<Int>;

struct V3 {
    double v[3];
    V3() : v{0, 0, 0} {}
    V3(double a, double b, double c) : v{a, b, c} {}
    double& operator[](int i) { return v[i]; }
    const double& operator[](int i) const { return v[i]; }
};
// Component order of a rank-2 tensor, as the reference stores it
// (src/tensor/tensor.h:452-454): XX YY ZZ XY YZ XZ YX ZY ZX.
enum { XX = 0, YY, ZZ, XY, YZ, XZ, YX, ZY, ZX };
struct T9 {
    double v[9];
    T9() : v{0, 0, 0, 0, 0, 0, 0, 0, 0} {}
    double& operator[](int i) { return v[i]; }
    const double& operator[](int i) const { return v[i]; }
};

constexpr Int kMaxInt = Int(1u << 31);  // the reference's Constants::MAX_INT

// ---------------------------------------------------------------- grid ----
// A grid as written by the reference's `mesh` tool (mesh.h:158-220): vertices,
// facets (vertex lists), cells (facet lists) and named boundary patches.
struct Grid {
    std::vector<V3> vertices;
    std::vector<IntVec> facets;
    std::vector<IntVec> cells;
    std::map<std::string, IntVec> boundaries;  // std::map: iterated by name
};

// Block spacing types of the hex generator (hexMesh.h:6-12).
enum Spacing { LINEAR = 0, GEOMETRIC = 1, WALL = 2 };

// One hexahedral block: 8 corner points in the reference's block order and
// per-edge spacing (12 edges) with division counts n[3].
struct HexBlock {
    V3 corners[8];
    Int n[3];
    double s[12];
    int type[12];
};
// A boundary patch request: name and three or more corner points on its plane.
struct PatchSpec {
    std::string name;
    std::vector<V3> points;
};

// Generate the grid of one straight-edged hex block plus named patches,
// equal to what the reference's mesh tool writes (hexMesh.cpp:48-361,
// meshApp.cpp:218-262).  Returns false on unsupported input.
bool generate_block_grid(const HexBlock& blk, const std::vector<PatchSpec>& patches,
                         Grid& g, std::string* err);

// Read the reference mesh tool's block-description file (one straight-edged
// block plus named patches).
bool read_block_spec(const std::string& path, HexBlock& blk, std::vector<PatchSpec>& patches,
                     std::string* err);

// Grid file I/O in the reference's text and binary token formats.
bool read_grid(const std::string& path, Grid& g, std::string* err);
bool write_grid(const std::string& path, const Grid& g, std::string* err);

// ----------------------------------------------------------- loaded mesh --
// The state the reference's Mesh::LoadMesh leaves behind (field.cpp:95-167).
struct LoadedMesh {
    std::vector<V3> vertices;
    std::vector<IntVec> facets;
    std::vector<IntVec> cells;      // interior cells first, then ghost cells
    std::vector<IntVec> faceid;     // local face group id per cell face
    std::map<std::string, IntVec> boundaries;
    IntVec FOC, FNC, FMC;           // facet owner / neighbour cell, mortar flag
    std::vector<uint8_t> has_boundary;
    Int nbcs = 0;                   // number of interior cells (gBCS)
    std::vector<V3> fc, fn, cc;     // facet centre, facet area normal, cell centre
    std::vector<double> cv;         // cell volume
};
bool load_mesh(const Grid& g, LoadedMesh& m, std::string* err);

// ------------------------------------------------------------ DG geometry --
// Basis tables for one direction (dg.cpp:32-163): LGL nodes/weights, Lagrange
// values psi[j*n+s] = l_j(x_s) and derivatives dpsi[s*n+i] = l'_i(x_s).
struct Basis1D {
    Int n = 1;
    std::vector<double> x, w, psi, dpsi;
    // 2:1 mortar matrices, n*n each, for the lower (0) and upper (1) half
    // of the interval: ref interpolates a parent's nodes to a half's nodes,
    // cor projects a half's values back (dg.cpp:516-583)
    std::vector<double> ref[2], cor[2];
};
void lgl_nodes(int n, double* x, double* w);
void build_basis(int n, Basis1D& b);

// Node-level geometry of the DG (or, with N=0, FV) discretisation: what
// initGeomMeshFields + DG::init_geom produce (field.cpp:171-384,
// dg.cpp:167-477).  Node index = cell*NP + i*NPY*NPZ + j*NPZ + k.
struct Geometry {
    Int npx = 1, npy = 1, npz = 1, np = 1, npf = 1;
    Int ncells = 0, nbcs = 0, nfacets = 0;
    uint64_t nbcs_field = 0, nall_field = 0;
    Basis1D basis[3];
    IntVec FO, FN;                    // per facet node (nfacets*npf)
    std::vector<V3> fN, fC;           // per facet node
    std::vector<double> fI, fD;       // per facet node
    std::vector<V3> cC;               // per node (nall_field)
    std::vector<double> cV;           // per node
    std::vector<T9> Jinv;             // per node (DG only)
    IntVec allFaces, faceIndices0, faceIndices1;
};
// order[3] = polynomial order per direction (npx = order+1).  For FV pass 0s.
// nonortho: 0 NONE, 1 MINIMUM, 2 ORTHOGONAL, 3 OVER_RELAXED (FV fD only).
bool build_geometry(const LoadedMesh& m, const int order[3], int nonortho,
                    Geometry& geo, std::string* err);

}  // namespace nsem
