// Run-time of the source-compatible field API: the global state the
// reference keeps in src/field/field.cpp, src/util/util.cpp, src/mp/mp.cpp
// and apps/utils (controls defaults and keys, parameter reader, the
// process object, Solver::Initialize, Fluid constants), the mesh loaded
// through the library's C-ABI with its face-node tables uploaded to the
// device, field files, boundary-condition tables, device memory, and
// reductions.
#include <sys/stat.h>
#include <sys/time.h>
#include <unistd.h>

#include <cstdarg>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>

#include "field.h"
#include "field_mesh.h"
#include "iteration.h"
#include "nebulasem_amd.h"
#include "properties.h"
#include "calc_walldist.h"
#include "solve.h"
#include "turbulence.h"
#include "wrapper.h"

// ------------------------------------------------------------- globals --
std::list<BaseField*> BaseField::allFields;
std::map<std::string, Util::ParamList*> Util::ParamList::list;
int MP::n_hosts = 1;
int MP::host_id = 0;
bool MP::printOn = true;
bool MP::Terminated = false;

namespace DG {
Int Nop[3] = {1, 1, 1};
Int NPX = 1, NPY = 1, NPZ = 1, NP = 1, NPF = 1, NPMAT = 0;
}  // namespace DG

// defaults as src/field/field.cpp:51-75
namespace Controls {
Scheme convection_scheme = HYBRID;
Int TVDbruner = 0;
NonOrthoScheme nonortho_scheme = OVER_RELAXED;
TimeScheme time_scheme = BDF1;
Scalar blend_factor = Scalar(0.2);
Scalar tolerance = Scalar(1e-5f);
Scalar dt = Scalar(.1);
Scalar SOR_omega = Scalar(1.7);
Solvers Solver = PCG;
Preconditioners Preconditioner = SSOR;
State state = STEADY;
Int max_iterations = 500;
Int write_interval = 20;
Int start_step = 0;
Int end_step = 2;
Int current_step = 0;
Int amr_step = 0;
Int n_deferred = 0;
Int save_average = 0;
Int print_time = 0;
CommMethod parallel_method = BLOCKED;
Vector gravity = Vector(0, 0, -9.860616);
FILE_FORMAT write_format = BINARY;
static Vector refine_dir(0, 0, 0);
static std::string refine_field;
static Scalar refine_min = 0, refine_max = 0;
static Int refine_level = 0, refine_buffer = 0, refine_limit = 0, decompose_type = 0;
static std::vector<Int> decompose_n, decompose_axis;
}  // namespace Controls

namespace Mesh {
std::string gMeshName = "grid";
Int gBCS = 0, gNCells = 0, gNFacets = 0, gBCSfield = 0, gALLfield = 0;
Int is_spherical = 0;
Scalar sphere_radius = 0, sphere_height = 0;
VectorCellField cC(false);
ScalarCellField cV(false);
VectorFacetField fN(false);
static std::vector<Vector> probePoints;
}  // namespace Mesh

namespace Fluid {
Scalar density, viscosity, Pr, Prt, T0, P0, cp, cv, beta;
// defaults and keys as apps/utils/properties.cpp:14-33
void enroll(Util::ParamList& params) {
    density = 1.177;
    viscosity = 1.568e-5;
    Pr = 0.9;
    Prt = 0.7;
    beta = 3.33e-3;
    T0 = 300;
    P0 = 101325;
    cp = 1004.67;
    cv = 715.5;
    params.enroll("rho", &density);
    params.enroll("viscosity", &viscosity);
    params.enroll("Pr", &Pr);
    params.enroll("Prt", &Prt);
    params.enroll("beta", &beta);
    params.enroll("T0", &T0);
    params.enroll("P0", &P0);
    params.enroll("cp", &cp);
    params.enroll("cv", &cv);
}
}  // namespace Fluid

// ---------------------------------------------------------- parameters --
Int Util::hash_function(std::string s) {
    Int h = 0;
    for (const char* p = s.c_str(); *p; p++) h = 31 * h + (Int)*p;
    return h;
}
int Util::nextc(std::istream& is) {
    char c;
    if (!(is >> c)) return 0;
    while (c == '#') {
        int ch;
        while ((ch = is.get()) != EOF && ch != '\n') {}
        if (!(is >> c)) return 0;
    }
    is.putback(c);
    return c;
}
Util::ParamList::ParamList(std::string n) : name(n) { list[name] = this; }
Util::ParamList::~ParamList() {
    auto it = list.find(name);
    if (it != list.end() && it->second == this) list.erase(it);
}
bool Util::ParamList::read(std::istream& is, const std::string& key, bool print) {
    auto show = [&](const auto& v) {
        if (print) std::cout << v;
    };
    if (auto it = ints.find(key); it != ints.end()) { is >> *it->second; show(*it->second); return true; }
    if (auto it = strings.find(key); it != strings.end()) { is >> *it->second; show(*it->second); return true; }
    if (auto it = options.find(key); it != options.end()) {
        std::string w;
        is >> w;
        *it->second->val = it->second->getID(w);
        show(w);
        return true;
    }
    if (auto it = scalars.find(key); it != scalars.end()) { is >> *it->second; show(*it->second); return true; }
    if (auto it = vectors.find(key); it != vectors.end()) { is >> *it->second; return true; }
    if (auto it = vints.find(key); it != vints.end()) { is >> *it->second; return true; }
    if (auto it = vscalars.find(key); it != vscalars.end()) { is >> *it->second; return true; }
    if (auto it = vvectors.find(key); it != vvectors.end()) { is >> *it->second; return true; }
    if (auto it = vstrings.find(key); it != vstrings.end()) { is >> *it->second; return true; }
    if (print) std::cout << "UNKNOWN";
    return false;
}

// read_params (util.cpp:32-79): blocks whose name has an enrolled list are
// read key by key; other blocks are skipped brace-balanced.  With `block`,
// only that block is read.  The stream is rewound afterwards.
void Util::read_params(std::istream& is, bool print, std::string block) {
    std::string str;
    char c;
    while (true) {
        if (!nextc(is)) break;
        is >> str;
        if (!(is >> c)) break;  // '{'
        auto it = ParamList::list.find(str);
        if (it == ParamList::list.end() || (!block.empty() && compare(str, block))) {
            int braces = 1;
            while (nextc(is)) {
                is >> c;
                if (c == '{') braces++;
                else if (c == '}' && --braces == 0) break;
            }
            continue;
        }
        if (print) std::cout << str << "\n{" << std::endl;
        ParamList* params = it->second;
        while (true) {
            c = (char)nextc(is);
            if (!c) break;
            if (c == '}') {
                is >> c;
                break;
            }
            is >> str;
            if (print) std::cout << str << " = ";
            params->read(is, str, print);
            if (print) std::cout << std::endl;
        }
        if (print) std::cout << "}\n";
        if (!block.empty()) break;
    }
    is.clear();
    is.seekg(0, std::ios::beg);
}

// ----------------------------------------------------------------- MP ---
static int g_start_ms = 0;
static int now_ms() {
    timeval tb;
    gettimeofday(&tb, nullptr);
    return int(tb.tv_sec * 1000 + tb.tv_usec / 1000);
}
MP::MP(int, char*[]) {
    g_start_ms = now_ms();
    printf("--------------------------------------------\n");
    printf("%d processes started on one GPU, pid %d\n", n_hosts, (int)getpid());
    fflush(stdout);
}
MP::~MP() {}
void MP::cleanup() {
    Terminated = true;
    printf("%d [%d] Exiting application run with %d processes\n", now_ms() - g_start_ms, host_id, n_hosts);
    fflush(stdout);
}
void MP::printH(const char* format, ...) {
    printf("%d [%d] ", now_ms() - g_start_ms, host_id);
    va_list ap;
    va_start(ap, format);
    vprintf(format, ap);
    va_end(ap);
    fflush(stdout);
}
void MP::print(const char* format, ...) {
    va_list ap;
    va_start(ap, format);
    vprintf(format, ap);
    va_end(ap);
    fflush(stdout);
}
bool MP::hasElapsed(const Int delta) {
    static int prev = 0;
    const int t = now_ms();
    if (t - prev < (int)delta) return false;
    prev = t;
    return true;
}

// ------------------------------------------------------------- Solver ---
namespace Solver {
std::ifstream input;
static std::string sname;
static std::vector<std::string> field_names;
void Initialize(int argc, char* argv[]) {
    atexit(MP::cleanup);
    MP::printOn = (MP::host_id == 0);
    if (argc < 2 || !strcmp(argv[1], "-h")) {
        std::cout << "Usage:\n  " << argv[0] << " <inputfile>\n";
        exit(0);
    }
    input.open(argv[1]);
    if (!input) nsem_field::fail(std::string("cannot open ") + argv[1]);
    {
        Util::ParamList params("general");
        params.enroll("solver", &sname);
        params.enroll("mesh", &Mesh::gMeshName);
        Mesh::enroll(params);
        params.read(input);
        std::string path(argv[0]);
        if (path.find(sname) == std::string::npos) {
            std::cout << "Incorrect solver " << path << " used instead of " << sname << "." << std::endl;
            exit(1);
        }
    }
    {
        Util::ParamList params("refinement");
        Controls::enrollRefine(params);
        params.read(input);
    }
    {
        Util::ParamList params("decomposition");
        Controls::enrollDecompose(params);
        params.read(input);
    }
    {
        Util::ParamList params("prepare");
        params.enroll("fields", &field_names);
        params.read(input);
    }
}
void Finalize() { input.close(); }
}  // namespace Solver

void Controls::enrollRefine(Util::ParamList& params) {
    params.enroll("direction", &refine_dir);
    params.enroll("field", &refine_field);
    params.enroll("field_max", &refine_max);
    params.enroll("field_min", &refine_min);
    params.enroll("max_level", &refine_level);
    params.enroll("buffer_zone", &refine_buffer);
    params.enroll("limit", &refine_limit);
}
void Controls::enrollDecompose(Util::ParamList& params) {
    params.enroll("n", &decompose_n);
    params.enroll("axis", &decompose_axis);
    params.enroll("type", new Util::Option(&decompose_type, {"METIS", "XYZ", "CELLID", "NONE"}));
}
// the general block's keys (field.cpp:496-552)
void Mesh::enroll(Util::ParamList& params) {
    using namespace Controls;
    using Util::BoolOption;
    using Util::Option;
    params.enroll("max_iterations", &max_iterations);
    params.enroll("write_interval", &write_interval);
    params.enroll("start_step", &start_step);
    params.enroll("end_step", &end_step);
    params.enroll("amr_step", &amr_step);
    params.enroll("n_deferred", &n_deferred);
    params.enroll("blend_factor", &blend_factor);
    params.enroll("tolerance", &tolerance);
    params.enroll("dt", &dt);
    params.enroll("SOR_omega", &SOR_omega);
    params.enroll("probe", &Mesh::probePoints);
    params.enroll("gravity", &gravity);
    params.enroll("is_spherical", new BoolOption(&is_spherical));
    params.enroll("sphere_radius", &sphere_radius);
    params.enroll("sphere_height", &sphere_height);
    params.enroll("convection_scheme",
                  new Option(&convection_scheme, {"CDS", "UDS", "BLENDED", "HYBRID", "RUSANOV", "LUD", "CDSS", "MUSCL",
                                                  "QUICK", "VANLEER", "VANALBADA", "MINMOD", "SUPERBEE", "SWEBY",
                                                  "QUICKL", "UMIST", "DDS", "FROMM"}));
    params.enroll("tvd_bruner", new BoolOption(&TVDbruner));
    params.enroll("nonortho_scheme", new Option(&nonortho_scheme, {"NONE", "MINIMUM", "ORTHOGONAL", "OVER_RELAXED"}));
    params.enroll("time_scheme", new Option(&time_scheme, {"BDF1", "BDF2", "BDF3", "BDF4", "BDF5", "BDF6", "AM1", "AM2",
                                                           "AM3", "AM4", "AM5", "AB1", "AB2", "AB3", "AB4", "AB5",
                                                           "RK1", "RK2", "RK3", "RK4"}));
    params.enroll("method", new Option(&Controls::Solver, {"JAC", "SOR", "PCG"}));
    params.enroll("preconditioner", new Option(&Preconditioner, {"NONE", "DIAG", "SSOR", "DILU"}));
    params.enroll("state", new Option(&state, {"STEADY", "TRANSIENT"}));
    params.enroll("parallel_method", new Option(&parallel_method, {"BLOCKED", "ASYNCHRONOUS"}));
    params.enroll("write_format", new Option(&write_format, {"TEXT", "BINARY"}));
    params.enroll("average", new BoolOption(&save_average));
    params.enroll("print_time", &print_time);
    params.enroll("npx", &DG::Nop[0]);
    params.enroll("npy", &DG::Nop[1]);
    params.enroll("npz", &DG::Nop[2]);
}

// --------------------------------------------------------- device memory --
namespace nsem_field {
hipStream_t stream() { return nullptr; }
void fail(const std::string& what) {
    fprintf(stderr, "nebulasem field API: %s\n", what.c_str());
    fflush(stderr);
    exit(1);
}
void check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(std::string(what) + ": " + hipGetErrorString(e));
}
static std::map<size_t, std::vector<void*>>& pool() {
    static auto* p = new std::map<size_t, std::vector<void*>>();
    return *p;
}
void* dev_alloc(size_t bytes) {
    bytes = (bytes + 255) / 256 * 256;
    auto& v = pool()[bytes];
    if (!v.empty()) {
        void* p = v.back();
        v.pop_back();
        return p;
    }
    void* p = nullptr;
    check(hipMalloc(&p, bytes ? bytes : 256), "hipMalloc");
    return p;
}
void dev_free(void* p, size_t bytes) {
    if (!p) return;
    bytes = (bytes + 255) / 256 * 256;
    pool()[bytes].push_back(p);
}
void upload(void* dst, const void* src, size_t bytes) { check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "upload"); }
void download(void* dst, const void* src, size_t bytes) {
    check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "download");
}
void copy_d2d(void* dst, const void* src, size_t bytes) {
    check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream()), "copy");
}

// ------------------------------------------------------------- the mesh --
static MeshTables g_mesh;
static nsem_mesh* g_handle = nullptr;
const MeshTables& mesh_tables() { return g_mesh; }
nsem_mesh* mesh_handle() { return g_handle; }
size_t entity_size(ENTITY e) {
    switch (e) {
        case CELL: return g_mesh.nall;
        case FACET: return (size_t)g_mesh.nfacets * g_mesh.npf;
        case VERTEX: return g_mesh.nvertices;
        default: return 0;
    }
}
bool dg() { return g_mesh.np > 1; }

template <class T>
static std::vector<T> mesh_array(const char* name) {
    uint64_t n = 0;
    int32_t kind = 0;
    if (nsem_mesh_array_size(g_handle, name, &n, &kind)) fail(std::string("mesh array ") + name + ": " + nsem_last_error());
    std::vector<T> v(n);
    if (n && nsem_mesh_get_array(g_handle, name, v.data(), n)) fail(std::string("mesh array ") + name);
    return v;
}
template <class T>
static T* put(const std::vector<T>& v) {
    T* d = static_cast<T*>(dev_alloc(v.size() * sizeof(T) + 8));
    if (!v.empty()) upload(d, v.data(), v.size() * sizeof(T));
    return d;
}

static void load_mesh_file(const std::string& path) {
    const int32_t order[3] = {(int32_t)DG::Nop[0], (int32_t)DG::Nop[1], (int32_t)DG::Nop[2]};
    if (g_handle) nsem_mesh_destroy(g_handle);
    g_handle = nullptr;
    if (nsem_mesh_load(path.c_str(), order, (int32_t)Controls::nonortho_scheme, &g_handle))
        fail(std::string("loading ") + path + ": " + nsem_last_error());
    nsem_mesh_info info;
    nsem_mesh_get_info(g_handle, &info);
    MeshTables& m = g_mesh;
    m.npx = info.npx; m.npy = info.npy; m.npz = info.npz; m.np = info.np; m.npf = info.npf;
    m.nbcs = info.nbcs; m.ncells = info.ncells; m.nfacets = info.nfacets; m.nvertices = info.nvertices;
    m.nbf = info.nbcs_field; m.nall = info.nall_field;
    DG::NPX = m.npx; DG::NPY = m.npy; DG::NPZ = m.npz; DG::NP = m.np; DG::NPF = m.npf;
    DG::NPMAT = m.np > 1 ? 1 : 0;
    Mesh::gBCS = m.nbcs; Mesh::gNCells = m.ncells; Mesh::gNFacets = m.nfacets;
    Mesh::gBCSfield = (Int)m.nbf; Mesh::gALLfield = (Int)m.nall;
    {
        std::vector<uint32_t> fmc = mesh_array<uint32_t>("FMC");
        for (uint32_t f : fmc)
            if (f) fail("the field API does not handle grids with 2:1 mortar faces");
    }
    std::vector<uint32_t> FO = mesh_array<uint32_t>("FO"), FN = mesh_array<uint32_t>("FN");
    std::vector<double> fNv = mesh_array<double>("fN"), fI = mesh_array<double>("fI"), fD = mesh_array<double>("fD");
    std::vector<double> cCv = mesh_array<double>("cC"), cVv = mesh_array<double>("cV");
    std::vector<uint32_t> allFaces = mesh_array<uint32_t>("allFaces"), fi0 = mesh_array<uint32_t>("faceIndices0"),
                          fi1 = mesh_array<uint32_t>("faceIndices1");
    std::vector<uint32_t> boff = mesh_array<uint32_t>("boff"), bfaces = mesh_array<uint32_t>("bfaces");
    const uint64_t nfn = (uint64_t)m.nfacets * m.npf;
    // per node (interior), the face nodes of its own cell that touch it, in
    // the order div_flux visits them (the cell's faces in allFaces order,
    // then the face's nodes), with the side: +1 owner (FO), -1 neighbour (FN)
    {
        std::vector<std::vector<int64_t>> ent(m.nall);
        for (uint64_t i = 0; i < m.ncells; i++)
            for (uint32_t f = fi0[i]; f < fi1[i]; f++)
                for (uint32_t n = 0; n < m.npf; n++) {
                    const uint64_t k = (uint64_t)allFaces[f] * m.npf + n;
                    const uint32_t c1 = FO[k], c2 = FN[k];
                    if (c1 >= i * m.np && c1 < (i + 1) * m.np) ent[c1].push_back((int64_t)(k << 1));
                    else if (c2 < m.nall) ent[c2].push_back((int64_t)((k << 1) | 1));
                }
        std::vector<uint32_t> off(m.nall + 1, 0);
        std::vector<uint64_t> lst;
        for (uint64_t t = 0; t < m.nall; t++) {
            for (int64_t e : ent[t]) lst.push_back((uint64_t)e);
            off[t + 1] = (uint32_t)lst.size();
        }
        m.ent_off = put(off);
        m.ent = put(lst);
        m.ent_off_h = std::move(off);
        m.ent_h = std::move(lst);
        m.FO_h = FO;
        m.FN_h = FN;
        std::vector<uint32_t> hb = mesh_array<uint32_t>("has_boundary");
        m.has_boundary_h.assign(hb.begin(), hb.end());
        m.has_boundary_h.resize(m.nbcs, 0);
        m.generation++;
    }
    // ghost nodes: owner node of each ghost node (fillBCs, field.h:2733-2743)
    {
        std::vector<uint32_t> go, gn;
        for (uint64_t i = m.nbcs; i < m.ncells; i++) {
            const uint32_t f = fi0[i];
            for (uint32_t n = 0; n < m.npf; n++) {
                const uint64_t k = (uint64_t)allFaces[f] * m.npf + n;
                if (FN[k] < m.nall) { go.push_back(FO[k]); gn.push_back(FN[k]); }
            }
        }
        m.nghostnodes = (uint32_t)go.size();
        m.ghost_owner = put(go);
        m.ghost_node = put(gn);
    }
    // boundary face nodes per patch: face node k, owner c1, ghost c2
    {
        m.npatch = (uint32_t)(boff.size() ? boff.size() - 1 : 0);
        m.patch_names.clear();
        for (uint32_t b = 0; b < m.npatch; b++) {
            char buf[256];
            nsem_mesh_boundary_name(g_handle, b, buf, sizeof buf);
            m.patch_names.push_back(buf);
        }
        std::vector<uint32_t> po(1, 0), pk, pj;
        for (uint32_t b = 0; b < m.npatch; b++) {
            for (uint32_t j = boff[b]; j < boff[b + 1]; j++)
                for (uint32_t n = 0; n < m.npf; n++) {
                    const uint64_t k = (uint64_t)bfaces[j] * m.npf + n;
                    if (FN[k] >= m.nall) continue;
                    pk.push_back((uint32_t)k);
                    pj.push_back((j - boff[b]) * m.npf + n);
                }
            po.push_back((uint32_t)pk.size());
        }
        m.patch_off_host = po;
        m.patch_k = put(pk);
        m.patch_slot = put(pj);
    }
    m.FO = put(FO);
    m.FN = put(FN);
    m.fN = put(fNv);
    m.fI = put(fI);
    m.fD = put(fD);
    m.cC = put(cCv);
    m.cV = put(cVv);
    if (m.np > 1) {
        std::vector<double> Jinv = mesh_array<double>("Jinv");
        // Jin = Jinv cV per node (the reference forms it so in divf/gradf)
        std::vector<double> jcv(9 * m.nall);
        for (uint64_t i = 0; i < m.nall; i++)
            for (int c = 0; c < 9; c++) jcv[9 * i + c] = Jinv[9 * i + c] * cVv[i];
        m.JcV = put(jcv);
        for (int d = 0; d < 3; d++) {
            char nm[8];
            snprintf(nm, sizeof nm, "dpsi%d", d);
            m.D[d] = put(mesh_array<double>(nm));
        }
    }
    // the mesh fields apps read: cC, cV, fN
    Mesh::cC.construct("");
    Mesh::cV.construct("");
    Mesh::fN.construct("");
    upload(Mesh::cC.dptr_w(), cCv.data(), cCv.size() * sizeof(double));
    upload(Mesh::cV.dptr_w(), cVv.data(), cVv.size() * sizeof(double));
    upload(Mesh::fN.dptr_w(), fNv.data(), fNv.size() * sizeof(double));
    (void)nfn;
    MP::printH("\t%d vertices\t%d facets\t%d cells\n", (int)m.nvertices, (int)m.nfacets, (int)m.ncells);
}

static bool exists(const std::string& p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}
}  // namespace nsem_field

// LoadMesh (field.cpp:95-167): the last refined grid <mesh>_<step> at or
// before step, text or binary
void Mesh::LoadMesh(Int step_) {
    int step = (int)step_;
    for (; step >= 0; --step) {
        const std::string b = gMeshName + "_" + std::to_string(step);
        if (nsem_field::exists(b + ".txt") || nsem_field::exists(b + ".bin")) break;
    }
    if (step < 0) step = 0;
    const std::string base = gMeshName + "_" + std::to_string(step);
    const std::string path = nsem_field::exists(base + ".txt") ? base + ".txt" : base + ".bin";
    if (!nsem_field::exists(path)) nsem_field::fail("no grid file " + base + ".txt or .bin");
    nsem_field::load_mesh_file(path);
}
// refineMesh (field.cpp:625-953): with max_level 0 no cell can be refined
// or coarsened and the grid stays as loaded; refinement itself is not built
void Prepare::refineMesh(Int step) {
    std::cout << "Refining grid at step " << step << std::endl;
    if (Controls::refine_level <= 0) return;
    nsem_field::fail("mesh refinement (max_level > 0) is not part of the field API");
}

// ---------------------------------------------------- boundary tables ---
namespace nsem_field {
static std::map<Int, FieldBCs>& bc_map() {
    static auto* m = new std::map<Int, FieldBCs>();
    return *m;
}
FieldBCs& bcs_of(Int fIndex, int ncomp) {
    FieldBCs& b = bc_map()[fIndex];
    const size_t np = g_mesh.npatch;
    if (b.kind.size() != np) {
        b.kind.assign(np, BC_NONE);
        b.value.assign(np * ncomp, 0.0);
        b.shape.assign(np, 0.0);
        b.tvalue.assign(np * ncomp, 0.0);
        b.fixed.assign(np, {});
    }
    return b;
}
bool has_bcs(Int fIndex) { return bc_map().count(fIndex) != 0; }

void copy_bcs(Int src, Int dst, int ncomp, int mode, Scalar psi) {
    if (!has_bcs(src)) return;
    FieldBCs s = bcs_of(src, ncomp);
    FieldBCs& d = bcs_of(dst, ncomp);
    for (size_t b = 0; b < s.kind.size(); b++) {
        if (s.kind[b] == BC_NONE) continue;
        if (mode == 0) {  // fixedBCs: CALC_DIRICHLET, values fixed at the first application
            d.kind[b] = BC_CALC_DIRICHLET;
            d.fixed[b].clear();
        } else {  // scaleBCs
            d.kind[b] = s.kind[b];
            d.shape[b] = s.shape[b];
            for (int c = 0; c < ncomp; c++) {
                d.value[b * ncomp + c] = s.value[b * ncomp + c];
                d.tvalue[b * ncomp + c] = s.tvalue[b * ncomp + c];
            }
            if (s.kind[b] == BC_NEUMANN) {
                for (int c = 0; c < ncomp; c++) d.value[b * ncomp + c] *= psi;
            } else if (s.kind[b] == BC_ROBIN) {
                for (int c = 0; c < ncomp; c++) {
                    d.value[b * ncomp + c] *= psi;
                    d.tvalue[b * ncomp + c] *= psi;
                }
            } else if (s.kind[b] != BC_SYMMETRY) {
                d.kind[b] = BC_CALC_DIRICHLET;
                d.fixed[b].clear();
            }
        }
    }
}

void read_field_file(const std::string& name, Int step, int ncomp, Scalar* host_all, Int fIndex) {
    const std::string base = name + std::to_string(step);
    std::string path = exists(base + ".bin") ? base + ".bin" : base + ".txt";
    if (!exists(path)) fail("cannot read field file " + base + ".txt or .bin");
    if (ncomp != 1 && ncomp != 3) fail("field files of " + std::to_string(ncomp) + " components are not read here");
    std::vector<int32_t> kind(g_mesh.npatch + 1);
    std::vector<double> val((g_mesh.npatch + 1) * ncomp), shape(g_mesh.npatch + 1), tval((g_mesh.npatch + 1) * ncomp);
    const double g[3] = {Controls::gravity[0], Controls::gravity[1], Controls::gravity[2]};
    if (nsem_field_read_bc(g_handle, path.c_str(), ncomp, g, host_all, kind.data(), val.data(), shape.data(),
                           tval.data()))
        fail("reading " + path + ": " + nsem_last_error());
    FieldBCs& b = bcs_of(fIndex, ncomp);
    for (size_t p = 0; p < g_mesh.npatch; p++) {
        b.kind[p] = kind[p] < 0 ? BC_NONE : kind[p];
        b.shape[p] = shape[p];
        for (int c = 0; c < ncomp; c++) {
            b.value[p * ncomp + c] = val[p * ncomp + c];
            b.tvalue[p * ncomp + c] = tval[p * ncomp + c];
        }
        b.fixed[p].clear();
    }
    MP::printH("Reading %s\n", base.c_str());
}
void write_field_file(const std::string& name, Int step, int ncomp, const Scalar* host_all, Int fIndex) {
    if (ncomp != 1 && ncomp != 3) return;
    const std::string path =
        name + std::to_string(step) + (Controls::write_format == Controls::BINARY ? ".bin" : ".txt");
    std::vector<int32_t> kind(g_mesh.npatch + 1, -1);
    std::vector<double> val((g_mesh.npatch + 1) * ncomp, 0.0), shape(g_mesh.npatch + 1, 0.0),
        tval((g_mesh.npatch + 1) * ncomp, 0.0);
    if (has_bcs(fIndex)) {
        FieldBCs& b = bcs_of(fIndex, ncomp);
        for (size_t p = 0; p < g_mesh.npatch; p++) {
            kind[p] = (b.kind[p] >= BC_NEUMANN && b.kind[p] <= BC_ROBIN) ? b.kind[p] : -1;
            shape[p] = b.shape[p];
            for (int c = 0; c < ncomp; c++) {
                val[p * ncomp + c] = b.value[p * ncomp + c];
                tval[p * ncomp + c] = b.tvalue[p * ncomp + c];
            }
        }
    }
    if (nsem_field_write_bc(g_handle, path.c_str(), ncomp, host_all, kind.data(), val.data(), shape.data(),
                            tval.data()))
        fail("writing " + path + ": " + nsem_last_error());
}
}  // namespace nsem_field

void Mesh::read_fields(Int step) {
    for (BaseField* f : BaseField::allFields)
        if (f->is_cell() && (f->access & READ)) f->read_file(step);
}
void Mesh::write_fields(Int step) {
    for (BaseField* f : BaseField::allFields)
        if (f->is_cell() && (f->access & WRITE)) f->write_file(step);
}

// calc_courant (field.cpp:440-448)
Vector Mesh::calc_courant(const VectorCellField& U, Scalar dt) {
    ScalarCellField Courant;
    Courant = mag(U) * dt / pow(cV, 1.0 / 3);
    const Scalar maxc = reduce_max(Courant);
    const Scalar minc = reduce_min(Courant);
    const Scalar avgc = reduce_sum(Courant) / Scalar(gBCSfield);
    return Vector(maxc, minc, avgc);
}

// ------------------------------------------------------------ turbulence --
// the reference's default selection is k-e (turbulence.cpp:8)
Int Turbulence_Model::turb_model = 2;
bool Turbulence_Model::bneedWallDist = false;
void Turbulence_Model::RegisterTable(Util::ParamList& params) {
    params.enroll("turbulence_model",
                  new Util::Option(&turb_model, {"NONE", "MIXING_LENGTH", "KE", "RNG_KE", "REALIZABLE_KE", "KW", "LES"}));
}
// Select (turbulence.cpp:21-57): NONE is the base (laminar) model
Turbulence_Model* Turbulence_Model::Select(VectorCellField& U, VectorCellField& Fc, ScalarFacetField& F,
                                           ScalarCellField& rho, ScalarCellField& mu) {
    bneedWallDist = false;
    if (turb_model != 0)
        nsem_field::fail("turbulence_model: only NONE (laminar) is part of this build; the RANS and LES models "
                         "are not");
    Turbulence_Model* t = new Turbulence_Model(U, Fc, F, rho, mu);
    t->enroll();
    return t;
}
void Mesh::calc_walldist(Int, Int) { nsem_field::fail("wall distance is not part of this build"); }
