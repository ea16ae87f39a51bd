// Geometry dump driver for the reference build (test infrastructure only).
//
// Links against the reference library built by oracle/ref.mk, loads a mesh
// exactly as the reference solvers do (Mesh::LoadMesh, field.cpp:95-167,
// which runs addBoundaryCells, fixHexCells, calcGeometry, DG::init_poly,
// removeBoundary and initGeomMeshFields), then writes the loaded-mesh arrays
// to a flat binary file so tests can compare this repository's mesh pipeline
// against the reference's, array by array.
//
// usage: geomdump <controls> <out.bin>   (run in the case directory)
//
// File format: repeated records
//   char name[32] | u32 kind (0 = u32, 1 = f64, 2 = i64) | u64 count | data
#include "field.h"
#include "wrapper.h"
#include <cstdio>
#include <cstdint>
#include <cstring>

static FILE* g_out = nullptr;

static void put(const char* name, uint32_t kind, uint64_t count, const void* data) {
    char nm[32];
    std::memset(nm, 0, sizeof(nm));
    std::strncpy(nm, name, sizeof(nm) - 1);
    std::fwrite(nm, 1, 32, g_out);
    std::fwrite(&kind, 4, 1, g_out);
    std::fwrite(&count, 8, 1, g_out);
    size_t esz = (kind == 0) ? 4 : 8;
    if (count) std::fwrite(data, esz, count, g_out);
}

static void put_u32(const char* name, const std::vector<uint32_t>& v) { put(name, 0, v.size(), v.data()); }
static void put_f64(const char* name, const std::vector<double>& v) { put(name, 1, v.size(), v.data()); }

int main(int argc, char* argv[]) {
    MP mp(argc, argv);
    atexit(MP::cleanup);
    if (argc < 3) { std::fprintf(stderr, "usage: geomdump <controls> <out.bin>\n"); return 2; }
    Solver::input.open(argv[1]);
    {
        std::string sname;
        Util::ParamList params("general");
        params.enroll("solver", &sname);
        params.enroll("mesh", &Mesh::gMeshName);
        Mesh::enroll(params);
        params.read(Solver::input);
    }
    Mesh::LoadMesh(0);

    using namespace Mesh;
    g_out = std::fopen(argv[2], "wb");
    if (!g_out) return 3;

    std::vector<uint32_t> hdr = {
        (uint32_t)DG::NPX, (uint32_t)DG::NPY, (uint32_t)DG::NPZ, (uint32_t)DG::NP,
        (uint32_t)DG::NPF, (uint32_t)gBCS, (uint32_t)gNCells, (uint32_t)gNFacets,
        (uint32_t)gNVertices, (uint32_t)gBCSfield, (uint32_t)gALLfield};
    put_u32("header", hdr);

    std::vector<double> vtx;
    for (auto& v : gVertices) { vtx.push_back(v[0]); vtx.push_back(v[1]); vtx.push_back(v[2]); }
    put_f64("vertices", vtx);

    auto csr = [](const std::vector<IntVector>& lists, std::vector<uint32_t>& off, std::vector<uint32_t>& val) {
        off.assign(1, 0);
        for (auto& l : lists) { for (auto x : l) val.push_back(x); off.push_back(val.size()); }
    };
    {
        std::vector<uint32_t> o, v;
        csr(gFacets, o, v); put_u32("facet_off", o); put_u32("facet_vtx", v);
    }
    {
        std::vector<uint32_t> o, v;
        csr(gCells, o, v); put_u32("cell_off", o); put_u32("cell_faces", v);
    }
    {
        std::vector<uint32_t> o, v;
        csr(gFaceID, o, v); put_u32("faceid_off", o); put_u32("faceid", v);
    }
    put_u32("FOC", std::vector<uint32_t>(gFOC.begin(), gFOC.end()));
    put_u32("FNC", std::vector<uint32_t>(gFNC.begin(), gFNC.end()));
    put_u32("FMC", std::vector<uint32_t>(gFMC.begin(), gFMC.end()));
    {
        std::vector<uint32_t> hb;
        for (size_t i = 0; i < gHasBoundary.size(); i++) hb.push_back(gHasBoundary[i] ? 1 : 0);
        put_u32("has_boundary", hb);
    }
    // boundaries in map order: names joined by '\n' as bytes in a u32 array
    {
        std::string names;
        std::vector<uint32_t> off(1, 0), faces;
        for (auto& kv : gBoundaries) {
            names += kv.first; names += "\n";
            for (auto f : kv.second) faces.push_back(f);
            off.push_back(faces.size());
        }
        std::vector<uint32_t> nb(names.begin(), names.end());
        put_u32("bnames", nb);
        put_u32("boff", off);
        put_u32("bfaces", faces);
    }
    {
        std::vector<double> fc, fn, cc, cv;
        for (auto& v : gFC) { fc.push_back(v[0]); fc.push_back(v[1]); fc.push_back(v[2]); }
        for (auto& v : gFN) { fn.push_back(v[0]); fn.push_back(v[1]); fn.push_back(v[2]); }
        for (auto& v : gCC) { cc.push_back(v[0]); cc.push_back(v[1]); cc.push_back(v[2]); }
        for (auto& v : gCV) cv.push_back(v);
        put_f64("gFC", fc); put_f64("gFN", fn); put_f64("gCC", cc); put_f64("gCV", cv);
    }
    {
        std::vector<uint32_t> fo(FO.size()), fnn(FN.size());
        for (Int i = 0; i < FO.size(); i++) { fo[i] = FO[i]; fnn[i] = FN[i]; }
        put_u32("FO", fo); put_u32("FN", fnn);
    }
    {
        std::vector<double> a;
        for (Int i = 0; i < fN.size(); i++) { a.push_back(fN[i][0]); a.push_back(fN[i][1]); a.push_back(fN[i][2]); }
        put_f64("fN", a); a.clear();
        for (Int i = 0; i < fC.size(); i++) { a.push_back(fC[i][0]); a.push_back(fC[i][1]); a.push_back(fC[i][2]); }
        put_f64("fC", a); a.clear();
        for (Int i = 0; i < fI.size(); i++) a.push_back(fI[i]);
        put_f64("fI", a); a.clear();
        for (Int i = 0; i < fD.size(); i++) a.push_back(fD[i]);
        put_f64("fD", a); a.clear();
        for (Int i = 0; i < cC.size(); i++) { a.push_back(cC[i][0]); a.push_back(cC[i][1]); a.push_back(cC[i][2]); }
        put_f64("cC", a); a.clear();
        for (Int i = 0; i < cV.size(); i++) a.push_back(cV[i]);
        put_f64("cV", a); a.clear();
        if (DG::NPMAT) {
            for (Int i = 0; i < DG::Jinv.size(); i++)
                for (int c = 0; c < 9; c++) a.push_back(DG::Jinv[i][c]);
            put_f64("Jinv", a); a.clear();
            for (int d = 0; d < 3; d++) {
                Int n = DG::Nop[d] + 1;
                char nm[16];
                std::vector<double> x(DG::xgl[d], DG::xgl[d] + n), w(DG::wgl[d], DG::wgl[d] + n);
                std::vector<double> p(DG::psi[d], DG::psi[d] + n * n), dp(DG::dpsi[d], DG::dpsi[d] + n * n);
                std::snprintf(nm, sizeof nm, "xgl%d", d); put_f64(nm, x);
                std::snprintf(nm, sizeof nm, "wgl%d", d); put_f64(nm, w);
                std::snprintf(nm, sizeof nm, "psi%d", d); put_f64(nm, p);
                std::snprintf(nm, sizeof nm, "dpsi%d", d); put_f64(nm, dp);
                // 2:1 mortar matrices, lower then upper half
                std::vector<double> rf, cr;
                for (int h = 0; h < 2; h++) {
                    rf.insert(rf.end(), DG::psiRef[d * 2 + h], DG::psiRef[d * 2 + h] + n * n);
                    cr.insert(cr.end(), DG::psiCor[d * 2 + h], DG::psiCor[d * 2 + h] + n * n);
                }
                std::snprintf(nm, sizeof nm, "psiRef%d", d); put_f64(nm, rf);
                std::snprintf(nm, sizeof nm, "psiCor%d", d); put_f64(nm, cr);
            }
        }
    }
    {
        Int nf = 0;
        for (auto& c : gCells) nf += c.size();
        std::vector<uint32_t> af(allFaces, allFaces + nf);
        std::vector<uint32_t> fi0(faceIndices[0], faceIndices[0] + gNCells);
        std::vector<uint32_t> fi1(faceIndices[1], faceIndices[1] + gNCells);
        put_u32("allFaces", af); put_u32("faceIndices0", fi0); put_u32("faceIndices1", fi1);
    }
    std::fclose(g_out);
    return 0;
}
