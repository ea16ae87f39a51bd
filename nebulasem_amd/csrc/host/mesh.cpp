// Host mesh pipeline: block grid generation, grid I/O, the load-time mesh
// fix-ups and the DG/FV node geometry.  See mesh.hpp for the contract.
//
// The floating-point helpers below evaluate exactly the same expressions, in
// the same order, as the reference's tensor library (src/tensor/tensor.h,
// tensor.cpp), because mesh numbering decisions (coplanarity, corner picking,
// normal orientation) and the node geometry are compared bit for bit.
#include "mesh.hpp"
#include "parallel.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <chrono>
#include <cstdio>
#include <cstdlib>

namespace nsem {

namespace {

constexpr double kPi = 3.14159265358979323846264;
constexpr double kEqualEps = 1e-7;  // tensor.h:460

inline V3 sub(const V3& a, const V3& b) { return V3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline V3 add(const V3& a, const V3& b) { return V3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
inline V3 smul(double s, const V3& a) { return V3(s * a[0], s * a[1], s * a[2]); }
inline V3 sdiv(const V3& a, double s) { return V3(a[0] / s, a[1] / s, a[2] / s); }
// Right-nested sum, as the reference's unrolled dot product (tensor.h:117-119).
inline double dot(const V3& p, const V3& q) {
    double r = p[2] * q[2] + 0.0;
    r = p[1] * q[1] + r;
    return p[0] * q[0] + r;
}
inline double mag(const V3& p) { return std::sqrt(dot(p, p)); }
inline V3 unit(const V3& p) { return sdiv(p, mag(p)); }
inline V3 cross(const V3& p, const V3& q) {
    return V3(p[1] * q[2] - p[2] * q[1], p[2] * q[0] - p[0] * q[2], p[0] * q[1] - p[1] * q[0]);
}
inline bool equal_s(double p, double q, double tol = kEqualEps) {
    double d = std::fabs(p - q);
    return d <= tol || d <= tol * std::fabs(p) || d <= tol * std::fabs(q);
}
inline bool equal_v(const V3& p, const V3& q) {
    return equal_s(p[0], q[0]) && equal_s(p[1], q[1]) && equal_s(p[2], q[2]);
}

// Point v strictly inside segment (v1,v2) (mesh.cpp:791-810).
bool point_in_line(const V3& v, const V3& v1, const V3& v2) {
    V3 p = sub(v, v1), q = sub(v, v2);
    double e = p[1] * q[2] - p[2] * q[1];
    if (!equal_s(e, 0.0)) return false;
    e = p[2] * q[0] - p[0] * q[2];
    if (!equal_s(e, 0.0)) return false;
    e = p[0] * q[1] - p[1] * q[0];
    if (!equal_s(e, 0.0)) return false;
    e = dot(sub(v, v2), sub(v1, v2));
    if (e > 0.0) {
        double e1 = dot(sub(v1, v2), sub(v1, v2));
        if (e < e1) return true;
    }
    return false;
}

// Transfinite interpolation on a face / in a cell; term order as
// tensor.h:522-584.
V3 interp_face(double r, double s, const V3& x00, const V3& x01, const V3& x10, const V3& x11,
               const V3& xr0, const V3& xr1, const V3& x0s, const V3& x1s) {
    V3 out;
    for (int c = 0; c < 3; c++) {
        double t = (-(1.0 - r)) * (1.0 - s) * x00[c];
        t = t + (1.0 - r) * x0s[c];
        t = t - (1.0 - r) * s * x01[c];
        t = t + (1.0 - s) * xr0[c];
        t = t + s * xr1[c];
        t = t - r * (1.0 - s) * x10[c];
        t = t + r * x1s[c];
        t = t - r * s * x11[c];
        out[c] = t;
    }
    return out;
}

V3 interp_cell(double r, double s, double t, const V3& x000, const V3& x001, const V3& x010,
               const V3& x011, const V3& x100, const V3& x101, const V3& x110, const V3& x111,
               const V3& xr00, const V3& xr01, const V3& xr10, const V3& xr11, const V3& x0s0,
               const V3& x0s1, const V3& x1s0, const V3& x1s1, const V3& x00t, const V3& x01t,
               const V3& x10t, const V3& x11t, const V3& x0st, const V3& x1st, const V3& xr0t,
               const V3& xr1t, const V3& xrs0, const V3& xrs1) {
    V3 out;
    for (int c = 0; c < 3; c++) {
        double a = (1.0 - r) * (1.0 - s) * (1.0 - t) * x000[c];
        a = a - (1.0 - r) * (1.0 - s) * x00t[c];
        a = a + (1.0 - r) * (1.0 - s) * t * x001[c];
        a = a - (1.0 - r) * (1.0 - t) * x0s0[c];
        a = a + (1.0 - r) * x0st[c];
        a = a - (1.0 - r) * t * x0s1[c];
        a = a + (1.0 - r) * s * (1.0 - t) * x010[c];
        a = a - (1.0 - r) * s * x01t[c];
        a = a + (1.0 - r) * s * t * x011[c];
        a = a - (1.0 - s) * (1.0 - t) * xr00[c];
        a = a + (1.0 - s) * xr0t[c];
        a = a - (1.0 - s) * t * xr01[c];
        a = a + (1.0 - t) * xrs0[c];
        a = a + t * xrs1[c];
        a = a - s * (1.0 - t) * xr10[c];
        a = a + s * xr1t[c];
        a = a - s * t * xr11[c];
        a = a + r * (1.0 - s) * (1.0 - t) * x100[c];
        a = a - r * (1.0 - s) * x10t[c];
        a = a + r * (1.0 - s) * t * x101[c];
        a = a - r * (1.0 - t) * x1s0[c];
        a = a + r * x1st[c];
        a = a - r * t * x1s1[c];
        a = a + r * s * (1.0 - t) * x110[c];
        a = a - r * s * x11t[c];
        a = a + r * s * t * x111[c];
        out[c] = a;
    }
    return out;
}

// Straight-edge hex interpolation shared by the grid generator and the DG
// node placement (hexMesh.cpp ADD(), dg.cpp:240-300).  ev[w] are the end
// points of the 12 block edges; m[w] the edge parameters.
static const int kSides[12][2] = {{0, 1}, {3, 2}, {7, 6}, {4, 5}, {0, 3}, {1, 2},
                                  {5, 6}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

V3 hex_point(const V3* vp, const double m[12], double rx, double ry, double rz) {
    V3 vd[12], vf[6];
    for (int w = 0; w < 12; w++) {
        const V3& a = vp[kSides[w][0]];
        const V3& b = vp[kSides[w][1]];
        vd[w] = add(smul(1 - m[w], a), smul(m[w], b));
    }
    vf[0] = interp_face(rx, ry, vp[0], vp[3], vp[1], vp[2], vd[0], vd[1], vd[4], vd[5]);
    vf[1] = interp_face(rx, ry, vp[4], vp[7], vp[5], vp[6], vd[3], vd[2], vd[7], vd[6]);
    vf[2] = interp_face(rx, rz, vp[0], vp[4], vp[1], vp[5], vd[0], vd[3], vd[8], vd[9]);
    vf[3] = interp_face(rx, rz, vp[3], vp[7], vp[2], vp[6], vd[1], vd[2], vd[11], vd[10]);
    vf[4] = interp_face(ry, rz, vp[0], vp[4], vp[3], vp[7], vd[4], vd[7], vd[8], vd[11]);
    vf[5] = interp_face(ry, rz, vp[1], vp[5], vp[2], vp[6], vd[5], vd[6], vd[9], vd[10]);
    return interp_cell(rx, ry, rz, vp[0], vp[4], vp[3], vp[7], vp[1], vp[5], vp[2], vp[6], vd[0],
                       vd[3], vd[1], vd[2], vd[4], vd[7], vd[5], vd[6], vd[8], vd[11], vd[9],
                       vd[10], vf[4], vf[5], vf[2], vf[3], vf[0], vf[1]);
}

template <typename T>
void erase_indices(std::vector<T>& data, const IntVec& idx) {
    // idx sorted ascending (as every caller builds it)
    std::vector<T> keep;
    keep.reserve(data.size() - idx.size());
    size_t p = 0;
    for (size_t i = 0; i < data.size(); i++) {
        if (p < idx.size() && idx[p] == i) { p++; continue; }
        keep.push_back(data[i]);
    }
    data.swap(keep);
}

struct PhaseTimer {
    bool on = std::getenv("NSEM_TIMING") != nullptr;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    void mark(const char* what) {
        if (!on) return;
        auto n = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[nsem] %-28s %8.3f s\n", what, std::chrono::duration<double>(n - t).count());
        t = n;
    }
};

bool set_err(std::string* err, const std::string& msg) {
    if (err) *err = msg;
    return false;
}

}  // namespace

// ===================================================== block generator ====
bool generate_block_grid(const HexBlock& blk, const std::vector<PatchSpec>& patches, Grid& g,
                         std::string* err) {
    Int n[3] = {blk.n[0], blk.n[1], blk.n[2]};
    double s[12];
    int type[12];
    for (int i = 0; i < 12; i++) { s[i] = blk.s[i]; type[i] = blk.type[i]; }
    const V3* vp = blk.corners;
    for (int i = 0; i < 8; i++)
        for (int j = i + 1; j < 8; j++)
            if (equal_v(vp[i], vp[j]))
                return set_err(err, "degenerate (collapsed) blocks are not supported");

    // Wall spacing doubles odd division counts; the reference walks edges
    // j, j+3, j+6, j+9 here (hexMesh.cpp:57-72).
    for (int j = 0; j < 3; j++) {
        bool found = false;
        for (int i = j; i < 12; i += 3)
            if (type[i] == WALL && (n[j] % 2) && n[j] != 1) { found = true; break; }
        if (found) {
            n[j]++;
            for (int i = j; i < 12; i += 3) s[i] = 1 / s[i];
        }
    }
    // Edge parameter tables (hexMesh.cpp:74-107).
    std::vector<double> sc[12];
    for (int i = 0; i < 12; i++) {
        Int nt = n[i / 4];
        sc[i].assign(nt + 1, 0.0);
        if (type[i] == WALL) s[i] = std::pow(s[i], double(1. / (nt / 2.)));
        else s[i] = std::pow(s[i], double(1. / nt));
    }
    for (int i = 0; i < 12; i++) {
        Int nt = n[i / 4];
        double r = s[i];
        if (nt == 1) {
            sc[i][0] = 0;
            sc[i][1] = 1;
        } else {
            if (type[i] == WALL) nt /= 2;
            for (Int j = 0; j <= nt; j++) {
                if (equal_s(r, 1.0)) sc[i][j] = double(j) / (nt);
                else sc[i][j] = (1 - std::pow(r, double(j))) / (1 - std::pow(s[i], double(nt)));
            }
            if (type[i] == WALL) {
                for (Int j = 0; j <= nt; j++) sc[i][j] /= 2;
                for (Int j = 0; j <= nt; j++) sc[i][j + nt] = 1.0 - sc[i][nt - j];
            }
        }
    }
    const Int nx = n[0] + 1, ny = n[1] + 1, nz = n[2] + 1;
    const uint64_t B1 = uint64_t(nx) * (ny - 1) * (nz - 1);
    const uint64_t B2 = uint64_t(nx - 1) * ny * (nz - 1);
    const uint64_t B3 = uint64_t(nx - 1) * (ny - 1) * nz;
    IntVec VI(uint64_t(nx) * ny * nz, 0);
    IntVec FI(B1 + B2 + B3, 0);
    auto I0 = [&](Int i, Int j, Int k) { return uint64_t(i) * ny * nz + uint64_t(j) * nz + k; };
    auto I1 = [&](Int i, Int j, Int k) { return uint64_t(i) * (ny - 1) * (nz - 1) + uint64_t(j) * (nz - 1) + k; };
    auto I2 = [&](Int i, Int j, Int k) { return uint64_t(i) * ny * (nz - 1) + uint64_t(j) * (nz - 1) + k + B1; };
    auto I3 = [&](Int i, Int j, Int k) { return uint64_t(i) * (ny - 1) * nz + uint64_t(j) * nz + k + B1 + B2; };

    g = Grid();
    auto add_vertex = [&](Int i, Int j, Int k) {
        double m[12];
        for (int w = 0; w < 4; w++) m[w] = sc[w][i];
        for (int w = 4; w < 8; w++) m[w] = sc[w][j];
        for (int w = 8; w < 12; w++) m[w] = sc[w][k];
        double rx = i / double(nx - 1), ry = j / double(ny - 1), rz = k / double(nz - 1);
        g.vertices.push_back(hex_point(vp, m, rx, ry, rz));
        VI[I0(i, j, k)] = g.vertices.size() - 1;
    };
    // vertices: interior first, then the block boundary (hexMesh.cpp:178-216)
    for (Int j = 1; j + 1 < ny; j++)
        for (Int i = 1; i + 1 < nx; i++)
            for (Int k = 1; k + 1 < nz; k++) add_vertex(i, j, k);
    for (Int i = 0; i < nx; i += (nx - 1))
        for (Int j = 0; j < ny; j++)
            for (Int k = 0; k < nz; k++) add_vertex(i, j, k);
    for (Int j = 0; j < ny; j += (ny - 1))
        for (Int i = 1; i + 1 < nx; i++)
            for (Int k = 0; k < nz; k++) add_vertex(i, j, k);
    for (Int k = 0; k < nz; k += (nz - 1))
        for (Int i = 1; i + 1 < nx; i++)
            for (Int j = 1; j + 1 < ny; j++) add_vertex(i, j, k);

    // facets: interior then boundary, three orientations (hexMesh.cpp:228-303)
    auto face = [&](uint64_t a1, uint64_t a2, uint64_t a3, uint64_t a4) {
        g.facets.push_back(IntVec{VI[a1], VI[a2], VI[a3], VI[a4]});
        return Int(g.facets.size() - 1);
    };
    const uint64_t nyz = uint64_t(ny) * nz;
    for (Int i = 0; i + 1 < nx; i++)
        for (Int j = 0; j + 1 < ny; j++)
            for (Int k = 1; k + 1 < nz; k++) {
                uint64_t m = I0(i, j, k);
                FI[I3(i, j, k)] = face(m, m + nyz, m + nyz + nz, m + nz);
            }
    for (Int i = 0; i + 1 < nx; i++)
        for (Int j = 1; j + 1 < ny; j++)
            for (Int k = 0; k + 1 < nz; k++) {
                uint64_t m = I0(i, j, k);
                FI[I2(i, j, k)] = face(m, m + nyz, m + nyz + 1, m + 1);
            }
    for (Int i = 1; i + 1 < nx; i++)
        for (Int j = 0; j + 1 < ny; j++)
            for (Int k = 0; k + 1 < nz; k++) {
                uint64_t m = I0(i, j, k);
                FI[I1(i, j, k)] = face(m, m + nz, m + nz + 1, m + 1);
            }
    struct PatchR { Int from, to; V3 N, C; };
    std::vector<PatchR> prs;
    for (Int k = 0; k < nz; k += (nz - 1)) {
        PatchR p; p.from = g.facets.size();
        for (Int i = 0; i + 1 < nx; i++)
            for (Int j = 0; j + 1 < ny; j++) {
                uint64_t m = I0(i, j, k);
                FI[I3(i, j, k)] = face(m, m + nyz, m + nyz + nz, m + nz);
            }
        p.to = g.facets.size(); prs.push_back(p);
    }
    for (Int j = 0; j < ny; j += (ny - 1)) {
        PatchR p; p.from = g.facets.size();
        for (Int i = 0; i + 1 < nx; i++)
            for (Int k = 0; k + 1 < nz; k++) {
                uint64_t m = I0(i, j, k);
                FI[I2(i, j, k)] = face(m, m + nyz, m + nyz + 1, m + 1);
            }
        p.to = g.facets.size(); prs.push_back(p);
    }
    for (Int i = 0; i < nx; i += (nx - 1)) {
        PatchR p; p.from = g.facets.size();
        for (Int j = 0; j + 1 < ny; j++)
            for (Int k = 0; k + 1 < nz; k++) {
                uint64_t m = I0(i, j, k);
                FI[I1(i, j, k)] = face(m, m + nz, m + nz + 1, m + 1);
            }
        p.to = g.facets.size(); prs.push_back(p);
    }
    static const int kPatchCorners[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4},
                                            {3, 2, 6, 7}, {0, 3, 7, 4}, {1, 2, 6, 5}};
    for (int q = 0; q < 6; q++) {
        const int* c = kPatchCorners[q];
        V3 N = cross(sub(vp[c[1]], vp[c[0]]), sub(vp[c[2]], vp[c[0]]));
        prs[q].N = sdiv(N, mag(N));
        prs[q].C = sdiv(add(add(add(vp[c[0]], vp[c[1]]), vp[c[2]]), vp[c[3]]), 4.0);
    }
    // cells (hexMesh.cpp:323-343)
    for (Int i = 0; i + 1 < nx; i++)
        for (Int j = 0; j + 1 < ny; j++)
            for (Int k = 0; k + 1 < nz; k++) {
                IntVec c;
                uint64_t m = I3(i, j, k);
                c.push_back(FI[m]); c.push_back(FI[m + 1]);
                m = I2(i, j, k);
                c.push_back(FI[m]); c.push_back(FI[m + (nz - 1)]);
                m = I1(i, j, k);
                c.push_back(FI[m]); c.push_back(FI[m + uint64_t(ny - 1) * (nz - 1)]);
                g.cells.push_back(c);
            }
    // Named patches pick every block patch lying in their plane
    // (meshApp.cpp:218-241).
    for (const auto& ps : patches) {
        if (ps.points.size() < 3) return set_err(err, "patch needs 3 points");
        V3 N = cross(sub(ps.points[1], ps.points[0]), sub(ps.points[2], ps.points[0]));
        N = sdiv(N, mag(N));
        IntVec list;
        for (const auto& p : prs) {
            V3 H = unit(sub(p.C, ps.points[0]));
            double d1 = mag(cross(N, p.N));
            double d2 = std::sqrt(std::fabs(dot(N, H)));
            if (d1 <= 10e-4 && d2 <= 10e-4)
                for (Int k = p.from; k < p.to; k++) list.push_back(k);
        }
        if (!list.empty()) {
            IntVec& gB = g.boundaries[ps.name];
            if (std::find(gB.begin(), gB.end(), list[0]) == gB.end())
                gB.insert(gB.end(), list.begin(), list.end());
        }
    }
    return true;
}

// ============================================================ grid I/O ====
namespace {

// Token reader over the reference's text or binary streams (util.h:191-258).
struct TokenIn {
    std::ifstream is;
    bool bin = false;
    bool good() const { return (bool)is; }
    void read_u32(Int& v) {
        if (bin) is.read((char*)&v, 4);
        else { long long t; is >> t; v = Int(t); }
    }
    void read_f64(double& v) {
        if (bin) is.read((char*)&v, 8);
        else is >> v;
    }
    void read_sym() {  // one structural character
        if (bin) { unsigned char n; is.read((char*)&n, 1); char c; is.read(&c, n ? 1 : 0); }
        else { char c; is >> c; }
    }
    void read_str(std::string& s) {
        if (bin) {
            unsigned char n; is.read((char*)&n, 1);
            s.resize(n); is.read(&s[0], n);
        } else is >> s;
    }
    void read_ivec(IntVec& v) {
        Int n; read_u32(n); read_sym();
        v.resize(n);
        for (Int i = 0; i < n; i++) read_u32(v[i]);
        read_sym();
    }
};

struct TokenOut {
    std::ofstream os;
    bool bin = false;
    void u32(Int v) { if (bin) os.write((const char*)&v, 4); else os << v; }
    void f64(double v) { if (bin) os.write((const char*)&v, 8); else os << v; }
    void str(const std::string& s) {  // whitespace is dropped in binary files
        if (bin) {
            std::string t;
            for (char c : s) if (!std::isspace((unsigned char)c)) t += c;
            if (!t.empty()) { unsigned char n = t.size(); os.write((const char*)&n, 1); os.write(t.data(), n); }
        } else os << s;
    }
    void ivec(const IntVec& p) {  // util.h:151-164
        Int sz = p.size();
        u32(sz);
        str(sz >= 16 ? "\n{ " : "{ ");
        for (Int i = 0; i < sz; i++) {
            if (sz >= 16 && (i % 16) == 0) str("\n");
            u32(p[i]); str(" ");
        }
        str(sz >= 16 ? "\n}" : "}");
    }
};

bool ends_with(const std::string& s, const std::string& e) {
    return s.size() >= e.size() && s.compare(s.size() - e.size(), e.size(), e) == 0;
}

}  // namespace

bool read_grid(const std::string& path, Grid& g, std::string* err) {
    TokenIn in;
    in.bin = ends_with(path, ".bin");
    in.is.open(path, in.bin ? std::ios::in | std::ios::binary : std::ios::in);
    if (!in.is) return set_err(err, "cannot open grid " + path);
    g = Grid();
    Int n;
    in.read_u32(n); in.read_sym();
    g.vertices.resize(n);
    for (Int i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) in.read_f64(g.vertices[i][c]);
    in.read_sym();
    in.read_u32(n); in.read_sym();
    g.facets.resize(n);
    for (Int i = 0; i < n; i++) in.read_ivec(g.facets[i]);
    in.read_sym();
    in.read_u32(n); in.read_sym();
    g.cells.resize(n);
    for (Int i = 0; i < n; i++) in.read_ivec(g.cells[i]);
    in.read_sym();
    in.read_u32(n); in.read_sym();
    for (Int i = 0; i < n; i++) {
        std::string name;
        IntVec idx;
        in.read_str(name);
        in.read_ivec(idx);
        IntVec& b = g.boundaries[name];
        b.insert(b.begin(), idx.begin(), idx.end());
    }
    if (!in.good()) return set_err(err, "truncated grid " + path);
    return true;
}

bool write_grid(const std::string& path, const Grid& g, std::string* err) {
    TokenOut out;
    out.bin = ends_with(path, ".bin");
    out.os.open(path, out.bin ? std::ios::out | std::ios::binary : std::ios::out);
    if (!out.os) return set_err(err, "cannot write grid " + path);
    out.os.precision(12);
    out.u32(g.vertices.size()); out.str("\n{\n");
    for (const auto& v : g.vertices) {
        for (int c = 0; c < 3; c++) { out.f64(v[c]); out.str(" "); }
        out.str("\n");
    }
    out.str("}\n");
    out.os.precision(6);
    out.u32(g.facets.size()); out.str("\n{\n");
    for (const auto& f : g.facets) { out.ivec(f); out.str("\n"); }
    out.str("}\n");
    out.u32(g.cells.size()); out.str("\n{\n");
    for (const auto& c : g.cells) { out.ivec(c); out.str("\n"); }
    out.str("}\n");
    out.u32(g.boundaries.size()); out.str("\n{\n");
    for (const auto& kv : g.boundaries) {
        out.str(kv.first); out.str(" "); out.ivec(kv.second); out.str("\n");
    }
    out.str("}\n");
    return (bool)out.os || set_err(err, "write failed " + path);
}

// ============================================================ load mesh ===
namespace {

struct MeshWork {
    std::vector<V3>& V;
    std::vector<IntVec>& F;

    void unit_normal(const IntVec& f, V3& N) const {  // mesh.cpp:840-852
        const V3& v1 = V[f[0]];
        const V3& v2 = V[f[1]];
        for (size_t j = 1; j < f.size(); j++) {
            const V3& v3 = V[f[f.size() - j]];
            if (!point_in_line(v2, v1, v3)) {
                N = unit(cross(sub(v2, v1), sub(v3, v1)));
                return;
            }
        }
    }
    bool coplanar(const IntVec& f1, const IntVec& f2) const {  // mesh.cpp:874-887
        V3 N1, N2;
        unit_normal(f1, N1);
        unit_normal(f2, N2);
        V3 c = cross(N1, N2);
        double e = dot(c, c);
        if (equal_s(e, 0.0)) {
            V3 v = sub(V[f2[1]], V[f1[0]]);
            e = dot(N1, v);
            if (equal_s(e, 0.0)) return true;
        }
        return false;
    }
    // Eight corners from two opposite faces (mesh.cpp:113-157).
    void hex_corners(const IntVec& f1, const IntVec& f2, IntVec& vp) const {
        IntVec fm[2];
        for (int f = 0; f < 2; f++) {
            const IntVec& fk = f == 0 ? f1 : f2;
            size_t i0 = fk.size() - 1;
            for (size_t i = 0; i < fk.size(); i++) {
                size_t i1 = (i == fk.size() - 1) ? 0 : i + 1;
                V3 v0 = unit(sub(V[fk[i]], V[fk[i0]]));
                V3 v1 = unit(sub(V[fk[i1]], V[fk[i]]));
                double dt = dot(v0, v1);
                if (dt > 1) dt = 1;
                else if (dt < -1) dt = -1;
                double angle = std::acos(dt);
                const double tol = kPi / 16;
                if (!(angle < tol || angle >= kPi - tol)) {
                    fm[f].push_back(fk[i]);
                    i0 = i;
                }
            }
        }
        for (int i = 0; i < 4; i++) vp.push_back(fm[0][i]);
        double mind = 1e20;
        int order[4] = {0, 1, 2, 3}, best[4] = {0, 1, 2, 3};
        do {
            double dist = 0;
            for (int i = 0; i < 4; i++) dist += mag(sub(V[fm[1][order[i]]], V[fm[0][i]]));
            if (dist < mind) { mind = dist; std::copy(order, order + 4, best); }
        } while (std::next_permutation(order, order + 4));
        for (int i = 0; i < 4; i++) vp.push_back(fm[1][best[i]]);
    }
    // Union of two polygons that share at least one edge (mesh.cpp:890-980):
    // both are turned to the same orientation and rotated to their first
    // vertex the other does not hold; the union walks f1 up to the first
    // shared vertex, then f2's unshared run, then f1 from the last shared
    // vertex, and starts again at a corner (not an edge midpoint).
    bool merge_facets(const IntVec& a, const IntVec& b, IntVec& f) const {
        IntVec f1 = a, f2 = b;
        {
            V3 N1, N2;
            unit_normal(a, N1);
            unit_normal(b, N2);
            if (dot(N1, N2) < 0) std::reverse(f2.begin() + 1, f2.end());
        }
        auto first_unshared = [](const IntVec& x, const IntVec& y) {
            for (size_t i = 0; i < x.size(); i++)
                if (std::find(y.begin(), y.end(), x[i]) == y.end()) return int(i);
            return -1;
        };
        const int v1 = first_unshared(f1, f2), v2 = first_unshared(f2, f1);
        if (v1 < 0) return false;
        std::rotate(f1.begin(), f1.begin() + v1, f1.end());
        const bool contained = v2 < 0;
        if (!contained) std::rotate(f2.begin(), f2.begin() + v2, f2.end());
        size_t sa[2] = {0, 0}, sb[2] = {0, 0};
        int count = 0;
        for (size_t i = 0; i < f1.size(); i++)
            for (size_t j = 0; j < f2.size(); j++)
                if (f1[i] == f2[j]) {
                    const int slot = count ? 1 : 0;
                    sa[slot] = i;
                    sb[slot] = j;
                    count++;
                }
        if (count < 2) return false;
        f.clear();
        for (size_t i = 0; i <= sa[0]; i++) f.push_back(f1[i]);
        if (contained) {
            for (size_t i = sb[0] + 1; i < sb[1]; i++) f.push_back(f2[i]);
        } else {
            for (size_t i = sb[0] + 1; i < f2.size(); i++) f.push_back(f2[i]);
            for (size_t i = 0; i < sb[1]; i++) f.push_back(f2[i]);
        }
        for (size_t i = sa[1]; i < f1.size(); i++) f.push_back(f1[i]);
        for (size_t turn = 0; turn < f.size() && point_in_line(V[f[0]], V[f.back()], V[f[1]]); turn++)
            std::rotate(f.begin(), f.begin() + 1, f.end());
        return true;
    }
    // The merged outline of a group of coplanar faces (mesh.cpp:984-1016):
    // faces that share no edge with the outline yet are retried after the
    // others.  False when a pass merges nothing (the group is not connected).
    bool merge_group(const IntVec& faces, IntVec& fn) const {
        fn = F[faces[0]];
        IntVec rest(faces.begin() + 1, faces.end());
        while (!rest.empty()) {
            IntVec left;
            for (Int fi : rest) {
                IntVec fm;
                if (merge_facets(fn, F[fi], fm)) fn = fm;
                else left.push_back(fi);
            }
            if (left.size() == rest.size()) return false;
            rest.swap(left);
        }
        return true;
    }
};

}  // namespace

bool load_mesh(const Grid& g, LoadedMesh& m, std::string* err) {
    m = LoadedMesh();
    m.vertices = g.vertices;
    m.facets = g.facets;
    m.cells = g.cells;
    m.boundaries = g.boundaries;
    std::vector<V3>& V = m.vertices;
    std::vector<IntVec>& F = m.facets;
    std::vector<IntVec>& C = m.cells;

    PhaseTimer tm;
    // ---- addBoundaryCells (mesh.cpp:55-109)
    m.nbcs = C.size();
    m.FOC.assign(F.size(), kMaxInt);
    m.FNC.assign(F.size(), kMaxInt);
    m.has_boundary.assign(C.size(), 0);
    for (size_t i = 0; i < C.size(); i++)
        for (Int fi : C[i]) {
            if (fi >= F.size()) return set_err(err, "cell references a missing facet");
            if (m.FOC[fi] == kMaxInt) m.FOC[fi] = i;
            else m.FNC[fi] = i;
        }
    IntVec faceInB(F.size(), 0);
    for (auto& kv : m.boundaries) {
        if (kv.first == "delete") continue;
        for (Int f : kv.second) faceInB[f] = 1;
    }
    {
        IntVec& del = m.boundaries["delete"];
        del.clear();
        for (size_t i = 0; i < F.size(); i++)
            if (m.FNC[i] == kMaxInt && !faceInB[i]) del.push_back(i);
    }
    for (auto& kv : m.boundaries)
        for (Int fi : kv.second)
            if (m.FNC[fi] == kMaxInt) {
                C.push_back(IntVec{fi});
                m.FNC[fi] = C.size() - 1;
                if (faceInB[fi]) m.has_boundary[m.FOC[fi]] = 1;
            }

    tm.mark("addBoundaryCells");
    // ---- fixHexCells, conforming hexahedra (mesh.cpp:161-446)
    MeshWork w{V, F};
    m.FMC.assign(F.size(), 0);
    m.faceid.clear();
    for (size_t cidx = 0; cidx < C.size(); cidx++) {
        IntVec& c = C[cidx];
        if (c.size() == 1) { m.faceid.push_back(IntVec{0}); continue; }
        if (c.size() < 6) return set_err(err, "cells must be hexahedra");
        int groupId[6];
        std::vector<IntVec> cng, fng;
        for (int i = 0; i < 6; i++) {
            IntVec indices{0}, ci{c[0]};
            for (size_t j = 1; j < c.size(); j++)
                if (w.coplanar(F[c[0]], F[c[j]])) { ci.push_back(c[j]); indices.push_back(j); }
            erase_indices(c, indices);
            // a side split into several faces (a 2:1 non-conforming side)
            // is represented by the outline of its faces
            IntVec outline;
            if (!w.merge_group(ci, outline)) return set_err(err, "could not merge the faces of a cell side");
            cng.push_back(ci);
            fng.push_back(outline);
            groupId[i] = -1;
        }
        const IntVec f0 = fng[0];
        for (int j = 0; j < 6; j++) {
            if (groupId[j] >= 0) continue;
            const IntVec& fj = fng[j];
            int id = j;
            if (j >= 1) {
                int local_id = -1;
                int count = 0;
                for (Int v : fj) if (v == f0[0] || v == f0[1]) count++;
                if (count >= 2) local_id = 2;
                else {
                    count = 0;
                    for (Int v : fj) if (v == f0[0] || v == f0[f0.size() - 1]) count++;
                    if (count >= 2) local_id = 4;
                }
                if (local_id < 0) continue;
                id = local_id;
            }
            groupId[j] = id;
            for (int k = 0; k < 6; k++) {
                if (groupId[k] >= 0) continue;
                const IntVec& fk = fng[k];
                bool shared = false;
                for (Int a : fj) for (Int b : fk) if (a == b) shared = true;
                if (!shared) { groupId[k] = groupId[j] ^ 1; break; }
            }
        }
        {
            V3 N = cross(sub(V[f0[1]], V[f0[0]]), sub(V[f0[f0.size() - 1]], V[f0[0]]));
            V3 e = sub(V[fng[1][0]], V[f0[0]]);
            if (dot(N, e) < 0)
                for (int i = 0; i < 6; i++) {
                    if (groupId[i] == 0) groupId[i] = 1;
                    else if (groupId[i] == 1) groupId[i] = 0;
                }
        }
        int i0 = -1, i1 = -1;
        for (int i = 0; i < 6; i++) {
            if (groupId[i] == 0) i0 = i;
            else if (groupId[i] == 1) i1 = i;
        }
        if (i0 < 0 || i1 < 0) return set_err(err, "could not pair opposite hex faces");
        for (int i = 0; i < 6; i++) if (groupId[i] < 0) return set_err(err, "could not group hex faces");
        Int i0n = (m.FNC[cng[i0][0]] != cidx) ? m.FNC[cng[i0][0]] : m.FOC[cng[i0][0]];
        Int i1n = (m.FNC[cng[i1][0]] != cidx) ? m.FNC[cng[i1][0]] : m.FOC[cng[i1][0]];
        bool flip = (i0n > i1n && (cidx >= i0n || cidx >= i1n));
        IntVec vp;
        if (!flip) w.hex_corners(fng[i0], fng[i1], vp);
        else {
            w.hex_corners(fng[i1], fng[i0], vp);
            for (int i = 0; i < 4; i++) std::swap(vp[i], vp[i + 4]);
        }
        const Int rots[6] = {vp[0], vp[4], vp[0], vp[3], vp[0], vp[1]};
        const Int rote[6] = {vp[1], vp[5], vp[1], vp[2], vp[3], vp[2]};
        for (int i = 0; i < 6; i++) {
            IntVec& fn = fng[i];
            const IntVec& cn = cng[i];
            int gid = groupId[i];
            Int rs = rots[gid], re = rote[gid];
            std::rotate(fn.begin(), std::find(fn.begin(), fn.end(), rs), fn.end());
            double dir = dot(unit(sub(V[fn[1]], V[fn[0]])), unit(sub(V[re], V[rs])));
            if (dir < 0.99) std::reverse(fn.begin() + 1, fn.end());
            for (size_t j = 0; j < cn.size(); j++) {
                IntVec& f = F[cn[j]];
                auto it1 = std::find_first_of(fn.begin(), fn.end(), f.begin(), f.end());
                auto it2 = std::find(f.begin(), f.end(), *it1);
                std::rotate(f.begin(), it2, f.end());
                if (j >= 2) std::rotate(f.rbegin(), f.rbegin() + (j - 1), f.rend());
                for (size_t k = 1; k < f.size(); k++) {
                    auto it3 = std::find(fn.begin(), fn.end(), f[k]);
                    if (it3 == fn.end()) continue;
                    if (it1 > it3) { std::reverse(f.begin() + 1, f.end()); break; }
                    it1 = it3;
                }
            }
        }
        IntVec b;
        for (int idx = 0; idx < 6; idx++)
            for (int i = 0; i < 6; i++) {
                if (groupId[i] != idx) continue;
                for (Int fc : cng[i]) { c.push_back(fc); b.push_back(idx); }
            }
        m.faceid.push_back(b);
        // mortar flags (mesh.cpp:433-444): a face that shares its side of
        // this cell with others is a mortar face; 2 when this cell owns it
        if (c.size() > 6)
            for (size_t j = 0; j < c.size(); j++) {
                const Int f = c[j];
                if (std::count(b.begin(), b.end(), b[j]) > 1) m.FMC[f] = (m.FOC[f] == Int(cidx)) ? 2 : 1;
            }
    }

    tm.mark("fixHexCells");
    // ---- calcGeometry (mesh.cpp:450-577), planar case
    m.fc.assign(F.size(), V3());
    m.cc.assign(C.size(), V3());
    m.fn.assign(F.size(), V3());
    m.cv.assign(C.size(), 0.0);
    for (size_t i = 0; i < F.size(); i++) {
        V3 s;
        for (Int v : F[i]) s = add(s, V[v]);
        m.fc[i] = sdiv(s, double(F[i].size()));
    }
    for (size_t i = 0; i < C.size(); i++) {
        V3 s;
        for (Int f : C[i]) s = add(s, m.fc[f]);
        m.cc[i] = sdiv(s, double(C[i].size()));
    }
    for (size_t i = 0; i < F.size(); i++) {
        const IntVec& f = F[i];
        V3 N, Cs;
        double Ntot = 0;
        const V3 v1 = m.fc[i];
        for (size_t j = 0; j < f.size(); j++) {
            const V3& v2 = V[f[j]];
            const V3& v3 = V[f[(j + 1 == f.size()) ? 0 : (j + 1)]];
            V3 Ni = cross(sub(v2, v1), sub(v3, v1));
            double magN = mag(Ni);
            V3 Ci = smul(magN, sdiv(add(add(v1, v2), v3), 3.0));
            Cs = add(Cs, Ci);
            Ntot += magN;
            N = add(N, Ni);
        }
        m.fc[i] = sdiv(Cs, Ntot);
        V3 v = sub(m.fc[i], m.cc[m.FOC[i]]);
        if (dot(v, N) < 0) N = V3(-N[0], -N[1], -N[2]);
        m.fn[i] = sdiv(N, 2.0);
    }
    for (Int i = 0; i < m.nbcs; i++) {
        double Vol = 0;
        V3 Cs;
        for (Int f : C[i]) {
            V3 v = sub(m.cc[i], m.fc[f]);
            double Vi = std::fabs(dot(v, m.fn[f]));
            Cs = add(Cs, sdiv(smul(Vi, add(smul(3, m.fc[f]), m.cc[i])), 4.0));
            Vol += Vi;
        }
        m.cc[i] = sdiv(Cs, Vol);
        m.cv[i] = Vol / 3.0;
    }
    for (size_t i = m.nbcs; i < C.size(); i++) {
        Int fi = C[i][0];
        m.cv[i] = m.cv[m.FOC[fi]];
        m.cc[i] = m.fc[fi];
    }

    tm.mark("calcGeometry");
    // ---- removeBoundary("delete") (mesh.cpp:581-669)
    {
        IntVec fs = m.boundaries["delete"];
        IntVec Idf(F.size(), 0), Idc(C.size(), 0);
        for (Int f : fs) {
            for (Int cell : {m.FOC[f], m.FNC[f]}) {
                IntVec& co = C[cell];
                IntVec& coid = m.faceid[cell];
                for (size_t j = 0; j < co.size(); j++)
                    if (co[j] == f) { co.erase(co.begin() + j); coid.erase(coid.begin() + j); break; }
            }
        }
        for (Int f : fs) Idf[f] = kMaxInt;
        Int count = 0;
        for (size_t i = 0; i < F.size(); i++) {
            if (Idf[i] != kMaxInt) Idf[i] = count++;
            else F[i].clear();
        }
        IntVec fz;
        for (size_t i = 0; i < F.size(); i++) if (F[i].empty()) fz.push_back(i);
        erase_indices(F, fz); erase_indices(m.FOC, fz); erase_indices(m.FNC, fz);
        erase_indices(m.FMC, fz); erase_indices(m.fc, fz); erase_indices(m.fn, fz);
        count = 0;
        for (size_t i = 0; i < C.size(); i++) Idc[i] = C[i].empty() ? kMaxInt : count++;
        IntVec cz;
        for (size_t i = 0; i < C.size(); i++) if (C[i].empty()) cz.push_back(i);
        erase_indices(C, cz); erase_indices(m.faceid, cz);
        erase_indices(m.cc, cz); erase_indices(m.cv, cz);
        for (auto& c : C) for (auto& f : c) f = Idf[f];
        for (size_t i = 0; i < F.size(); i++) {
            m.FOC[i] = Idc[m.FOC[i]];
            m.FNC[i] = Idc[m.FNC[i]];
            m.FMC[i] = Idc[m.FMC[i]];
        }
        for (auto& kv : m.boundaries) for (auto& f : kv.second) f = Idf[f];
        m.boundaries.erase("delete");
        // has_boundary is indexed by interior cell; interior cells are never
        // removed here, so it keeps its meaning.
    }
    tm.mark("removeBoundary");
    // ---- removeUnusedVertices (mesh.cpp:674-721)
    {
        IntVec used(V.size(), 0);
        for (const auto& f : F) {
            V3 v1 = V[f[f.size() - 1]];
            for (size_t j = 0; j < f.size(); j++) {
                V3 v2 = V[f[j]];
                V3 v3 = V[f[(j + 1 == f.size()) ? 0 : (j + 1)]];
                if (!point_in_line(v2, v1, v3)) used[f[j]] = 1;
                v1 = v2;
            }
        }
        for (auto& f : F) {
            IntVec rem;
            for (size_t j = 0; j < f.size(); j++) if (!used[f[j]]) rem.push_back(j);
            if (!rem.empty()) erase_indices(f, rem);
        }
        IntVec rv;
        Int cnt = 0;
        for (size_t i = 0; i < used.size(); i++) {
            if (!used[i]) rv.push_back(i);
            else used[i] = cnt++;
        }
        if (!rv.empty()) {
            for (auto& f : F) for (auto& v : f) v = used[v];
            erase_indices(V, rv);
        }
    }
    tm.mark("removeUnusedVertices");
    // ---- drop empty and "interior" patches (field.cpp:148-155)
    for (auto it = m.boundaries.begin(); it != m.boundaries.end();) {
        if (it->second.empty() || it->first.find("interior") != std::string::npos)
            it = m.boundaries.erase(it);
        else ++it;
    }
    return true;
}

// ===================================================== basis/geometry ====
// Legendre polynomial and first two derivatives by recurrence (dg.cpp:32-47).
static void legendre(int p, double x, double& L0, double& L0_1, double& L0_2) {
    double L1 = 0, L1_1 = 0, L1_2 = 0;
    L0 = 1; L0_1 = 0; L0_2 = 0;
    for (int i = 1; i <= p; i++) {
        double L2 = L1, L2_1 = L1_1, L2_2 = L1_2;
        L1 = L0; L1_1 = L0_1; L1_2 = L0_2;
        double a = (2 * i - 1.0) / i;
        double b = (i - 1.0) / i;
        L0 = a * x * L1 - b * L2;
        L0_1 = a * (L1 + x * L1_1) - b * L2_1;
        L0_2 = a * (2 * L1_1 + x * L1_2) - b * L2_2;
    }
}

// LGL nodes by Newton iteration on (1-x^2)P'(x) from Chebyshev guesses
// (dg.cpp:52-99).
void lgl_nodes(int N, double* x, double* w) {
    double L0, L0_1, L0_2;
    int p = N - 1, ph = N / 2;
    if (N == 1) { x[0] = 0; w[0] = 2; return; }
    for (int i = 0; i < ph; i++) {
        double xi = std::cos((2 * i + 1) * kPi / (2 * N));
        for (int k = 1; k <= 20; k++) {
            legendre(p, xi, L0, L0_1, L0_2);
            double dx = -(1 - xi * xi) * L0_1 / (-2 * xi * L0_1 + (1 - xi * xi) * L0_2);
            xi += dx;
            if (std::fabs(dx) < 1.0e-20) break;
        }
        x[p - i] = xi;
        w[p - i] = 2 / (p * (p + 1) * L0 * L0);
    }
    if (N != 2 * ph) {
        legendre(p, 0.0, L0, L0_1, L0_2);
        x[ph] = 0;
        w[ph] = 2 / (p * (p + 1) * L0 * L0);
    }
    for (int i = 0; i < ph; i++) { x[i] = -x[p - i]; w[i] = +w[p - i]; }
}

// Lagrange basis of the n nodes x at the ns points xs: out[j*ns+s] = l_j(xs[s]).
static void lagrange_at(int n, const double* x, int ns, const double* xs, double* out) {
    for (int s = 0; s < ns; s++)
        for (int j = 0; j < n; j++) {
            double prod = 1;
            for (int k = 0; k < n; k++)
                if (k != j) prod *= ((xs[s] - x[k]) / (x[j] - x[k]));
            out[j * ns + s] = prod;
        }
}

// Gauss-Jordan inverse of an n x n matrix after one bubble pass that moves
// larger first-column entries up (tensor.cpp:170-222; the same operation
// order, so the mortar matrices agree bit for bit).
static void gauss_jordan_inverse(const std::vector<double>& a_in, std::vector<double>& x, int n) {
    std::vector<double> a = a_in;
    x.assign(size_t(n) * n, 0.0);
    for (int i = 0; i < n; i++) x[i * n + i] = 1.0;
    for (int i = n - 1; i > 0; i--)
        if (a[(i - 1) * n] < a[i * n])
            for (int k = 0; k < n; k++) {
                std::swap(a[i * n + k], a[(i - 1) * n + k]);
                std::swap(x[i * n + k], x[(i - 1) * n + k]);
            }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            if (j == i) continue;
            const double t = a[j * n + i] / a[i * n + i];
            for (int k = 0; k < n; k++) {
                a[j * n + k] -= a[i * n + k] * t;
                x[j * n + k] -= x[i * n + k] * t;
            }
        }
    for (int i = 0; i < n; i++) {
        const double t = a[i * n + i];
        for (int j = 0; j < n; j++) {
            a[i * n + j] = a[i * n + j] / t;
            x[i * n + j] = x[i * n + j] / t;
        }
    }
}

// 2:1 mortar matrices (dg.cpp:516-583).  With an (n+1)-point LGL rule on
// the parent interval, M = (phi_j, phi_k), S_c = (phi_j, phi_k(half c)) and
// its transpose G_c; ref[c] = (M^-1 S_c)^T, cor[c] = (M^-1 G_c)^T.
static void build_mortar(Basis1D& b) {
    const int n = b.n, ne = n + 1;
    std::vector<double> xe(ne), we(ne), xr(2 * ne, 0.0), xc(2 * ne, 0.0);
    lgl_nodes(ne, xe.data(), we.data());
    if (n != 1)
        for (int j = 0; j < ne; j++) {
            xr[j] = -0.5 + xe[j] / 2;
            xr[j + ne] = 0.5 + xe[j] / 2;
            xc[j] = 2 * xe[j] + 1;
            xc[j + ne] = 2 * xe[j] - 1;
        }
    std::vector<double> pr(2 * ne * ne, 0.0), pe(size_t(n) * ne);
    for (int c = 0; c < 2; c++) lagrange_at(n, b.x.data(), ne, &xr[c * ne], &pr[c * ne * ne]);
    lagrange_at(n, b.x.data(), ne, xe.data(), pe.data());
    std::vector<double> M(n * n, 0.0), S(2 * n * n, 0.0), G(2 * n * n, 0.0);
    for (int j = 0; j < n; j++)
        for (int k = 0; k < n; k++)
            for (int q = 0; q < ne; q++) {
                M[j * n + k] += (we[q] / 2) * pe[j * ne + q] * pe[k * ne + q];
                for (int c = 0; c < 2; c++) {
                    const double vs = (we[q] / 2) * pe[j * ne + q] * pr[c * ne * ne + k * ne + q];
                    S[c * n * n + j * n + k] += vs;
                    G[c * n * n + k * n + j] += vs;
                }
            }
    std::vector<double> Mi;
    gauss_jordan_inverse(M, Mi, n);
    auto product_transposed = [&](const double* B, std::vector<double>& out) {
        std::vector<double> t(n * n);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                double sum = 0.0;
                for (int k = 0; k < n; k++) sum += Mi[i * n + k] * B[k * n + j];
                t[i * n + j] = sum;
            }
        out.assign(n * n, 0.0);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) out[i * n + j] = t[j * n + i];
    };
    for (int c = 0; c < 2; c++) {
        product_transposed(&S[c * n * n], b.ref[c]);
        product_transposed(&G[c * n * n], b.cor[c]);
    }
}

void build_basis(int n, Basis1D& b) {
    b.n = n;
    b.x.assign(n, 0); b.w.assign(n, 0);
    b.psi.assign(n * n, 0); b.dpsi.assign(n * n, 0);
    lgl_nodes(n, b.x.data(), b.w.data());
    // psi[j*n+s] = l_j(x_s)  (dg.cpp:103-119)
    for (int s = 0; s < n; s++)
        for (int j = 0; j < n; j++) {
            double prod = 1;
            for (int k = 0; k < n; k++)
                if (k != j) prod *= ((b.x[s] - b.x[k]) / (b.x[j] - b.x[k]));
            b.psi[j * n + s] = prod;
        }
    // dpsi[s*n+i] = l'_i(x_s)  (dg.cpp:122-143)
    for (int s = 0; s < n; s++)
        for (int i = 0; i < n; i++) {
            double acc = 0;
            for (int j = 0; j < n; j++) {
                if (i == j) continue;
                double prod = 1;
                for (int k = 0; k < n; k++)
                    if (k != i && k != j) prod *= ((b.x[s] - b.x[k]) / (b.x[i] - b.x[k]));
                acc += prod / (b.x[i] - b.x[j]);
            }
            b.dpsi[s * n + i] = acc;
        }
    build_mortar(b);
}

namespace {
inline T9 outer(const V3& p, const V3& q) {  // tensor.cpp mul(Vector,Vector)
    T9 r;
    r[XX] = p[0] * q[0]; r[YY] = p[1] * q[1]; r[ZZ] = p[2] * q[2];
    r[XY] = p[0] * q[1]; r[YZ] = p[1] * q[2]; r[XZ] = p[0] * q[2];
    r[YX] = p[1] * q[0]; r[ZY] = p[2] * q[1]; r[ZX] = p[2] * q[0];
    return r;
}
inline T9 tmul(const T9& p, const T9& q) {  // tensor.cpp mul(Tensor,Tensor)
    T9 r;
    r[XX] = p[XX] * q[XX] + p[XY] * q[YX] + p[XZ] * q[ZX];
    r[XY] = p[XX] * q[XY] + p[XY] * q[YY] + p[XZ] * q[ZY];
    r[XZ] = p[XX] * q[XZ] + p[XY] * q[YZ] + p[XZ] * q[ZZ];
    r[YX] = p[YX] * q[XX] + p[YY] * q[YX] + p[YZ] * q[ZX];
    r[YY] = p[YX] * q[XY] + p[YY] * q[YY] + p[YZ] * q[ZY];
    r[YZ] = p[YX] * q[XZ] + p[YY] * q[YZ] + p[YZ] * q[ZZ];
    r[ZX] = p[ZX] * q[XX] + p[ZY] * q[YX] + p[ZZ] * q[ZX];
    r[ZY] = p[ZX] * q[XY] + p[ZY] * q[YY] + p[ZZ] * q[ZY];
    r[ZZ] = p[ZX] * q[XZ] + p[ZY] * q[YZ] + p[ZZ] * q[ZZ];
    return r;
}
inline T9 trn(const T9& p) {
    T9 r = p;
    r[XY] = p[YX]; r[YZ] = p[ZY]; r[XZ] = p[ZX];
    r[YX] = p[XY]; r[ZY] = p[YZ]; r[ZX] = p[XZ];
    return r;
}
inline T9 inv(const T9& p) {  // adjugate / determinant (tensor.cpp)
    T9 r;
    r[XX] = p[YY] * p[ZZ] - p[YZ] * p[ZY];
    r[YY] = p[XX] * p[ZZ] - p[XZ] * p[ZX];
    r[ZZ] = p[XX] * p[YY] - p[XY] * p[YX];
    r[XY] = p[XZ] * p[ZY] - p[XY] * p[ZZ];
    r[XZ] = p[XY] * p[YZ] - p[XZ] * p[YY];
    r[YX] = p[YZ] * p[ZX] - p[YX] * p[ZZ];
    r[YZ] = p[XZ] * p[YX] - p[XX] * p[YZ];
    r[ZX] = p[YX] * p[ZY] - p[YY] * p[ZX];
    r[ZY] = p[XY] * p[ZX] - p[XX] * p[ZY];
    double d = p[XX] * r[XX] + p[XY] * r[YX] + p[XZ] * r[ZX];
    if (d == 0) return T9();
    for (int i = 0; i < 9; i++) r[i] /= d;
    return r;
}
inline V3 tdot(const T9& q, const V3& p) {  // tensor.cpp dot(Tensor,Vector)
    return V3(q[XX] * p[0] + q[XY] * p[1] + q[XZ] * p[2],
              q[YX] * p[0] + q[YY] * p[1] + q[YZ] * p[2],
              q[ZX] * p[0] + q[ZY] * p[1] + q[ZZ] * p[2]);
}
}  // namespace

bool build_geometry(const LoadedMesh& m, const int order[3], int nonortho, Geometry& geo,
                    std::string* err) {
    geo = Geometry();
    const Int NPX = order[0] + 1, NPY = order[1] + 1, NPZ = order[2] + 1;
    const Int NP = NPX * NPY * NPZ;
    Int NPF;  // dg.cpp:147-163
    if (NPX <= NPY && NPX <= NPZ) NPF = NPY * NPZ;
    else if (NPY <= NPX && NPY <= NPZ) NPF = NPX * NPZ;
    else NPF = NPX * NPY;
    const bool dg = NP > 1;
    geo.npx = NPX; geo.npy = NPY; geo.npz = NPZ; geo.np = NP; geo.npf = NPF;
    geo.ncells = m.cells.size();
    geo.nbcs = m.nbcs;
    geo.nfacets = m.facets.size();
    geo.nbcs_field = uint64_t(m.nbcs) * NP;
    geo.nall_field = uint64_t(geo.ncells) * NP;
    if (geo.nall_field >= (uint64_t(1) << 32) - 1 || uint64_t(geo.nfacets) * NPF >= (uint64_t(1) << 32))
        return set_err(err, "mesh too large for 32-bit node indices");
    const Int gALL = Int(geo.nall_field), gBCSf = Int(geo.nbcs_field);
    for (int d = 0; d < 3; d++) build_basis(order[d] + 1, geo.basis[d]);
    PhaseTimer tm;

    // cell -> face CSR (field.cpp:178-193)
    geo.faceIndices0.resize(geo.ncells);
    geo.faceIndices1.resize(geo.ncells);
    for (Int i = 0; i < geo.ncells; i++) {
        geo.faceIndices0[i] = geo.allFaces.size();
        for (Int f : m.cells[i]) geo.allFaces.push_back(f);
        geo.faceIndices1[i] = geo.allFaces.size();
    }
    // expanded cell/facet quantities (field.cpp:213-230, dg.h:98-112)
    const uint64_t nfn = uint64_t(geo.nfacets) * NPF;
    geo.cC.resize(geo.nall_field);
    geo.cV.resize(geo.nall_field);
    for (Int c = 0; c < geo.ncells; c++)
        for (Int j = 0; j < NP; j++) {
            geo.cC[uint64_t(c) * NP + j] = m.cc[c];
            geo.cV[uint64_t(c) * NP + j] = m.cv[c];
        }
    geo.fC.resize(nfn); geo.fN.resize(nfn);
    geo.FO.resize(nfn); geo.FN.resize(nfn);
    for (Int f = 0; f < geo.nfacets; f++)
        for (Int j = 0; j < NPF; j++) {
            geo.fC[uint64_t(f) * NPF + j] = m.fc[f];
            geo.fN[uint64_t(f) * NPF + j] = m.fn[f];
            geo.FO[uint64_t(f) * NPF + j] = m.FOC[f];
            geo.FN[uint64_t(f) * NPF + j] = m.FNC[f];
        }
    if (dg) {
        const Basis1D& b0 = geo.basis[0];
        const Basis1D& b1 = geo.basis[1];
        const Basis1D& b2 = geo.basis[2];
        auto INDEX4 = [&](uint64_t c, Int i, Int j, Int k) {
            return c * NP + uint64_t(i) * NPY * NPZ + uint64_t(j) * NPZ + k;
        };
        std::fill(geo.FO.begin(), geo.FO.end(), gALL);
        std::fill(geo.FN.begin(), geo.FN.end(), gALL);
        MeshWork w{const_cast<std::vector<V3>&>(m.vertices), const_cast<std::vector<IntVec>&>(m.facets)};
        // node coordinates by transfinite interpolation (dg.cpp:175-325)
        for (Int ci = 0; ci < m.nbcs; ci++) {
            const IntVec& c = m.cells[ci];
            const IntVec& fids = m.faceid[ci];
            if (c.size() < 4 || fids.size() != c.size())
                return set_err(err, "a cell lost its faces (every boundary face unnamed?)");
        }
        // the first side of each cell and its opposite, as outlines when
        // split (dg.cpp:180-201: merged face by face in cell order)
        std::vector<IntVec> side0(m.nbcs), side1(m.nbcs);
        for (Int ci = 0; ci < m.nbcs; ci++) {
            const IntVec& c = m.cells[ci];
            const IntVec& fids = m.faceid[ci];
            const Int id0 = fids[0], id1 = id0 ^ 1;
            for (size_t i = 0; i < c.size(); i++) {
                IntVec* side = fids[i] == id0 ? &side0[ci] : fids[i] == id1 ? &side1[ci] : nullptr;
                if (!side) continue;
                if (side->empty()) { *side = m.facets[c[i]]; continue; }
                IntVec fm;
                if (!w.merge_facets(*side, m.facets[c[i]], fm))
                    return set_err(err, "could not merge the faces of a cell side");
                *side = fm;
            }
            if (side0[ci].empty() || side1[ci].empty()) return set_err(err, "a cell has no opposite sides");
        }
        parallel_for(m.nbcs, [&](uint64_t cb, uint64_t ce) {
        for (Int ci = (Int)cb; ci < (Int)ce; ci++) {
            const IntVec& c = m.cells[ci];
            const IntVec& fids = m.faceid[ci];
            const Int id0 = fids[0];
            (void)c;
            IntVec vpi;
            w.hex_corners(side0[ci], side1[ci], vpi);
            V3 vp1[8], vp[8];
            for (int i = 0; i < 8; i++) vp1[i] = m.vertices[vpi[i]];
            if (id0 == 2) {
                const int ord[8] = {0, 1, 5, 4, 3, 2, 6, 7};
                for (int i = 0; i < 8; i++) vp[ord[i]] = vp1[i];
            } else if (id0 == 4) {
                const int ord[8] = {0, 3, 7, 4, 1, 2, 6, 5};
                for (int i = 0; i < 8; i++) vp[ord[i]] = vp1[i];
            } else {
                for (int i = 0; i < 8; i++) vp[i] = vp1[i];
            }
            for (Int i = 0; i < NPX; i++)
                for (Int j = 0; j < NPY; j++)
                    for (Int k = 0; k < NPZ; k++) {
                        double rx = (b0.x[i] + 1) / 2, ry = (b1.x[j] + 1) / 2, rz = (b2.x[k] + 1) / 2;
                        double mm[12];
                        for (int q = 0; q < 4; q++) mm[q] = rx;
                        for (int q = 4; q < 8; q++) mm[q] = ry;
                        for (int q = 8; q < 12; q++) mm[q] = rz;
                        V3 v = hex_point(vp, mm, rx, ry, rz);
                        double wgt = b0.w[i] * b1.w[j] * b2.w[k] / 8;
                        uint64_t index = INDEX4(ci, i, j, k);
                        geo.cC[index] = v;
                        geo.cV[index] *= wgt;
                    }
        }
        });
        tm.mark("node coordinates");
        // face-node maps (dg.cpp:327-410)
        const Int face_map[6] = {0, NPZ - 1, 0, NPY - 1, 0, NPX - 1};
        for (Int ci = 0; ci < m.nbcs; ci++) {
            const IntVec& c = m.cells[ci];
            const IntVec& face_ids = m.faceid[ci];
            for (size_t mm = 0; mm < c.size(); mm++) {
                Int face_o = face_ids[mm];
                Int fi = c[mm];
                Int cj = m.FNC[fi];
                Int fm = m.FMC[fi];
                if (cj == ci) continue;
                Int face_n = face_o ^ 1;
                if (cj < m.nbcs) {
                    const IntVec& cn = m.cells[cj];
                    for (size_t i = 0; i < cn.size(); i++)
                        if (cn[i] == fi) { face_n = m.faceid[cj][i]; break; }
                }
                Int vo = face_map[face_o], vn = face_map[face_n];
                auto put = [&](uint64_t indf, uint64_t index0, uint64_t index1, double wgt) {
                    geo.FO[indf] = Int(index0);
                    geo.FN[indf] = Int(index1);
                    if (index1 >= gBCSf) {
                        geo.cC[index1] = geo.cC[index0];
                        geo.cV[index1] = geo.cV[index0];
                    }
                    geo.fC[indf] = (fm <= 1) ? geo.cC[index0] : geo.cC[index1];
                    for (int q = 0; q < 3; q++) geo.fN[indf][q] *= wgt;
                };
                if (face_o == 0 || face_o == 1) {
                    for (Int i = 0; i < NPX; i++)
                        for (Int j = 0; j < NPY; j++) {
                            double wgt = b0.w[i] * b1.w[j] / 4;
                            uint64_t indf = uint64_t(fi) * NPF + i * NPY + j;
                            uint64_t index1;
                            if (face_n == 0 || face_n == 1) index1 = INDEX4(cj, i, j, vn);
                            else if (face_n == 2 || face_n == 3) index1 = INDEX4(cj, i, vn, j);
                            else index1 = INDEX4(cj, vn, i, j);
                            put(indf, INDEX4(ci, i, j, vo), index1, wgt);
                        }
                } else if (face_o == 2 || face_o == 3) {
                    for (Int i = 0; i < NPX; i++)
                        for (Int k = 0; k < NPZ; k++) {
                            double wgt = b0.w[i] * b2.w[k] / 4;
                            uint64_t indf = uint64_t(fi) * NPF + i * NPZ + k;
                            uint64_t index1;
                            if (face_n == 0 || face_n == 1) index1 = INDEX4(cj, i, k, vn);
                            else if (face_n == 2 || face_n == 3) index1 = INDEX4(cj, i, vn, k);
                            else index1 = INDEX4(cj, vn, i, k);
                            put(indf, INDEX4(ci, i, vo, k), index1, wgt);
                        }
                } else {
                    for (Int j = 0; j < NPY; j++)
                        for (Int k = 0; k < NPZ; k++) {
                            double wgt = b1.w[j] * b2.w[k] / 4;
                            uint64_t indf = uint64_t(fi) * NPF + j * NPZ + k;
                            uint64_t index1;
                            if (face_n == 0 || face_n == 1) index1 = INDEX4(cj, j, k, vn);
                            else if (face_n == 2 || face_n == 3) index1 = INDEX4(cj, j, vn, k);
                            else index1 = INDEX4(cj, vn, j, k);
                            put(indf, INDEX4(ci, vo, j, k), index1, wgt);
                        }
                }
            }
        }
        tm.mark("face-node maps");
        // inverse-transposed Jacobian, pseudo-inverse for collapsed
        // directions (dg.cpp:412-476)
        geo.Jinv.assign(geo.nall_field, T9());
        tm.mark("Jacobian storage");
        parallel_for(m.nbcs, [&](uint64_t cb, uint64_t ce) {
        for (Int ci = (Int)cb; ci < (Int)ce; ci++)
            for (Int ii = 0; ii < NPX; ii++)
                for (Int jj = 0; jj < NPY; jj++)
                    for (Int kk = 0; kk < NPZ; kk++) {
                        uint64_t index = INDEX4(ci, ii, jj, kk);
                        T9 Ji;
                        auto jacd = [&](Int im, Int jm, Int km) {
                            uint64_t index1 = INDEX4(ci, im, jm, km);
                            V3 d(b0.dpsi[ii * NPX + im] * b1.psi[jj * NPY + jm] * b2.psi[kk * NPZ + km],
                                 b0.psi[ii * NPX + im] * b1.dpsi[jj * NPY + jm] * b2.psi[kk * NPZ + km],
                                 b0.psi[ii * NPX + im] * b1.psi[jj * NPY + jm] * b2.dpsi[kk * NPZ + km]);
                            T9 o = outer(geo.cC[index1], d);
                            for (int q = 0; q < 9; q++) Ji[q] += o[q];
                        };
                        for (Int i = 0; i < NPX; i++) jacd(i, jj, kk);
                        for (Int j = 0; j < NPY; j++) if (j != jj) jacd(ii, j, kk);
                        for (Int k = 0; k < NPZ; k++) if (k != kk) jacd(ii, jj, k);
                        T9 JiT = trn(Ji);
                        Ji = tmul(JiT, Ji);
                        if (NPX == 1) Ji[XX] = 1;
                        if (NPY == 1) Ji[YY] = 1;
                        if (NPZ == 1) Ji[ZZ] = 1;
                        Ji = inv(Ji);
                        if (NPX == 1) Ji[XX] = 0;
                        if (NPY == 1) Ji[YY] = 0;
                        if (NPZ == 1) Ji[ZZ] = 0;
                        Ji = tmul(Ji, JiT);
                        geo.Jinv[index] = trn(Ji);
                    }
        });
    }
    tm.mark("Jacobians");
    // facet interpolation factor (field.cpp:254-269), single domain
    geo.fI.assign(nfn, 0.0);
    for (uint64_t k = 0; k < nfn; k++) {
        Int c1 = geo.FO[k], c2 = geo.FN[k];
        if (c2 >= gBCSf) geo.fI[k] = 0;
        else if (NP > 1) geo.fI[k] = 0.5;
        else {
            V3 a = sub(geo.fC[k], geo.cC[c1]);
            V3 b = sub(geo.cC[c2], geo.cC[c1]);
            geo.fI[k] = 1 - dot(a, geo.fN[k]) / dot(b, geo.fN[k]);
        }
    }
    // diffusivity factor (field.cpp:280-352); the DG penalty form is only
    // needed by the implicit DG laplacian, which this path does not use.
    geo.fD.assign(nfn, 0.0);
    if (!dg) {
        for (uint64_t k = 0; k < nfn; k++) {
            Int c1 = geo.FO[k], c2 = geo.FN[k];
            V3 dv = sub(geo.cC[c2], geo.cC[c1]);
            const V3& n = geo.fN[k];
            if (nonortho == 3) geo.fD[k] = dot(n, n) / dot(n, dv);
            else if (nonortho == 1) geo.fD[k] = dot(n, dv) / dot(dv, dv);
            else geo.fD[k] = std::sqrt(dot(n, n) / dot(dv, dv));
        }
    }
    (void)tdot;
    return true;
}

}  // namespace nsem
