// Reader for the reference's field files (<name><step>.txt / .bin): the
// internal values (explicit list or up to four analytic generators) and the
// per-patch boundary conditions.  Follows MeshField::readInternal_ and
// readBoundary_ (src/field/field.h:1407-1546) and BCondition::operator>>
// (field.h:215-263).
#include "fields.hpp"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace nsem {
namespace {

constexpr double kPi = 3.14159265358979323846264;

inline double dot3(const double* p, const double* q) {
    double r = p[2] * q[2] + 0.0;
    r = p[1] * q[1] + r;
    return p[0] * q[0] + r;
}
inline bool equal_s(double p, double q, double tol = 1e-7) {
    double d = std::fabs(p - q);
    return d <= tol || d <= tol * std::fabs(p) || d <= tol * std::fabs(q);
}
std::string upper(std::string s) {
    for (auto& c : s) c = std::toupper((unsigned char)c);
    return s;
}

// Token stream over either format; the binary format stores strings with a
// one-byte length and numbers raw (util.h:191-258).
struct In {
    std::ifstream is;
    bool bin = false;
    bool str(std::string& s) {
        if (bin) {
            unsigned char n;
            if (!is.read((char*)&n, 1)) return false;
            s.resize(n);
            return (bool)is.read(&s[0], n);
        }
        return (bool)(is >> s);
    }
    bool u32(uint32_t& v) {
        if (bin) return (bool)is.read((char*)&v, 4);
        long long t;
        if (!(is >> t)) return false;
        v = (uint32_t)t;
        return true;
    }
    bool f64(double& v) {
        if (bin) return (bool)is.read((char*)&v, 8);
        return (bool)(is >> v);
    }
    bool sym() {
        if (bin) {
            unsigned char n;
            if (!is.read((char*)&n, 1)) return false;
            char c;
            return n ? (bool)is.read(&c, 1) : true;
        }
        char c;
        return (bool)(is >> c);
    }
    // A "word" token in the text format may be glued to a brace ("top{").
    bool word(std::string& s) {
        if (!str(s)) return false;
        return true;
    }
};

}  // namespace

bool read_field_file(const std::string& path, const FieldContext& ctx, int ncomp, FieldData& out,
                     std::string* err) {
    auto fail = [&](const std::string& m) { if (err) *err = path + ": " + m; return false; };
    In in;
    in.bin = path.size() > 4 && path.compare(path.size() - 4, 4, ".bin") == 0;
    in.is.open(path, in.bin ? std::ios::in | std::ios::binary : std::ios::in);
    if (!in.is) return fail("cannot open");
    std::string s;
    uint32_t size;
    if (!in.str(s) || !in.u32(size)) return fail("bad header");
    if ((int)size != ncomp) return fail("component count mismatch");
    std::string name;
    if (!in.str(name) || !in.u32(size) || !in.sym()) return fail("bad internal header");
    if (name != "internal") return fail("internal field not found");
    const uint64_t nall = ctx.nall;
    out.values.assign(nall * ncomp, 0.0);
    if (size <= 4) {
        for (uint32_t g = 0; g < size; g++) {
            std::string kind;
            if (!in.str(kind)) return fail("bad generator");
            auto rd = [&](double* v, int n) {
                for (int c = 0; c < n; c++) if (!in.f64(v[c])) return false;
                return true;
            };
            double value[3] = {0, 0, 0}, pert[3] = {0, 0, 0}, center[3], radius[3];
            if (kind == "uniform") {
                if (!rd(value, ncomp)) return fail("bad uniform");
                for (uint64_t i = 0; i < nall; i++)
                    for (int c = 0; c < ncomp; c++) out.values[i * ncomp + c] += value[c];
            } else if (kind == "cosine" || kind == "cosine2" || kind == "gaussian" || kind == "linear") {
                if (!rd(value, ncomp) || !rd(pert, ncomp) || !rd(center, 3) || !rd(radius, 3))
                    return fail("bad " + kind);
                double pw = kind == "cosine2" ? 2.0 : 1.0;
                for (uint64_t i = 0; i < nall; i++) {
                    const double* x = &ctx.cC[3 * i];
                    double d[3] = {(x[0] - center[0]) / radius[0], (x[1] - center[1]) / radius[1],
                                   (x[2] - center[2]) / radius[2]};
                    double R = std::sqrt(dot3(d, d));
                    for (int c = 0; c < ncomp; c++) {
                        double val = value[c];
                        if (kind == "gaussian") {
                            double v = std::exp(-R * R);
                            if (equal_s(v, 0.0)) v = 0;
                            val += pert[c] * v;
                        } else if (kind == "linear") {
                            double Rm = R <= 1.0 ? R : 1.0;
                            val += pert[c] * (1.0 - Rm);
                        } else {
                            double Rm = R <= 1.0 ? R : 1.0;
                            val += (pert[c] / 2) * std::pow(1.0 + std::cos(Rm * kPi), pw);
                        }
                        out.values[i * ncomp + c] += val;
                    }
                }
            } else if (kind == "gaussian-outside") {
                double r1, r2;
                if (!rd(value, ncomp) || !rd(pert, ncomp) || !rd(center, 3) || !in.f64(r1) || !in.f64(r2))
                    return fail("bad gaussian-outside");
                for (uint64_t i = 0; i < nall; i++) {
                    const double* x = &ctx.cC[3 * i];
                    double d[3] = {x[0] - center[0], x[1] - center[1], x[2] - center[2]};
                    double R = (std::sqrt(dot3(d, d)) - r1) / r2;
                    for (int c = 0; c < ncomp; c++) {
                        double val = value[c];
                        if (R <= 0) val += pert[c];
                        else {
                            double v = std::exp(-R * R);
                            if (equal_s(v, 0.0)) v = 0;
                            val += pert[c] * v;
                        }
                        out.values[i * ncomp + c] += val;
                    }
                }
            } else if (kind == "hydrostatic") {
                double p0[3], scale, expon;
                if (!rd(p0, ncomp) || !in.f64(scale) || !in.f64(expon)) return fail("bad hydrostatic");
                for (uint64_t i = 0; i < nall; i++) {
                    double gh = dot3(&ctx.cC[3 * i], ctx.gravity);
                    for (int c = 0; c < ncomp; c++)
                        out.values[i * ncomp + c] += p0[c] * std::pow(1.0 + scale * gh, expon);
                }
            } else {
                return fail("unknown initialization " + kind);
            }
        }
        if (!in.sym()) return fail("bad internal terminator");
    } else {
        if (size > ctx.nbcs) return fail("explicit list longer than the interior");
        for (uint64_t i = 0; i < uint64_t(size) * ncomp; i++)
            if (!in.f64(out.values[i])) return fail("short explicit list");
        if (!in.sym()) return fail("bad internal terminator");
    }
    // boundary block
    uint32_t nb;
    if (!in.str(name) || !in.u32(nb) || !in.sym()) return fail("bad boundary header");
    out.bcs.clear();
    for (uint32_t b = 0; b < nb; b++) {
        BoundaryCondition bc;
        bc.value.assign(ncomp, 0.0);
        bc.tvalue.assign(ncomp, 0.0);
        if (!in.str(bc.patch)) return fail("bad patch name");
        if (!in.bin) {
            // text: the name may carry the brace ("top{")
            size_t br = bc.patch.find('{');
            if (br != std::string::npos) bc.patch = bc.patch.substr(0, br);
            else if (!in.sym()) return fail("bad patch brace");
        } else if (!in.sym()) return fail("bad patch brace");
        while (true) {
            std::string key;
            if (!in.str(key)) return fail("unterminated patch");
            if (key == "}") break;
            std::string K = upper(key);
            if (K == "TYPE") {
                std::string t;
                in.str(t);
                bc.type = upper(t);
            } else if (K == "VALUE") {
                for (int c = 0; c < ncomp; c++) in.f64(bc.value[c]);
            } else if (K == "SHAPE") {
                in.f64(bc.shape);
            } else if (K == "TVALUE") {
                for (int c = 0; c < ncomp; c++) in.f64(bc.tvalue[c]);
            } else {
                return fail("unsupported boundary keyword " + key);
            }
        }
        out.bcs.push_back(bc);
    }
    return true;
}

namespace {

// Token sink of either format.  The binary stream (Util::ofstream_bin,
// util.h:191-222) writes a string as a length byte and its characters with
// all whitespace removed (nothing when that leaves it empty), an Int as 4
// bytes and a Scalar as 8, little-endian; the text stream writes them as they
// are, numbers as a C++ stream of precision 12 does (%.12g).
struct Sink {
    bool bin = false;
    std::string out;
    void str(const std::string& s) {
        if (!bin) { out += s; return; }
        std::string t;
        for (char c : s)
            if (!std::isspace((unsigned char)c)) t += c;
        if (t.empty()) return;
        out += (char)(unsigned char)t.size();
        out += t;
    }
    void u32(uint32_t v) {
        if (!bin) { out += std::to_string(v); return; }
        out.append(reinterpret_cast<const char*>(&v), 4);
    }
    void f64(double v) {
        if (!bin) {
            char buf[64];
            std::snprintf(buf, sizeof(buf), "%.12g", v);
            out += buf;
            return;
        }
        out.append(reinterpret_cast<const char*>(&v), 8);
    }
};

// A field component block: a Scalar as "v", a Vector as "x y z " (the
// tensor writer puts a space after every component, tensor.h:386-390).
void put_value(Sink& o, const double* v, int ncomp) {
    if (ncomp == 1) { o.f64(v[0]); return; }
    for (int c = 0; c < ncomp; c++) { o.f64(v[c]); o.str(" "); }
}

}  // namespace

bool write_field_file(const std::string& path, int ncomp, uint64_t n, const double* values,
                      const std::vector<BoundaryCondition>& bcs, std::string* err) {
    Sink o;
    o.bin = path.size() > 4 && path.compare(path.size() - 4, 4, ".bin") == 0;
    o.str("size "); o.u32((uint32_t)ncomp); o.str("\n");
    o.str("internal "); o.u32((uint32_t)n); o.str("\n{\n");
    for (uint64_t i = 0; i < n; i++) {
        put_value(o, values + i * ncomp, ncomp);
        o.str("\n");
    }
    o.str("}\n");
    o.str("boundary "); o.u32((uint32_t)bcs.size()); o.str("\n{\n");
    for (const BoundaryCondition& bc : bcs) {
        o.str(bc.patch); o.str("\n{\n");
        o.str("\ttype "); o.str(bc.type); o.str("\n");
        // written unless equal(mag(value), 0) (tensor.h:471-476, EqualEpsilon 1e-7)
        double m2 = 0;
        for (int c = 0; c < ncomp; c++) m2 += bc.value[c] * bc.value[c];
        if (!(std::sqrt(m2) <= 1e-7)) {
            o.str("\tvalue "); put_value(o, bc.value.data(), ncomp); o.str("\n");
        }
        if (!(std::fabs(bc.shape) <= 1e-7)) {
            o.str("\tshape "); o.f64(bc.shape); o.str("\n");
        }
        double t2 = 0;
        for (int c = 0; c < (int)bc.tvalue.size() && c < ncomp; c++) t2 += bc.tvalue[c] * bc.tvalue[c];
        if (!(std::sqrt(t2) <= 1e-7)) {
            o.str("\ttvalue "); put_value(o, bc.tvalue.data(), ncomp); o.str("\n");
        }
        o.str("}\n");
    }
    o.str("}\n");
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        if (err) *err = "cannot write " + path;
        return false;
    }
    const bool ok = std::fwrite(o.out.data(), 1, o.out.size(), f) == o.out.size();
    std::fclose(f);
    if (!ok && err) *err = "short write to " + path;
    return ok;
}

// ------------------------------------------------------------ amrTree ----
// Node::operator<< (mesh.h:63-67) writes "id nchildren cid level"; the
// vector writer (util.h:129-135) puts the count, "\n{\n", one node per line
// and "}\n".  In the binary stream the separators reduce to the braces.
bool read_amr_tree(const std::string& path, std::vector<uint32_t>& nodes, std::string* err) {
    auto fail = [&](const std::string& m) { if (err) *err = path + ": " + m; return false; };
    In in;
    in.bin = path.size() > 4 && path.compare(path.size() - 4, 4, ".bin") == 0;
    in.is.open(path, in.bin ? std::ios::in | std::ios::binary : std::ios::in);
    if (!in.is) return fail("cannot open");
    uint32_t n;
    if (!in.u32(n) || !in.sym()) return fail("bad header");
    nodes.assign((size_t)n * 4, 0);
    for (size_t i = 0; i < (size_t)n * 4; i++)
        if (!in.u32(nodes[i])) return fail("short node list");
    if (!in.sym()) return fail("bad terminator");
    return true;
}

bool write_amr_tree(const std::string& path, const uint32_t* nodes, uint64_t n, std::string* err) {
    Sink o;
    o.bin = path.size() > 4 && path.compare(path.size() - 4, 4, ".bin") == 0;
    o.u32((uint32_t)n);
    o.str("\n{\n");
    for (uint64_t i = 0; i < n; i++) {
        const uint32_t* p = nodes + 4 * i;
        o.u32(p[0]); o.str(" "); o.u32(p[1]); o.str(" "); o.u32(p[2]); o.str(" "); o.u32(p[3]);
        o.str("\n");
    }
    o.str("}\n");
    std::ofstream f(path, std::ios::out | std::ios::binary);
    if (!f || !f.write(o.out.data(), (std::streamsize)o.out.size())) {
        if (err) *err = path + ": cannot write";
        return false;
    }
    return true;
}

}  // namespace nsem
