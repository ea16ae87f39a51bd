// Reader for the reference mesh tool's block-description file (the input of
// apps/mesh/meshApp.cpp:16-262): key points, hex blocks with division counts
// and spacing, and named boundary planes.  Supports straight-edged blocks;
// one block per file (the configurations this path runs are single blocks).
#include <cctype>
#include <fstream>
#include <sstream>

#include "mesh.hpp"

namespace nsem {
namespace {

// Next non-blank, non-comment character without consuming it
// (util.cpp:18-30); 0 at end of input.
int next_char(std::istream& is) {
    char c;
    if (!(is >> c)) return 0;
    while (c == '#') {
        while (is.get(c) && c != '\n') {}
        if (!(is >> c)) return 0;
    }
    is.putback(c);
    return c;
}

template <typename T>
bool read_list(std::istream& is, std::vector<T>& out) {
    long long n;
    char sym;
    if (!(is >> n >> sym)) return false;
    out.resize(n);
    for (long long i = 0; i < n; i++)
        if (!(is >> out[i])) return false;
    is >> sym;
    return (bool)is;
}

bool ieq(const std::string& a, const char* b) {
    std::string t = a;
    for (auto& c : t) c = std::toupper((unsigned char)c);
    std::string u = b;
    for (auto& c : u) c = std::toupper((unsigned char)c);
    return t == u;
}

}  // namespace

bool read_block_spec(const std::string& path, HexBlock& blk, std::vector<PatchSpec>& patches,
                     std::string* err) {
    std::ifstream is(path);
    if (!is) { if (err) *err = "cannot open " + path; return false; }
    std::vector<V3> keys;
    if (next_char(is)) {
        long long n;
        char sym;
        is >> n >> sym;
        keys.resize(n);
        for (long long i = 0; i < n; i++) is >> keys[i][0] >> keys[i][1] >> keys[i][2];
        is >> sym;
    }
    int nblocks = 0;
    int c;
    while ((c = next_char(is)) != 0) {
        if (std::isdigit(c)) {
            if (nblocks++) { if (err) *err = "multi-block grids are not supported"; return false; }
            std::vector<long long> index;
            if (!read_list(is, index) || index.size() != 8) { if (err) *err = "bad block"; return false; }
            for (int i = 0; i < 8; i++) blk.corners[i] = keys.at(index[i]);
            std::string str;
            is >> str;
            for (int i = 0; i < 12; i++) blk.s[i] = 1.0;
            std::vector<long long> n;
            if (ieq(str, "linear")) {
                read_list(is, n);
                for (int i = 0; i < 12; i++) blk.type[i] = LINEAR;
            } else {
                int type;
                if (ieq(str, "geometric")) type = GEOMETRIC;
                else if (ieq(str, "wall")) type = WALL;
                else { if (err) *err = "unsupported spacing " + str; return false; }
                read_list(is, n);
                std::vector<double> ts;
                read_list(is, ts);
                int sz = ts.size();
                int r = 12 / sz;
                for (int i = 0; i < sz; i++)
                    for (int j = 0; j < r; j++)
                        if (i * r + j < 12) { blk.s[i * r + j] = ts[i]; }
                for (int i = 0; i < 12; i++) blk.type[i] = type;
            }
            if (n.size() != 3) { if (err) *err = "bad division counts"; return false; }
            for (int i = 0; i < 3; i++) blk.n[i] = (Int)n[i];
            if ((c = next_char(is)) && c == 'e') {
                std::string kw;
                std::streampos pos = is.tellg();
                is >> kw;
                if (ieq(kw, "edges")) { if (err) *err = "curved edges are not supported"; return false; }
                is.seekg(pos);
            }
        } else {
            PatchSpec ps;
            is >> ps.name;
            if (ps.name == "default") { if (err) *err = "default patch is not supported"; return false; }
            while ((c = next_char(is)) && std::isdigit(c)) {
                std::vector<long long> idx;
                read_list(is, idx);
                PatchSpec one;
                one.name = ps.name;
                for (auto k : idx) one.points.push_back(keys.at(k));
                patches.push_back(one);
            }
        }
    }
    if (!nblocks) { if (err) *err = "no block in " + path; return false; }
    return true;
}

}  // namespace nsem
