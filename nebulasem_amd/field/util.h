// Parameters of the controls file (src/util/util.h ParamList / Option /
// BoolOption / read_params, util.cpp:32-79): named blocks `name { key
// value ... }`, each key enrolled with the address it is read into.  An
// enrolled block reads its keys where the file has them; other blocks and
// unknown keys are skipped.  Values: Int, Scalar, Vector (three numbers),
// strings, options (one word of a list, case-insensitive) and the
// `n { ... }` lists of the reference's vector format.
#ifndef NSEM_FIELD_UTIL_H
#define NSEM_FIELD_UTIL_H

#include <algorithm>
#include <initializer_list>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "tensor.h"

#define forEach(field, i) for (Int i = 0; i < Int((field).size()); i++)
#define forEachIt(container, it) for (auto it = (container).begin(); it != (container).end(); ++it)

namespace Util {

Int hash_function(std::string s);
// next non-blank character, skipping '#' comments; 0 at the end
int nextc(std::istream& is);
// 0 when equal ignoring case (the reference's convention)
inline int compare(std::string s1, std::string s2) {
    std::transform(s1.begin(), s1.end(), s1.begin(), ::toupper);
    std::transform(s2.begin(), s2.end(), s2.begin(), ::toupper);
    return s1 != s2;
}

struct Option {
    Int* val;
    std::vector<std::string> list;
    Option(void* v, std::initializer_list<std::string> names) : val((Int*)v), list(names) {}
    virtual ~Option() {}
    Int getID(const std::string& s) const {
        for (Int i = 0; i < list.size(); i++)
            if (!compare(list[i], s)) return i;
        std::cout << "Unknown parameter : " << s << std::endl;
        return 0;
    }
};
struct BoolOption : public Option {
    explicit BoolOption(void* v) : Option(v, {"NO", "YES"}) {}
};

template <class T>
std::istream& operator>>(std::istream& is, std::vector<T>& v) {
    Int n;
    char c;
    is >> n >> c;
    v.resize(n);
    for (Int i = 0; i < n; i++) is >> v[i];
    is >> c;
    return is;
}

void read_params(std::istream& is, bool print, std::string block = "");

struct ParamList {
    std::string name;
    static std::map<std::string, ParamList*> list;
    explicit ParamList(std::string n);
    ~ParamList();
    void enroll(std::string key, Int* p) { ints[key] = p; }
    void enroll(std::string key, Scalar* p) { scalars[key] = p; }
    void enroll(std::string key, Vector* p) { vectors[key] = p; }
    void enroll(std::string key, std::string* p) { strings[key] = p; }
    void enroll(std::string key, Option* p) { options[key] = p; }
    void enroll(std::string key, std::vector<Int>* p) { vints[key] = p; }
    void enroll(std::string key, std::vector<Scalar>* p) { vscalars[key] = p; }
    void enroll(std::string key, std::vector<Vector>* p) { vvectors[key] = p; }
    void enroll(std::string key, std::vector<std::string>* p) { vstrings[key] = p; }
    // read this block from the stream (from its start)
    void read(std::istream& is) { read_params(is, false, name); }
    // read the value of `key` if enrolled; false when not
    bool read(std::istream& is, const std::string& key, bool print);

   private:
    std::map<std::string, Int*> ints;
    std::map<std::string, Scalar*> scalars;
    std::map<std::string, Vector*> vectors;
    std::map<std::string, std::string*> strings;
    std::map<std::string, Option*> options;
    std::map<std::string, std::vector<Int>*> vints;
    std::map<std::string, std::vector<Scalar>*> vscalars;
    std::map<std::string, std::vector<Vector>*> vvectors;
    std::map<std::string, std::vector<std::string>*> vstrings;
};

}  // namespace Util

#endif
