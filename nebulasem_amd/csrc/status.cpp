// Thread-local error message behind nsem_last_error().
#include "status.hpp"

namespace {
thread_local std::string g_last_error = "";
}

int nsem_fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

extern "C" {
const char* nsem_last_error(void) { return g_last_error.c_str(); }
const char* nsem_version(void) { return "nebulasem_amd 0.1 (gfx950)"; }
}
