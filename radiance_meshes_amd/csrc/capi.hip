// Library-wide C ABI entry points: error reporting and version.
#include "common.h"

namespace rmesh {
namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string& msg) { g_last_error = msg; }
}  // namespace rmesh

extern "C" const char* rm_last_error(void) { return rmesh::g_last_error.c_str(); }

extern "C" int rm_version(void) { return 1; }
