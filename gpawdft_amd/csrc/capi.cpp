// Library-wide C entry points: error reporting and version.
#include <string>

#include "common.hpp"

namespace gpd {
static thread_local std::string g_last_error;
void set_last_error(const char* msg) { g_last_error = msg ? msg : ""; }
}  // namespace gpd

extern "C" const char* gpd_last_error(void) { return gpd::g_last_error.c_str(); }

extern "C" const char* gpd_version(void) { return "gpawdft_amd 0.1 (gfx950)"; }
