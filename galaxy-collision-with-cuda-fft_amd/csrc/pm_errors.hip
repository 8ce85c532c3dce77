// Thread-local error state of the C ABI.
#include "pm_common.h"

namespace {
thread_local std::string g_last_error;
}

namespace pm {
void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }
}  // namespace pm

extern "C" const char* pm_last_error(void) { return g_last_error.c_str(); }

extern "C" int32_t pm_abi_version(void) { return 2; }
