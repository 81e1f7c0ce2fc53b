// Per-thread last-error message for the C ABI.
#include <string>

#include "../../include/dcvc_codec.h"
#include "capi_util.h"

namespace dcvc {
static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace dcvc

extern "C" const char* dcvc_last_error(void) { return dcvc::g_last_error.c_str(); }
