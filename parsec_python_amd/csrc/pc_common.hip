// Error text and version for the C-ABI.
#include "pc_common.h"

namespace pc {
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }
}  // namespace pc

extern "C" {

const char* pc_last_error(void) { return pc::last_error(); }

int pc_version(void) { return 1; }

}  // extern "C"
