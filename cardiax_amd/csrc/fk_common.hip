// Thread-local error message of the C-ABI (fk_last_error).
#include <string>

#include "cardiax_amd.h"
#include "fk_host.h"

namespace fk {

static thread_local std::string g_last_error;
static thread_local int32_t g_last_code = FK_OK;

void set_error_text(int32_t code, const char* msg) {
  g_last_error = msg;
  g_last_code = code;
}
void clear_error() {
  g_last_error.clear();
  g_last_code = FK_OK;
}
int32_t last_code() { return g_last_code; }

}  // namespace fk

extern "C" const char* fk_last_error(void) { return fk::g_last_error.c_str(); }
