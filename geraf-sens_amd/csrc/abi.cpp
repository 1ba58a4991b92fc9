// Library identification and the thread-local error slot of the C ABI.
#include <hip/hip_runtime.h>

#include <string>

#include "abi_util.hpp"

namespace geraf {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace geraf

extern "C" const char* geraf_rf_last_error(void) { return geraf::g_last_error.c_str(); }

#define GERAF_STR2(x) #x
#define GERAF_STR(x) GERAF_STR2(x)

extern "C" const char* geraf_rf_version(void) {
  return "geraf_rf 0.1 gfx950 (clang " __clang_version__ ", HIP " GERAF_STR(
      HIP_VERSION_MAJOR) "." GERAF_STR(HIP_VERSION_MINOR) ")";
}
