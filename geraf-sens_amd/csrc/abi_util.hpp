// Host-side argument checking and error reporting for the C ABI.
// Every extern "C" entry point wraps its body in GERAF_TRY_BEGIN/END, so a
// failed check becomes a status code plus a thread-local message instead of
// an exception crossing the C boundary.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <initializer_list>
#include <stdexcept>
#include <string>

#include "geraf_rf.h"

namespace geraf {

struct ArgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void set_last_error(const std::string& msg);

inline void check_mf_sizes(int64_t A, int64_t K, int64_t N) {
  if (A < 0 || K < 0 || N < 0) throw ArgError("sizes must be non-negative");
  if (A > INT32_MAX || K > INT32_MAX || N > INT32_MAX)
    throw ArgError("sizes must fit in int32");
  if (A * K > INT64_C(1) << 40) throw ArgError("A*K too large");
}

inline void check_params(const geraf_fmcw_params* p) {
  if (!p) throw ArgError("params is NULL");
  if (!std::isfinite(p->f0) || !std::isfinite(p->slope) || !std::isfinite(p->fs) ||
      !std::isfinite(p->c) || !std::isfinite(p->path_bias))
    throw ArgError("params must be finite");
  if (p->fs == 0.0) throw ArgError("adcSampleRate must be non-zero");
  if (p->c == 0.0) throw ArgError("speed of light must be non-zero");
}

inline void require_ptrs(std::initializer_list<const void*> ptrs, bool needed) {
  if (!needed) return;
  for (const void* p : ptrs)
    if (!p) throw ArgError("required pointer is NULL");
}

inline int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw HipError(std::string(what) + ": " + hipGetErrorString(e));
  return GERAF_RF_OK;
}

}  // namespace geraf

#define GERAF_TRY_BEGIN try {
#define GERAF_TRY_END                      \
  }                                        \
  catch (const geraf::ArgError& e) {       \
    geraf::set_last_error(e.what());       \
    return GERAF_RF_EINVAL;                \
  }                                        \
  catch (const geraf::HipError& e) {       \
    geraf::set_last_error(e.what());       \
    return GERAF_RF_EHIP;                  \
  }                                        \
  catch (const std::exception& e) {        \
    geraf::set_last_error(e.what());       \
    return GERAF_RF_EINVAL;                \
  }
