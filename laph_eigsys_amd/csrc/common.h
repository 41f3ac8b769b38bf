// Internal helpers shared by the liblapheig sources: error reporting across
// the C-ABI (thread-local message + negative code) and HIP call checking.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/laph.h"

namespace laph {

// Records a printf-style message for laph_last_error() and returns code.
int set_error(int code, const char* fmt, ...);

// Runs f(); converts exceptions into error codes so nothing crosses the ABI.
template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return set_error(LAPH_ERR_NOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return set_error(LAPH_ERR_INVALID, "%s", e.what());
  } catch (...) {
    return set_error(LAPH_ERR_INVALID, "unknown exception");
  }
}

}  // namespace laph

#define LAPH_HIP(call)                                                                    \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return laph::set_error(e_ == hipErrorOutOfMemory ? LAPH_ERR_NOMEM : LAPH_ERR_HIP,   \
                             "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),       \
                             __FILE__, __LINE__);                                         \
  } while (0)

#define LAPH_CHECK(cond, code, ...)                       \
  do {                                                    \
    if (!(cond)) return laph::set_error(code, __VA_ARGS__); \
  } while (0)

#define LAPH_TRY(call)          \
  do {                          \
    int rc_ = (call);           \
    if (rc_ != LAPH_OK) return rc_; \
  } while (0)
