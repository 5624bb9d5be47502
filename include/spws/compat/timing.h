/* Shared helpers of the reference-compatible headers: wall-clock seconds
 * (the reference times with clock(), i.e. CPU time) and status checking. */
#ifndef SPWS_COMPAT_TIMING_H
#define SPWS_COMPAT_TIMING_H
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "spws/spws.h"

namespace spws_compat {

inline double now_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

/* The reference's compute() always returns 0; a failing device call is
 * reported on stderr and returned as the nonzero spws_status. */
inline int report(int rc, const char* what) {
  if (rc != SPWS_OK) std::fprintf(stderr, "%s: %s\n", what, spws_last_error());
  return rc;
}

}  // namespace spws_compat

#endif
