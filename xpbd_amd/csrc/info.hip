// Build identification and error strings of libxpbd_hip.so.
#include <hip/hip_runtime.h>

#include "xpbd.h"

#ifndef XPBD_BUILD_FLAGS
#define XPBD_BUILD_FLAGS "unknown"
#endif

extern "C" const char* xpbd_build_info(void) {
  return "libxpbd_hip gfx950 flags=" XPBD_BUILD_FLAGS;
}

extern "C" const char* xpbd_error_string(int code) {
  if (code == -1) return "invalid argument to a host routine";
  if (code == XPBD_ERR_NOALLOC_TOO_LARGE)
    return "more bodies than the allocation-free broad phase's fixed scratch holds "
           "(XPBD_NOALLOC_MAX_BODIES = 65536)";
  return hipGetErrorString((hipError_t)code);
}
