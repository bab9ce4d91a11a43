// Error state, version and memory release of the dsplat C ABI.
#include <stdio.h>
#include <string.h>

#include "dsplat_common.hpp"

namespace {
thread_local char g_last_error[512] = "";
}

int dsplat_set_error(int code, const char* msg) {
  snprintf(g_last_error, sizeof(g_last_error), "%s", msg);
  return code;
}

int dsplat_set_hip_error(hipError_t e, const char* what) {
  snprintf(g_last_error, sizeof(g_last_error), "%s: %s (%d)", what, hipGetErrorString(e), int(e));
  return DSPLAT_ERR_HIP;
}

extern "C" const char* dsplat_last_error(void) { return g_last_error; }

extern "C" int dsplat_version(void) { return 100; }

extern "C" int dsplat_free(void* stream, void* ptr) {
  if (!ptr) return DSPLAT_OK;
  DSPLAT_HIP(hipFreeAsync(ptr, dsplat::as_stream(stream)));
  return DSPLAT_OK;
}
