// C-ABI housekeeping: error messages, device query, library version.
#include "common.h"

#include <stdarg.h>
#include <stdio.h>

static thread_local char g_last_error[512] = "";

void gis_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

GIS_API const char* gis_last_error(void) { return g_last_error; }

GIS_API int gis_version(void) { return 1; }

int gis_cu_count() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cached[dev] == 0) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
    cached[dev] = cus;
  }
  return cached[dev];
}

// Makes `device` current on the calling thread and checks that it is usable.
GIS_API int gis_init(int device) {
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    gis_set_error("gis_init: hipSetDevice(%d): %s", device, hipGetErrorString(e));
    return (int)e;
  }
  (void)gis_cu_count();
  return gis_split_scratch_setup();
}
