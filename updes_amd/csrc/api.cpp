// Library-level entry points: ABI version, error reporting, device queries.
#include <stdarg.h>
#include <stdio.h>

#include "common.h"

namespace rbfx {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char* last_error() { return g_err; }

static DeviceInfo g_info[64];

const DeviceInfo& device_info() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  DeviceInfo& d = g_info[dev];
  if (d.device != dev) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
      d.num_cus = prop.multiProcessorCount;
      d.lds_per_block = prop.sharedMemPerBlock;
    } else {
      d.num_cus = 256;
      d.lds_per_block = 65536;
    }
    d.device = dev;
  }
  return d;
}

}  // namespace rbfx

extern "C" int rbfx_abi_version(void) { return RBFX_ABI_VERSION; }

extern "C" const char* rbfx_last_error(void) { return rbfx::last_error(); }

extern "C" int rbfx_init(void) {
  int n = 0;
  RBFX_HIP(hipGetDeviceCount(&n));
  RBFX_REQUIRE(n > 0, "no HIP device");
  (void)rbfx::device_info();
  return RBFX_OK;
}
