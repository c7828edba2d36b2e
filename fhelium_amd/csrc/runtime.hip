// Device runtime helpers for the C ABI: per-thread error text, version and
// device description (the gfx950 counterpart of the reference's CUDA
// inspection module csrc/runtime/cuda_info.cpp).
#include <cstdarg>

#include "fh_common.h"

namespace fh {
namespace {
thread_local char g_error[512] = {0};
}  // namespace

void set_error(const char* fmt, ...) {
  va_list args;
  va_start(args, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, args);
  va_end(args);
}

const char* last_error() { return g_error; }

}  // namespace fh

FH_API const char* fh_last_error(void) { return fh::last_error(); }

FH_API int fh_abi_version(void) { return 4; }

// Writes "<arch> cu=<n> lds=<bytes> l2=<bytes> mem=<bytes>" for the current
// device into buf.
FH_API int fh_device_info(char* buf, int64_t buf_len) {
  FH_CHECK_ARG(buf && buf_len > 0, "device info: empty buffer");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    fh::set_error("hipGetDevice: %s", hipGetErrorString(e));
    return fh::kLaunchFailed;
  }
  hipDeviceProp_t p;
  e = hipGetDeviceProperties(&p, dev);
  if (e != hipSuccess) {
    fh::set_error("hipGetDeviceProperties: %s", hipGetErrorString(e));
    return fh::kLaunchFailed;
  }
  snprintf(buf, static_cast<size_t>(buf_len), "%s cu=%d lds=%zu l2=%d mem=%zu clock_khz=%d",
           p.gcnArchName, p.multiProcessorCount, p.sharedMemPerMultiprocessor,
           p.l2CacheSize, p.totalGlobalMem, p.clockRate);
  return fh::kOk;
}
