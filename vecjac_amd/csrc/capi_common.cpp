// Error reporting and host<->device staging shared by the C-ABI entry points.
#include "capi_common.hpp"

#include <mutex>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

namespace vj {

static thread_local char g_err[512] = "";

int set_error(int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int check_launch()
{
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error(VJ_EHIP - (int)e, "kernel launch failed: %s", hipGetErrorString(e));
  return 0;
}

hipStream_t lib_stream()
{
  static std::mutex mu;
  static hipStream_t streams[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
    return (hipStream_t)0;
  std::lock_guard<std::mutex> g(mu);
  if (!streams[dev]) {
    hipStream_t s = (hipStream_t)0;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
      return (hipStream_t)0;
    streams[dev] = s;
  }
  return streams[dev];
}

const char *last_error() { return g_err; }

} // namespace vj
