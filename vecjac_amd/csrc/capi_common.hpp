// Helpers for the C-ABI layer: HIP status checks and stream-ordered staging
// buffers for the host-pointer (reference-compatible) entry points.
#pragma once

#include "vj_internal.hpp"

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <vector>

namespace vj {

hipStream_t lib_stream();
const char *last_error();

#define VJ_HIP(call)                                                                       \
  do {                                                                                     \
    const hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                                  \
      return ::vj::set_error(::vj::VJ_EHIP - (int)e_, "%s failed: %s", #call,              \
                             hipGetErrorString(e_));                                       \
  } while (0)

// A set of stream-ordered device allocations released (in stream order) on
// scope exit.
class Staging {
public:
  explicit Staging(hipStream_t s) : s_(s) {}
  ~Staging()
  {
    for (void *p : ptrs_)
      (void)hipFreeAsync(p, s_);
  }
  template <typename T> int alloc(T **out, size_t count)
  {
    void *p = nullptr;
    const size_t bytes = (count ? count : 1) * sizeof(T);
    const hipError_t e = hipMallocAsync(&p, bytes, s_);
    if (e != hipSuccess)
      return set_error(VJ_EHIP - (int)e, "hipMallocAsync(%zu) failed: %s", bytes, hipGetErrorString(e));
    ptrs_.push_back(p);
    *out = (T *)p;
    return 0;
  }
  hipStream_t stream() const { return s_; }

private:
  hipStream_t s_;
  std::vector<void *> ptrs_;
};

} // namespace vj
