// Internal declarations shared by the HIP translation units and the C-ABI layer.
#pragma once

#include "vj_dev.hpp"

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace vj {

// Records a message for vjh_last_error() and returns a negative status.
int set_error(int code, const char *fmt, ...);
// Maps the last launch error to a status (0 or a negative code).
int check_launch();
// Negative status used for failures of the HIP runtime itself.
constexpr int VJ_EHIP = -100000;

template <typename T>
int launch_rbjac2(long long n, const T *a11, const T *a22, const T *a21, T *c, T *at, T *l1, T *l2, unsigned *p, hipStream_t s);
template <typename T>
int launch_cbjac2(long long n, const T *a11, const T *a22, const T *a21r, const T *a21i, T *c, T *cat, T *sat, T *l1, T *l2, unsigned *p, hipStream_t s);
// internal forms: blocks gated by the low bits of p (in and out)
template <typename T>
int launch_rbjac2i(long long n, const T *a11, const T *a22, const T *a21, T *c, T *at, T *l1, T *l2, unsigned *p, hipStream_t s);
template <typename T>
int launch_cbjac2i(long long n, const T *a11, const T *a22, const T *a21r, const T *a21i, T *c, T *cat, T *sat, T *l1, T *l2, unsigned *p, hipStream_t s);

} // namespace vj
