// The batched Jacobi SVD kernels (FP32 complex); see svd_batched_impl.hpp.
#include "svd_batched_impl.hpp"

namespace vj {
template int batched_svd<float, true>(const BatchArgs<float> &, int, int, int, bool, hipStream_t);
} // namespace vj
