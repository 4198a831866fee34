// The batched Jacobi SVD kernels (FP64 real); see svd_batched_impl.hpp.
#include "svd_batched_impl.hpp"

namespace vj {
template int batched_svd<double, false>(const BatchArgs<double> &, int, int, int, bool, hipStream_t);
} // namespace vj
