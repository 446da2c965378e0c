// Instantiation of the cooperative panel leaf: float registers over __half storage,
// 6 rows per thread.
#include "leaf.hpp"

namespace hplx {
template <>
const void* leaf_kernel_ptr<LeafKind::kF32H>() {
  return reinterpret_cast<const void*>(&leaf::leaf_kernel<float, __half, 6>);
}
}  // namespace hplx
