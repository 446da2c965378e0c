// Instantiation of the cooperative panel leaf: float registers over float storage,
// 6 rows per thread.
#include "leaf.hpp"

namespace hplx {
template <>
const void* leaf_kernel_ptr<LeafKind::kF32>() {
  return reinterpret_cast<const void*>(&leaf::leaf_kernel<float, float, 6>);
}
}  // namespace hplx
