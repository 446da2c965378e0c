// Instantiation of the cooperative panel leaf: double registers over double storage,
// 3 rows per thread.
#include "leaf.hpp"

namespace hplx {
template <>
const void* leaf_kernel_ptr<LeafKind::kF64>() {
  return reinterpret_cast<const void*>(&leaf::leaf_kernel<double, double, 3>);
}
}  // namespace hplx
