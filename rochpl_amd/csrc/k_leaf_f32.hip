// Instantiation of the cooperative panel leaf: float registers over float
// storage (the FP32 panel of the mixed-precision factorisation), 3 rows per
// thread.
#include "leaf.hpp"

namespace hplx {
template <>
const void* leaf_kernel_ptr<LeafKind::kF32>() {
  return reinterpret_cast<const void*>(&leaf::leaf_kernel<float, float, 3>);
}
}  // namespace hplx
