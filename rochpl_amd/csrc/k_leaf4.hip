// Instantiation of the cooperative panel leaf with 4 row(s) per thread.
#include "leaf.hpp"

namespace hplx {
template <>
const void* leaf_kernel_ptr<4>() {
  return reinterpret_cast<const void*>(&leaf::leaf_kernel<4>);
}
}  // namespace hplx
