// Instantiation of the cooperative panel leaf with 2 row(s) per thread.
#include "leaf.hpp"

namespace hplx {
template <>
const void* leaf_kernel_ptr<2>() {
  return reinterpret_cast<const void*>(&leaf::leaf_kernel<2>);
}
}  // namespace hplx
