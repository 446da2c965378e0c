// Instantiation of the cooperative panel leaf with 1 row(s) per thread.
#include "leaf.hpp"

namespace hplx {
template <>
const void* leaf_kernel_ptr<1>() {
  return reinterpret_cast<const void*>(&leaf::leaf_kernel<1>);
}
}  // namespace hplx
