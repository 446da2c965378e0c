// Instantiation of the cooperative panel leaf (3 rows per thread).
#include "leaf.hpp"

namespace hplx {
template <>
const void* leaf_kernel_ptr<kLeafRPT>() {
  return reinterpret_cast<const void*>(&leaf::leaf_kernel<kLeafRPT>);
}
}  // namespace hplx
