// Direct global -> LDS loads for the streamed GEMMs (k_update.hip,
// k_mxp_gemm.hip).
//
// GLOBAL_LOAD_LDS_DWORDX4 moves 16 bytes per lane from the lane's own global
// address to LDS at M0 + 16 * lane, without passing through VGPRs.  The
// loads are issued from inline assembly: with the compiler's builtin the
// waitcnt pass cannot tell which LDS buffer a load writes and waits for all
// of them before every LDS read, which would serialise a ring of stages.
// The kernels therefore count their own loads and wait explicitly (wait_vm)
// before the workgroup barrier that publishes a stage.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hplx {

// 16 bytes per lane from g into lds + 16 * lane (lds: this wave's 1 KB
// destination, the same for every lane).
__device__ inline void lds_load16(const void* g, void* lds) {
  const uint32_t l = static_cast<uint32_t>(reinterpret_cast<uintptr_t>(
      (__attribute__((address_space(3))) uint8_t*)(lds)));
  const uint32_t u = __builtin_amdgcn_readfirstlane(l);
  asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off"
               :
               : "s"(u), "v"(g)
               : "memory", "m0");
}

// s_waitcnt vmcnt(n) with the other counters left alone (gfx9 encoding).
template <int N>
__device__ inline void wait_vm() {
  static_assert(N >= 0 && N < 64, "vmcnt is 6 bits");
  __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) |
                             (0xF << 8));
}

}  // namespace hplx
