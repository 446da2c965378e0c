// Tile claiming of the look-ahead's update kernels (FP64 and low-precision):
// a persistent "claimed" launch beside the panel factorisation takes tiles
// from per-XCD counters until a stop flag is set, and a "resume" launch
// finishes every XCD's share from where the claims ended.
#pragma once

#include <hip/hip_runtime.h>

namespace hplx {

constexpr int kClaimXcds  = 8;
constexpr int kClaimChunk = 1;  // tiles per counter update

struct UpdSched {
  int*                      claim;   // per-XCD counters, claim[32 x]
  int                       per;     // tiles of each XCD share
  const unsigned long long* vacate;  // 8 words: __smid() values whose CUs
                                     // the launch leaves (null: none)
  const int*                stop;    // claims stop once *stop != 0
};

__device__ inline bool on_vacated_cu(const UpdSched& s) {
  if(!s.vacate) return false;
  const unsigned id = __smid() & 511u;
  return (s.vacate[id >> 6] >> (id & 63)) & 1ull;
}

// XCC_ID[3:0] of the hardware register: the XCD this wave runs on.
__device__ inline int xcd_id() {
  return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & (kClaimXcds - 1);
}

// First claim on the first XCD share not yet found exhausted, starting with
// this workgroup's XCD: returns (XCD << 24 | counter value), or -1.
__device__ inline int claim_issue(int* claim, int x, unsigned dead) {
  for(int d = 0; d < kClaimXcds; ++d) {
    const int y = (x + d) & (kClaimXcds - 1);
    if(!((dead >> y) & 1u))
      return (y << 24) | atomicAdd(&claim[32 * y], kClaimChunk);
  }
  return -1;
}

// The tile a claim stands for, or further claims on the other shares until
// one succeeds; -1 once every share is exhausted.
__device__ inline int claim_resolve(int* claim, int x, int v, int per, int T,
                                    unsigned& dead) {
  while(v >= 0) {
    const int y = v >> 24, k = v & 0xFFFFFF;
    if(k < per && y * per + k < T) return y * per + k;
    dead |= 1u << y;
    v = claim_issue(claim, x, dead);
  }
  return -1;
}

// Thread 0's next claim for a workgroup: the tile, or -1 once *stop is set
// or every share is exhausted.
__device__ inline int claim_next_tile(const UpdSched& sch, int T,
                                      unsigned& dead) {
  if(__hip_atomic_load(sch.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    return -1;
  const int x = xcd_id();
  return claim_resolve(sch.claim, x, claim_issue(sch.claim, x, dead), sch.per,
                       T, dead);
}

}  // namespace hplx
