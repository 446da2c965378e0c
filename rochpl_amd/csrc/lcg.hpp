// 64-bit LCG of the HPL matrix generator, shared by host and device code.
//
// Reference: include/hpl_pmatgen.hpp:35-39 (MULT, IADD), the jump constants
// of src/matgen/HPL_xjumpm.cpp:81-91 (here computed by squaring in O(log k)
// instead of a linear loop) and the value formula of
// src/matgen/HPL_pdrandmat_device.cpp:65-73.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hplx {

constexpr uint64_t kLcgMult = 6364136223846793005ULL;
constexpr uint64_t kLcgAdd  = 1ULL;

struct Affine {
  uint64_t a, c;  // s -> a*s + c (mod 2^64)
};

// Map of k generator steps.
__host__ __device__ inline Affine lcg_jump(uint64_t k) {
  Affine r{1, 0}, p{kLcgMult, kLcgAdd};
  while(k) {
    if(k & 1) {
      r.c = p.a * r.c + p.c;
      r.a = p.a * r.a;
    }
    p.c = p.a * p.c + p.c;
    p.a = p.a * p.a;
    k >>= 1;
  }
  return r;
}

// The generator's first state: one step from the seed.
__host__ __device__ inline uint64_t lcg_start(uint64_t seed) {
  return kLcgMult * seed + kLcgAdd;
}

// State of element index k (= i + j*M for an M-row matrix).
__host__ __device__ inline uint64_t lcg_state(uint64_t seed, uint64_t k) {
  const Affine j = lcg_jump(k);
  return j.a * lcg_start(seed) + j.c;
}

// Matrix value of a state: 0.5 - s/2^64, rounded as RN(0.5 - RN(s*2^-32)*2^-32).
// Both products by powers of two are exact, so contracting either into an
// FMA gives the same bits.
__host__ __device__ inline double lcg_value(uint64_t s) {
  const double hi = static_cast<double>(static_cast<uint32_t>(s >> 32));
  const double lo = static_cast<double>(static_cast<uint32_t>(s));
  const double d  = hi + lo * 0x1p-32;
  return 0.5 - d * 0x1p-32;
}

}  // namespace hplx
