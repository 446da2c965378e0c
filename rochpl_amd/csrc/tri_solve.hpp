// Triangular solve of one right-hand-side column on one workgroup (the
// device routine of k_trsm.hip's kernels, whose header describes the
// 64-row chunk scheme), shared with the persistent panel kernel
// (k_panel_kernel.hip), which runs the panel recursion's solves with it.
// Any workgroup size that is a multiple of 64 works; each row's sum is
// formed by one thread in the same order whatever the size.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "hplx_internal.hpp"

namespace hplx {
namespace trs {

template <typename V, typename MT>
__device__ inline V to_v(MT x) {
  if constexpr(sizeof(MT) == 2)
    return static_cast<V>(__half2float(x));
  else
    return static_cast<V>(x);
}

template <bool TRANS, typename V, typename MT>
__device__ inline V op_at(const MT* T, int ldt, int i, int j) {
  return to_v<V>(TRANS ? T[j + static_cast<size_t>(i) * ldt]
                       : T[i + static_cast<size_t>(j) * ldt]);
}

// LOWER: op(T) is lower triangular (forward sweep), else upper (backward).
// UNIT: unit diagonal (not read).  IDENTITY: b is column `col` of I.
// V: element type of X and of the arithmetic (double, or float for the FP32
// panel); MT: element type of T (V, or the FP32 / FP16 factors whose
// diagonal blocks are inverted in FP64).
// One workgroup solves for the column xc (column `col` of the identity when
// IDENTITY).
template <typename MT, typename V, bool LOWER, bool TRANS, bool UNIT,
          bool IDENTITY>
__device__ inline void tri_solve_col(int k, const MT* __restrict__ T, int ldt,
                                     V* __restrict__ xc, int col) {
  __shared__ V sx[kMaxNB];
  const int t    = threadIdx.x;
  const int lane = t & 63;
  for(int i = t; i < k; i += blockDim.x)
    sx[i] = IDENTITY ? (i == col ? V(1) : V(0)) : xc[i];
  __syncthreads();
  const int nchunk = (k + 63) / 64;
  // with the identity, the rows before `col` in sweep order stay zero
  const int first = IDENTITY ? (LOWER ? col / 64 : nchunk - 1 - col / 64) : 0;
  for(int s = first; s < nchunk; ++s) {
    const int c0 = LOWER ? s * 64 : (nchunk - 1 - s) * 64;
    const int w  = min(64, k - c0);
    if(t < 64) {
      const int i  = c0 + lane;
      V         xi = lane < w ? sx[i] : V(0);
      const V   d  = (UNIT || lane >= w) ? V(1) : op_at<TRANS, V>(T, ldt, i, i);
      V         e[64];
#pragma unroll
      for(int j = 0; j < 64; ++j) {
        const bool in = j < w && lane < w && (LOWER ? j < lane : j > lane);
        e[j]          = in ? op_at<TRANS, V>(T, ldt, i, c0 + j) : V(0);
      }
      if(LOWER) {
#pragma unroll
        for(int j = 0; j < 64; ++j) {
          if(j < w) {
            if(!UNIT && lane == j) xi = xi / d;
            const V xj = __shfl(xi, j);
            if(lane > j) xi = fma(-e[j], xj, xi);
          }
        }
      } else {
#pragma unroll
        for(int j = 63; j >= 0; --j) {
          if(j < w) {
            if(!UNIT && lane == j) xi = xi / d;
            const V xj = __shfl(xi, j);
            if(lane < j) xi = fma(-e[j], xj, xi);
          }
        }
      }
      if(lane < w) sx[i] = xi;
    }
    __syncthreads();
    const int r0 = LOWER ? c0 + w : 0;
    const int r1 = LOWER ? k : c0;
    if(w == 64) {
      // full chunk: the row's 64 entries are loaded before the first use, so
      // their latencies overlap
      for(int r = r0 + t; r < r1; r += blockDim.x) {
        MT tv[64];
#pragma unroll
        for(int j = 0; j < 64; ++j)
          tv[j] = TRANS ? T[c0 + j + static_cast<size_t>(r) * ldt]
                        : T[r + static_cast<size_t>(c0 + j) * ldt];
        V acc = V(0);
#pragma unroll
        for(int j = 0; j < 64; ++j) acc = fma(to_v<V>(tv[j]), sx[c0 + j], acc);
        sx[r] -= acc;
      }
    } else {
      for(int r = r0 + t; r < r1; r += blockDim.x) {
        V acc = V(0);
        for(int j = 0; j < w; ++j)
          acc = fma(op_at<TRANS, V>(T, ldt, r, c0 + j), sx[c0 + j], acc);
        sx[r] -= acc;
      }
    }
    __syncthreads();
  }
  for(int i = t; i < k; i += blockDim.x) xc[i] = sx[i];
}

}  // namespace trs
}  // namespace hplx
