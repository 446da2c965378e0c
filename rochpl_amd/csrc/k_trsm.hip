// Triangular solves on one diagonal block (k <= 1024), one workgroup per
// right-hand-side column:  op(T) x = b in place, op(T) = T or T^T.
//
// Used for the unit-lower L11 solves of the panel recursion and of the
// trailing update (rocblas_dtrsm at src/pfact/HPL_pdrpanrlT.cpp:124-135 and
// src/pgesv/HPL_pdupdateTT.cpp:95-106), for the explicit inverse of L11, for
// the diagonal blocks of the back-solve (rocblas_dtrsv at
// src/pgesv/HPL_pdtrsv_device.cpp:170-178,279-287), and for the transposed
// solves of the adjoint.
//
// x lives in LDS.  The block is processed in 64-row chunks along the sweep
// direction: wave 0 solves the chunk's 64 x 64 triangle with the chunk's rows
// of op(T) in registers (one row per lane) and x broadcast by lane shuffles;
// then every thread applies the chunk's 64 columns to the rows still to be
// solved.  One workgroup barrier per 64 rows.
#include "hplx_internal.hpp"

namespace hplx {
namespace {

template <bool TRANS>
__device__ inline double op_at(const double* T, int ldt, int i, int j) {
  return TRANS ? T[j + static_cast<size_t>(i) * ldt]
               : T[i + static_cast<size_t>(j) * ldt];
}

// LOWER: op(T) is lower triangular (forward sweep), else upper (backward).
// UNIT: unit diagonal (not read).  IDENTITY: b is column `col` of I.
template <bool LOWER, bool TRANS, bool UNIT, bool IDENTITY>
__global__ __launch_bounds__(256) void tri_solve_kernel(
    int k, const double* __restrict__ T, int ldt, double* __restrict__ X,
    int ldx) {
  __shared__ double sx[kMaxNB];
  const int col  = blockIdx.x;
  const int t    = threadIdx.x;
  const int lane = t & 63;
  double*   xc   = X + static_cast<size_t>(col) * ldx;
  for(int i = t; i < k; i += blockDim.x)
    sx[i] = IDENTITY ? (i == col ? 1.0 : 0.0) : xc[i];
  __syncthreads();
  const int nchunk = (k + 63) / 64;
  // with the identity and a forward sweep, rows above `col` stay zero
  const int first = (LOWER && IDENTITY) ? col / 64 : 0;
  for(int s = first; s < nchunk; ++s) {
    const int c0 = LOWER ? s * 64 : (nchunk - 1 - s) * 64;
    const int w  = min(64, k - c0);
    if(t < 64) {
      const int i  = c0 + lane;
      double    xi = lane < w ? sx[i] : 0.0;
      const double d = (UNIT || lane >= w) ? 1.0 : op_at<TRANS>(T, ldt, i, i);
      double    e[64];
#pragma unroll
      for(int j = 0; j < 64; ++j) {
        const bool in = j < w && lane < w && (LOWER ? j < lane : j > lane);
        e[j]          = in ? op_at<TRANS>(T, ldt, i, c0 + j) : 0.0;
      }
      if(LOWER) {
#pragma unroll
        for(int j = 0; j < 64; ++j) {
          if(j < w) {
            if(!UNIT && lane == j) xi = xi / d;
            const double xj = __shfl(xi, j);
            if(lane > j) xi = fma(-e[j], xj, xi);
          }
        }
      } else {
#pragma unroll
        for(int j = 63; j >= 0; --j) {
          if(j < w) {
            if(!UNIT && lane == j) xi = xi / d;
            const double xj = __shfl(xi, j);
            if(lane < j) xi = fma(-e[j], xj, xi);
          }
        }
      }
      if(lane < w) sx[i] = xi;
    }
    __syncthreads();
    const int r0 = LOWER ? c0 + w : 0;
    const int r1 = LOWER ? k : c0;
    for(int r = r0 + t; r < r1; r += blockDim.x) {
      double acc = 0.0;
      for(int j = 0; j < w; ++j)
        acc = fma(op_at<TRANS>(T, ldt, r, c0 + j), sx[c0 + j], acc);
      sx[r] -= acc;
    }
    __syncthreads();
  }
  for(int i = t; i < k; i += blockDim.x) xc[i] = sx[i];
}

}  // namespace

hipError_t launch_tri_solve(hipStream_t s, bool lower, bool trans, bool unit,
                            int k, int nrhs, const double* T, int ldt,
                            double* X, int ldx) {
  if(k <= 0 || nrhs <= 0) return hipSuccess;
  if(k > kMaxNB) return hipErrorInvalidValue;
  const dim3 g(nrhs), b(256);
  if(lower && !trans && unit)
    hipLaunchKernelGGL((tri_solve_kernel<true, false, true, false>), g, b, 0, s, k, T, ldt, X, ldx);
  else if(!lower && !trans && !unit)
    hipLaunchKernelGGL((tri_solve_kernel<false, false, false, false>), g, b, 0, s, k, T, ldt, X, ldx);
  else if(lower && trans && !unit)  // U^T: U stored upper
    hipLaunchKernelGGL((tri_solve_kernel<true, true, false, false>), g, b, 0, s, k, T, ldt, X, ldx);
  else if(!lower && trans && unit)  // L^T: L stored unit lower
    hipLaunchKernelGGL((tri_solve_kernel<false, true, true, false>), g, b, 0, s, k, T, ldt, X, ldx);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_linv(hipStream_t s, int jb, const double* L, int ldl,
                       double* Linv, int ldi) {
  if(jb <= 0) return hipSuccess;
  if(jb > kMaxNB) return hipErrorInvalidValue;
  hipLaunchKernelGGL((tri_solve_kernel<true, false, true, true>), dim3(jb),
                     dim3(256), 0, s, jb, L, ldl, Linv, ldi);
  return hipGetLastError();
}

hipError_t launch_trsm_colsweep(hipStream_t s, int k, int n, const double* L,
                                int ldl, double* B, int ldb) {
  return launch_tri_solve(s, true, false, true, k, n, L, ldl, B, ldb);
}

}  // namespace hplx
