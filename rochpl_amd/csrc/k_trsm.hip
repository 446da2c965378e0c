// Triangular solves on one diagonal block (k <= 1024), one workgroup per
// right-hand-side column:  op(T) x = b in place, op(T) = T or T^T.
//
// Used for the unit-lower L11 solves of the panel recursion and of the
// trailing update (rocblas_dtrsm at src/pfact/HPL_pdrpanrlT.cpp:124-135 and
// src/pgesv/HPL_pdupdateTT.cpp:95-106), for the explicit inverse of L11, for
// the diagonal blocks of the back-solve (rocblas_dtrsv at
// src/pgesv/HPL_pdtrsv_device.cpp:170-178,279-287), for the transposed
// solves of the adjoint, and (diag_inv_kernel) for the FP32 or FP64
// inverses of the diagonal blocks of the mixed-precision factors.
//
// x lives in LDS.  The block is processed in 64-row chunks along the sweep
// direction: wave 0 solves the chunk's 64 x 64 triangle with the chunk's rows
// of op(T) in registers (one row per lane) and x broadcast by lane shuffles;
// then every thread applies the chunk's 64 columns to the rows still to be
// solved, with all 64 loads of a row issued before the first product (the
// block is latency bound: one workgroup per column).  One workgroup barrier
// per 64 rows.
#include "mxp.hpp"
#include "tri_solve.hpp"

namespace hplx {
namespace {

using namespace trs;

template <typename MT, typename V, bool LOWER, bool TRANS, bool UNIT,
          bool IDENTITY>
__global__ __launch_bounds__(256) void tri_solve_kernel(
    int k, const MT* __restrict__ T, int ldt, V* __restrict__ X, int ldx) {
  tri_solve_col<MT, V, LOWER, TRANS, UNIT, IDENTITY>(
      k, T, ldt, X + static_cast<size_t>(blockIdx.x) * ldx, blockIdx.x);
}

// Inverses (in V: FP64 or FP32) of consecutive kb x kb diagonal blocks of T:
// block J (blockIdx.y) starts at T[J kb, J kb]; column blockIdx.x of its
// inverse goes to Inv + (J kb + col) kb (each inverse column-major with
// leading dimension kb).  LOWER: unit lower triangle; else the upper
// triangle.
template <typename MT, typename V, bool LOWER>
__global__ __launch_bounds__(256) void diag_inv_kernel(int kb,
                                                       const MT* __restrict__ T,
                                                       int ldt,
                                                       V* __restrict__ Inv) {
  const size_t J = blockIdx.y;
  tri_solve_col<MT, V, LOWER, false, LOWER, true>(
      kb, T + J * kb * (1 + static_cast<size_t>(ldt)), ldt,
      Inv + (J * kb + blockIdx.x) * kb, blockIdx.x);
}

template <typename V>
hipError_t tri_solve_t(hipStream_t s, bool lower, bool trans, bool unit, int k,
                       int nrhs, const V* T, int ldt, V* X, int ldx) {
  if(k <= 0 || nrhs <= 0) return hipSuccess;
  if(k > kMaxNB) return hipErrorInvalidValue;
  const dim3 g(nrhs), b(256);
  if(lower && !trans && unit)
    hipLaunchKernelGGL((tri_solve_kernel<V, V, true, false, true, false>), g, b, 0, s, k, T, ldt, X, ldx);
  else if(!lower && !trans && !unit)
    hipLaunchKernelGGL((tri_solve_kernel<V, V, false, false, false, false>), g, b, 0, s, k, T, ldt, X, ldx);
  else if(lower && trans && !unit)  // U^T: U stored upper
    hipLaunchKernelGGL((tri_solve_kernel<V, V, true, true, false, false>), g, b, 0, s, k, T, ldt, X, ldx);
  else if(!lower && trans && unit)  // L^T: L stored unit lower
    hipLaunchKernelGGL((tri_solve_kernel<V, V, false, true, true, false>), g, b, 0, s, k, T, ldt, X, ldx);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

template <typename V>
hipError_t linv_t(hipStream_t s, int jb, const V* L, int ldl, V* Linv,
                  int ldi) {
  if(jb <= 0) return hipSuccess;
  if(jb > kMaxNB) return hipErrorInvalidValue;
  hipLaunchKernelGGL((tri_solve_kernel<V, V, true, false, true, true>), dim3(jb),
                     dim3(256), 0, s, jb, L, ldl, Linv, ldi);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_tri_solve(hipStream_t s, bool lower, bool trans, bool unit,
                            int k, int nrhs, const double* T, int ldt,
                            double* X, int ldx) {
  return tri_solve_t(s, lower, trans, unit, k, nrhs, T, ldt, X, ldx);
}

hipError_t launch_tri_solve(hipStream_t s, bool lower, bool trans, bool unit,
                            int k, int nrhs, const float* T, int ldt, float* X,
                            int ldx) {
  return tri_solve_t(s, lower, trans, unit, k, nrhs, T, ldt, X, ldx);
}

hipError_t launch_linv(hipStream_t s, int jb, const double* L, int ldl,
                       double* Linv, int ldi) {
  return linv_t(s, jb, L, ldl, Linv, ldi);
}

hipError_t launch_linv(hipStream_t s, int jb, const float* L, int ldl,
                       float* Linv, int ldi) {
  return linv_t(s, jb, L, ldl, Linv, ldi);
}

template <typename MT, typename V>
hipError_t diag_inverse_t(hipStream_t s, bool lower, int nblk, int kb,
                          const MT* T, int ldt, V* Inv) {
  if(nblk <= 0 || kb <= 0) return hipSuccess;
  if(kb > kMaxNB || nblk > 65535) return hipErrorInvalidValue;
  const dim3 g(kb, nblk);
  if(lower)
    hipLaunchKernelGGL((diag_inv_kernel<MT, V, true>), g, dim3(256), 0, s, kb, T, ldt, Inv);
  else
    hipLaunchKernelGGL((diag_inv_kernel<MT, V, false>), g, dim3(256), 0, s, kb, T, ldt, Inv);
  return hipGetLastError();
}

hipError_t launch_diag_inverse(hipStream_t s, bool lower, int nblk, int kb,
                               const float* T, int ldt, double* Inv) {
  return diag_inverse_t(s, lower, nblk, kb, T, ldt, Inv);
}

hipError_t launch_diag_inverse(hipStream_t s, bool lower, int nblk, int kb,
                               const __half* T, int ldt, double* Inv) {
  return diag_inverse_t(s, lower, nblk, kb, T, ldt, Inv);
}

hipError_t launch_diag_inverse(hipStream_t s, bool lower, int nblk, int kb,
                               const float* T, int ldt, float* Inv) {
  return diag_inverse_t(s, lower, nblk, kb, T, ldt, Inv);
}

hipError_t launch_diag_inverse(hipStream_t s, bool lower, int nblk, int kb,
                               const __half* T, int ldt, float* Inv) {
  return diag_inverse_t(s, lower, nblk, kb, T, ldt, Inv);
}

hipError_t launch_trsm_colsweep(hipStream_t s, int k, int n, const double* L,
                                int ldl, double* B, int ldb) {
  return launch_tri_solve(s, true, false, true, k, n, L, ldl, B, ldb);
}

hipError_t launch_trsm_colsweep(hipStream_t s, int k, int n, const float* L,
                                int ldl, float* B, int ldb) {
  return launch_tri_solve(s, true, false, true, k, n, L, ldl, B, ldb);
}

}  // namespace hplx
