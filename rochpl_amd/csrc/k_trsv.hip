// Back-solve U x = b' (HPL_pdtrsv at P=Q=1, src/pgesv/HPL_pdtrsv_device.cpp:29-346).
//
// The reference walks the NB-blocks from the bottom with rocblas_dtrsv on the
// diagonal block and two rocblas_dgemv calls for the rows above
// (:169-180,195-332).  Here each block is one single-workgroup blocked solve
// of the diagonal block (k_trsm.hip) followed by one HBM-bound GEMV over all
// rows above,
// with the block's columns split across the 4 waves of a workgroup so that
// enough loads are in flight.  Every element of the upper triangle is read
// exactly once.
#include "mxp.hpp"

namespace hplx {
namespace {

template <typename MT>
__device__ inline double to_d(MT x) {
  if constexpr(sizeof(MT) == 2)
    return static_cast<double>(__half2float(x));
  else
    return static_cast<double>(x);
}

// y[0:m] -= U[0:m, 0:kb] * xs[0:kb] in FP64, U of element type MT (double, or
// the FP32 / FP16 factors of the mixed-precision solves).  Workgroup = 64
// rows; wave w takes a quarter of the columns; partial sums meet in LDS.
// xback (optional): workgroup 0 also copies xs[0:kb] there.
template <typename MT>
__global__ __launch_bounds__(256) void gemv_sub_kernel(
    int m, int kb, const MT* __restrict__ U, int ldu,
    const double* __restrict__ xs, double* __restrict__ y,
    double* __restrict__ xback) {
  __shared__ double sxs[kMaxNB];
  __shared__ double part[4][64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for(int i = t; i < kb; i += blockDim.x) sxs[i] = xs[i];
  __syncthreads();
  if(xback && blockIdx.x == 0)
    for(int i = t; i < kb; i += blockDim.x) xback[i] = sxs[i];
  const int row = blockIdx.x * 64 + lane;
  const int ck  = (kb + 3) / 4;
  const int c0  = wave * ck;
  const int c1  = min(kb, c0 + ck);
  double    acc = 0.0;
  if(row < m) {
    const MT* u = U + row + static_cast<size_t>(c0) * ldu;
    int c = c0;
    for(; c + 4 <= c1; c += 4) {
      const double u0 = to_d(u[0]);
      const double u1 = to_d(u[static_cast<size_t>(ldu)]);
      const double u2 = to_d(u[2 * static_cast<size_t>(ldu)]);
      const double u3 = to_d(u[3 * static_cast<size_t>(ldu)]);
      acc = fma(u0, sxs[c], acc);
      acc = fma(u1, sxs[c + 1], acc);
      acc = fma(u2, sxs[c + 2], acc);
      acc = fma(u3, sxs[c + 3], acc);
      u += 4 * static_cast<size_t>(ldu);
    }
    for(; c < c1; ++c) {
      acc = fma(to_d(u[0]), sxs[c], acc);
      u += ldu;
    }
  }
  part[wave][lane] = acc;
  __syncthreads();
  if(wave == 0 && row < m)
    y[row] -= (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
}

// y[0:k] = Inv * x[0:k] for a k x k FP64 matrix (column-major, leading
// dimension k), k <= kMaxNB.  Workgroup = 64 rows with 16 waves, each wave a
// sixteenth of the columns in chunks of 32 whose loads are all in flight at
// once (the matrix is small: the kernel is latency bound).
__global__ __launch_bounds__(1024) void inv_gemv_kernel(
    int k, const double* __restrict__ Inv, const double* __restrict__ x,
    double* __restrict__ y) {
  __shared__ double sx[kMaxNB];
  __shared__ double part[16][64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for(int i = t; i < k; i += blockDim.x) sx[i] = x[i];
  __syncthreads();
  const int row = blockIdx.x * 64 + lane;
  const int ck  = (k + 15) / 16;
  const int c0  = wave * ck;
  const int c1  = min(k, c0 + ck);
  double    acc = 0.0;
  if(row < k) {
    for(int cc = c0; cc < c1; cc += 32) {
      const double* p = Inv + row + static_cast<size_t>(cc) * k;
      double        v[32];
#pragma unroll
      for(int j = 0; j < 32; ++j)
        v[j] = cc + j < c1 ? p[static_cast<size_t>(j) * k] : 0.0;
#pragma unroll
      for(int j = 0; j < 32; ++j)
        if(cc + j < c1) acc = fma(v[j], sx[cc + j], acc);
    }
  }
  part[wave][lane] = acc;
  __syncthreads();
  if(wave == 0 && row < k) {
    double sum = 0.0;
#pragma unroll
    for(int w = 0; w < 16; ++w) sum += part[w][lane];
    y[row] = sum;
  }
}

template <typename MT>
hipError_t gemv_sub_t(hipStream_t s, int m, int k, const MT* M, int ldm,
                      const double* x, double* y, double* xback) {
  if(m <= 0 || k <= 0) return hipSuccess;
  if(k > kMaxNB) return hipErrorInvalidValue;
  hipLaunchKernelGGL((gemv_sub_kernel<MT>), dim3((m + 63) / 64), dim3(256), 0,
                     s, m, k, M, ldm, x, y, xback);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_gemv_sub(hipStream_t s, int m, int k, const float* M,
                           int ldm, const double* x, double* y,
                           double* xback) {
  return gemv_sub_t(s, m, k, M, ldm, x, y, xback);
}

hipError_t launch_gemv_sub(hipStream_t s, int m, int k, const __half* M,
                           int ldm, const double* x, double* y,
                           double* xback) {
  return gemv_sub_t(s, m, k, M, ldm, x, y, xback);
}

hipError_t launch_gemv_sub(hipStream_t s, int m, int k, const double* M,
                           int ldm, const double* x, double* y,
                           double* xback) {
  return gemv_sub_t(s, m, k, M, ldm, x, y, xback);
}

hipError_t launch_inv_gemv(hipStream_t s, int k, const double* Inv,
                           const double* x, double* y) {
  if(k <= 0) return hipSuccess;
  if(k > kMaxNB) return hipErrorInvalidValue;
  hipLaunchKernelGGL(inv_gemv_kernel, dim3((k + 63) / 64), dim3(1024), 0, s, k,
                     Inv, x, y);
  return hipGetLastError();
}

hipError_t launch_trsv(hipStream_t s, int n, int nb, const double* A, int lda,
                       double* x) {
  if(n <= 0) return hipSuccess;
  if(nb <= 0 || nb > kMaxNB) nb = nb <= 0 ? 1 : kMaxNB;
  const int nblk = (n + nb - 1) / nb;
  for(int J = nblk - 1; J >= 0; --J) {
    const int s0 = J * nb;
    const int kb = (n - s0) < nb ? (n - s0) : nb;
    hipError_t e = launch_tri_solve(s, false, false, false, kb, 1,
                                    A + s0 + static_cast<size_t>(s0) * lda,
                                    lda, x + s0, n);
    if(e != hipSuccess) return e;
    if(s0 > 0)
      hipLaunchKernelGGL(gemv_sub_kernel<double>, dim3((s0 + 63) / 64),
                         dim3(256), 0, s, s0, kb,
                         A + static_cast<size_t>(s0) * lda, lda, x + s0, x,
                         nullptr);
  }
  return hipGetLastError();
}

}  // namespace hplx
