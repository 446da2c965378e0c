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

// Split-K partial products of C (m x n) -= A^T B for the transposed solve,
// where A (K x m) and B (K x n) are column-major with K far larger than m and
// n: split s sums k in [s * kc, min(K, (s + 1) * kc)) for one 64 x 64 tile of
// C into P + s * m * n (column-major, leading dimension m).  Each of the 256
// threads owns a 4 x 4 block of the tile; 32-deep K steps through LDS, each
// column of A and B read as 256 contiguous bytes.
__global__ __launch_bounds__(256) void tn_partial_kernel(
    int K, int m, int n, int kc, const double* __restrict__ A, int lda,
    const double* __restrict__ B, int ldb, double* __restrict__ P, int tiles_m,
    int tiles_n) {
  __shared__ double sa[32][65];
  __shared__ double sb[32][65];
  const int tiles = tiles_m * tiles_n, split = blockIdx.x / tiles;
  const int tile = blockIdx.x - split * tiles;
  const int m0 = (tile % tiles_m) * 64, n0 = (tile / tiles_m) * 64;
  const int k0 = split * kc, k1 = min(K, k0 + kc);
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  double acc[4][4] = {};
  for(int k = k0; k < k1; k += 32) {
#pragma unroll
    for(int r = 0; r < 8; ++r) {
      const int e = t + 256 * r, kk = e & 31, c = e >> 5;
      const bool kin = k + kk < k1;
      sa[kk][c] = kin && m0 + c < m
                      ? A[(k + kk) + static_cast<size_t>(m0 + c) * lda] : 0.0;
      sb[kk][c] = kin && n0 + c < n
                      ? B[(k + kk) + static_cast<size_t>(n0 + c) * ldb] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for(int kk = 0; kk < 32; ++kk) {
      double av[4], bv[4];
#pragma unroll
      for(int i = 0; i < 4; ++i) av[i] = sa[kk][ty * 4 + i];
#pragma unroll
      for(int j = 0; j < 4; ++j) bv[j] = sb[kk][tx * 4 + j];
#pragma unroll
      for(int i = 0; i < 4; ++i)
#pragma unroll
        for(int j = 0; j < 4; ++j) acc[i][j] = fma(av[i], bv[j], acc[i][j]);
    }
    __syncthreads();
  }
  double* ps = P + static_cast<size_t>(split) * m * n;
#pragma unroll
  for(int i = 0; i < 4; ++i)
#pragma unroll
    for(int j = 0; j < 4; ++j) {
      const int row = m0 + ty * 4 + i, col = n0 + tx * 4 + j;
      if(row < m && col < n) ps[row + static_cast<size_t>(col) * m] = acc[i][j];
    }
}

// C[i, j] -= sum over the splits of P, in split order (the result does not
// depend on how the splits were scheduled).
__global__ __launch_bounds__(256) void tn_reduce_kernel(
    int m, int n, int splits, const double* __restrict__ P,
    double* __restrict__ C, int ldc) {
  const long e = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
  if(e >= static_cast<long>(m) * n) return;
  const int i = static_cast<int>(e % m), j = static_cast<int>(e / m);
  double sum = 0.0;
  for(int sp = 0; sp < splits; ++sp) sum += P[static_cast<size_t>(sp) * m * n + e];
  C[i + static_cast<size_t>(j) * ldc] -= sum;
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

hipError_t launch_tn_sub(hipStream_t s, int K, int m, int n, const double* A,
                         int lda, const double* B, int ldb, double* C, int ldc,
                         double* work, size_t work_doubles) {
  if(K <= 0 || m <= 0 || n <= 0) return hipSuccess;
  const size_t mn = static_cast<size_t>(m) * n;
  if(mn > work_doubles) return hipErrorInvalidValue;
  const int tiles_m = (m + 63) / 64, tiles_n = (n + 63) / 64;
  // about 256 k per split, at most 64 splits and what the work buffer holds
  long splits = (K + 255) / 256;
  splits = splits < 64 ? splits : 64;
  const long cap = static_cast<long>(work_doubles / mn);
  splits = splits < cap ? splits : cap;
  const int kc = static_cast<int>(((K + splits - 1) / splits + 31) / 32 * 32);
  splits = (K + kc - 1) / kc;
  const long grid = static_cast<long>(tiles_m) * tiles_n * splits;
  if(grid > 0x7fffffffL || static_cast<long>(mn) > 0x7fffffffL * 256L)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(tn_partial_kernel, dim3(static_cast<unsigned>(grid)),
                     dim3(256), 0, s, K, m, n, kc, A, lda, B, ldb, work,
                     tiles_m, tiles_n);
  hipLaunchKernelGGL(tn_reduce_kernel,
                     dim3(static_cast<unsigned>((mn + 255) / 256)), dim3(256),
                     0, s, m, n, static_cast<int>(splits), work, C, ldc);
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
