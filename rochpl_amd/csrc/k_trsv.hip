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

typedef double ud4 __attribute__((ext_vector_type(4)));

template <typename V, typename MT>
__device__ inline V to_v(MT x) {
  if constexpr(sizeof(MT) == 2)
    return static_cast<V>(__half2float(x));
  else
    return static_cast<V>(x);
}

// y[0:m] -= U[0:m, 0:kb] * xs[0:kb] in the vectors' type V (FP64, or FP32 in
// the mixed-precision correction solves), U of element type MT (double, or
// the FP32 / FP16 factors of the mixed forms).  Workgroup = 64 rows; wave w
// takes a quarter of the columns; partial sums meet in LDS.  xback
// (optional): workgroup 0 also copies xs[0:kb] there.
template <typename MT, typename V>
__global__ __launch_bounds__(256) void gemv_sub_kernel(
    int m, int kb, const MT* __restrict__ U, int ldu,
    const V* __restrict__ xs, V* __restrict__ y,
    V* __restrict__ xback) {
  __shared__ V sxs[kMaxNB];
  __shared__ V part[4][64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for(int i = t; i < kb; i += blockDim.x) sxs[i] = xs[i];
  __syncthreads();
  if(xback && blockIdx.x == 0)
    for(int i = t; i < kb; i += blockDim.x) xback[i] = sxs[i];
  const int row = blockIdx.x * 64 + lane;
  const int ck  = (kb + 3) / 4;
  const int c0  = wave * ck;
  const int c1  = min(kb, c0 + ck);
  V         acc = 0;
  if(row < m) {
    const MT* u = U + row + static_cast<size_t>(c0) * ldu;
    int c = c0;
    for(; c + 4 <= c1; c += 4) {
      const V u0 = to_v<V>(u[0]);
      const V u1 = to_v<V>(u[static_cast<size_t>(ldu)]);
      const V u2 = to_v<V>(u[2 * static_cast<size_t>(ldu)]);
      const V u3 = to_v<V>(u[3 * static_cast<size_t>(ldu)]);
      acc = fma(u0, sxs[c], acc);
      acc = fma(u1, sxs[c + 1], acc);
      acc = fma(u2, sxs[c + 2], acc);
      acc = fma(u3, sxs[c + 3], acc);
      u += 4 * static_cast<size_t>(ldu);
    }
    for(; c < c1; ++c) {
      acc = fma(to_v<V>(u[0]), sxs[c], acc);
      u += ldu;
    }
  }
  part[wave][lane] = acc;
  __syncthreads();
  if(wave == 0 && row < m)
    y[row] -= (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
}

// y[0:k] = Inv * x[0:k] for a k x k matrix of type V (FP64 or FP32;
// column-major, leading dimension k), k <= kMaxNB.  Workgroup = 64 rows with
// 16 waves, each wave a sixteenth of the columns in chunks of 32 whose loads
// are all in flight at once (the matrix is small: the kernel is latency
// bound).
template <typename V>
__global__ __launch_bounds__(1024) void inv_gemv_kernel(
    int k, const V* __restrict__ Inv, const V* __restrict__ x,
    V* __restrict__ y) {
  __shared__ V sx[kMaxNB];
  __shared__ V part[16][64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for(int i = t; i < k; i += blockDim.x) sx[i] = x[i];
  __syncthreads();
  const int row = blockIdx.x * 64 + lane;
  const int ck  = (k + 15) / 16;
  const int c0  = wave * ck;
  const int c1  = min(k, c0 + ck);
  V         acc = 0;
  if(row < k) {
    for(int cc = c0; cc < c1; cc += 32) {
      const V* p = Inv + row + static_cast<size_t>(cc) * k;
      V             v[32];
#pragma unroll
      for(int j = 0; j < 32; ++j)
        v[j] = cc + j < c1 ? p[static_cast<size_t>(j) * k] : V(0);
#pragma unroll
      for(int j = 0; j < 32; ++j)
        if(cc + j < c1) acc = fma(v[j], sx[cc + j], acc);
    }
  }
  part[wave][lane] = acc;
  __syncthreads();
  if(wave == 0 && row < k) {
    V sum = 0;
#pragma unroll
    for(int w = 0; w < 16; ++w) sum += part[w][lane];
    y[row] = sum;
  }
}

// C (m x n) -= op(A) op(B) for the solves with many right-hand sides:
// op(A) = A^T with A (K x m) when TA, else A (m x K); op(B) = B^T with
// B (n x K) when TB, else B (K x n); all column-major.  BETA0: C = -op(A)
// op(B) instead (the adjoint's rank-nrhs gradA).
// One 64 x 64 tile of the result per workgroup over
// k in [s * kc, min(K, (s + 1) * kc)).  DIRECT (one split): C is updated in
// place; otherwise split s writes its partial tile to P + s * m * n
// (column-major, leading dimension m) for tn_reduce_kernel.  The tile is
// 4 waves of 32 x 32 (2 x 2 blocks of V_MFMA_F64_16X16X4_F64); 32-deep K
// steps go global -> registers -> LDS (k-major, as the trailing update's
// stages), the next step's loads in flight during the current step's MFMAs.
template <bool DIRECT, bool TA, bool TB = false, bool BETA0 = false>
__global__ __launch_bounds__(256) void tn_kernel(
    int K, int m, int n, int kc, const double* __restrict__ A, int lda,
    const double* __restrict__ B, int ldb, double* __restrict__ out, int ldo,
    int tiles_m, int tiles_n) {
  __shared__ double sa[32][65];
  __shared__ double sb[32][65];
  const int tiles = tiles_m * tiles_n, split = blockIdx.x / tiles;
  const int tile = blockIdx.x - split * tiles;
  const int m0 = (tile % tiles_m) * 64, n0 = (tile / tiles_m) * 64;
  const int k0 = split * kc, k1 = min(K, k0 + kc);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm0 = (wave & 1) * 32, wn0 = (wave >> 1) * 32;
  const int l16 = lane & 15, lk = lane >> 4;
  double ra[8], rb[8];
  // element r of a thread's share: k row kk, column c of the 32 x 64 step
  // (k-contiguous for A^T and B, m-contiguous for A)
  auto akk = [&](int e) { return TA ? (e & 31) : (e >> 6); };
  auto ac  = [&](int e) { return TA ? (e >> 5) : (e & 63); };
  auto load = [&](int k) {
#pragma unroll
    for(int r = 0; r < 8; ++r) {
      const int e = t + 256 * r;
      const int ka = akk(e), ca = ac(e);
      const int kb = TB ? (e >> 6) : (e & 31), cb = TB ? (e & 63) : (e >> 5);
      ra[r] = k + ka < k1 && m0 + ca < m
                  ? (TA ? A[(k + ka) + static_cast<size_t>(m0 + ca) * lda]
                        : A[(m0 + ca) + static_cast<size_t>(k + ka) * lda])
                  : 0.0;
      rb[r] = k + kb < k1 && n0 + cb < n
                  ? (TB ? B[(n0 + cb) + static_cast<size_t>(k + kb) * ldb]
                        : B[(k + kb) + static_cast<size_t>(n0 + cb) * ldb])
                  : 0.0;
    }
  };
  ud4 acc[2][2];
#pragma unroll
  for(int i = 0; i < 2; ++i)
#pragma unroll
    for(int j = 0; j < 2; ++j) acc[i][j] = ud4{0.0, 0.0, 0.0, 0.0};
  if(k0 < k1) load(k0);
  for(int k = k0; k < k1; k += 32) {
#pragma unroll
    for(int r = 0; r < 8; ++r) {
      const int e = t + 256 * r;
      sa[akk(e)][ac(e)] = ra[r];
      if(TB)
        sb[e >> 6][e & 63] = rb[r];
      else
        sb[e & 31][e >> 5] = rb[r];
    }
    __syncthreads();
    if(k + 32 < k1) load(k + 32);
#pragma unroll
    for(int ks = 0; ks < 8; ++ks) {
      double fa[2], fb[2];
#pragma unroll
      for(int i = 0; i < 2; ++i) fa[i] = sa[ks * 4 + lk][wm0 + i * 16 + l16];
#pragma unroll
      for(int j = 0; j < 2; ++j) fb[j] = sb[ks * 4 + lk][wn0 + j * 16 + l16];
#pragma unroll
      for(int i = 0; i < 2; ++i)
#pragma unroll
        for(int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[j], fa[i],
                                                           acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // acc[i][j][r]: row wm0 + 16 i + l16, column wn0 + 16 j + 4 r + lk
  double* po = DIRECT ? out : out + static_cast<size_t>(split) * m * n;
#pragma unroll
  for(int i = 0; i < 2; ++i) {
    const int row = m0 + wm0 + i * 16 + l16;
    if(row >= m) continue;
#pragma unroll
    for(int j = 0; j < 2; ++j)
#pragma unroll
      for(int r = 0; r < 4; ++r) {
        const int col = n0 + wn0 + j * 16 + 4 * r + lk;
        if(col >= n) continue;
        if(DIRECT && BETA0)
          po[row + static_cast<size_t>(col) * ldo] = -acc[i][j][r];
        else if(DIRECT)
          po[row + static_cast<size_t>(col) * ldo] -= acc[i][j][r];
        else
          po[row + static_cast<size_t>(col) * m] = acc[i][j][r];
      }
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

template <typename MT, typename V>
hipError_t gemv_sub_t(hipStream_t s, int m, int k, const MT* M, int ldm,
                      const V* x, V* y, V* xback) {
  if(m <= 0 || k <= 0) return hipSuccess;
  if(k > kMaxNB) return hipErrorInvalidValue;
  hipLaunchKernelGGL((gemv_sub_kernel<MT, V>), dim3((m + 63) / 64), dim3(256), 0,
                     s, m, k, M, ldm, x, y, xback);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_gemv_sub(hipStream_t s, int m, int k, const float* M,
                           int ldm, const double* x, double* y,
                           double* xback) {
  return gemv_sub_t(s, m, k, M, ldm, x, y, xback);
}

hipError_t launch_gemv_sub(hipStream_t s, int m, int k, const float* M,
                           int ldm, const float* x, float* y, float* xback) {
  return gemv_sub_t(s, m, k, M, ldm, x, y, xback);
}

hipError_t launch_gemv_sub(hipStream_t s, int m, int k, const __half* M,
                           int ldm, const float* x, float* y, float* xback) {
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

template <typename V>
hipError_t inv_gemv_t(hipStream_t s, int k, const V* Inv, const V* x, V* y) {
  if(k <= 0) return hipSuccess;
  if(k > kMaxNB) return hipErrorInvalidValue;
  hipLaunchKernelGGL(inv_gemv_kernel<V>, dim3((k + 63) / 64), dim3(1024), 0, s,
                     k, Inv, x, y);
  return hipGetLastError();
}

hipError_t launch_inv_gemv(hipStream_t s, int k, const double* Inv,
                           const double* x, double* y) {
  return inv_gemv_t(s, k, Inv, x, y);
}

hipError_t launch_inv_gemv(hipStream_t s, int k, const float* Inv,
                           const float* x, float* y) {
  return inv_gemv_t(s, k, Inv, x, y);
}

hipError_t launch_tn_sub(hipStream_t s, int K, int m, int n, const double* A,
                         int lda, const double* B, int ldb, double* C, int ldc,
                         double* work, size_t work_doubles) {
  if(K <= 0 || m <= 0 || n <= 0) return hipSuccess;
  const size_t mn = static_cast<size_t>(m) * n;
  const int tiles_m = (m + 63) / 64, tiles_n = (n + 63) / 64;
  const long tiles = static_cast<long>(tiles_m) * tiles_n;
  // split K when the tiles alone leave the GPU short of work: about 256 k
  // per split, at most 64 splits and what the work buffer holds
  long splits = tiles >= 256 ? 1 : (K + 255) / 256;
  splits = splits < 64 ? splits : 64;
  const long cap = work ? static_cast<long>(work_doubles / mn) : 0;
  splits = splits < cap ? splits : (cap < 1 ? 1 : cap);
  const int kc = static_cast<int>(((K + splits - 1) / splits + 31) / 32 * 32);
  splits = (K + kc - 1) / kc;
  if(tiles * splits > 0x7fffffffL) return hipErrorInvalidValue;
  const dim3 grid(static_cast<unsigned>(tiles * splits));
  if(splits == 1) {
    hipLaunchKernelGGL((tn_kernel<true, true>), grid, dim3(256), 0, s, K, m,
                       n, kc, A, lda, B, ldb, C, ldc, tiles_m, tiles_n);
    return hipGetLastError();
  }
  hipLaunchKernelGGL((tn_kernel<false, true>), grid, dim3(256), 0, s, K, m,
                     n, kc, A, lda, B, ldb, work, m, tiles_m, tiles_n);
  hipLaunchKernelGGL(tn_reduce_kernel,
                     dim3(static_cast<unsigned>((mn + 255) / 256)), dim3(256),
                     0, s, m, n, static_cast<int>(splits), work, C, ldc);
  return hipGetLastError();
}

hipError_t launch_nn_sub(hipStream_t s, int K, int m, int n, const double* A,
                         int lda, const double* B, int ldb, double* C,
                         int ldc) {
  if(K <= 0 || m <= 0 || n <= 0) return hipSuccess;
  const int  tiles_m = (m + 63) / 64, tiles_n = (n + 63) / 64;
  const long tiles   = static_cast<long>(tiles_m) * tiles_n;
  if(tiles > 0x7fffffffL || K > 0x7fffffff - 32) return hipErrorInvalidValue;
  hipLaunchKernelGGL((tn_kernel<true, false>),
                     dim3(static_cast<unsigned>(tiles)), dim3(256), 0, s, K, m,
                     n, K, A, lda, B, ldb, C, ldc, tiles_m, tiles_n);
  return hipGetLastError();
}

hipError_t launch_outer_neg(hipStream_t s, int K, int m, int n,
                            const double* A, int lda, const double* B, int ldb,
                            double* C, int ldc) {
  if(m <= 0 || n <= 0) return hipSuccess;
  if(K <= 0) return hipErrorInvalidValue;
  const int  tiles_m = (m + 63) / 64, tiles_n = (n + 63) / 64;
  const long tiles   = static_cast<long>(tiles_m) * tiles_n;
  if(tiles > 0x7fffffffL || K > 0x7fffffff - 32) return hipErrorInvalidValue;
  hipLaunchKernelGGL((tn_kernel<true, false, true, true>),
                     dim3(static_cast<unsigned>(tiles)), dim3(256), 0, s, K, m,
                     n, K, A, lda, B, ldb, C, ldc, tiles_m, tiles_n);
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
      hipLaunchKernelGGL((gemv_sub_kernel<double, double>), dim3((s0 + 63) / 64),
                         dim3(256), 0, s, s0, kb,
                         A + static_cast<size_t>(s0) * lda, lda, x + s0, x,
                         nullptr);
  }
  return hipGetLastError();
}

}  // namespace hplx
