// FP64 MFMA GEMM for the trailing update, the U12 = L11^{-1} A12 product and
// the updates inside the panel recursion.
//
// Replaces rocblas_dgemm(N,T) at src/pgesv/HPL_pdupdateTT.cpp:131-144 and
// src/pfact/HPL_pdrpanrlT.cpp:142-155, and rocblas_dtrsm at
// src/pgesv/HPL_pdupdateTT.cpp:95-106 (as a product with the explicit
// inverse of the unit lower L11).  The reference multiplies by a transposed
// copy of U (HPL_dlatcpy); here U12 stays in place in A: each of its
// columns is k-contiguous, which is the layout the B operand wants.
//
// C (m x n) = beta*C + alpha*A*B, all column-major, no transposes.
//   - workgroup tile 128 x 128, 4 waves in 2 x 2, each wave 64 x 64 as
//     4 x 4 blocks of V_MFMA_F64_16X16X4_F64;
//   - K streamed in 16-deep stages through double-buffered LDS (66 KB per
//     workgroup, two workgroups per CU), one barrier per stage;
//   - the MFMA computes the transposed block (D = B^T A^T) so that each
//     accumulator register of 16 lanes is 16 consecutive rows of one C
//     column: C is read and written in 128-byte runs;
//   - for C -= A*B the accumulators start from C and A is negated on its
//     way into LDS, so the epilogue is a plain store;
//   - workgroup -> tile mapping is XCD aware: the dispatcher deals
//     workgroups to the 8 XCDs round robin, so workgroup ids are remapped to
//     give each XCD a contiguous range of tiles, walked in groups of
//     kGroupM tile rows so that concurrently running tiles on one XCD share
//     A and B strips in that XCD's L2.
#include "hplx_internal.hpp"

namespace hplx {
namespace {

constexpr int BM      = 128;
constexpr int BN      = 128;
constexpr int BK      = 16;
constexpr int NT      = 256;
constexpr int WM      = 64;
constexpr int WN      = 64;
constexpr int TM      = WM / 16;
constexpr int TN      = WN / 16;
constexpr int LDS_LD  = BM + 1;  // odd stride: conflict-free transposed B stores
constexpr int kGroupM = 8;
constexpr int kXcds   = 8;

typedef double double4_t __attribute__((ext_vector_type(4)));

// V_MFMA_F64_16X16X4_F64 operand layout (lane l of 64):
//   a-operand: X[l % 16][l / 16]   (16 x 4)
//   b-operand: Y[l / 16][l % 16]   (4 x 16)
//   result:    D[4*r + l / 16][l % 16], r = 0..3
__device__ inline int dl_row(int r, int lane) { return 4 * r + (lane >> 4); }

template <bool BETA0, bool NEG, bool LOWER_A>
__global__ __launch_bounds__(NT, 2) void dgemm_nn_kernel(
    int M, int N, int K, const double* __restrict__ A, int lda,
    const double* __restrict__ B, int ldb, double* __restrict__ C, int ldc,
    int tiles_m, int tiles_n) {
  __shared__ double sA[2][BK * LDS_LD];
  __shared__ double sB[2][BK * LDS_LD];

  // ---- XCD-aware tile mapping
  const int T     = tiles_m * tiles_n;
  const int per   = gridDim.x / kXcds;
  const int bid   = blockIdx.x;
  const int tile  = (bid % kXcds) * per + bid / kXcds;
  if(tile >= T) return;
  const int width  = kGroupM * tiles_n;
  const int group  = tile / width;
  const int firstm = group * kGroupM;
  const int gsz    = min(tiles_m - firstm, kGroupM);
  const int within = tile - group * width;
  const int tm     = firstm + within % gsz;
  const int tn     = within / gsz;
  const int m0     = tm * BM;
  const int n0     = tn * BN;
  const int kend   = LOWER_A ? min(K, m0 + BM) : K;

  const int t    = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm0  = (wave & 1) * WM;
  const int wn0  = (wave >> 1) * WN;
  const int l16  = lane & 15;
  const int lk   = lane >> 4;

  // ---- accumulators
  double4_t acc[TM][TN];
#pragma unroll
  for(int i = 0; i < TM; ++i)
#pragma unroll
    for(int j = 0; j < TN; ++j) {
      if(BETA0) {
        acc[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};
      } else {
        const int row = m0 + wm0 + i * 16 + l16;
#pragma unroll
        for(int r = 0; r < 4; ++r) {
          const int col = n0 + wn0 + j * 16 + dl_row(r, lane);
          acc[i][j][r]  = (row < M && col < N)
                              ? C[row + static_cast<size_t>(col) * ldc]
                              : 0.0;
        }
      }
    }

  // ---- global -> register staging (8 doubles of A and of B per thread)
  double ra[8], rb[8];
  auto load_stage = [&](int k0) {
#pragma unroll
    for(int q = 0; q < 8; ++q) {
      const int idx = t + NT * q;
      // A: 128 rows x 16 k, rows fastest
      const int r  = idx & (BM - 1);
      const int kk = idx >> 7;
      const int gr = m0 + r, gk = k0 + kk;
      ra[q] = (gr < M && gk < kend) ? A[gr + static_cast<size_t>(gk) * lda]
                                    : 0.0;
      // B: 16 k x 128 columns, k fastest (16 threads read one 128-byte run)
      const int bk = idx & (BK - 1);
      const int bc = idx >> 4;
      const int gk2 = k0 + bk, gc = n0 + bc;
      rb[q] = (gk2 < kend && gc < N) ? B[gk2 + static_cast<size_t>(gc) * ldb]
                                     : 0.0;
    }
  };
  auto store_stage = [&](int buf) {
#pragma unroll
    for(int q = 0; q < 8; ++q) {
      const int idx = t + NT * q;
      const int r   = idx & (BM - 1);
      const int kk  = idx >> 7;
      sA[buf][kk * LDS_LD + r] = NEG ? -ra[q] : ra[q];
      const int bk = idx & (BK - 1);
      const int bc = idx >> 4;
      sB[buf][bk * LDS_LD + bc] = rb[q];
    }
  };

  if(kend > 0) {
    load_stage(0);
    store_stage(0);
  }
  __syncthreads();

  int buf = 0;
  for(int k0 = 0; k0 < kend; k0 += BK) {
    const bool more = (k0 + BK) < kend;
    if(more) load_stage(k0 + BK);
    const double* a_s = sA[buf];
    const double* b_s = sB[buf];
#pragma unroll
    for(int ks = 0; ks < BK / 4; ++ks) {
      double fa[TM], fb[TN];
      const int krow = (ks * 4 + lk) * LDS_LD;
#pragma unroll
      for(int i = 0; i < TM; ++i) fa[i] = a_s[krow + wm0 + i * 16 + l16];
#pragma unroll
      for(int j = 0; j < TN; ++j) fb[j] = b_s[krow + wn0 + j * 16 + l16];
#pragma unroll
      for(int i = 0; i < TM; ++i)
#pragma unroll
        for(int j = 0; j < TN; ++j)
          // D = B^T-block x A^T-block: D[n][m] accumulates C[m][n]
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[j], fa[i],
                                                           acc[i][j], 0, 0, 0);
    }
    if(more) store_stage(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

  // ---- epilogue: 16 lanes store 16 consecutive rows of one column
#pragma unroll
  for(int i = 0; i < TM; ++i) {
    const int row = m0 + wm0 + i * 16 + l16;
    if(row >= M) continue;
#pragma unroll
    for(int j = 0; j < TN; ++j)
#pragma unroll
      for(int r = 0; r < 4; ++r) {
        const int col = n0 + wn0 + j * 16 + dl_row(r, lane);
        if(col < N) {
          C[row + static_cast<size_t>(col) * ldc] = acc[i][j][r];
        }
      }
  }
}

template <bool BETA0, bool NEG, bool LOWER_A>
hipError_t run(hipStream_t s, int m, int n, int k, const double* A, int lda,
               const double* B, int ldb, double* C, int ldc) {
  const int  tiles_m = (m + BM - 1) / BM;
  const int  tiles_n = (n + BN - 1) / BN;
  const long T       = static_cast<long>(tiles_m) * tiles_n;
  const long grid    = ((T + kXcds - 1) / kXcds) * kXcds;
  if(grid > 0x7fffffffL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((dgemm_nn_kernel<BETA0, NEG, LOWER_A>),
                     dim3(static_cast<unsigned>(grid)), dim3(NT), 0, s, m, n,
                     k, A, lda, B, ldb, C, ldc, tiles_m, tiles_n);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_dgemm(hipStream_t s, int m, int n, int k, double alpha,
                        const double* A, int lda, const double* B, int ldb,
                        double beta, double* C, int ldc, bool lower_a) {
  if(m <= 0 || n <= 0) return hipSuccess;
  if(!(alpha == 1.0 || alpha == -1.0) || !(beta == 0.0 || beta == 1.0))
    return hipErrorInvalidValue;
  const bool beta0 = beta == 0.0;
  const bool neg   = alpha == -1.0;
  if(k <= 0) {
    if(!beta0) return hipSuccess;
    // C = 0: run with K = 0 (accumulators stay zero)
  }
  if(beta0) {
    if(neg)
      return lower_a ? run<true, true, true>(s, m, n, k, A, lda, B, ldb, C, ldc)
                     : run<true, true, false>(s, m, n, k, A, lda, B, ldb, C, ldc);
    return lower_a ? run<true, false, true>(s, m, n, k, A, lda, B, ldb, C, ldc)
                   : run<true, false, false>(s, m, n, k, A, lda, B, ldb, C, ldc);
  }
  if(neg)
    return lower_a ? run<false, true, true>(s, m, n, k, A, lda, B, ldb, C, ldc)
                   : run<false, true, false>(s, m, n, k, A, lda, B, ldb, C, ldc);
  return lower_a ? run<false, false, true>(s, m, n, k, A, lda, B, ldb, C, ldc)
                 : run<false, false, false>(s, m, n, k, A, lda, B, ldb, C, ldc);
}

}  // namespace hplx
