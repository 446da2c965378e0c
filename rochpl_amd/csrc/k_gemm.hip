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
//     workgroup, two workgroups per CU), one barrier per stage; the next
//     stage's global loads are in flight during the current stage's MFMAs
//     and the MFMA operands of the next k-step are read from LDS while the
//     current k-step's MFMAs run;
//   - global loads are unconditional (edge tiles clamp the address and zero
//     the value), so no divergent branch forces an early wait on them;
//     16-byte loads are used when A and B are 16-byte aligned with even
//     leading dimensions and M, K even (always true for HPL's even NB);
//   - the MFMA computes the transposed block (D = B^T A^T) so that each
//     accumulator register of 16 lanes is 16 consecutive rows of one C
//     column: C is read and written in 128-byte runs;
//   - for C -= A*B, A is negated on its way into LDS and the epilogue adds
//     the accumulators to C;
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
constexpr int LDS_LD  = BM + 2;  // even (16-byte rows), staggers LDS banks
constexpr int kGroupM = 8;
constexpr int kXcds   = 8;

template <typename V>
using vec4_t = V __attribute__((ext_vector_type(4)));
template <typename V>
using vec2_t = V __attribute__((ext_vector_type(2)));

// V_MFMA_F64_16X16X4_F64 / V_MFMA_F32_16X16X4_F32 operand layout (lane l):
//   a-operand: X[l % 16][l / 16]   (16 x 4)
//   b-operand: Y[l / 16][l % 16]   (4 x 16)
//   result:    D[4*r + l / 16][l % 16], r = 0..3         (f64)
//              D[4*(l / 16) + r][l % 16], r = 0..3       (f32)
template <typename V>
__device__ inline int dl_row(int r, int lane) {
  if constexpr(sizeof(V) == 8)
    return 4 * r + (lane >> 4);
  else
    return 4 * (lane >> 4) + r;
}

template <typename V>
__device__ inline vec4_t<V> mfma16x16x4(V a, V b, vec4_t<V> c) {
  if constexpr(sizeof(V) == 8)
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Stages one 128 x 16 tile of an operand (outer index o = row of A or
// column of B, inner index k) from global memory into LDS as S[k][o].
//   KFAST = false: o is the contiguous index (A, or B^T when B is transposed)
//   KFAST = true:  k is the contiguous index (B, or A^T when A is transposed)
// Loads are unconditional: out-of-range elements are read from a clamped
// address and replaced by zero at the LDS store.
template <typename V, bool KFAST, bool VEC2>
struct TileLoader {
  static constexpr int NQ = VEC2 ? 4 : 8;
  V    r[8];
  bool v[NQ];
  // Fast path (a full tile and a full stage): per-lane byte offsets from a
  // uniform stage base, fixed for the whole tile, so a stage's loads need no
  // address arithmetic, clamping or masking.
  uint32_t    off[NQ];
  const char* base;    // the tile's first element (uniform)
  size_t      kbytes;  // bytes per unit of k in the stage base (uniform)

  __device__ void init_fast(const V* __restrict__ P, int ld, int o0) {
    const int t = threadIdx.x;
    if(!KFAST) {
      base   = reinterpret_cast<const char*>(P + o0);
      kbytes = static_cast<size_t>(ld) * sizeof(V);
    } else {
      base   = reinterpret_cast<const char*>(P + static_cast<size_t>(o0) * ld);
      kbytes = sizeof(V);
    }
#pragma unroll
    for(int q = 0; q < NQ; ++q) {
      int o, k;
      if(!KFAST && VEC2) {
        o = 2 * (t & 63);
        k = (t >> 6) + 4 * q;
      } else if(!KFAST) {
        o = t & 127;
        k = (t >> 7) + 2 * q;
      } else if(VEC2) {
        o = (t >> 3) + 32 * q;
        k = 2 * (t & 7);
      } else {
        o = (t >> 4) + 16 * q;
        k = t & 15;
      }
      const size_t e = KFAST ? static_cast<size_t>(k) + static_cast<size_t>(o) * ld
                             : static_cast<size_t>(o) + static_cast<size_t>(k) * ld;
      off[q] = static_cast<uint32_t>(e * sizeof(V));
    }
  }

  __device__ void load_fast(int k0) {
    const char* s = base + static_cast<size_t>(k0) * kbytes;
#pragma unroll
    for(int q = 0; q < NQ; ++q) {
      if(VEC2) {
        const vec2_t<V> x = *reinterpret_cast<const vec2_t<V>*>(s + off[q]);
        r[2 * q]          = x.x;
        r[2 * q + 1]      = x.y;
      } else {
        r[q] = *reinterpret_cast<const V*>(s + off[q]);
      }
    }
  }

  __device__ void store_fast(V* S, bool neg) const {
    const int t = threadIdx.x;
    if(!KFAST && VEC2) {
#pragma unroll
      for(int q = 0; q < 4; ++q) {
        const int k = (t >> 6) + 4 * q;
        vec2_t<V> x;
        x.x = neg ? -r[2 * q] : r[2 * q];
        x.y = neg ? -r[2 * q + 1] : r[2 * q + 1];
        *reinterpret_cast<vec2_t<V>*>(S + k * LDS_LD + 2 * (t & 63)) = x;
      }
    } else if(!KFAST) {
#pragma unroll
      for(int q = 0; q < 8; ++q) {
        const int k = (t >> 7) + 2 * q;
        S[k * LDS_LD + (t & 127)] = neg ? -r[q] : r[q];
      }
    } else if(VEC2) {
#pragma unroll
      for(int q = 0; q < 4; ++q) {
        const int o = (t >> 3) + 32 * q;
        const int k = 2 * (t & 7);
        S[k * LDS_LD + o]       = neg ? -r[2 * q] : r[2 * q];
        S[(k + 1) * LDS_LD + o] = neg ? -r[2 * q + 1] : r[2 * q + 1];
      }
    } else {
#pragma unroll
      for(int q = 0; q < 8; ++q) {
        const int o = (t >> 4) + 16 * q;
        S[(t & 15) * LDS_LD + o] = neg ? -r[q] : r[q];
      }
    }
  }

  __device__ void load(const V* __restrict__ P, int ld, int o0, int O,
                       int k0, int kend) {
    const int t = threadIdx.x;
    if(!KFAST && VEC2) {  // pairs of o at k = (t>>6) + 4q
      const int op = o0 + 2 * (t & 63);
      const int oc = min(op, O - 2);
#pragma unroll
      for(int q = 0; q < 4; ++q) {
        const int k  = k0 + (t >> 6) + 4 * q;
        v[q]         = op < O && k < kend;
        const int kc = min(k, kend - 1);
        const vec2_t<V> x = *reinterpret_cast<const vec2_t<V>*>(
            P + oc + static_cast<size_t>(kc) * ld);
        r[2 * q]     = x.x;
        r[2 * q + 1] = x.y;
      }
    } else if(!KFAST) {  // o = t&127 at k = (t>>7) + 2q
      const int o  = o0 + (t & 127);
      const int oc = min(o, O - 1);
#pragma unroll
      for(int q = 0; q < 8; ++q) {
        const int k  = k0 + (t >> 7) + 2 * q;
        v[q]         = o < O && k < kend;
        const int kc = min(k, kend - 1);
        r[q]         = P[oc + static_cast<size_t>(kc) * ld];
      }
    } else if(VEC2) {  // pairs of k (t&7) in o = (t>>3) + 32q
      const int k  = k0 + 2 * (t & 7);
      const int kc = min(k, kend - 2);
#pragma unroll
      for(int q = 0; q < 4; ++q) {
        const int o  = o0 + (t >> 3) + 32 * q;
        v[q]         = k < kend && o < O;
        const int oc = min(o, O - 1);
        const vec2_t<V> x = *reinterpret_cast<const vec2_t<V>*>(
            P + kc + static_cast<size_t>(oc) * ld);
        r[2 * q]     = x.x;
        r[2 * q + 1] = x.y;
      }
    } else {  // k = t&15 in o = (t>>4) + 16q
      const int k  = k0 + (t & 15);
      const int kc = min(k, kend - 1);
#pragma unroll
      for(int q = 0; q < 8; ++q) {
        const int o  = o0 + (t >> 4) + 16 * q;
        v[q]         = k < kend && o < O;
        const int oc = min(o, O - 1);
        r[q]         = P[kc + static_cast<size_t>(oc) * ld];
      }
    }
  }

  __device__ void store(V* S, bool neg) const {
    const int t = threadIdx.x;
    if(!KFAST && VEC2) {
#pragma unroll
      for(int q = 0; q < 4; ++q) {
        const int k = (t >> 6) + 4 * q;
        vec2_t<V> x;
        x.x = v[q] ? (neg ? -r[2 * q] : r[2 * q]) : V(0);
        x.y = v[q] ? (neg ? -r[2 * q + 1] : r[2 * q + 1]) : V(0);
        *reinterpret_cast<vec2_t<V>*>(S + k * LDS_LD + 2 * (t & 63)) = x;
      }
    } else if(!KFAST) {
#pragma unroll
      for(int q = 0; q < 8; ++q) {
        const int k = (t >> 7) + 2 * q;
        S[k * LDS_LD + (t & 127)] = v[q] ? (neg ? -r[q] : r[q]) : V(0);
      }
    } else if(VEC2) {
#pragma unroll
      for(int q = 0; q < 4; ++q) {
        const int o = (t >> 3) + 32 * q;
        const int k = 2 * (t & 7);
        S[k * LDS_LD + o]       = v[q] ? (neg ? -r[2 * q] : r[2 * q]) : V(0);
        S[(k + 1) * LDS_LD + o] = v[q] ? (neg ? -r[2 * q + 1] : r[2 * q + 1])
                                       : V(0);
      }
    } else {
#pragma unroll
      for(int q = 0; q < 8; ++q) {
        const int o = (t >> 4) + 16 * q;
        S[(t & 15) * LDS_LD + o] = v[q] ? (neg ? -r[q] : r[q]) : V(0);
      }
    }
  }
};

// KIND only names the instance: 1 = the trailing update A22 -= L21*U12, so
// profiles report it apart from the panel-internal products.
// TA / TB: the operand is stored transposed (op(A) = A^T, op(B) = B^T).
// V: double (FP64 MFMA) or float (FP32 MFMA, the FP32 panel of the
// mixed-precision factorisation).
template <typename V, bool BETA0, bool NEG, bool LOWER_A, bool TA, bool TB,
          bool VA, bool VB, int KIND>
__global__ __launch_bounds__(NT, 2) void dgemm_kernel(
    int M, int N, int K, const V* __restrict__ A, int lda,
    const V* __restrict__ B, int ldb, V* __restrict__ C, int ldc,
    int tiles_m, int tiles_n) {
  __shared__ __attribute__((aligned(16))) V sA[2][BK * LDS_LD];
  __shared__ __attribute__((aligned(16))) V sB[2][BK * LDS_LD];

  // ---- XCD-aware tile mapping
  const int T    = tiles_m * tiles_n;
  const int per  = gridDim.x / kXcds;
  const int bid  = blockIdx.x;
  const int tile = (bid % kXcds) * per + bid / kXcds;
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

  TileLoader<V, TA, VA>  la;   // A: o = row of C; k-contiguous when TA
  TileLoader<V, !TB, VB> lb;   // B: o = column of C; k-contiguous unless TB
  // interior tiles take the fast path for every full stage
  const bool full_tile = m0 + BM <= M && n0 + BN <= N;
  if(full_tile) {
    la.init_fast(A, lda, m0);
    lb.init_fast(B, ldb, n0);
  }
  auto fast_stage = [&](int k0) { return full_tile && k0 + BK <= kend; };
  auto load_stage = [&](int k0) {
    if(fast_stage(k0)) {
      la.load_fast(k0);
      lb.load_fast(k0);
    } else {
      la.load(A, lda, m0, M, k0, kend);
      lb.load(B, ldb, n0, N, k0, kend);
    }
  };
  auto store_stage = [&](int buf, int k0) {
    if(fast_stage(k0)) {
      la.store_fast(sA[buf], NEG);
      lb.store_fast(sB[buf], false);
    } else {
      la.store(sA[buf], NEG);
      lb.store(sB[buf], false);
    }
  };

  if(kend > 0) load_stage(0);

  // ---- accumulators start at zero; C is added in the epilogue, so no load
  // of C is outstanding while the main loop runs
  vec4_t<V> acc[TM][TN];
#pragma unroll
  for(int i = 0; i < TM; ++i)
#pragma unroll
    for(int j = 0; j < TN; ++j) acc[i][j] = vec4_t<V>{0, 0, 0, 0};

  if(kend > 0) store_stage(0, 0);
  __syncthreads();

  int buf = 0;
  for(int k0 = 0; k0 < kend; k0 += BK) {
    const bool more = (k0 + BK) < kend;
    if(more) load_stage(k0 + BK);
    const V* a_s = sA[buf] + lk * LDS_LD + wm0 + l16;
    const V* b_s = sB[buf] + lk * LDS_LD + wn0 + l16;
    V        fa[2][TM], fb[2][TN];
#pragma unroll
    for(int i = 0; i < TM; ++i) fa[0][i] = a_s[i * 16];
#pragma unroll
    for(int j = 0; j < TN; ++j) fb[0][j] = b_s[j * 16];
#pragma unroll
    for(int ks = 0; ks < BK / 4; ++ks) {
      const int cur = ks & 1;
      if(ks + 1 < BK / 4) {
        const int off = (ks + 1) * 4 * LDS_LD;
#pragma unroll
        for(int i = 0; i < TM; ++i) fa[cur ^ 1][i] = a_s[off + i * 16];
#pragma unroll
        for(int j = 0; j < TN; ++j) fb[cur ^ 1][j] = b_s[off + j * 16];
      }
#pragma unroll
      for(int i = 0; i < TM; ++i)
#pragma unroll
        for(int j = 0; j < TN; ++j)
          // D = B^T-block x A^T-block: D[n][m] accumulates C[m][n]
          acc[i][j] = mfma16x16x4<V>(fb[cur][j], fa[cur][i], acc[i][j]);
    }
    if(more) store_stage(buf ^ 1, k0 + BK);
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
        const int col = n0 + wn0 + j * 16 + dl_row<V>(r, lane);
        if(col < N) {
          V* c = C + row + static_cast<size_t>(col) * ldc;
          *c        = BETA0 ? acc[i][j][r] : *c + acc[i][j][r];
        }
      }
  }
}

// a pair of elements (the VEC2 load) is naturally aligned
template <typename V>
bool pair_aligned(const V* p) {
  return reinterpret_cast<uintptr_t>(p) % (2 * sizeof(V)) == 0;
}

template <bool BETA0, bool NEG, bool LOWER_A, bool TA, bool TB, int KIND,
          typename V>
hipError_t run(hipStream_t s, int m, int n, int k, const V* A, int lda,
               const V* B, int ldb, V* C, int ldc) {
  const int  tiles_m = (m + BM - 1) / BM;
  const int  tiles_n = (n + BN - 1) / BN;
  const long T       = static_cast<long>(tiles_m) * tiles_n;
  const long grid    = ((T + kXcds - 1) / kXcds) * kXcds;
  if(grid > 0x7fffffffL) return hipErrorInvalidValue;
  // 16-byte staging: aligned operand, even leading dimension and an even
  // extent along the pair direction (o for o-contiguous tiles, K for
  // k-contiguous ones; a LOWER_A row tile's K bound m0 + BM is even too)
  const bool va = pair_aligned(A) && lda % 2 == 0 && (TA ? k % 2 == 0 : m % 2 == 0);
  const bool vb = pair_aligned(B) && ldb % 2 == 0 && (TB ? n % 2 == 0 : k % 2 == 0);
  const dim3 g(static_cast<unsigned>(grid)), b(NT);
  if(va && vb)
    hipLaunchKernelGGL((dgemm_kernel<V, BETA0, NEG, LOWER_A, TA, TB, true, true, KIND>),
                       g, b, 0, s, m, n, k, A, lda, B, ldb, C, ldc, tiles_m, tiles_n);
  else if(va)
    hipLaunchKernelGGL((dgemm_kernel<V, BETA0, NEG, LOWER_A, TA, TB, true, false, KIND>),
                       g, b, 0, s, m, n, k, A, lda, B, ldb, C, ldc, tiles_m, tiles_n);
  else if(vb)
    hipLaunchKernelGGL((dgemm_kernel<V, BETA0, NEG, LOWER_A, TA, TB, false, true, KIND>),
                       g, b, 0, s, m, n, k, A, lda, B, ldb, C, ldc, tiles_m, tiles_n);
  else
    hipLaunchKernelGGL((dgemm_kernel<V, BETA0, NEG, LOWER_A, TA, TB, false, false, KIND>),
                       g, b, 0, s, m, n, k, A, lda, B, ldb, C, ldc, tiles_m, tiles_n);
  return hipGetLastError();
}

template <typename V>
hipError_t gemm_t(hipStream_t s, int m, int n, int k, double alpha,
                  const V* A, int lda, const V* B, int ldb, double beta, V* C,
                  int ldc, bool lower_a, int kind, bool trans_a, bool trans_b) {
  if(m <= 0 || n <= 0) return hipSuccess;
  if(!(alpha == 1.0 || alpha == -1.0) || !(beta == 0.0 || beta == 1.0))
    return hipErrorInvalidValue;
  const bool beta0 = beta == 0.0;
  const bool neg   = alpha == -1.0;
  if(k < 0) return hipErrorInvalidValue;
  if(k == 0 && !beta0) return hipSuccess;  // C unchanged
  if(trans_a || trans_b) {
    // products of the adjoint solves: C -= op(A) op(B) or C = -op(A) op(B)
    if constexpr(sizeof(V) == 8) {
      if(lower_a || !neg) return hipErrorInvalidValue;
      if(trans_a && !trans_b)
        return beta0 ? run<true, true, false, true, false, 0>(s, m, n, k, A, lda, B, ldb, C, ldc)
                     : run<false, true, false, true, false, 0>(s, m, n, k, A, lda, B, ldb, C, ldc);
      if(!trans_a && trans_b)
        return beta0 ? run<true, true, false, false, true, 0>(s, m, n, k, A, lda, B, ldb, C, ldc)
                     : run<false, true, false, false, true, 0>(s, m, n, k, A, lda, B, ldb, C, ldc);
    }
    return hipErrorInvalidValue;
  }
  if(beta0) {
    if(neg)
      return lower_a ? run<true, true, true, false, false, 0>(s, m, n, k, A, lda, B, ldb, C, ldc)
                     : run<true, true, false, false, false, 0>(s, m, n, k, A, lda, B, ldb, C, ldc);
    return lower_a ? run<true, false, true, false, false, 0>(s, m, n, k, A, lda, B, ldb, C, ldc)
                   : run<true, false, false, false, false, 0>(s, m, n, k, A, lda, B, ldb, C, ldc);
  }
  if constexpr(sizeof(V) == 8)
    if(neg && !lower_a && kind == 1)
      return run<false, true, false, false, false, 1>(s, m, n, k, A, lda, B, ldb, C, ldc);
  if(neg)
    return lower_a ? run<false, true, true, false, false, 0>(s, m, n, k, A, lda, B, ldb, C, ldc)
                   : run<false, true, false, false, false, 0>(s, m, n, k, A, lda, B, ldb, C, ldc);
  return lower_a ? run<false, false, true, false, false, 0>(s, m, n, k, A, lda, B, ldb, C, ldc)
                 : run<false, false, false, false, false, 0>(s, m, n, k, A, lda, B, ldb, C, ldc);
}

}  // namespace

hipError_t launch_dgemm(hipStream_t s, int m, int n, int k, double alpha,
                        const double* A, int lda, const double* B, int ldb,
                        double beta, double* C, int ldc, bool lower_a,
                        int kind, bool trans_a, bool trans_b) {
  return gemm_t(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, lower_a, kind,
                trans_a, trans_b);
}

hipError_t launch_sgemm(hipStream_t s, int m, int n, int k, double alpha,
                        const float* A, int lda, const float* B, int ldb,
                        double beta, float* C, int ldc, bool lower_a) {
  return gemm_t(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, lower_a, 0,
                false, false);
}

}  // namespace hplx
