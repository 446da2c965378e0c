// Low-precision trailing update of the mixed-precision factorisations:
// C -= L21 * U12 with C in the factor storage (FP32 or FP16), L21 and U12
// rounded to BF16 or to per-tile-scaled FP8 e4m3, products on MFMA with FP32
// accumulation.  It stands in for rocblas_dgemm of HPL_pdupdateTT
// (src/pgesv/HPL_pdupdateTT.cpp:131-144) in the BF16 / FP8 forms; the
// reference itself has no low-precision path.
//
// Operands are packed once per panel into the order the GEMM consumes them:
//   A (L21, m x K):  [term][K stage][row][k in stage]
//   B (U12, K x n):  [term][K stage][column][k in stage]
// so that one workgroup's stage of A (or B) is one contiguous block of
// rows x stage bytes, whatever the tile height.  A K stage is 32 (BF16, one
// V_MFMA_F32_16X16X32_BF16 deep) or 128 (FP8, one
// V_MFMA_SCALE_F32_16X16X128_F8F6F4 deep); K and the row/column counts are
// padded with zeros to whole stages and whole 256-row tiles.
//
// BF16 forms.  Each FP32 operand x is split into `terms` BF16 numbers,
// x ~ h0 + h1 + h2 with h0 = bf16(x), h1 = bf16(x - h0), h2 = bf16(x - h0 - h1)
// (each difference is exact in FP32), and the product keeps the pairs whose
// term indices add up to less than `terms`: 1 product for terms = 1 (plain
// BF16), 3 for terms = 2 (about 16 significant bits), 6 for terms = 3 (about
// FP32).  All products accumulate into the same FP32 accumulators.
//
// FP8 form.  L21 and U12 are cut into 128 x 128 tiles (row tile x K stage,
// K stage x column tile); every tile gets a power-of-two scale 2^e with the
// largest e such that max|tile| * 2^e <= 448 (the e4m3 maximum), is rounded
// to e4m3 (round to nearest even) and keeps e.  Each 128-deep MFMA product
// of an A tile with a B tile is multiplied by 2^-(eA + eB) and added to the
// FP32 accumulators: the MFMA applies the factor itself through its E8M0
// operand scales (127 - eA, 127 - eB) and accumulates in place.
//
// Tile: WM x WN waves, each wave 64 rows x 16*NI columns as 4 x NI blocks of
// 16 x 16 (BF16: 128 x 128 workgroup tiles of 64 x 64 waves; FP8: 256 x 128
// tiles of four 64 x 128 waves, whose stages carry 25 % fewer operand bytes
// per flop).  The MFMA computes the transposed block
// (D = B^T A^T) so that 16 lanes hold 16 consecutive rows of one C column;
// the C update of full tiles goes through LDS as 4-row vectors per lane
// (epilogue_vec), with non-temporal C accesses.
// The next stage's global loads are in flight while the current stage's
// MFMAs run; LDS is single-buffered (two barriers per stage).  Workgroup ids
// are remapped XCD-aware as in the FP64 GEMM.
#include <type_traits>
#include <utility>
#include <vector>

#include "claim.hpp"
#include "hplx_internal.hpp"
#include "mxp.hpp"

namespace hplx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float  f32x4 __attribute__((ext_vector_type(4)));
typedef int    i32x8 __attribute__((ext_vector_type(8)));
typedef int    i32x4 __attribute__((ext_vector_type(4)));

constexpr int kXcds   = 8;
constexpr int kGroupM = 8;

// ---------------------------------------------------------------------------
// Rounding (bit-level, so that the rounding does not depend on conversion
// instructions and matches a software model exactly).
// ---------------------------------------------------------------------------

// FP32 -> BF16, round to nearest even (finite inputs).
__device__ inline uint16_t bf16_bits(float x) {
  const uint32_t u = __float_as_uint(x);
  return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline float bf16_val(uint16_t h) {
  return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

// FP32 -> OCP FP8 e4m3 (bias 7, max 448, subnormals 2^-9 apart), round to
// nearest even; |x| <= 448 is guaranteed by the tile scale.
__device__ inline uint8_t e4m3_bits(float x) {
  const uint32_t u    = __float_as_uint(x);
  const uint32_t sign = (u >> 24) & 0x80u;
  const float    a    = fabsf(x);
  if(a < 0x1p-10f) return static_cast<uint8_t>(sign);  // rounds to zero
  int e = static_cast<int>((__float_as_uint(a) >> 23) & 0xff) - 127;
  uint32_t q;
  if(e < -6) {
    // subnormal: value = m * 2^-9, m = 0..7 (8 rounds up into the normals)
    const float m = a * 0x1p9f;
    q             = static_cast<uint32_t>(__builtin_rintf(m));  // RNE
    return static_cast<uint8_t>(sign | q);
  }
  // normal: 3 mantissa bits; round the FP32 mantissa to 3 bits, RNE
  const uint32_t mant  = __float_as_uint(a) & 0x7fffffu;
  const uint32_t keep  = mant >> 20;
  const uint32_t rest  = mant & 0xfffffu;
  uint32_t       r     = keep;
  const uint32_t half  = 0x80000u;
  if(rest > half || (rest == half && (keep & 1u))) ++r;
  int ee = e + 7;
  if(r == 8u) {
    r = 0;
    ++ee;
  }
  q = (static_cast<uint32_t>(ee) << 3) | r;
  return static_cast<uint8_t>(sign | q);
}

// Largest e with amax * 2^e <= 448 (0 for an all-zero tile), clamped to the
// range the scale arithmetic uses.
__device__ inline int fp8_tile_exp(float amax) {
  if(!(amax > 0.0f)) return 0;
  int e = static_cast<int>(floorf(log2f(448.0f / amax)));
  // guard the log2 rounding at exact powers
  while(ldexpf(amax, e) > 448.0f) --e;
  while(ldexpf(amax, e + 1) <= 448.0f) ++e;
  return e < -100 ? -100 : (e > 100 ? 100 : e);
}

// ---------------------------------------------------------------------------
// Packing.
// ---------------------------------------------------------------------------

// A (m x K, column-major FP32) -> BF16 terms, [term][ks][row][32].  Block:
// 64 rows x one 32-deep stage, 256 threads.
template <int TERMS>
__global__ __launch_bounds__(256) void pack_a_bf16_kernel(
    int m, int K, const float* __restrict__ L, int ldl,
    uint16_t* __restrict__ Ap, int m_pad, size_t plane) {
  __shared__ float tile[32][65];
  const int t  = threadIdx.x;
  const int r0 = blockIdx.x * 64;
  const int ks = blockIdx.y;
  const int k0 = ks * 32;
#pragma unroll
  for(int j = 0; j < 8; ++j) {
    const int kk = (t >> 6) + 4 * j;
    const int r  = r0 + (t & 63);
    const int k  = k0 + kk;
    tile[kk][t & 63] =
        (r < m && k < K) ? L[r + static_cast<size_t>(k) * ldl] : 0.0f;
  }
  __syncthreads();
  const int row = t >> 2, seg = t & 3;
  uint16_t  h[TERMS][8];
#pragma unroll
  for(int i = 0; i < 8; ++i) {
    float x = tile[seg * 8 + i][row];
#pragma unroll
    for(int q = 0; q < TERMS; ++q) {
      h[q][i] = bf16_bits(x);
      x -= bf16_val(h[q][i]);
    }
  }
  const size_t off =
      (static_cast<size_t>(ks) * m_pad + r0 + row) * 32 + seg * 8;
#pragma unroll
  for(int q = 0; q < TERMS; ++q) {
    i32x4 v;
#pragma unroll
    for(int i = 0; i < 4; ++i)
      v[i] = static_cast<int>(h[q][2 * i] | (static_cast<uint32_t>(h[q][2 * i + 1]) << 16));
    *reinterpret_cast<i32x4*>(Ap + q * plane + off) = v;
  }
}

// B (K x n, column-major FP32) -> BF16 terms, [term][ks][col][32].  Thread =
// (column, 8-deep segment of one stage).
template <int TERMS>
__global__ __launch_bounds__(256) void pack_b_bf16_kernel(
    int K, int n, const float* __restrict__ U, int ldu,
    uint16_t* __restrict__ Bp, int n_pad, size_t plane) {
  const int t   = threadIdx.x;
  const int col = blockIdx.x * 64 + (t >> 2);
  const int seg = t & 3;
  const int ks  = blockIdx.y;
  const int k0  = ks * 32 + seg * 8;
  if(col >= n_pad) return;
  uint16_t h[TERMS][8];
#pragma unroll
  for(int i = 0; i < 8; ++i) {
    const int k = k0 + i;
    float     x = (col < n && k < K) ? U[k + static_cast<size_t>(col) * ldu] : 0.0f;
#pragma unroll
    for(int q = 0; q < TERMS; ++q) {
      h[q][i] = bf16_bits(x);
      x -= bf16_val(h[q][i]);
    }
  }
  const size_t off = (static_cast<size_t>(ks) * n_pad + col) * 32 + seg * 8;
#pragma unroll
  for(int q = 0; q < TERMS; ++q) {
    i32x4 v;
#pragma unroll
    for(int i = 0; i < 4; ++i)
      v[i] = static_cast<int>(h[q][2 * i] | (static_cast<uint32_t>(h[q][2 * i + 1]) << 16));
    *reinterpret_cast<i32x4*>(Bp + q * plane + off) = v;
  }
}

__device__ inline float block_amax(float v, float* sh) {
  for(int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  const int w = threadIdx.x >> 6;
  if((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  float r = sh[0];
  for(int k = 1; k < static_cast<int>(blockDim.x >> 6); ++k) r = fmaxf(r, sh[k]);
  return r;
}

// A (m x K FP32) -> e4m3 [ks][row][128] + exponents ea[row tile][ks].
// Block: one 128 x 128 tile, 256 threads; thread = row (t & 127) and every
// other column.
__global__ __launch_bounds__(256) void pack_a_fp8_kernel(
    int m, int K, const float* __restrict__ L, int ldl,
    uint8_t* __restrict__ Ap, int m_pad, int nks, int* __restrict__ ea) {
  __shared__ float   sh[4];
  __shared__ uint8_t q8[128][132];
  const int t  = threadIdx.x;
  const int mt = blockIdx.x, ks = blockIdx.y;
  const int r  = mt * 128 + (t & 127);
  float     v[64];
  float     amax = 0.0f;
#pragma unroll
  for(int j = 0; j < 64; ++j) {
    const int k = ks * 128 + 2 * j + (t >> 7);
    v[j]        = (r < m && k < K) ? L[r + static_cast<size_t>(k) * ldl] : 0.0f;
    amax        = fmaxf(amax, fabsf(v[j]));
  }
  const int   e = fp8_tile_exp(block_amax(amax, sh));
  const float s = ldexpf(1.0f, e);
#pragma unroll
  for(int j = 0; j < 64; ++j) q8[t & 127][2 * j + (t >> 7)] = e4m3_bits(v[j] * s);
  if(t == 0) ea[mt * nks + ks] = e;
  __syncthreads();
  // write rows: thread = (row t/2, half t%2), 64 bytes each
  const int row = t >> 1, half = t & 1;
  uint8_t*  dst = Ap + (static_cast<size_t>(ks) * m_pad + mt * 128 + row) * 128 + half * 64;
#pragma unroll
  for(int c = 0; c < 4; ++c) {
    i32x4 w;
#pragma unroll
    for(int i = 0; i < 4; ++i) {
      const int b = half * 64 + c * 16 + i * 4;
      w[i]        = static_cast<int>(q8[row][b] | (q8[row][b + 1] << 8) |
                                     (q8[row][b + 2] << 16) |
                                     (static_cast<uint32_t>(q8[row][b + 3]) << 24));
    }
    *reinterpret_cast<i32x4*>(dst + c * 16) = w;
  }
}

// B (K x n FP32) -> e4m3 [ks][col][128] + exponents eb[ks][col tile].
// Block: one 128 x 128 tile; thread = (column t/2, half of the stage t%2).
__global__ __launch_bounds__(256) void pack_b_fp8_kernel(
    int K, int n, const float* __restrict__ U, int ldu,
    uint8_t* __restrict__ Bp, int n_pad, int ntn, int* __restrict__ eb) {
  __shared__ float sh[4];
  const int t   = threadIdx.x;
  const int nt  = blockIdx.x, ks = blockIdx.y;
  const int col = nt * 128 + (t >> 1);
  const int k0  = ks * 128 + (t & 1) * 64;
  float     v[64];
  float     amax = 0.0f;
#pragma unroll
  for(int j = 0; j < 64; ++j) {
    const int k = k0 + j;
    v[j]        = (col < n && k < K) ? U[k + static_cast<size_t>(col) * ldu] : 0.0f;
    amax        = fmaxf(amax, fabsf(v[j]));
  }
  const int   e = fp8_tile_exp(block_amax(amax, sh));
  const float s = ldexpf(1.0f, e);
  if(t == 0) eb[ks * ntn + nt] = e;
  uint8_t* dst = Bp + (static_cast<size_t>(ks) * n_pad + col) * 128 + (t & 1) * 64;
#pragma unroll
  for(int c = 0; c < 4; ++c) {
    i32x4 w;
#pragma unroll
    for(int i = 0; i < 4; ++i) {
      const int j = c * 16 + i * 4;
      w[i]        = static_cast<int>(e4m3_bits(v[j] * s) |
                                     (e4m3_bits(v[j + 1] * s) << 8) |
                                     (e4m3_bits(v[j + 2] * s) << 16) |
                                     (static_cast<uint32_t>(e4m3_bits(v[j + 3] * s)) << 24));
    }
    *reinterpret_cast<i32x4*>(dst + c * 16) = w;
  }
}

// ---------------------------------------------------------------------------
// GEMM
// ---------------------------------------------------------------------------

template <typename S>
__device__ inline float ld_c(const S* p) {
  return static_cast<float>(*p);
}
template <typename S>
__device__ inline void st_c(S* p, float v) {
  *p = static_cast<S>(v);
}

// Tile `tile` of the launch's order: groups of kGroupM tile rows walked
// column by column.  Every schedule below hands XCD x the tiles
// [x per, (x + 1) per) of this order (per = ceil(T / 8)).
__device__ inline void tile_coords(int tile, int tiles_m, int tiles_n, int& tm,
                                   int& tn) {
  const int width  = kGroupM * tiles_n;
  const int group  = tile / width;
  const int firstm = group * kGroupM;
  const int gsz    = min(tiles_m - firstm, kGroupM);
  const int within = tile - group * width;
  tm               = firstm + within % gsz;
  tn               = within / gsz;
}

// Launch modes of the update kernels: one tile per workgroup (XCD-aware),
// the look-ahead's claimed mode (a persistent grid beside the panel
// factorisation: workgroups on the vacated CUs exit, the others claim tiles
// from their XCD's counter until the stop flag is set) and its resume mode
// (one tile per workgroup from where each XCD's claims ended).
enum : int { kModePlain = 0, kModeClaimed = 1, kModeResume = 2 };

template <int MODE, typename F>
__device__ __forceinline__ void run_tiles(int tiles_m, int tiles_n,
                                          const UpdSched& sch,
                                          const int* __restrict__ claimed,
                                          F&& tile_fn) {
  const int T = tiles_m * tiles_n;
  int       tm, tn;
  if constexpr(MODE == kModePlain) {
    const int per  = gridDim.x / kXcds, bid = blockIdx.x;
    const int tile = (bid % kXcds) * per + bid / kXcds;
    if(tile >= T) return;
    tile_coords(tile, tiles_m, tiles_n, tm, tn);
    tile_fn(tm, tn);
  } else if constexpr(MODE == kModeResume) {
    const int x = blockIdx.x % kXcds, i = blockIdx.x / kXcds;
    const int done = min(claimed[32 * x], sch.per);
    const int tile = x * sch.per + done + i;
    if(done + i >= sch.per || tile >= T) return;
    tile_coords(tile, tiles_m, tiles_n, tm, tn);
    tile_fn(tm, tn);
  } else {
    __shared__ int      s_tile;
    __shared__ unsigned s_dead;
    // all waves of a workgroup share one CU, so this exit is uniform
    if(on_vacated_cu(sch)) return;
    if(threadIdx.x == 0) {
      unsigned dead = 0u;
      s_tile        = claim_next_tile(sch, T, dead);
      s_dead        = dead;
    }
    __syncthreads();
    for(;;) {
      const int tile = __builtin_amdgcn_readfirstlane(s_tile);
      if(tile < 0) break;
      tile_coords(tile, tiles_m, tiles_n, tm, tn);
      tile_fn(tm, tn);
      __syncthreads();  // every wave is done with the tile's LDS and s_tile
      if(threadIdx.x == 0) {
        unsigned dead = s_dead;
        s_tile        = claim_next_tile(sch, T, dead);
        s_dead        = dead;
      }
      __syncthreads();
    }
  }
}

// Epilogue C -= acc for one wave's 64 x 64 sub-tile.
//
// The MFMA leaves lane l with rows l % 16 of four different columns, so a
// direct update touches every column in 64-byte (FP32) or 32-byte (FP16)
// pieces with 128 instructions per lane.  Instead the sub-tile goes through
// a per-wave LDS scratch E (32 columns x kEpiPitch floats, two passes of 32
// columns): each lane then reads 4 consecutive rows of one column, and C is
// loaded and stored as 16-byte (FP32) / 8-byte (FP16) vectors, 256 (128)
// contiguous bytes per column per instruction.  Full, aligned tiles only;
// the others take the element-wise path.
constexpr int kEpiPitch      = 68;                // floats per LDS column
constexpr int kEpiWaveFloats = 32 * kEpiPitch;   // per wave
constexpr int epi_bytes(int waves) { return waves * kEpiWaveFloats * 4; }

typedef float f32x4v __attribute__((ext_vector_type(4)));
typedef _Float16 f16x4v __attribute__((ext_vector_type(4)));

template <int NI, typename S>
__device__ inline void epilogue_vec(f32x4 (&acc)[4][NI], S* __restrict__ C,
                                    int ldc, int r0, int c0, float* E) {
  typedef typename std::conditional<sizeof(S) == 4, f32x4v, f16x4v>::type CV;
  const int lane = threadIdx.x & 63, l16 = lane & 15, lk = lane >> 4;
  const int rl   = 4 * l16;  // local rows rl .. rl+3
#pragma unroll
  for(int p = 0; p < NI / 2; ++p) {
    // the pass's 8 C vectors in flight at once, before the LDS transpose:
    // one memory latency per pass instead of one per vector (a load issued
    // after a store to C would wait for it).  C is read and written once per
    // update, so its accesses are non-temporal: streamed through L2 without
    // evicting the operand stages the other workgroups of the XCD reuse
    CV x[8];
#pragma unroll
    for(int q = 0; q < 8; ++q)
      x[q] = __builtin_nontemporal_load(reinterpret_cast<const CV*>(
          C + (r0 + rl) + static_cast<size_t>(c0 + 32 * p + 4 * q + lk) * ldc));
#pragma unroll
    for(int jj = 0; jj < 2; ++jj)
#pragma unroll
      for(int i = 0; i < 4; ++i)
#pragma unroll
        for(int r = 0; r < 4; ++r)
          E[(jj * 16 + 4 * lk + r) * kEpiPitch + 16 * i + l16] =
              acc[i][2 * p + jj][r];
    __syncthreads();
#pragma unroll
    for(int q = 0; q < 8; ++q) {
      const int    cl = 4 * q + lk;  // local column 0..31
      const f32x4v v  = *reinterpret_cast<const f32x4v*>(&E[cl * kEpiPitch + rl]);
      S* c = C + (r0 + rl) + static_cast<size_t>(c0 + 32 * p + cl) * ldc;
      if constexpr(sizeof(S) == 4) {
        __builtin_nontemporal_store(x[q] - v, reinterpret_cast<f32x4v*>(c));
      } else {
        f16x4v y;
#pragma unroll
        for(int e = 0; e < 4; ++e)
          y[e] = static_cast<_Float16>(static_cast<float>(x[q][e]) - v[e]);
        __builtin_nontemporal_store(y, reinterpret_cast<f16x4v*>(c));
      }
    }
    __syncthreads();
  }
}

// Pairs (a-term, b-term) of the split product, largest first.
__device__ constexpr int kPairA[6] = {0, 0, 1, 0, 1, 2};
__device__ constexpr int kPairB[6] = {0, 1, 0, 2, 1, 0};

constexpr int kBf16RS = 48;  // LDS row stride in bf16 (96 bytes; 3 % faster than 80)

// One LDS arena per workgroup: operand stages during the K loop, epilogue
// scratch after.
template <int TERMS, int WM, int WN>
constexpr int bf16_smem_bytes() {
  constexpr int ops = TERMS * (64 * WM + 64 * WN) * kBf16RS * 2;
  return ops > epi_bytes(WM * WN) ? ops : epi_bytes(WM * WN);
}

// Tile (tm, tn) of C -= A B on BF16 MFMA.
template <int TERMS, int WM, int WN, typename S>
__device__ __forceinline__ void bf16_tile(
    int tm, int tn, int M, int N, int nks, const uint16_t* __restrict__ Ap,
    int m_pad, size_t aplane, const uint16_t* __restrict__ Bp, int n_pad,
    size_t bplane, S* __restrict__ C, int ldc, int vec_ok, uint8_t* smem) {
  constexpr int NT     = 64 * WM * WN;
  constexpr int BM     = 64 * WM;
  constexpr int BN     = 64 * WN;
  constexpr int RS     = kBf16RS;
  constexpr int NPROD  = TERMS * (TERMS + 1) / 2;
  constexpr int ACH    = BM * 4 / NT;  // 16-byte chunks per thread, A, per term
  constexpr int BCH    = BN * 4 / NT;
  uint16_t(*sA)[BM * RS] = reinterpret_cast<uint16_t(*)[BM * RS]>(smem);
  uint16_t(*sB)[BN * RS] =
      reinterpret_cast<uint16_t(*)[BN * RS]>(smem + TERMS * BM * RS * 2);

  const int m0 = tm * BM, n0 = tn * BN;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm0 = (wave % WM) * 64, wn0 = (wave / WM) * 64;
  const int l16 = lane & 15, lk = lane >> 4;

  i32x4 ra[TERMS][ACH], rb[TERMS][BCH];
  auto load = [&](int ks) {
#pragma unroll
    for(int q = 0; q < TERMS; ++q) {
      const uint16_t* a = Ap + q * aplane + (static_cast<size_t>(ks) * m_pad + m0) * 32;
      const uint16_t* b = Bp + q * bplane + (static_cast<size_t>(ks) * n_pad + n0) * 32;
#pragma unroll
      for(int c = 0; c < ACH; ++c)
        ra[q][c] = *reinterpret_cast<const i32x4*>(a + (t + c * NT) * 8);
#pragma unroll
      for(int c = 0; c < BCH; ++c)
        rb[q][c] = *reinterpret_cast<const i32x4*>(b + (t + c * NT) * 8);
    }
  };
  auto store = [&]() {
#pragma unroll
    for(int q = 0; q < TERMS; ++q) {
#pragma unroll
      for(int c = 0; c < ACH; ++c) {
        const int ch = t + c * NT;
        *reinterpret_cast<i32x4*>(&sA[q][(ch >> 2) * RS + (ch & 3) * 8]) = ra[q][c];
      }
#pragma unroll
      for(int c = 0; c < BCH; ++c) {
        const int ch = t + c * NT;
        *reinterpret_cast<i32x4*>(&sB[q][(ch >> 2) * RS + (ch & 3) * 8]) = rb[q][c];
      }
    }
  };

  f32x4 acc[4][4];
#pragma unroll
  for(int i = 0; i < 4; ++i)
#pragma unroll
    for(int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  load(0);
  for(int ks = 0; ks < nks; ++ks) {
    store();
    __syncthreads();
    if(ks + 1 < nks) load(ks + 1);
    bf16x8 fa[TERMS][4], fb[TERMS][4];
#pragma unroll
    for(int q = 0; q < TERMS; ++q) {
#pragma unroll
      for(int i = 0; i < 4; ++i)
        fa[q][i] = *reinterpret_cast<const bf16x8*>(&sA[q][(wm0 + i * 16 + l16) * RS + lk * 8]);
#pragma unroll
      for(int j = 0; j < 4; ++j)
        fb[q][j] = *reinterpret_cast<const bf16x8*>(&sB[q][(wn0 + j * 16 + l16) * RS + lk * 8]);
    }
#pragma unroll
    for(int p = NPROD - 1; p >= 0; --p)  // smallest products first
#pragma unroll
      for(int i = 0; i < 4; ++i)
#pragma unroll
        for(int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              fb[kPairB[p]][j], fa[kPairA[p]][i], acc[i][j], 0, 0, 0);
    __syncthreads();
  }

  // epilogue: C -= acc; lane l16 is the row, 4*(l/16)+r the column
  if(vec_ok && m0 + BM <= M && n0 + BN <= N) {
    epilogue_vec(acc, C, ldc, m0 + wm0, n0 + wn0,
                 reinterpret_cast<float*>(smem) + wave * kEpiWaveFloats);
    return;
  }
#pragma unroll
  for(int i = 0; i < 4; ++i) {
    const int row = m0 + wm0 + i * 16 + l16;
    if(row >= M) continue;
#pragma unroll
    for(int j = 0; j < 4; ++j)
#pragma unroll
      for(int r = 0; r < 4; ++r) {
        const int col = n0 + wn0 + j * 16 + lk * 4 + r;
        if(col < N) {
          S* c = C + row + static_cast<size_t>(col) * ldc;
          st_c(c, ld_c(c) - acc[i][j][r]);
        }
      }
  }
}

template <int TERMS, int WM, int WN, int MODE, typename S>
__global__ __launch_bounds__(64 * WM * WN) void gemm_bf16_kernel(
    int M, int N, int nks, const uint16_t* __restrict__ Ap, int m_pad,
    size_t aplane, const uint16_t* __restrict__ Bp, int n_pad, size_t bplane,
    S* __restrict__ C, int ldc, int tiles_m, int tiles_n, int vec_ok,
    UpdSched sch, const int* __restrict__ claimed) {
  __shared__ __attribute__((aligned(16)))
      uint8_t smem[bf16_smem_bytes<TERMS, WM, WN>()];
  run_tiles<MODE>(tiles_m, tiles_n, sch, claimed, [&](int tm, int tn) {
    bf16_tile<TERMS, WM, WN, S>(tm, tn, M, N, nks, Ap, m_pad, aplane, Bp,
                                n_pad, bplane, C, ldc, vec_ok, smem);
  });
}

// LDS row of the FP8 stages: KS stages of 128 bytes, padded by 16 (the 16
// rows of a fragment read start 4 banks apart).
template <int WM, int WN, int NI, int KS>
constexpr int fp8_smem_bytes() {
  constexpr int ops = (64 * WM + 16 * NI * WN) * (KS * 128 + 16);
  return ops > epi_bytes(WM * WN) ? ops : epi_bytes(WM * WN);
}

// Tile (tm, tn) of C -= A B on FP8 MFMA: WM x WN waves, each 64 rows x
// 16*NI columns (4 x NI MFMA blocks); the workgroup tile is 64*WM x
// 16*NI*WN.
template <int WM, int WN, int NI, int KS, typename S>
__device__ __forceinline__ void fp8_tile(
    int tm, int tn, int M, int N, int nks, const uint8_t* __restrict__ Ap,
    int m_pad, const int* __restrict__ ea, const uint8_t* __restrict__ Bp,
    int n_pad, const int* __restrict__ eb, int ntn128, S* __restrict__ C,
    int ldc, int vec_ok, uint8_t* smem) {
  constexpr int NT  = 64 * WM * WN;
  constexpr int BM  = 64 * WM;
  constexpr int WNC = 16 * NI;  // columns per wave
  constexpr int BN  = WNC * WN;
  constexpr int RS  = KS * 128 + 16;
  constexpr int ACH = BM * 8 / NT;  // 16-byte chunks per thread per stage
  constexpr int BCH = BN * 8 / NT;
  uint8_t* sA = smem;
  uint8_t* sB = smem + BM * RS;

  const int m0 = tm * BM, n0 = tn * BN;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm0 = (wave % WM) * 64, wn0 = (wave / WM) * WNC;
  const int l16 = lane & 15, lk = lane >> 4;
  const int atile = (m0 + wm0) / 128;  // this wave's scale tiles
  const int btile = (n0 + wn0) / 128;
  const int nit   = (nks + KS - 1) / KS;

  i32x4 ra[KS][ACH], rb[KS][BCH];
  auto load = [&](int it) {
#pragma unroll
    for(int h = 0; h < KS; ++h) {
      const int ks = it * KS + h;
      if(ks >= nks) break;
      const uint8_t* a = Ap + (static_cast<size_t>(ks) * m_pad + m0) * 128;
      const uint8_t* b = Bp + (static_cast<size_t>(ks) * n_pad + n0) * 128;
#pragma unroll
      for(int c = 0; c < ACH; ++c)
        ra[h][c] = *reinterpret_cast<const i32x4*>(a + (t + c * NT) * 16);
#pragma unroll
      for(int c = 0; c < BCH; ++c)
        rb[h][c] = *reinterpret_cast<const i32x4*>(b + (t + c * NT) * 16);
    }
  };
  auto store = [&]() {
#pragma unroll
    for(int h = 0; h < KS; ++h) {
#pragma unroll
      for(int c = 0; c < ACH; ++c) {
        const int ch = t + c * NT;
        *reinterpret_cast<i32x4*>(&sA[(ch >> 3) * RS + h * 128 + (ch & 7) * 16]) = ra[h][c];
      }
#pragma unroll
      for(int c = 0; c < BCH; ++c) {
        const int ch = t + c * NT;
        *reinterpret_cast<i32x4*>(&sB[(ch >> 3) * RS + h * 128 + (ch & 7) * 16]) = rb[h][c];
      }
    }
  };

  f32x4 acc[4][NI];
#pragma unroll
  for(int i = 0; i < 4; ++i)
#pragma unroll
    for(int j = 0; j < NI; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  load(0);
  for(int it = 0; it < nit; ++it) {
    store();
    __syncthreads();
    if(it + 1 < nit) load(it + 1);
#pragma unroll
    for(int h = 0; h < KS; ++h) {
      const int ks = it * KS + h;
      if(ks >= nks) break;
      // the tiles' power-of-two scales as the MFMA's E8M0 block scales
      // (2^(e - 127) per operand): 2^-(eA + eB) on every product
      const int sa = 127 - ea[atile * nks + ks];
      const int sb = 127 - eb[ks * ntn128 + btile];
      auto frag = [&](const uint8_t* base, int row) {
        const uint8_t* p  = base + (row + l16) * RS + h * 128 + lk * 32;
        const i32x4    lo = *reinterpret_cast<const i32x4*>(p);
        const i32x4    hi = *reinterpret_cast<const i32x4*>(p + 16);
        return i32x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
      };
      i32x8 fa[4];
#pragma unroll
      for(int i = 0; i < 4; ++i) fa[i] = frag(sA, wm0 + i * 16);
      // B fragments one block column at a time (a wave of NI = 8 cannot
      // hold all of them beside 4 x 8 accumulators)
#pragma unroll
      for(int j = 0; j < NI; ++j) {
        const i32x8 fb = frag(sB, wn0 + j * 16);
#pragma unroll
        for(int i = 0; i < 4; ++i)
          // format 0 = e4m3 for both operands
          acc[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
              fb, fa[i], acc[i][j], 0, 0, 0, sb, 0, sa);
      }
    }
    __syncthreads();
  }

  if(vec_ok && m0 + BM <= M && n0 + BN <= N) {
    epilogue_vec(acc, C, ldc, m0 + wm0, n0 + wn0,
                 reinterpret_cast<float*>(smem) + wave * kEpiWaveFloats);
    return;
  }
#pragma unroll
  for(int i = 0; i < 4; ++i) {
    const int row = m0 + wm0 + i * 16 + l16;
    if(row >= M) continue;
#pragma unroll
    for(int j = 0; j < NI; ++j)
#pragma unroll
      for(int r = 0; r < 4; ++r) {
        const int col = n0 + wn0 + j * 16 + lk * 4 + r;
        if(col < N) {
          S* c = C + row + static_cast<size_t>(col) * ldc;
          st_c(c, ld_c(c) - acc[i][j][r]);
        }
      }
  }
}

// MINW: waves per SIMD the register allocation must allow.
template <int WM, int WN, int NI, int KS, int MINW, int MODE, typename S>
__global__ __launch_bounds__(64 * WM * WN, MINW) void gemm_fp8_kernel(
    int M, int N, int nks, const uint8_t* __restrict__ Ap, int m_pad,
    const int* __restrict__ ea, const uint8_t* __restrict__ Bp, int n_pad,
    const int* __restrict__ eb, int ntn128, S* __restrict__ C, int ldc,
    int tiles_m, int tiles_n, int vec_ok, UpdSched sch,
    const int* __restrict__ claimed) {
  __shared__ __attribute__((aligned(16)))
      uint8_t smem[fp8_smem_bytes<WM, WN, NI, KS>()];
  run_tiles<MODE>(tiles_m, tiles_n, sch, claimed, [&](int tm, int tn) {
    fp8_tile<WM, WN, NI, KS, S>(tm, tn, M, N, nks, Ap, m_pad, ea, Bp, n_pad,
                                eb, ntn128, C, ldc, vec_ok, smem);
  });
}

constexpr int kWM = 2, kWN = 2;  // 128 x 128 workgroup tiles (BF16)
// FP8: 256 x 128 workgroup tiles of 4 waves, each 64 x 128 (two workgroups
// per CU); 128-deep FP8 stages, one per LDS fill
constexpr int kFp8WM = 4, kFp8WN = 1, kFp8NI = 8, kFp8MinW = 2;
constexpr int kFp8Stages = 1;

// The vector epilogue needs C 16-byte aligned with ldc a multiple of 4
// (every 4-row group of a column starts on a 4-element boundary).
template <typename S>
int vec_epilogue_ok(const S* C, int ldc) {
  return reinterpret_cast<uintptr_t>(C) % 16 == 0 && ldc % 4 == 0;
}

// Grid and LDS reservation of a claimed launch of kernel fn (threads per
// workgroup): one workgroup per slot of the device (occupancy x CUs), each
// asking for its share of the CU's LDS, so that no workgroup of another
// kernel that uses LDS (every kernel of the panel factorisation does) is
// placed beside it: those run on the vacated CUs only.
struct ClaimedShape {
  int grid = 0, pad = 0;
};
hipError_t claimed_shape(const void* fn, int threads, ClaimedShape* out) {
  static std::vector<std::pair<const void*, ClaimedShape>> cache;
  for(const auto& c : cache)
    if(c.first == fn) {
      *out = c.second;
      return hipSuccess;
    }
  int        dev = 0, cus = 0, per_cu = 0, occ = 0;
  hipError_t e   = hipGetDevice(&dev);
  if(e == hipSuccess)
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if(e == hipSuccess)
    e = hipDeviceGetAttribute(
        &per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev);
  if(e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, threads, 0);
  hipFuncAttributes fa;
  if(e == hipSuccess) e = hipFuncGetAttributes(&fa, fn);
  if(e != hipSuccess) return e;
  if(occ < 1 || cus < 1) return hipErrorInvalidValue;
  ClaimedShape cs;
  cs.grid = occ * cus;
  cs.pad  = per_cu / occ - static_cast<int>(fa.sharedSizeBytes);
  if(cs.pad < 0) cs.pad = 0;
  if(cs.pad > 0)
    e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                            cs.pad);
  if(e != hipSuccess) return e;
  cache.emplace_back(fn, cs);
  *out = cs;
  return hipSuccess;
}

// Tiles per XCD share of a T-tile launch (the claim counters' range).
inline hipError_t share_of(long T, int* per) {
  const long p = (T + kXcds - 1) / kXcds;
  if(p * kXcds > 0x7fffffffL || p >= (1L << 24)) return hipErrorInvalidValue;
  *per = static_cast<int>(p);
  return hipSuccess;
}

template <int TERMS, int WM, int WN, typename S>
hipError_t gemm_bf16_launch(hipStream_t s, int m, int n, int K,
                            const MxpPacked& P, S* C, int ldc,
                            const MxpLaunch& L) {
  constexpr int BM = 64 * WM, BN = 64 * WN, NT = 64 * WM * WN;
  const int  tiles_m = (m + BM - 1) / BM, tiles_n = (n + BN - 1) / BN;
  int        per     = 0;
  hipError_t e       = share_of(static_cast<long>(tiles_m) * tiles_n, &per);
  if(e != hipSuccess) return e;
  const UpdSched sch{L.claim, per, L.vacate, L.stop};
  const int      nks = (K + 31) / 32, vec = vec_epilogue_ok(C, ldc);
  const auto*    A   = static_cast<const uint16_t*>(P.a);
  const auto*    B   = static_cast<const uint16_t*>(P.b);
  const dim3     plain(static_cast<unsigned>(per) * kXcds);
  if(L.mode == kModePlain) {
    hipLaunchKernelGGL((gemm_bf16_kernel<TERMS, WM, WN, kModePlain, S>), plain,
                       dim3(NT), 0, s, m, n, nks, A, P.m_pad, P.aplane, B,
                       P.n_pad, P.bplane, C, ldc, tiles_m, tiles_n, vec, sch,
                       L.claim);
  } else if(L.mode == kModeResume) {
    if(!L.claim) return hipErrorInvalidValue;
    hipLaunchKernelGGL((gemm_bf16_kernel<TERMS, WM, WN, kModeResume, S>), plain,
                       dim3(NT), 0, s, m, n, nks, A, P.m_pad, P.aplane, B,
                       P.n_pad, P.bplane, C, ldc, tiles_m, tiles_n, vec, sch,
                       L.claim);
  } else if(L.mode == kModeClaimed) {
    if(!L.claim || !L.stop) return hipErrorInvalidValue;
    ClaimedShape cs;
    e = claimed_shape(reinterpret_cast<const void*>(
                          gemm_bf16_kernel<TERMS, WM, WN, kModeClaimed, S>),
                      NT, &cs);
    if(e != hipSuccess) return e;
    hipLaunchKernelGGL((gemm_bf16_kernel<TERMS, WM, WN, kModeClaimed, S>),
                       dim3(cs.grid), dim3(NT), cs.pad, s, m, n, nks, A,
                       P.m_pad, P.aplane, B, P.n_pad, P.bplane, C, ldc, tiles_m,
                       tiles_n, vec, sch, L.claim);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Tiles: 128 x 128 for 2 and 3 terms; 256 x 128 (8 waves) for plain BF16,
// whose stages are small enough that the taller tile's operand reuse pays
// (2.17 against 2.60 ms at 32768 x 32768 x 512; for 3 terms the 128 x 128
// tile is the faster, 6.5 against 6.9 ms).  Rows are padded to 256.
template <typename S>
hipError_t gemm_bf16_t(hipStream_t s, int terms, int m, int n, int K,
                       const MxpPacked& P, S* C, int ldc, const MxpLaunch& L) {
  if(m <= 0 || n <= 0 || K <= 0) return hipSuccess;
  switch(terms) {
    case 1: return gemm_bf16_launch<1, 4, 2>(s, m, n, K, P, C, ldc, L);
    case 2: return gemm_bf16_launch<2, kWM, kWN>(s, m, n, K, P, C, ldc, L);
    case 3: return gemm_bf16_launch<3, kWM, kWN>(s, m, n, K, P, C, ldc, L);
    default: return hipErrorInvalidValue;
  }
}

template <typename S>
hipError_t gemm_fp8_t(hipStream_t s, int m, int n, int K, const MxpPacked& P,
                      S* C, int ldc, const MxpLaunch& L) {
  if(m <= 0 || n <= 0 || K <= 0) return hipSuccess;
  constexpr int BM = 64 * kFp8WM, BN = 16 * kFp8NI * kFp8WN;
  constexpr int NT = 64 * kFp8WM * kFp8WN;
  static_assert(kMxpRowPad % BM == 0 && kMxpRowPad % BN == 0,
                "packed operands are padded to whole tiles");
  const int  tiles_m = (m + BM - 1) / BM, tiles_n = (n + BN - 1) / BN;
  int        per     = 0;
  hipError_t e       = share_of(static_cast<long>(tiles_m) * tiles_n, &per);
  if(e != hipSuccess) return e;
  const UpdSched sch{L.claim, per, L.vacate, L.stop};
  const int      nks = (K + 127) / 128, vec = vec_epilogue_ok(C, ldc);
  const auto*    A   = static_cast<const uint8_t*>(P.a);
  const auto*    B   = static_cast<const uint8_t*>(P.b);
  const dim3     plain(static_cast<unsigned>(per) * kXcds);
  if(L.mode == kModePlain) {
    hipLaunchKernelGGL((gemm_fp8_kernel<kFp8WM, kFp8WN, kFp8NI, kFp8Stages,
                                        kFp8MinW, kModePlain, S>),
                       plain, dim3(NT), 0, s, m, n, nks, A, P.m_pad, P.ea, B,
                       P.n_pad, P.eb, P.n_pad / 128, C, ldc, tiles_m, tiles_n,
                       vec, sch, L.claim);
  } else if(L.mode == kModeResume) {
    if(!L.claim) return hipErrorInvalidValue;
    hipLaunchKernelGGL((gemm_fp8_kernel<kFp8WM, kFp8WN, kFp8NI, kFp8Stages,
                                        kFp8MinW, kModeResume, S>),
                       plain, dim3(NT), 0, s, m, n, nks, A, P.m_pad, P.ea, B,
                       P.n_pad, P.eb, P.n_pad / 128, C, ldc, tiles_m, tiles_n,
                       vec, sch, L.claim);
  } else if(L.mode == kModeClaimed) {
    if(!L.claim || !L.stop) return hipErrorInvalidValue;
    ClaimedShape cs;
    e = claimed_shape(
        reinterpret_cast<const void*>(
            gemm_fp8_kernel<kFp8WM, kFp8WN, kFp8NI, kFp8Stages, kFp8MinW,
                            kModeClaimed, S>),
        NT, &cs);
    if(e != hipSuccess) return e;
    hipLaunchKernelGGL((gemm_fp8_kernel<kFp8WM, kFp8WN, kFp8NI, kFp8Stages,
                                        kFp8MinW, kModeClaimed, S>),
                       dim3(cs.grid), dim3(NT), cs.pad, s, m, n, nks, A,
                       P.m_pad, P.ea, B, P.n_pad, P.eb, P.n_pad / 128, C, ldc,
                       tiles_m, tiles_n, vec, sch, L.claim);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

size_t mxp_pack_bytes(int update, int terms, int rows, int K) {
  const size_t rp = static_cast<size_t>((rows + kMxpRowPad - 1) / kMxpRowPad) * kMxpRowPad;
  if(update == HPLX_UPDATE_FP8) {
    const size_t kp = static_cast<size_t>((K + 127) / 128) * 128;
    return rp * kp;  // one byte per element
  }
  const size_t kp = static_cast<size_t>((K + 31) / 32) * 32;
  return rp * kp * 2 * static_cast<size_t>(terms);
}

size_t mxp_scale_count(int rows, int K) {
  return static_cast<size_t>((rows + kMxpRowPad - 1) / kMxpRowPad) *
         (kMxpRowPad / 128) * static_cast<size_t>((K + 127) / 128);
}

hipError_t launch_mxp_pack(hipStream_t s, int update, int terms, int m, int K,
                           const float* L, int ldl, int n, const float* U,
                           int ldu, MxpPacked& P) {
  P.m_pad = (m + kMxpRowPad - 1) / kMxpRowPad * kMxpRowPad;
  P.n_pad = (n + kMxpRowPad - 1) / kMxpRowPad * kMxpRowPad;
  if(m <= 0 || n <= 0 || K <= 0) return hipSuccess;
  if(update == HPLX_UPDATE_FP8) {
    const int nks = (K + 127) / 128;
    P.aplane = P.bplane = 0;
    hipLaunchKernelGGL(pack_a_fp8_kernel, dim3(P.m_pad / 128, nks), dim3(256), 0, s,
                       m, K, L, ldl, static_cast<uint8_t*>(P.a), P.m_pad, nks, P.ea);
    hipLaunchKernelGGL(pack_b_fp8_kernel, dim3(P.n_pad / 128, nks), dim3(256), 0, s,
                       K, n, U, ldu, static_cast<uint8_t*>(P.b), P.n_pad, P.n_pad / 128,
                       P.eb);
    return hipGetLastError();
  }
  if(update != HPLX_UPDATE_BF16) return hipErrorInvalidValue;
  const int nks = (K + 31) / 32;
  P.aplane      = static_cast<size_t>(nks) * P.m_pad * 32;
  P.bplane      = static_cast<size_t>(nks) * P.n_pad * 32;
  uint16_t* A   = static_cast<uint16_t*>(P.a);
  uint16_t* B   = static_cast<uint16_t*>(P.b);
  const dim3 ga(P.m_pad / 64, nks), gb(P.n_pad / 64, nks), b(256);
  switch(terms) {
    case 1:
      hipLaunchKernelGGL(pack_a_bf16_kernel<1>, ga, b, 0, s, m, K, L, ldl, A, P.m_pad, P.aplane);
      hipLaunchKernelGGL(pack_b_bf16_kernel<1>, gb, b, 0, s, K, n, U, ldu, B, P.n_pad, P.bplane);
      break;
    case 2:
      hipLaunchKernelGGL(pack_a_bf16_kernel<2>, ga, b, 0, s, m, K, L, ldl, A, P.m_pad, P.aplane);
      hipLaunchKernelGGL(pack_b_bf16_kernel<2>, gb, b, 0, s, K, n, U, ldu, B, P.n_pad, P.bplane);
      break;
    case 3:
      hipLaunchKernelGGL(pack_a_bf16_kernel<3>, ga, b, 0, s, m, K, L, ldl, A, P.m_pad, P.aplane);
      hipLaunchKernelGGL(pack_b_bf16_kernel<3>, gb, b, 0, s, K, n, U, ldu, B, P.n_pad, P.bplane);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_mxp_gemm(hipStream_t s, int update, int terms, int m, int n,
                           int K, const MxpPacked& P, float* C, int ldc,
                           const MxpLaunch& L) {
  return update == HPLX_UPDATE_FP8 ? gemm_fp8_t(s, m, n, K, P, C, ldc, L)
                                   : gemm_bf16_t(s, terms, m, n, K, P, C, ldc, L);
}

hipError_t launch_mxp_gemm(hipStream_t s, int update, int terms, int m, int n,
                           int K, const MxpPacked& P, __half* C, int ldc,
                           const MxpLaunch& L) {
  return update == HPLX_UPDATE_FP8 ? gemm_fp8_t(s, m, n, K, P, C, ldc, L)
                                   : gemm_bf16_t(s, terms, m, n, K, P, C, ldc, L);
}

void mxp_tile(int update, int terms, int* bm, int* bn) {
  if(update == HPLX_UPDATE_FP8) {
    *bm = 64 * kFp8WM;
    *bn = 16 * kFp8NI * kFp8WN;
  } else if(terms == 1) {
    *bm = 64 * 4;
    *bn = 64 * 2;
  } else {
    *bm = 64 * kWM;
    *bn = 64 * kWN;
  }
}

MxpPacked mxp_cols(const MxpPacked& P, int update, int c0) {
  MxpPacked Q = P;
  if(update == HPLX_UPDATE_FP8) {
    Q.b  = static_cast<uint8_t*>(P.b) + static_cast<size_t>(c0) * 128;
    Q.eb = P.eb + c0 / 128;
  } else {
    Q.b = static_cast<uint16_t*>(P.b) + static_cast<size_t>(c0) * 32;
  }
  return Q;
}

}  // namespace hplx
