// The FP64 trailing update A22 -= L21 * U12 (rocblas_dgemm of
// HPL_pdupdateTT, src/pgesv/HPL_pdupdateTT.cpp:131-144), with U12 given
// transposed (n-contiguous), as the reference also multiplies by a
// transposed copy of U (HPL_dlatcpy at :107).
//
// Both operands are then contiguous along the tile edge, which lets every
// stage move from HBM straight into LDS without passing through registers:
//   - 128 x 128 workgroup tile, 8 waves each 64 x 32 = 4 x 2 blocks of
//     V_MFMA_F64_16X16X4_F64 (64 accumulator VGPRs, so two workgroups per
//     CU give four waves per SIMD to cover each other's waits), the
//     transposed block computed so that 16 lanes store 16 consecutive rows
//     of a C column;
//   - K in 8-deep stages through a ring of 4 LDS slots (66.5 KB per
//     workgroup); each stage is 16 GLOBAL_LOAD_LDS_DWORDX4 (one 128-double
//     LDS row each, rows padded to 130 doubles; wave w loads K row w of both
//     operands), issued two stages before the barrier that publishes it;
//   - per stage one wait for this wave's loads of the next stage and one
//     workgroup barrier, placed between the stage's two 4-deep halves:
//     the second half's fragments are read before it and the next stage's
//     first fragments right after it, so no MFMA block waits for an LDS
//     read that the barrier held back; after the barrier the slot of the
//     previous stage is refilled;
//   - rows of L21 / columns of U12^T beyond the matrix are read from a
//     clamped (in-bounds) address: they only reach C entries outside the
//     matrix, which are not stored, so no masking is needed; K must be a
//     multiple of 8 (HPL's NB);
//   - workgroup -> tile mapping XCD-aware (8 XCDs, groups of 8 tile rows);
//   - interior tiles update C with up to 16 loads per lane in flight.
//
// Claimed mode (UpdSched.claim set; the look-ahead's split update): a fixed
// grid of workgroups loops over tiles, taking them from one counter per XCD
// (its share of the tile order above; an XCD whose share is exhausted takes
// tiles of the others), so two launches can drain one update together.  The
// next claim is issued before a tile's K loop and read after it.  A
// workgroup that lands on a CU of UpdSched.vacate (a __smid() bitmap) leaves
// at once: while such a launch holds the other CUs, those CUs stay free for
// the panel factorisation queued beside it.
//
// The stage loads come from lds_stream.hpp; the kernel waits explicitly, no
// compiler-issued vector memory access is in flight inside the K loop, and
// all stage loads are drained before the epilogue.
#include "claim.hpp"
#include "hplx_internal.hpp"
#include "update_tile.hpp"

namespace hplx {
namespace {

using namespace upd;

// kPanel marks the panel products (HPL_pdrpanrlT's GEMMs) so that kernel
// traces tell them from the trailing update; the code is the same.
template <bool kPanel>
__global__ __launch_bounds__(512, 2) void update_kernel(
    int M, int N, int K, const double* __restrict__ A, int lda,
    const double* __restrict__ Bt, int ldb, double* __restrict__ C, int ldc,
    int tiles_m, int tiles_n, int a_rows, int b_cols) {
  HPLX_USLOT(0) HPLX_USLOT(1) HPLX_USLOT(2) HPLX_USLOT(3)
  const int T = tiles_m * tiles_n, per = gridDim.x / kUXcds, bid = blockIdx.x;
  const int tile = (bid % kUXcds) * per + bid / kUXcds;
  if(tile >= T) return;
  update_tile(tile, M, N, K, A, lda, Bt, ldb, C, ldc, tiles_n, tiles_m, a_rows,
              b_cols, sA0, sB0, sA1, sB1, sA2, sB2, sA3, sB3);
}

// The rest of an update after a claimed launch has stopped: the same
// one-tile-per-workgroup kernel, starting each XCD share where the claims
// ended (claimed[32 x], read after the claimed launch has finished; grid
// 8 x per, workgroups past their share's end exit).
__global__ __launch_bounds__(512, 2) void update_resume_kernel(
    int M, int N, int K, const double* __restrict__ A, int lda,
    const double* __restrict__ Bt, int ldb, double* __restrict__ C, int ldc,
    int tiles_m, int tiles_n, int a_rows, int b_cols,
    const int* __restrict__ claimed, int per) {
  HPLX_USLOT(0) HPLX_USLOT(1) HPLX_USLOT(2) HPLX_USLOT(3)
  const int x = blockIdx.x % kUXcds, i = blockIdx.x / kUXcds;
  const int done = min(claimed[32 * x], per);
  const int tile = x * per + done + i;
  if(done + i >= per || tile >= tiles_m * tiles_n) return;
  update_tile(tile, M, N, K, A, lda, Bt, ldb, C, ldc, tiles_n, tiles_m, a_rows,
              b_cols, sA0, sB0, sA1, sB1, sA2, sB2, sA3, sB3);
}

// The claimed mode: a persistent grid whose workgroups take kClaimChunk
// consecutive tiles per counter update (thread 0, after the epilogue of the
// chunk's last tile) until the tiles run out or *stop becomes non-zero.  The
// claim state lives in LDS.  Like update_kernel it is held to 128 VGPRs (four
// waves per SIMD, two workgroups per CU); the tile loop then spills a few
// registers once per tile, outside the K loop.
struct UpdClaim {
  int      tile, end;  // current chunk [tile, end), tile < 0: none left
  unsigned dead;
};

// The launch also reserves the rest of the CU's LDS (dynamic shared memory it
// does not use; each of the two workgroups on a CU takes half), so that no
// other kernel's workgroup that uses LDS (every kernel of the panel
// factorisation does) is placed beside them: those run on the vacated CUs
// only, instead of competing with the update for a SIMD.
__global__ __launch_bounds__(512, 4) void update_claimed_kernel(
    int M, int N, int K, const double* __restrict__ A, int lda,
    const double* __restrict__ Bt, int ldb, double* __restrict__ C, int ldc,
    int tiles_m, int tiles_n, int a_rows, int b_cols, UpdSched sch) {
  HPLX_USLOT(0) HPLX_USLOT(1) HPLX_USLOT(2) HPLX_USLOT(3)
  __shared__ UpdClaim sc;
  // all waves of a workgroup share one CU, so this exit is uniform
  if(on_vacated_cu(sch)) return;
  const int T = tiles_m * tiles_n;
  // thread 0: claim the next chunk into sc, or none once *stop is set
  auto claim_next = [&](unsigned dead) {
    const int t0 = claim_next_tile(sch, T, dead);
    sc.tile = t0;
    // a chunk ends at its XCD share's end and at the last tile
    sc.end  = t0 < 0 ? t0 : min(min(t0 + kClaimChunk, T),
                                (t0 / sch.per + 1) * sch.per);
    sc.dead = dead;
  };
  if(threadIdx.x == 0) claim_next(0u);
  __syncthreads();
  for(;;) {
    const int t0 = __builtin_amdgcn_readfirstlane(sc.tile);
    if(t0 < 0) break;
    const int end = __builtin_amdgcn_readfirstlane(sc.end);
    for(int tile = t0; tile < end; ++tile) {
      if(tile > t0) __syncthreads();  // the LDS ring is free again
      update_tile(tile, M, N, K, A, lda, Bt, ldb, C, ldc, tiles_n, tiles_m,
                  a_rows, b_cols, sA0, sB0, sA1, sB1, sA2, sB2, sA3, sB3);
    }
    __syncthreads();  // every wave is done with sc and the LDS ring
    if(threadIdx.x == 0) claim_next(sc.dead);
    __syncthreads();
  }
}
#undef HPLX_USLOT

// B (cols x rows, row-major with ldb) = A^T for A (rows x cols, column-major
// with lda); 32 x 32 tiles through LDS.
__global__ __launch_bounds__(256) void transpose_kernel(
    int rows, int cols, const double* __restrict__ A, int lda,
    double* __restrict__ B, int ldb, int gx) {
  __shared__ double tile[32][33];
  const int bx = blockIdx.x % gx, by = blockIdx.x / gx;
  const int r0 = bx * 32, c0 = by * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
  for(int q = 0; q < 4; ++q) {
    const int c = c0 + ty + 8 * q, r = r0 + tx;
    if(r < rows && c < cols) tile[ty + 8 * q][tx] = A[r + static_cast<size_t>(c) * lda];
  }
  __syncthreads();
#pragma unroll
  for(int q = 0; q < 4; ++q) {
    const int r = r0 + ty + 8 * q, c = c0 + tx;
    if(r < rows && c < cols) B[c + static_cast<size_t>(r) * ldb] = tile[tx][ty + 8 * q];
  }
}

}  // namespace

bool update_fast_ok(int m, int n, int k, const double* A, int lda,
                    const double* Bt, int ldb, int a_rows, int b_cols) {
  const auto al16 = [](const void* p) {
    return reinterpret_cast<uintptr_t>(p) % 16 == 0;
  };
  // the stage loads move row (column) pairs: an edge pair is clamped to
  // start at a_rows - 2 (b_cols - 2), which keeps its place in the pair grid
  // only when that extent is even
  return m > 0 && n > 0 && k > 0 && k % UBK == 0 && al16(A) && al16(Bt) &&
         lda % 2 == 0 && ldb % 2 == 0 && a_rows >= 2 && b_cols >= 2 &&
         a_rows % 2 == 0 && b_cols % 2 == 0 && a_rows >= m && b_cols >= n;
}

namespace {

struct UpdGrid {
  int  tiles_m, tiles_n;
  long T;
};

hipError_t update_grid(int m, int n, int k, const double* A, int lda,
                       const double* Bt, int ldb, int a_rows, int b_cols,
                       UpdGrid* g) {
  if(!update_fast_ok(m, n, k, A, lda, Bt, ldb, a_rows, b_cols))
    return hipErrorInvalidValue;
  g->tiles_m = (m + 127) / 128;
  g->tiles_n = (n + 127) / 128;
  g->T       = static_cast<long>(g->tiles_m) * g->tiles_n;
  // the claimed mode's counters hold a share index below 2^24
  if(g->T > 0x7fffffffL || (g->T + kUXcds - 1) / kUXcds >= (1L << 24))
    return hipErrorInvalidValue;
  return hipSuccess;
}

__global__ void claim_log_kernel(const int* claim, int per, int* out) {
  const int x = threadIdx.x;
  if(x < kUXcds) out[x] = min(claim[32 * x], per);
}

// *p = 1, visible to workgroups on every XCD while they run.
__global__ void set_flag_kernel(int* p) {
  __hip_atomic_store(p, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void zero_ints_kernel(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) p[i] = 0;
}

// Each workgroup records the __smid() of the CU it ran on, then waits about
// `spin` clock ticks so that the grid spreads over every CU.
__global__ void cu_census_kernel(int* out, long long spin) {
  if(threadIdx.x == 0) out[blockIdx.x] = static_cast<int>(__smid() & 511u);
  const long long t0 = clock64();
  while(clock64() - t0 < spin) {
  }
}

}  // namespace

hipError_t launch_update(hipStream_t s, int m, int n, int k, const double* A,
                         int lda, const double* Bt, int ldb, double* C, int ldc,
                         int a_rows, int b_cols, bool panel) {
  UpdGrid    g;
  hipError_t e = update_grid(m, n, k, A, lda, Bt, ldb, a_rows, b_cols, &g);
  if(e != hipSuccess) return e;
  const long grid = (g.T + kUXcds - 1) / kUXcds * kUXcds;
  if(panel)
    hipLaunchKernelGGL(update_kernel<true>, dim3(static_cast<unsigned>(grid)),
                       dim3(512), 0, s, m, n, k, A, lda, Bt, ldb, C, ldc,
                       g.tiles_m, g.tiles_n, a_rows, b_cols);
  else
    hipLaunchKernelGGL(update_kernel<false>, dim3(static_cast<unsigned>(grid)),
                       dim3(512), 0, s, m, n, k, A, lda, Bt, ldb, C, ldc,
                       g.tiles_m, g.tiles_n, a_rows, b_cols);
  return hipGetLastError();
}

hipError_t launch_update_claimed(hipStream_t s, int m, int n, int k,
                                 const double* A, int lda, const double* Bt,
                                 int ldb, double* C, int ldc, int a_rows,
                                 int b_cols, int* claim,
                                 const unsigned long long* vacate,
                                 const int* stop, int grid) {
  if(!claim || !stop || grid < 1) return hipErrorInvalidValue;
  UpdGrid    g;
  hipError_t e = update_grid(m, n, k, A, lda, Bt, ldb, a_rows, b_cols, &g);
  if(e != hipSuccess) return e;
  UpdSched sch{};
  sch.claim  = claim;
  sch.per    = static_cast<int>((g.T + kUXcds - 1) / kUXcds);
  sch.vacate = vacate;
  sch.stop   = stop;
  // LDS left on a CU beside two workgroups of this kernel, reserved by them
  static int pad = -1;
  if(pad < 0) {
    int dev = 0, per_cu = 0;
    hipFuncAttributes fa;
    e = hipGetDevice(&dev);
    if(e == hipSuccess)
      e = hipDeviceGetAttribute(
          &per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev);
    if(e == hipSuccess)
      e = hipFuncGetAttributes(&fa,
                               reinterpret_cast<const void*>(update_claimed_kernel));
    if(e != hipSuccess) return e;
    const int rest = per_cu / 2 - static_cast<int>(fa.sharedSizeBytes);
    if(rest > 0)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(update_claimed_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, rest);
    if(e != hipSuccess) return e;
    pad = rest > 0 ? rest : 0;
  }
  hipLaunchKernelGGL(update_claimed_kernel, dim3(grid), dim3(512), pad, s, m,
                     n, k, A, lda, Bt, ldb, C, ldc, g.tiles_m, g.tiles_n,
                     a_rows, b_cols, sch);
  return hipGetLastError();
}

hipError_t launch_update_resume(hipStream_t s, int m, int n, int k,
                                const double* A, int lda, const double* Bt,
                                int ldb, double* C, int ldc, int a_rows,
                                int b_cols, const int* claimed) {
  if(!claimed) return hipErrorInvalidValue;
  UpdGrid    g;
  hipError_t e = update_grid(m, n, k, A, lda, Bt, ldb, a_rows, b_cols, &g);
  if(e != hipSuccess) return e;
  const long per  = (g.T + kUXcds - 1) / kUXcds;
  const long grid = per * kUXcds;
  hipLaunchKernelGGL(update_resume_kernel, dim3(static_cast<unsigned>(grid)),
                     dim3(512), 0, s, m, n, k, A, lda, Bt, ldb, C, ldc,
                     g.tiles_m, g.tiles_n, a_rows, b_cols, claimed,
                     static_cast<int>(per));
  return hipGetLastError();
}

hipError_t launch_claim_log(hipStream_t s, const int* claim, int m, int n,
                            int* out, int bm, int bn) {
  const long T   = static_cast<long>((m + bm - 1) / bm) * ((n + bn - 1) / bn);
  const int  per = static_cast<int>((T + kUXcds - 1) / kUXcds);
  hipLaunchKernelGGL(claim_log_kernel, dim3(1), dim3(64), 0, s, claim, per,
                     out);
  return hipGetLastError();
}

double claimed_flops(int m, int n, int k, const int* done, int bm, int bn) {
  const int  tiles_m = (m + bm - 1) / bm, tiles_n = (n + bn - 1) / bn;
  const long T       = static_cast<long>(tiles_m) * tiles_n;
  const long per     = (T + kUXcds - 1) / kUXcds;
  const long width   = static_cast<long>(kUGroupM) * tiles_n;
  double     f       = 0.0;
  for(int x = 0; x < kUXcds; ++x)
    for(long t = x * per; t < x * per + done[x] && t < T; ++t) {
      const long group = t / width, firstm = group * kUGroupM;
      const long gsz   = tiles_m - firstm < kUGroupM ? tiles_m - firstm : kUGroupM;
      const long within = t - group * width;
      const long tm = firstm + within % gsz, tn = within / gsz;
      const long h = m - tm * bm < bm ? m - tm * bm : bm;
      const long w = n - tn * bn < bn ? n - tn * bn : bn;
      f += 2.0 * static_cast<double>(h * w) * k;
    }
  return f;
}

hipError_t launch_set_flag(hipStream_t s, int* p) {
  hipLaunchKernelGGL(set_flag_kernel, dim3(1), dim3(1), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_zero_ints(hipStream_t s, int* p, int n) {
  if(n <= 0) return hipSuccess;
  hipLaunchKernelGGL(zero_ints_kernel, dim3((n + 255) / 256), dim3(256), 0, s,
                     p, n);
  return hipGetLastError();
}

hipError_t launch_cu_census(hipStream_t s, int* out, int grid) {
  if(grid < 1) return hipSuccess;
  hipLaunchKernelGGL(cu_census_kernel, dim3(grid), dim3(64), 0, s, out,
                     20000LL);
  return hipGetLastError();
}

hipError_t launch_transpose(hipStream_t s, int rows, int cols, const double* A,
                            int lda, double* B, int ldb) {
  if(rows <= 0 || cols <= 0) return hipSuccess;
  const long gx = (rows + 31) / 32, gy = (cols + 31) / 32;
  if(gx * gy > 0x7fffffffL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(transpose_kernel, dim3(static_cast<unsigned>(gx * gy)),
                     dim3(256), 0, s, rows, cols, A, lda, B, ldb,
                     static_cast<int>(gx));
  return hipGetLastError();
}

}  // namespace hplx
