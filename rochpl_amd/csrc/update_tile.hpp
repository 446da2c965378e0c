// One 128 x 128 tile of the streamed FP64 update C -= A * Bt^T (the tile
// routine of update_kernel, update_claimed_kernel and update_resume_kernel
// in k_update.hip, whose header describes the stage pipeline), shared with
// the persistent panel kernel (k_panel_kernel.hip), which runs the panel
// recursion's products with it.
#pragma once

#include <hip/hip_runtime.h>

#include "lds_stream.hpp"

namespace hplx {
namespace upd {

constexpr int UBK  = 8;    // K per stage
constexpr int UNS  = 4;    // stages in the LDS ring
constexpr int UROW = 130;  // LDS row pitch in doubles (128 + 2)
constexpr int kUXcds = 8, kUGroupM = 8;

typedef double ud4 __attribute__((ext_vector_type(4)));

// Tiles of the claimed mode (claim.hpp): counters claim[32 x] for XCD x,
// tile ranges [x per, (x + 1) per) of the order below.

#define HPLX_USLOT(n)                                             \
  __shared__ __attribute__((aligned(16))) double sA##n[UBK][UROW]; \
  __shared__ __attribute__((aligned(16))) double sB##n[UBK][UROW];

// One 128 x 128 tile of C (M x N) -= A (M x K, column-major) * Bt^T (Bt: K
// rows of N, row-major, i.e. column-major N x K).  a_rows / b_cols: readable
// extent of A's columns / Bt's rows from the given base (clamp bound for
// edge tiles).
__device__ __forceinline__ void update_tile(
    int tile, int M, int N, int K, const double* __restrict__ A, int lda,
    const double* __restrict__ Bt, int ldb, double* __restrict__ C, int ldc,
    int tiles_n, int tiles_m, int a_rows, int b_cols,
    double (*sA0)[UROW], double (*sB0)[UROW], double (*sA1)[UROW],
    double (*sB1)[UROW], double (*sA2)[UROW], double (*sB2)[UROW],
    double (*sA3)[UROW], double (*sB3)[UROW]) {
  const int width = kUGroupM * tiles_n, group = tile / width;
  const int firstm = group * kUGroupM, gsz = min(tiles_m - firstm, kUGroupM);
  const int within = tile - group * width;
  const int m0 = (firstm + within % gsz) * 128, n0 = (within / gsz) * 128;
  // the thread index passes through an empty volatile asm so that a caller's
  // tile loop cannot hoist the per-lane addressing out of the loop and hold
  // it in registers across the K loop
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  const int lane = t & 63, wave = t >> 6;
  const int wm0 = (wave & 1) * 64, wn0 = (wave >> 1) * 32;
  const int l16 = lane & 15, lk = lane >> 4;

  // wave w loads k row w of every stage, lane l the pair 2l
  const int oa = min(m0 + 2 * lane, a_rows - 2);
  const int ob = min(n0 + 2 * lane, b_cols - 2);
  const double* pa0 = A + oa + static_cast<size_t>(wave) * lda;
  const double* pb0 = Bt + ob + static_cast<size_t>(wave) * ldb;
  const int nst = K / UBK;
  // past the last stage the last one is loaded again, so every iteration
  // waits for the same number of outstanding loads
  auto issue = [&](int st, double (*a)[UROW], double (*b)[UROW]) {
    const int s = st < nst ? st : nst - 1;
    const double* pa = pa0 + static_cast<size_t>(s) * UBK * lda;
    const double* pb = pb0 + static_cast<size_t>(s) * UBK * ldb;
    lds_load16(pa, &a[wave][0]);
    lds_load16(pb, &b[wave][0]);
  };
  ud4 acc[4][2];
#pragma unroll
  for(int i = 0; i < 4; ++i)
#pragma unroll
    for(int j = 0; j < 2; ++j) acc[i][j] = ud4{0.0, 0.0, 0.0, 0.0};
  // fragments of one 4-deep half stage, and the 16 MFMAs that consume them
  auto frag = [&](double (*a)[UROW], double (*b)[UROW], int ks, double* fa,
                  double* fb) {
#pragma unroll
    for(int i = 0; i < 4; ++i) fa[i] = a[ks * 4 + lk][wm0 + i * 16 + l16];
#pragma unroll
    for(int j = 0; j < 2; ++j) fb[j] = b[ks * 4 + lk][wn0 + j * 16 + l16];
  };
  auto mma = [&](const double* fa, const double* fb) {
#pragma unroll
    for(int i = 0; i < 4; ++i)
#pragma unroll
      for(int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[j], fa[i],
                                                         acc[i][j], 0, 0, 0);
  };

  issue(0, sA0, sB0);
  issue(1, sA1, sB1);
  issue(2, sA2, sB2);
  double f0a[4], f0b[2], f1a[4], f1b[2];
  wait_vm<2 * (UNS - 2)>();
  __builtin_amdgcn_s_barrier();
  frag(sA0, sB0, 0, f0a, f0b);
  // stage s lives in slot s % 4; unrolled by 4 so every slot is static.
  // Per stage: read the second half's fragments, multiply the first half,
  // wait for this wave's loads of stage s + 1 and pass the barrier that
  // publishes them (and frees the slot of stage s - 1), refill that slot,
  // read stage s + 1's first fragments, multiply the second half.
  for(int s = 0; s < nst; s += UNS) {
#define HPLX_USTEP(q, ca, cb, na, nb, ra, rb)                                \
  if(s + q < nst) {                                                          \
    frag(ca, cb, 1, f1a, f1b);                                               \
    mma(f0a, f0b);                                                           \
    wait_vm<2 * (UNS - 3)>(); /* this wave's loads of stage s+q+1 are in */  \
    __builtin_amdgcn_s_barrier(); /* every wave's; slot of s+q-1 free */     \
    issue(s + q + UNS - 1, ra, rb);                                          \
    if(s + q + 1 < nst) frag(na, nb, 0, f0a, f0b);                           \
    mma(f1a, f1b);                                                           \
  }
    HPLX_USTEP(0, sA0, sB0, sA1, sB1, sA3, sB3)
    HPLX_USTEP(1, sA1, sB1, sA2, sB2, sA0, sB0)
    HPLX_USTEP(2, sA2, sB2, sA3, sB3, sA1, sB1)
    HPLX_USTEP(3, sA3, sB3, sA0, sB0, sA2, sB2)
#undef HPLX_USTEP
  }
  wait_vm<0>();  // the trailing reloads

  // D[4r + l/16][l%16] of block (i, j) is C[row = l16][col = 4r + lk]
  if(m0 + 128 <= M && n0 + 128 <= N) {
    // interior tile: the 8 values of block row i + 1 are loaded before
    // those of block row i are stored, so up to 16 loads per lane are in
    // flight instead of one load-store round trip per element
    double* cb = C + (m0 + wm0 + l16) + static_cast<size_t>(n0 + wn0 + lk) * ldc;
    double cv[2][8];
#pragma unroll
    for(int e = 0; e < 8; ++e)
      cv[0][e] = cb[static_cast<size_t>((e >> 2) * 16 + 4 * (e & 3)) * ldc];
#pragma unroll
    for(int i = 0; i < 4; ++i) {
      if(i + 1 < 4) {
#pragma unroll
        for(int e = 0; e < 8; ++e)
          cv[(i + 1) & 1][e] =
              cb[(i + 1) * 16 + static_cast<size_t>((e >> 2) * 16 + 4 * (e & 3)) * ldc];
      }
#pragma unroll
      for(int e = 0; e < 8; ++e)
        cb[i * 16 + static_cast<size_t>((e >> 2) * 16 + 4 * (e & 3)) * ldc] =
            cv[i & 1][e] - acc[i][e >> 2][e & 3];
    }
    return;
  }
#pragma unroll
  for(int i = 0; i < 4; ++i) {
    const int row = m0 + wm0 + i * 16 + l16;
    if(row >= M) continue;
#pragma unroll
    for(int j = 0; j < 2; ++j)
#pragma unroll
      for(int r = 0; r < 4; ++r) {
        const int col = n0 + wn0 + j * 16 + 4 * r + lk;
        if(col < N) {
          double* c = C + row + static_cast<size_t>(col) * ldc;
          *c        = *c - acc[i][j][r];
        }
      }
  }
}

}  // namespace upd
}  // namespace hplx
