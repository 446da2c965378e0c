// Internal declarations of the mixed-precision factorisation and solve
// (FP32 or FP16 factor storage, BF16 / FP8 trailing update, FP64 iterative
// refinement).
#pragma once

#include "hplx_internal.hpp"

namespace hplx {

// Packed operands are padded to whole multiples of this many rows/columns
// (the largest GEMM tile edge).
constexpr int kMxpRowPad = 256;

// One panel's packed update operands (see k_mxp_gemm.hip for the layout).
struct MxpPacked {
  void*  a      = nullptr;  // L21, m_pad rows
  void*  b      = nullptr;  // U12, n_pad columns
  int*   ea     = nullptr;  // FP8: exponent per (128-row tile, 128-deep stage)
  int*   eb     = nullptr;  // FP8: exponent per (stage, 128-column tile)
  int    m_pad  = 0;
  int    n_pad  = 0;
  size_t aplane = 0;  // BF16: elements per split term
  size_t bplane = 0;
};

// Bytes of one packed operand of `rows` rows (or columns) and depth K.
size_t mxp_pack_bytes(int update, int terms, int rows, int K);
// FP8 scale exponents of one packed operand.
size_t mxp_scale_count(int rows, int K);

// Pack L (m x K, column-major FP32) and U (K x n, column-major FP32); sets
// P.m_pad / n_pad / planes.  P.a, P.b (and P.ea, P.eb for FP8) must be sized
// by mxp_pack_bytes / mxp_scale_count.
hipError_t launch_mxp_pack(hipStream_t s, int update, int terms, int m, int K,
                           const float* L, int ldl, int n, const float* U,
                           int ldu, MxpPacked& P);
// How an update launch is scheduled: plain (one tile per workgroup), or the
// look-ahead's claimed / resume pair (claim: per-XCD counters claim[32 x];
// vacate: the CUs a claimed launch leaves to the panel factorisation; stop:
// the flag that ends its claims; the resume launch continues from claim).
struct MxpLaunch {
  int                       mode   = 0;  // 0 plain, 1 claimed, 2 resume
  int*                      claim  = nullptr;
  const unsigned long long* vacate = nullptr;
  const int*                stop   = nullptr;
};
// C (m x n, storage type) -= L * U from the packed operands.
hipError_t launch_mxp_gemm(hipStream_t s, int update, int terms, int m, int n,
                           int K, const MxpPacked& P, float* C, int ldc,
                           const MxpLaunch& L = MxpLaunch());
hipError_t launch_mxp_gemm(hipStream_t s, int update, int terms, int m, int n,
                           int K, const MxpPacked& P, __half* C, int ldc,
                           const MxpLaunch& L = MxpLaunch());
// The packed U12 from column c0 on (c0 a multiple of 128): the same packing
// seen as starting at column c0, for an update of a column range.
MxpPacked mxp_cols(const MxpPacked& P, int update, int c0);
// The C tile of the update kernel a form runs (BF16 / FP8).
void mxp_tile(int update, int terms, int* bm, int* bn);

// Storage <-> FP32 copies of an m x n block (panel staging of FP16 storage,
// U rows back into storage).
hipError_t launch_copy_block(hipStream_t s, int m, int n, const __half* A,
                             int lda, float* B, int ldb);
hipError_t launch_copy_block(hipStream_t s, int m, int n, const float* A,
                             int lda, __half* B, int ldb);
hipError_t launch_copy_block(hipStream_t s, int m, int n, const double* A,
                             int lda, float* B, int ldb);
hipError_t launch_copy_block(hipStream_t s, int m, int n, const double* A,
                             int lda, __half* B, int ldb);

// ---- solves with low-precision factors and FP64 vectors
// The interchanges y[i] <-> y[ipiv[i]] of one panel (i = k0 .. k0+jb-1 in
// order) given as their net effect, cnt (dst, src) pairs:
// y[dst] = y_old[src] (cnt <= kMaxSwapPairs, all reads before any write).
constexpr int kMaxSwapPairs = 2 * kMaxNB;
hipError_t launch_apply_swaps(hipStream_t s, int cnt, const int* pairs,
                              double* y);
hipError_t launch_apply_swaps(hipStream_t s, int cnt, const int* pairs,
                              float* y);
// y[0:m] -= M[0:m, 0:k] * x[0:k], M in storage type, accumulation in the
// vectors' type; if xback is not null, x[0:k] is also copied there.
hipError_t launch_gemv_sub(hipStream_t s, int m, int k, const float* M,
                           int ldm, const double* x, double* y,
                           double* xback = nullptr);
hipError_t launch_gemv_sub(hipStream_t s, int m, int k, const __half* M,
                           int ldm, const double* x, double* y,
                           double* xback = nullptr);
hipError_t launch_gemv_sub(hipStream_t s, int m, int k, const double* M,
                           int ldm, const double* x, double* y,
                           double* xback = nullptr);
hipError_t launch_gemv_sub(hipStream_t s, int m, int k, const float* M,
                           int ldm, const float* x, float* y,
                           float* xback = nullptr);
hipError_t launch_gemv_sub(hipStream_t s, int m, int k, const __half* M,
                           int ldm, const float* x, float* y,
                           float* xback = nullptr);
// Inverses (FP64 or FP32) of nblk consecutive kb x kb diagonal blocks of T
// (unit lower triangle or upper triangle), block J at Inv + J kb kb, leading
// dimension kb.
hipError_t launch_diag_inverse(hipStream_t s, bool lower, int nblk, int kb,
                               const float* T, int ldt, double* Inv);
hipError_t launch_diag_inverse(hipStream_t s, bool lower, int nblk, int kb,
                               const __half* T, int ldt, double* Inv);
hipError_t launch_diag_inverse(hipStream_t s, bool lower, int nblk, int kb,
                               const float* T, int ldt, float* Inv);
hipError_t launch_diag_inverse(hipStream_t s, bool lower, int nblk, int kb,
                               const __half* T, int ldt, float* Inv);
// y[0:k] = Inv x[0:k], Inv k x k (leading dimension k).
hipError_t launch_inv_gemv(hipStream_t s, int k, const double* Inv,
                           const double* x, double* y);
hipError_t launch_inv_gemv(hipStream_t s, int k, const float* Inv,
                           const float* x, float* y);

// The generator's matrix rounded to FP32 / FP16 storage.
hipError_t launch_matgen(hipStream_t s, int m, int n, float* A, int lda,
                         int seed, double diag_shift);
hipError_t launch_matgen(hipStream_t s, int m, int n, __half* A, int lda,
                         int seed, double diag_shift);

// ---- FP64 refinement
// b (the generator's column n) into y.
hipError_t launch_gen_b(hipStream_t s, int n, int seed, double* y);
// r = b - A x with A regenerated in FP64; out[0] = scaled residual,
// out[1] = ||r||_oo, out[2] = ||A||_oo, out[3] = ||x||_oo, out[4] = ||b||_oo.
size_t     ir_part_len(int n);
hipError_t launch_ir_residual(hipStream_t s, int n, int seed,
                              double diag_shift, const double* x, double* r,
                              double* part, double* out);
// x += d
hipError_t launch_axpy(hipStream_t s, int n, const double* d, double* x);
// x += d with an FP32 correction d (the sum in FP64)
hipError_t launch_axpy(hipStream_t s, int n, const float* d, double* x);
// y = (FP32) x, rounded to nearest
hipError_t launch_to_f32(hipStream_t s, int n, const double* x, float* y);
// x = (FP64) y
hipError_t launch_to_f64(hipStream_t s, int n, const float* y, double* x);

}  // namespace hplx
