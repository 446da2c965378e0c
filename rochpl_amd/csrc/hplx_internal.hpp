// Internal declarations shared by the host ABI layer and the HIP kernels.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hplx.h"

namespace hplx {

// ---------------------------------------------------------------------------
// Limits of the kernels.
// ---------------------------------------------------------------------------
constexpr int kLeafMax     = 32;    // widest leaf the register-resident kernel takes
constexpr int kLeafThreads = 512;   // threads per leaf workgroup
constexpr int kLeafRPT     = 3;     // rows per thread in the leaf kernel

// Look-ahead: columns a step needs right of the next panel before it overlaps
// the rest of the update with that panel (fewer go serially), and the CUs
// per XCD a panel of `rows` rows is given -- one per leaf workgroup (512 x
// RPT rows each), taken evenly from the 8 XCDs and, within each, from all 4
// shader engines (a workgroup waits for a CU of the engine it was handed to,
// so an engine without a free CU would hold up the panel's kernels).
constexpr int kMinOverlapCols = 256;
inline int vacate_set_size(int rows) {
  const int g = (rows + kLeafThreads * kLeafRPT - 1) / (kLeafThreads * kLeafRPT);
  return (g + 31) / 32 * 4;
}
constexpr int kMaxLeafGrid = 256;   // workgroups a cooperative leaf may use
constexpr int kMaxGrid     = kMaxLeafGrid;  // candidate slots per parity
constexpr int kMaxNB       = 1024;  // widest panel (pivot composition, Linv)
constexpr int kSlotStride  = 72;    // doubles per pivot candidate slot

// One step of a panel factorisation as the persistent panel kernel runs it
// (k_panel_kernel.hip): the steps k_panel.hip's recursion would launch, in
// its order, recorded once per (panel width, PanelAlgo) in panel-relative
// coordinates.  kind: kPanelLeaf (columns [ac, ac + n) of rows [ac, m)),
// kPanelTrsm (X (k x n at br, bc) = L^{-1} X, L unit lower at ar, ac),
// kPanelGemm (C (rows x n at cr, cc) -= A (rows x k at ar, ac) B (k x n at
// br, bc)), rows = m_fix if m_fix >= 0, else m - m_sub.
enum : int { kPanelLeaf = 0, kPanelTrsm = 1, kPanelGemm = 2 };
struct PanelOp {
  int kind;
  int m_fix, m_sub;
  int n, k;
  int ar, ac, br, bc, cr, cc;
  int pad;
};
constexpr int kMaxPanelOps = 512;  // steps of one recorded panel
// Panel height from which HPLX_OPT_PANEL_KERNEL = 2 takes the persistent
// kernel: below it the launch sequence is faster (measured: N = 16384 and
// 32768 solves 21 % and 11 % slower with the kernel on every panel; 65536,
// 98304 and 193280 within run-to-run spread).
constexpr int kPanelKernelMinRows = 65536;

// Per-context device state.  The small workspace (candidate slots, barrier
// counter, flags, L11 inverse, composed pivots) lives with the context; the
// n-sized workspace lives with the matrix (MatWs).
struct Ctx {
  int         device     = 0;
  int         num_cus    = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream     = nullptr;  // where work is queued
  std::string last_error;
  uint32_t    leaf_launches = 0;     // sequence numbers of the leaf slots
  int         leaf_resident = 0;     // leaf workgroups resident at once

  void*     small_base = nullptr;
  double*   Linv       = nullptr;    // kMaxNB x kMaxNB
  int*      permU      = nullptr;    // 2*kMaxNB
  double*   slots      = nullptr;    // 2 x (kMaxGrid+1) x kSlotStride
  int*      flags      = nullptr;    // [0] first zero pivot (1-based), [1] error
  double*   dout       = nullptr;    // 16 doubles of small results
  // second L11^{-1} / composed-pivot buffers: with the look-ahead, panel k+1
  // is factored (and its L11^{-1} / permU built) while update k still reads
  // those of panel k
  double*   Linv2      = nullptr;
  int*      permU2     = nullptr;
  int*      claim      = nullptr;    // 2 sets of kClaimInts update-claim counters
  int*      stop       = nullptr;    // 2 flags: claims of update k stop
  unsigned long long* vacate = nullptr;  // kVacateSets x 8 words (see cu_sets)

  // Where panel_factor queues its work and which kMaxNB^2 buffer its wide
  // products use as scratch (null: stream / Linv).
  hipStream_t panel_stream_ = nullptr;
  double*     panel_scratch_ = nullptr;

  // The persistent panel kernel (hplx_set_option HPLX_OPT_PANEL_KERNEL):
  // panel_kernel 1 runs each FP64 panel as one launch, 2 (default) those of
  // at least kPanelKernelMinRows rows, 0 none.  panel_rec, while
  // set, makes the recursion record its steps instead of launching them.
  // Two recorded step lists live on the device (pops, kMaxPanelOps each),
  // keyed by panel width and PanelAlgo; pbar is the kernel's grid barrier
  // counter.
  int                   panel_kernel = 2;
  std::vector<PanelOp>* panel_rec    = nullptr;
  PanelOp*              pops         = nullptr;
  unsigned*             pbar         = nullptr;
  int                   pop_key[2][5] = {{0, 0, 0, 0, 0}, {0, 0, 0, 0, 0}};
  int                   pop_count[2]  = {0, 0};   // 0: slot empty
  int                   pop_leaves[2] = {0, 0};
  int                   pop_next      = 0;
  std::vector<PanelOp>  pop_host[2];
  int                   pk_resident   = 0;        // its workgroups resident at once
  int                   panel_cus     = 0;        // CUs the next panel has (0: all)
  int                   panel_kernel_runs = 0;    // panels of the last solve

  // Look-ahead (hplx_set_option HPLX_OPT_LOOKAHEAD): internal streams of the
  // update and the panel, and the CU sets a panel may be given: vacate set r
  // holds the first r CUs of every XCD (CU ids from a census at context
  // creation; r = 1 .. cus_per_xcd).
  int                     lookahead   = 1;
  hipStream_t             ustream     = nullptr;
  hipStream_t             pstream     = nullptr;
  int                     cus_per_xcd = 0;  // 0: census incomplete, no look-ahead
  std::vector<hipEvent_t> ev_sync;         // timing-disabled, cross-stream order

  // progress lines from the panel loop (HPLX_PROGRESS=<panels>, the
  // counterpart of the reference's HPL_PROGRESS_REPORT build option)
  int                     progress_every = 0;
  std::vector<hipEvent_t> ev_prog;
  // timing of the trailing-update GEMM launches of the last solve: one
  // event pair per launch, its kind (1: a whole update in one launch, 2: the
  // claimed part of a split update, 3: its resume part) and its flops (for
  // kinds 2 and 3 filled from the claim log when the stats are read)
  std::vector<hipEvent_t> ev;
  int                     ev_pairs = 0;
  // the FP64 loop's row interchanges (swap_gather launches): one event pair
  // per launch and their algorithmic bytes (4 x 8 B per interchanged element)
  std::vector<hipEvent_t> ev_swap;
  int                     swap_pairs = 0;
  double                  swap_bytes = 0.0;
  // the FP64 loop's panel factorisations: one event pair per panel (on the
  // stream it runs on) and their algorithmic bytes (the m x jb panel read
  // and written once: 16 B per element)
  std::vector<hipEvent_t> ev_panel;
  int                     panel_pairs = 0;
  double                  panel_bytes = 0.0;
  double                  gemm_flops = 0.0;
  std::vector<int>        ev_kind;
  std::vector<double>     ev_flops;
  // split updates of the last solve: shape and the pair indices; the tiles
  // each XCD share had claimed (8 per split, copied back at the end)
  struct Split {
    int m, n, k, pair_claimed, pair_resume;
    int bm = 128, bn = 128;  // the update kernel's tile
  };
  std::vector<Split> splits;
  std::vector<int>   claim_host;
  int*               claim_log = nullptr;  // device, 8 x kMaxClaimLog ints
};
constexpr int kMaxClaimLog = 8192;  // split updates whose claims are logged

inline hipStream_t panel_stream(const Ctx& c) {
  return c.panel_stream_ ? c.panel_stream_ : c.stream;
}
inline double* panel_scratch(const Ctx& c) {
  return c.panel_scratch_ ? c.panel_scratch_ : c.Linv;
}
constexpr int kVacateSets = 32;  // CUs per XCD on MI355X

struct MatWs {
  double* W     = nullptr;  // max(nb*(n+1), check scratch) doubles
  size_t  W_len = 0;
};

// ---------------------------------------------------------------------------
// Kernel launchers (defined in the .hip files).  All return hipError_t.
// ---------------------------------------------------------------------------
hipError_t launch_matgen(hipStream_t s, int m, int n, double* A, int lda,
                         int seed, double diag_shift);

// Largest panel height the register-resident leaf holds on this device.
inline int max_leaf_rows(const Ctx& c) {
  const int g = c.leaf_resident < kMaxLeafGrid ? c.leaf_resident : kMaxLeafGrid;
  return g * kLeafThreads * kLeafRPT;
}

// Workgroups of the leaf kernel (and of the persistent panel kernel) one CU
// holds (occupancy query).
hipError_t leaf_occupancy(int* blocks_per_cu);


// Zero the flags (queued on c.stream).
hipError_t reset_panel_state(Ctx& c);

// The panel factorisation's parameters (HPL_T_palg's pfact, rfact, nbmin,
// nbdiv; pfact and rfact are HPLX_LEFT_LOOKING, HPLX_CROUT or
// HPLX_RIGHT_LOOKING).
struct PanelAlgo {
  int nbmin = 32, nbdiv = 2;
  int pfact = HPLX_RIGHT_LOOKING, rfact = HPLX_RIGHT_LOOKING;
};
inline bool valid_fact(int f) {
  return f == HPLX_LEFT_LOOKING || f == HPLX_CROUT || f == HPLX_RIGHT_LOOKING;
}
inline PanelAlgo panel_algo(const hplx_palg& p) {
  PanelAlgo a;
  a.nbmin = p.nbmin;
  a.nbdiv = p.nbdiv;
  a.pfact = p.pfact;
  a.rfact = p.rfact;
  return a;
}
inline PanelAlgo panel_algo(int nbmin, int nbdiv) {
  PanelAlgo a;
  a.nbmin = nbmin;
  a.nbdiv = nbdiv;
  return a;
}

// Panel factorisation of an m x jb panel in place (recursive, in the
// variants' orders, register-resident cooperative leaves).  ipiv receives
// panel-relative pivot rows + row_offset; zero pivots are reported as
// row_offset + column + 1.
hipError_t panel_factor(Ctx& c, const PanelAlgo& a, double* P, int lda,
                        int m, int jb, int* ipiv, int row_offset);
// The same recursion in FP32 (FP32 leaves, solves and MFMA GEMMs).
hipError_t panel_factor(Ctx& c, const PanelAlgo& a, float* P, int lda,
                        int m, int jb, int* ipiv, int row_offset);

// Record the steps of an FP64 panel of width jb under algorithm a and keep
// them in one of the context's two device slots (synchronous; call it
// before queuing the panels that use it, never while a panel kernel that
// reads the slot being replaced may still run).  Returns HPLX_SUCCESS or a
// negative HPLX_ERR_*; a panel that cannot run as one kernel (too many
// steps) is left to the launch sequence and is not an error.
int panel_kernel_prepare(Ctx& c, const PanelAlgo& a, int jb);
// The recorded steps of a panel of width jb under a (host only, no context):
// kPanelStepInts ints per step (kind, m_fix, m_sub, n, k, ar, ac, br, bc,
// cr, cc) into out, at most max_steps of them; returns the step count or a
// negative HPLX_ERR_*.
constexpr int kPanelStepInts = 11;
int panel_steps(const PanelAlgo& a, int jb, int* out, int max_steps);
// Workgroups of the persistent panel kernel one CU holds.
hipError_t panel_kernel_occupancy(int* blocks_per_cu);
// The panel from recorded slot `slot` as one launch on the panel stream;
// *ran = false (nothing queued) when this panel does not qualify (grid, or
// a product the streamed tile does not take), and the caller then queues
// the launch sequence.
hipError_t launch_panel_kernel(Ctx& c, int slot, double* P, int lda, int m,
                               int jb, int* ipiv, int row_offset, bool* ran);

hipError_t launch_pipid(hipStream_t s, int jb, const int* ipiv, int ipiv_base,
                        int* permU);
hipError_t launch_swap_gather(hipStream_t s, int jb, int n, double* A,
                              int lda, const int* permU, double* W, int ldw);
hipError_t launch_swap_gather(hipStream_t s, int jb, int n, float* A, int lda,
                              const int* permU, float* W, int ldw);
hipError_t launch_swap_gather(hipStream_t s, int jb, int n, __half* A,
                              int lda, const int* permU, float* W, int ldw);
hipError_t launch_swap_inplace(hipStream_t s, int jb, int n, double* A,
                               int lda, const int* permU);
hipError_t launch_swap_inverse(hipStream_t s, int jb, int n, double* A,
                               int lda, const int* permU);
hipError_t launch_linv(hipStream_t s, int jb, const double* L, int ldl,
                       double* Linv, int ldi);
hipError_t launch_linv(hipStream_t s, int jb, const float* L, int ldl,
                       float* Linv, int ldi);
// op(T) x = b for nrhs columns of X; op(T) lower or upper as `lower` says
// (T itself is upper when lower && trans, lower when !lower && trans).
hipError_t launch_tri_solve(hipStream_t s, bool lower, bool trans, bool unit,
                            int k, int nrhs, const double* T, int ldt,
                            double* X, int ldx);
hipError_t launch_tri_solve(hipStream_t s, bool lower, bool trans, bool unit,
                            int k, int nrhs, const float* T, int ldt, float* X,
                            int ldx);
hipError_t launch_trsm_colsweep(hipStream_t s, int k, int n, const double* L,
                                int ldl, double* B, int ldb);
hipError_t launch_trsm_colsweep(hipStream_t s, int k, int n, const float* L,
                                int ldl, float* B, int ldb);

// C = beta*C + alpha*op(A)*op(B); alpha in {+1,-1}, beta in {0,1}.
// lower_a: A is lower triangular (only k < row-tile end contributes).  kind 1
// selects the separately named instance used for the trailing update.
// trans_a / trans_b (one at a time, alpha = -1): op(X) = X^T.
hipError_t launch_dgemm(hipStream_t s, int m, int n, int k, double alpha,
                        const double* A, int lda, const double* B, int ldb,
                        double beta, double* C, int ldc, bool lower_a,
                        int kind = 0, bool trans_a = false,
                        bool trans_b = false);
// The FP32 MFMA instance (no transposes, no trailing-update kind).
hipError_t launch_sgemm(hipStream_t s, int m, int n, int k, double alpha,
                        const float* A, int lda, const float* B, int ldb,
                        double beta, float* C, int ldc, bool lower_a);

// The trailing update C -= A * Bt^T with Bt = U12^T stored n-contiguous
// (k_update.hip): K a multiple of 8, 16-byte aligned operands, even leading
// dimensions; a_rows / b_cols bound the (clamped) reads of A's columns and
// Bt's rows.  update_fast_ok says whether a call qualifies.  panel = true
// launches the instance kept for the panel products: a separate kernel name
// in traces, so that the trailing update's launches stand alone; same code.
bool       update_fast_ok(int m, int n, int k, const double* A, int lda,
                          const double* Bt, int ldb, int a_rows, int b_cols);
hipError_t launch_update(hipStream_t s, int m, int n, int k, const double* A,
                         int lda, const double* Bt, int ldb, double* C,
                         int ldc, int a_rows, int b_cols, bool panel = false);
// The same update split between two launches for the look-ahead:
// launch_update_claimed runs `grid` persistent workgroups (one per CU) that
// take tiles through the per-XCD counters claim[0], claim[32], ...,
// claim[224] (zeroed before the launch) until none are left or *stop is
// non-zero; a workgroup on a CU whose __smid() bit is set in vacate (8 words
// in device memory; null: none) exits at once, leaving that CU to the panel
// factorisation.  launch_update_resume, queued after the claimed launch has
// finished, updates every tile the claims did not reach.
hipError_t launch_update_claimed(hipStream_t s, int m, int n, int k,
                                 const double* A, int lda, const double* Bt,
                                 int ldb, double* C, int ldc, int a_rows,
                                 int b_cols, int* claim,
                                 const unsigned long long* vacate,
                                 const int* stop, int grid);
hipError_t launch_update_resume(hipStream_t s, int m, int n, int k,
                                const double* A, int lda, const double* Bt,
                                int ldb, double* C, int ldc, int a_rows,
                                int b_cols, const int* claimed);
constexpr int kClaimInts = 8 * 32;  // ints of one set of claim counters
hipError_t launch_zero_ints(hipStream_t s, int* p, int n);
// *p = 1 with an agent-scope store (read by running kernels on any XCD).
hipError_t launch_set_flag(hipStream_t s, int* p);
// out[x] = min(claim[32 x], per) for the 8 XCD shares of a claimed update of
// an m x n C in bm x bn tiles (queued after the claimed launch).
hipError_t launch_claim_log(hipStream_t s, const int* claim, int m, int n,
                            int* out, int bm = 128, int bn = 128);
// Algorithmic flops of the tiles that the claimed launch of an m x n x k
// update in bm x bn tiles took (the tile order of every update kernel:
// groups of 8 tile rows walked column by column, XCD x taking the share
// [x per, (x + 1) per)), given the 8 counts launch_claim_log recorded.
double claimed_flops(int m, int n, int k, const int* done, int bm = 128,
                     int bn = 128);
// out[b] = __smid() of the CU that ran workgroup b (grid workgroups that
// each stay ~10 us, so that they spread over the device).
hipError_t launch_cu_census(hipStream_t s, int* out, int grid);
// B (cols x rows, ldb) = A^T for A (rows x cols, lda), column-major both.
hipError_t launch_transpose(hipStream_t s, int rows, int cols, const double* A,
                            int lda, double* B, int ldb);
// Leading dimension of the transposed U12 of n columns (even, with room for
// the clamped pair reads).
__host__ __device__ inline int ut_ld(int n) { return (n + 3) & ~1; }

// C (m x n, ldc) -= A^T B for A (K x m, lda) and B (K x n, ldb) (the
// transposed solve's U12^T z and L21^T z steps).  When the 64 x 64 tiles of
// C are too few to fill the GPU, K is split and the partial products go to
// work (work_doubles of them; none needed otherwise), summed in a fixed
// order.
hipError_t launch_tn_sub(hipStream_t s, int K, int m, int n, const double* A,
                         int lda, const double* B, int ldb, double* C, int ldc,
                         double* work, size_t work_doubles);
// C (m x n, ldc) -= A B for A (m x K, lda) and B (K x n, ldb): the solves'
// products with many right-hand sides (same kernel as launch_tn_sub, A read
// m-contiguous, never split).
hipError_t launch_nn_sub(hipStream_t s, int K, int m, int n, const double* A,
                         int lda, const double* B, int ldb, double* C,
                         int ldc);
// C (m x n, ldc) = -A B^T for A (m x K, lda) and B (n x K, ldb): the
// adjoint's gradA = -gradB X^T (write-bound; same kernel as launch_tn_sub).
hipError_t launch_outer_neg(hipStream_t s, int K, int m, int n,
                            const double* A, int lda, const double* B, int ldb,
                            double* C, int ldc);
hipError_t launch_trsv(hipStream_t s, int n, int nb, const double* A, int lda,
                       double* x);

// The solve with the factors of hplx_pdgetrf / hplx_pdgesv (hplx_pdgetrs)
// queued on s: permU is 2 kMaxNB ints of scratch, work holds the split-K
// partial products of the transposed solve (work_doubles doubles; a smaller
// buffer only means fewer splits).  Neither needs to start zeroed.  No
// allocation and no synchronisation.
hipError_t getrs_enqueue(hipStream_t s, const double* A, int N, int nb,
                         int lda, const int* ipiv, int trans, int nrhs,
                         double* B, int ldb, int* permU, double* work,
                         size_t work_doubles);

size_t     check_part_len(int n);  // doubles of scratch launch_check needs
hipError_t launch_check(hipStream_t s, int n, int seed, double diag_shift,
                        const double* X, double* part, size_t part_len,
                        double* out_dev);

}  // namespace hplx
