/*
 * hplx.h -- C ABI of rochpl_amd, an MI355X-native LU-with-partial-pivoting
 * solver that drops in for rocHPL's P=Q=1 HPL_pdgesv path.
 *
 * Conventions
 *   - Every entry point returns int: HPLX_SUCCESS (0), a negative HPLX_ERR_*
 *     code, or (hplx_pdgesv / hplx_pdfact) a positive LAPACK-style column
 *     index of the first exactly-zero pivot.  The library never calls exit();
 *     the text of the last error is kept per context (hplx_last_error).
 *   - Matrices are column-major doubles in device memory (HBM); sizes and
 *     leading dimensions are plain ints as in the reference, offsets are
 *     computed in 64 bits.
 *   - All device work is queued on the context's stream (hplx_set_stream);
 *     entry points that return host values synchronise that stream.
 *   - Indices are 0-based.  A pivot vector ipiv[i] = r means "row i was
 *     interchanged with row r" (LAPACK semantics, 0-based), exactly what the
 *     reference's HPL_pdmxswp writes into ipiv (src/pfact/HPL_pdmxswp.cpp:121).
 */
#ifndef HPLX_H
#define HPLX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define HPLX_SUCCESS 0
#define HPLX_ERR_INVALID -1     /* bad argument                            */
#define HPLX_ERR_UNSUPPORTED -2 /* valid HPL setting the library lacks    */
#define HPLX_ERR_HIP -3         /* HIP runtime error                       */
#define HPLX_ERR_ALLOC -4       /* device or pinned allocation failed      */

/* ---- enums with the reference's values --------------------------------- */
/* HPL_T_FACT, include/hpl_pgesv.hpp:39-43 */
#define HPLX_LEFT_LOOKING 301
#define HPLX_CROUT 302
#define HPLX_RIGHT_LOOKING 303
/* HPL_T_TOP, include/hpl_comm.hpp:31-38 (accepted, no-op on one GPU) */
#define HPLX_1RING 401
#define HPLX_1RING_M 402
#define HPLX_2RING 403
#define HPLX_2RING_M 404
#define HPLX_BLONG 405
#define HPLX_BLONG_M 406
/* HPL_T_SWAP, include/hpl_pgesv.hpp:45-51 */
#define HPLX_SWAP00 451
#define HPLX_SWAP01 452
#define HPLX_SW_MIX 453
#define HPLX_NO_SWP 499

/* ---- algorithm parameters: mirrors HPL_T_palg (include/hpl_pgesv.hpp:63-81)
 * The function-pointer plug points pffun/rffun/upfun are selected by the
 * pfact/rfact/L1notran values (src/HPL_pddriver.cpp:164-200): every pfact
 * and rfact (left-looking, Crout, right-looking) is built, and L1notran 0
 * or 1 (it only picks how the reference stores its copy of the panel's top
 * block; this library keeps no such copy, so both run the same kernels).
 * As in the reference (src/HPL_pdinfo.cpp:925-1004), depth must be 1, the
 * swap long, U transposed and equil 0: anything else returns
 * HPLX_ERR_UNSUPPORTED from hplx_pdgesv.
 * frac is accepted and not read: at P=1 the reference itself ignores it
 * (HPL_pdpanel_init sets nu0 = JB look-ahead columns, nu1 = 0 and
 * nu2 = the rest when nprow == 1, src/panel/HPL_pdpanel_init.cpp:176-184;
 * only its nprow > 1 branch, :185-197, splits by frac). */
typedef struct hplx_palg {
  int    btopo;    /* row broadcast topology (no-op at Q=1)        */
  int    depth;    /* look-ahead depth (must be 1, src/HPL_pdinfo.cpp:930-935) */
  int    nbdiv;    /* recursive division factor (>= 2)            */
  int    nbmin;    /* recursion stopping criterium (>= 1)         */
  int    pfact;    /* panel fact variant (HPLX_LEFT_LOOKING ...)  */
  int    rfact;    /* recursive fact variant (HPLX_LEFT_LOOKING ...) */
  int    fswap;    /* swapping algorithm                          */
  int    fsthr;    /* swapping threshold                          */
  int    equil;    /* equilibration (must be 0)                   */
  int    L1notran; /* L1 form, 0 or 1 (same kernels either way)   */
  int    Unotran;  /* 0: U transposed (required)                  */
  int    align;    /* memory alignment in doubles                 */
  double frac;     /* update split fraction (accepted, not read at P=1) */
  int    its;      /* iterations of the timed solve               */
} hplx_palg;

/* Fill *palg with the reference defaults (src/HPL_pdinfo.cpp:487-562):
 * right/right, nbmin 32, nbdiv 2, depth 1, long swap, thresh 64, frac 0.3. */
void hplx_palg_default(hplx_palg* palg);

/* The checks HPL_pdinfo applies to these settings (src/HPL_pdinfo.cpp:
 * 925-1004: depth 1, long swap, U transposed, no equilibration) plus the
 * variants this library builds, without a device: HPLX_SUCCESS, or the code
 * hplx_pdgesv / hplx_pdgetrf would return for *palg. */
int hplx_palg_check(const hplx_palg* palg);

/* ---- distributed matrix: mirrors HPL_T_pmat (include/hpl_pgesv.hpp:83-109)
 * A is the augmented [A|b], n x (n+1), leading dimension ld, b = column n.
 * ipiv and info are new: the reference returns neither. */
typedef struct hplx_pmat {
  double* A;    /* device: ld x (n+1)                                    */
  double* X;    /* device: n, solution after hplx_pdgesv                  */
  int*    ipiv; /* device: n, global 0-based pivot rows                   */
  int     n;    /* global problem size                                   */
  int     nb;   /* blocking factor                                       */
  int     ld;   /* leading dimension = ((n+95)/128)*128+32                */
  int     mp;   /* local rows (= n at P=Q=1)                              */
  int     nq;   /* local columns (= n+1)                                  */
  int     info; /* 0, or 1-based column of the first exactly-zero pivot  */
  void*   ws;   /* library-owned device workspace (opaque)                */
} hplx_pmat;

typedef struct hplx_ctx hplx_ctx;

/* ---- context (replaces HPL_InitGPU / HPL_FreeGPU, src/HPL_InitGPU.cpp:29-130:
 * device selection, streams, events; no rocBLAS handle) ------------------ */
int         hplx_ctx_create(int device, hplx_ctx** ctx);
int         hplx_ctx_destroy(hplx_ctx* ctx);
const char* hplx_last_error(const hplx_ctx* ctx);
/* Queue all work on this stream (a hipStream_t; NULL = the context's own). */
int         hplx_set_stream(hplx_ctx* ctx, void* stream);
int         hplx_synchronize(hplx_ctx* ctx);
/* Number of compute units of the context's device. */
int         hplx_num_cus(hplx_ctx* ctx);

/* Scheduling options of a context.  HPLX_OPT_LOOKAHEAD: the reference's
 * DEPTH=1 look-ahead (src/pgesv/HPL_pdgesv.cpp:168-245), which at P=Q=1 it
 * runs on one in-order stream (src/HPL_InitGPU.cpp:67,93).  1 (default):
 * hplx_pdgesv / hplx_pdgetrf factor panel k+1 on a second stream while the
 * update of step k runs, on CUs the update leaves free for it; 0: every
 * step in order on the context's stream.  Pivots and results are the same
 * bit for bit either way.  The environment variable HPLX_LOOKAHEAD=0|1 sets
 * the default of new contexts.  hplx_set_option returns HPLX_ERR_INVALID for
 * an unknown option or value; hplx_get_option writes the current value. */
#define HPLX_OPT_LOOKAHEAD 1
/* HPLX_OPT_PANEL_KERNEL: how an FP64 panel factorisation (HPL_pdfact with
 * the recursive HPL_pdrpan*T / HPL_pdpan*T of src/pfact/) runs.  1: every
 * panel as one persistent kernel that walks the recursion with a grid
 * barrier between its steps (a panel it cannot take, e.g. one with more
 * than 512 recursion steps, a product depth that is not a multiple of 8 or
 * an odd leading dimension, runs as the sequence).  0: every panel as a
 * sequence of kernels, one per leaf, triangular solve and product of the
 * recursion.  2 (default): the persistent kernel for panels of at least
 * 65536 rows, the sequence below (where it is faster).  Pivots and factors
 * are the same bit for bit in all three.  The environment variable
 * HPLX_PANEL_KERNEL=0|1|2 sets the default of new contexts. */
#define HPLX_OPT_PANEL_KERNEL 2
int         hplx_set_option(hplx_ctx* ctx, int option, int value);
int         hplx_get_option(hplx_ctx* ctx, int option, int* value);
/* Panels of the last hplx_pdgesv / hplx_pdgetrf / hplx_pdfact* call that
 * ran as the persistent panel kernel (HPLX_OPT_PANEL_KERNEL); no
 * counterpart in the reference. */
int         hplx_panel_kernel_runs(hplx_ctx* ctx);
/* The steps the persistent panel kernel walks for a panel of width jb under
 * palg's NBMIN, NDIV, PFACT and RFACT: the recursion of HPL_pdrpan*T /
 * HPL_pdpan*T (src/pfact/HPL_pdrpanrlT.cpp:93-160) recorded on the host, no
 * device call.  11 ints per step: kind (0 leaf, 1 triangular solve, 2
 * product), m_fix, m_sub, n, k, ar, ac, br, bc, cr, cc (panel-relative
 * coordinates; a leaf factors columns [ac, ac + n); a solve sets the k x n
 * block at (br, bc) to L^-1 times it, L the unit lower triangle at (ar, ac);
 * a product subtracts A (rows x k at ar, ac) times B (k x n at br, bc) from
 * C (rows x n at cr, cc), rows = m_fix if m_fix >= 0, else m - m_sub).
 * Writes at most max_steps steps; returns the step count or a negative
 * HPLX_ERR_*.  No counterpart in the reference. */
int         hplx_panel_steps(const hplx_palg* palg, int jb, int* steps,
                             int max_steps);

/* ---- plain device memory helpers (for callers without a runtime of their
 * own, e.g. ctypes tests) ------------------------------------------------- */
int hplx_malloc(hplx_ctx* ctx, size_t bytes, void** dptr);
int hplx_free(hplx_ctx* ctx, void* dptr);
int hplx_memcpy_h2d(hplx_ctx* ctx, void* dst, const void* src, size_t bytes);
int hplx_memcpy_d2h(hplx_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- allocation: HPL_pdmatgen / HPL_pdmatfree
 * (src/matgen/HPL_pdmatgen.cpp:85-436).  Allocates A|b with the reference's
 * ld padding, X, ipiv and the solver workspace in one place; the solver
 * itself never allocates. */
int hplx_pmat_alloc(hplx_ctx* ctx, int n, int nb, hplx_pmat* mat);
int hplx_pmat_free(hplx_ctx* ctx, hplx_pmat* mat);

/* ---- HPL_pdrandmat (include/hpl_pmatgen.hpp:53-59,
 * src/matgen/HPL_pdrandmat_device.cpp:94-201): fill the m x n matrix A (lda)
 * from the 64-bit LCG with seed `seed`; element (i,j) uses state i + j*m.
 * diag_shift (new) is added to every element with i == j. */
int hplx_pdrandmat(hplx_ctx* ctx, int m, int n, int nb, double* A, int lda,
                   int seed, double diag_shift);

/* ---- HPL_pdgesv (include/hpl_pgesv.hpp:162, src/pgesv/HPL_pdgesv.cpp:19-368)
 * Factors [A|b] in place (upper triangle = U, strictly lower = the panel
 * multipliers L, left unpivoted by later panels as in the reference), writes
 * x to mat->X and the pivots to mat->ipiv.  Returns 0, an error, or a
 * positive zero-pivot column (then X is not computed). */
int hplx_pdgesv(hplx_ctx* ctx, const hplx_palg* palg, hplx_pmat* mat);

/* ---- factor only: the panel loop of hplx_pdgesv over the n x n part of
 * mat->A (column n is not touched, no back-solve).  Keeps the factors and
 * pivots for hplx_pdgetrs / hplx_pdgesv_adjoint.  Returns like
 * hplx_pdgesv.  No counterpart in the reference (it always factors [A|b]). */
int hplx_pdgetrf(hplx_ctx* ctx, const hplx_palg* palg, hplx_pmat* mat);

/* ---- solve with the factors of hplx_pdgetrf / hplx_pdgesv: trans = 0
 * solves A X = B, trans = 1 solves A^T X = B, for nrhs columns of B (device,
 * ldb >= n) in place.  The L solve applies each panel's interchanges the way
 * HPL_pdgesv carries b through the factorisation (HPL's L is left unpivoted
 * by later panels); the U solve is HPL_pdtrsv's.  No allocation: the
 * transposed solve keeps split-K partial sums in the context's own panel
 * buffers, so calls on one context must not overlap. */
int hplx_pdgetrs(hplx_ctx* ctx, const hplx_pmat* mat, int trans, int nrhs,
                 double* B, int ldb);

/* ---- hplx_pdgetrs for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Needs no context: the same kernels as
 * hplx_pdgetrs, queued on `stream` (a hipStream_t of the current device;
 * NULL = its null stream) in the same order, so the result is bit-identical;
 * nothing is allocated and nothing is synchronised (the caller waits on the
 * stream).  Scratch: one file-scope device array of 8 MiB + 8 KiB per
 * device, reserved when the library's code object loads and held until the
 * process ends; every call on a device shares it, so calls that overlap on
 * one device (on different streams, or from several threads) race.
 * mat->nb above 1024 (the scratch's pivot slots) returns
 * HPLX_ERR_UNSUPPORTED; bad arguments HPLX_ERR_INVALID; a launch failure
 * HPLX_ERR_HIP (no error string: there is no context to hold one). */
int hplx_pdgetrs_noalloc(const hplx_pmat* mat, int trans, int nrhs, double* B,
                         int ldb, void* stream);

/* ---- adjoint of X = A^{-1} B (new): with the factors of A in mat, the
 * forward solution X (n x nrhs, ldx) and the upstream gradient G (ldg):
 *   gradB = A^{-T} G          (n x nrhs, ldgb)
 *   gradA = -gradB X^T        (n x n, ldga; skipped when gradA is NULL)
 * All device pointers. */
int hplx_pdgesv_adjoint(hplx_ctx* ctx, const hplx_pmat* mat, int nrhs,
                        const double* X, int ldx, const double* G, int ldg,
                        double* gradB, int ldgb, double* gradA, int ldga);

/* ---- HPL_pdtest residual check (src/HPL_pdtest.cpp:303-432) with [A|b]
 * regenerated on the fly from `seed` (A is not re-read; its storage may hold
 * the factors).  Uses mat->X and mat's workspace.  out[9]:
 *   [0] ||Ax-b||_oo/(eps*(||A||_oo*||x||_oo+||b||_oo)*N), true inf-norm
 *   [1] ||Ax-b||_oo
 *   [2] the same ratio with |r_0| in place of ||r||_oo (what the reference
 *       prints, src/pauxil/HPL_pdlange_device.cpp:284)
 *   [3] ||A||_oo [4] ||A||_1 [5] ||x||_oo [6] ||x||_1 [7] ||b||_oo [8] |r_0| */
int hplx_pdtest_check(hplx_ctx* ctx, const hplx_pmat* mat, int seed,
                      double diag_shift, double* out);

/* ---- hot-path components, callable one at a time (parity tests) -------- */

/* HPL_pdfact (src/pfact/HPL_pdfact.cpp:20-110) with rffun = HPL_pdrpanrlT
 * (src/pfact/HPL_pdrpanrlT.cpp:19-161) and pffun = HPL_pdpanrlT
 * (src/pfact/HPL_pdpanrlT_device.cpp:23-468): factor the m x jb panel P (lda)
 * in place; row interchanges span the jb panel columns.  ipiv (device, jb)
 * receives panel-relative pivot rows plus row_offset. */
int hplx_pdfact(hplx_ctx* ctx, int nbmin, int nbdiv, double* P, int lda,
                int m, int jb, int* ipiv, int row_offset);

/* HPL_pdfact with the panel functions palg selects (src/HPL_pddriver.cpp:
 * 164-200): rffun = HPL_pdrpan{ll,cr,rl}T by palg->rfact
 * (src/pfact/HPL_pdrpanllT.cpp:116-159, HPL_pdrpancrT.cpp:116-175,
 * HPL_pdrpanrlT.cpp:116-160), pffun = HPL_pdpan{ll,cr,rl}T by palg->pfact,
 * nbmin and nbdiv from palg; otherwise as hplx_pdfact.  The three leaf
 * variants give every element the same products in the same order, so one
 * leaf kernel computes all three; with nbmin above 32 the 32-column leaves
 * are blocked in the pfact variant's order. */
int hplx_pdfact_palg(hplx_ctx* ctx, const hplx_palg* palg, double* P, int lda,
                     int m, int jb, int* ipiv, int row_offset);

/* HPL_pipid + HPL_pdpanel_swapids P=1 branch (src/pgesv/HPL_pipid.cpp:19-164,
 * src/panel/HPL_pdpanel_swapids.cpp:22-41), on the device: from jb
 * panel-relative interchanges ipiv (device) build permU (device, 2*jb ints):
 * permU[i] = source row of top row i, permU[jb+i] = destination of the
 * displaced top row i or -1. */
int hplx_pipid(hplx_ctx* ctx, int jb, const int* ipiv, int* permU);

/* HPL_dlaswp00N (src/pauxil/HPL_dlaswp00N_device.cpp:22-112): apply permU to
 * the n columns of A (lda): rows [0,jb) gathered, displaced rows scattered. */
int hplx_dlaswp00N(hplx_ctx* ctx, int jb, int n, double* A, int lda,
                   const int* permU);

/* HPL_pdupdateTT P=1 branch (src/pgesv/HPL_pdupdateTT.cpp:91-165) fused
 * with HPL_pdlaswp_end (src/pgesv/HPL_pdlaswp.cpp:386-389): for the panel
 * whose top-left element is A[j + j*lda] (width jb, m = rows from j down)
 * and its n trailing columns: swap rows with permU, U12 = L11^{-1} A12,
 * A22 -= L21 * U12.  W is caller scratch of jb*(n+4) doubles (device): the
 * gathered rows, then U12 transposed for the update. */
int hplx_pdupdate(hplx_ctx* ctx, double* A, int lda, int m, int jb, int n,
                  const int* permU, double* W);

/* C = beta*C + alpha*A*B (A m x k, B k x n, column-major, no transposes) on
 * the FP64 MFMA kernel that carries the trailing update
 * (rocblas_dgemm at src/pgesv/HPL_pdupdateTT.cpp:131-144).  alpha must be
 * +1 or -1 and beta 0 or 1. */
int hplx_dgemm(hplx_ctx* ctx, int m, int n, int k, double alpha,
               const double* A, int lda, const double* B, int ldb,
               double beta, double* C, int ldc);

/* HPL_pdtrsv (src/pgesv/HPL_pdtrsv_device.cpp:29-346): solve U x = b where
 * U is the upper triangle of the n x n matrix A (lda) and b its column n,
 * blocked by nb (<= 1024); x (device, n). */
int hplx_pdtrsv(hplx_ctx* ctx, int n, int nb, const double* A, int lda,
                double* x);

/* Device time in ms of the FP64 trailing-update GEMM kernels queued since
 * the last hplx_pdgesv started (HIP events around each launch), with the
 * count of launches and their algorithmic flops. */
int hplx_update_stats(hplx_ctx* ctx, double* gemm_ms, int* launches,
                      double* gemm_flops);

/* The same, restricted to one kind of launch: 1 a whole update in one launch
 * (the look-ahead's columns of the next panel, and every update of the
 * serial schedule), 2 the claimed part of a split update (beside the next
 * panel's factorisation, on the CUs it leaves), 3 the resume part of a split
 * update (on every CU); 0 all.  Flops of kinds 2 and 3 are those of the
 * tiles each part took (from the per-XCD claim counts the solve logs). */
int hplx_update_timing(hplx_ctx* ctx, int kind, double* ms, int* launches,
                       double* flops);

/* Timing of the row interchanges of the trailing columns (the gather /
 * scatter kernel of HPL_pdlaswp_end / HPL_dlaswp00N at P=1,
 * src/pauxil/HPL_dlaswp00N_device.cpp:22-112) in the last hplx_pdgesv /
 * hplx_pdgetrf: summed launch time (ms, HIP events), launches, and their
 * algorithmic bytes (4 x 8 B per interchanged element: the pivot row read,
 * the U buffer written, the displaced top row read and written).  No
 * counterpart in the reference, which times none. */
int hplx_swap_timing(hplx_ctx* ctx, double* ms, int* launches, double* bytes);

/* Timing of the panel factorisations (HPL_pdfact, src/pfact/HPL_pdfact.cpp)
 * in the last hplx_pdgesv / hplx_pdgetrf: summed time (ms, HIP events
 * around each panel on the stream it runs on, beside the update with the
 * look-ahead), panels, and their algorithmic bytes (each m x jb panel read
 * and written once, 16 B per element).  No counterpart in the reference. */
int hplx_panel_timing(hplx_ctx* ctx, double* ms, int* panels, double* bytes);

/* ======================================================================
 * Mixed precision (no counterpart in the reference, which is FP64 only).
 *
 * The same right-looking recursive LU with partial pivoting as hplx_pdgesv
 * (same panel recursion, leaves, pivot rule and pivot composition), with
 *   - the factors stored in FP32 or FP16 (storage),
 *   - every panel factorised in FP32 (FP16-stored panels are staged in an
 *     FP32 copy),
 *   - U12 = L11^{-1} P A12 in FP32,
 *   - the trailing update A22 -= L21 U12 in BF16 (optionally split into 2 or
 *     3 BF16 terms per operand) or in per-128x128-tile power-of-two-scaled
 *     FP8 e4m3, on MFMA with FP32 accumulation,
 * followed by FP64 iterative refinement: r = b - A x with A and b
 * regenerated tile by tile in FP64 from the generator's seed (never stored),
 * the correction d = LU \ r solved with the low-precision factors in FP32
 * (solve_fp64 = 1: in FP64), and x += d in FP64, until the HPL scaled
 * residual is below the threshold.
 * ==================================================================== */

/* storage of the factors */
#define HPLX_STORE_FP32 32
#define HPLX_STORE_FP16 16
/* arithmetic of the trailing update */
#define HPLX_UPDATE_FP32 0 /* FP32 MFMA (reference form for tests)          */
#define HPLX_UPDATE_BF16 1 /* BF16 MFMA, `terms` BF16 numbers per operand     */
#define HPLX_UPDATE_FP8 2  /* e4m3 MFMA, power-of-two scale per 128x128 tile */

typedef struct hplx_mxp_opts {
  int    update;    /* HPLX_UPDATE_*                                        */
  int    terms;     /* BF16: 1 (plain), 2 (3 products) or 3 (6 products)  */
  int    max_its;   /* refinement solves allowed after the first (<= 64)    */
  int    solve_fp64; /* 0: solves with the factors in FP32 (FP64 only for
                        r = b - A x, x += d and the norms); 1: in FP64      */
  double threshold; /* target HPL scaled residual (16.0)                    */
} hplx_mxp_opts;

typedef struct hplx_mxp_pmat {
  void*   A;          /* ld x n factors, storage type, column-major          */
  double* X;          /* solution (n)                                        */
  int*    ipiv;       /* pivots (n), as hplx_pmat                            */
  int     n, nb, ld;
  int     storage;    /* HPLX_STORE_*                                        */
  int     seed;       /* generator of A and b (set by hplx_mxp_matgen)       */
  int     info;       /* first exact zero pivot (1-based) or 0               */
  double  diag_shift; /* A = generator + diag_shift * I                      */
  void*   ws;         /* library workspace                                   */
} hplx_mxp_pmat;

typedef struct hplx_mxp_result {
  int    its;          /* refinement corrections applied                    */
  int    converged;    /* 1 when resid < threshold                          */
  double resid;        /* final scaled residual (true ||r||_oo)             */
  double factor_ms;    /* device time of the factorisation                   */
  double solve_ms;     /* first solve + refinement                           */
  double update_ms;    /* device time of the low-precision update GEMMs      */
  double update_flops; /* their FP64-equivalent flops (2 m n k per launch)   */
  int    update_launches;
  int    pad;
  double history[65];  /* scaled residual after each solve (its + 1 values) */
} hplx_mxp_result;

void hplx_mxp_opts_default(hplx_mxp_opts* opts);

/* Allocate the n x n factor storage and every workspace of the mixed solve
 * (nb <= 1024). */
int hplx_mxp_alloc(hplx_ctx* ctx, int n, int nb, int storage,
                   hplx_mxp_pmat* mat);
int hplx_mxp_free(hplx_ctx* ctx, hplx_mxp_pmat* mat);

/* A = generator(seed) + diag_shift * I, generated in FP64 (the
 * HPL_pdrandmat values) and rounded to the storage type. */
int hplx_mxp_matgen(hplx_ctx* ctx, hplx_mxp_pmat* mat, int seed,
                    double diag_shift);

/* Factor mat->A in place (no refinement).  Returns 0, an error, or a
 * positive zero-pivot column. */
int hplx_mxp_pdgetrf(hplx_ctx* ctx, const hplx_palg* palg,
                     const hplx_mxp_opts* opts, hplx_mxp_pmat* mat);

/* x := A^{-1} x with the stored low-precision factors, FP64 vector x
 * (device, n), solved in the precision the factorisation's opts chose
 * (FP32: x is rounded to FP32, solved, and widened back). */
int hplx_mxp_pdgetrs(hplx_ctx* ctx, const hplx_mxp_pmat* mat, double* x);

/* The mixed-precision counterpart of hplx_pdgesv: factor, solve for the
 * generator's b, refine in FP64 until the scaled residual is below
 * opts->threshold or opts->max_its corrections were applied.  mat->X holds
 * x; res (may be NULL) the refinement record.  Returns 0 when converged,
 * HPLX_ERR_UNSUPPORTED-style errors on bad input, a positive zero-pivot
 * column, or 1 + n when refinement did not converge. */
int hplx_mxp_pdgesv(hplx_ctx* ctx, const hplx_palg* palg,
                    const hplx_mxp_opts* opts, hplx_mxp_pmat* mat,
                    hplx_mxp_result* res);

/* The refinement residual of mat->X: out[5] = scaled residual,
 * ||b - A x||_oo, ||A||_oo, ||x||_oo, ||b||_oo (A, b regenerated in FP64). */
int hplx_mxp_residual(hplx_ctx* ctx, const hplx_mxp_pmat* mat,
                      const double* x, double* out);

/* Components (parity tests).
 * The low-precision trailing update: C (m x n, storage type, ldc) -= L U with
 * L (m x k) and U (k x n) FP32 column-major, rounded to the update
 * arithmetic exactly as the factorisation rounds them. */
int hplx_mxp_update(hplx_ctx* ctx, int update, int terms, int storage, int m,
                    int n, int k, const float* L, int ldl, const float* U,
                    int ldu, void* C, int ldc);
/* The FP32 panel factorisation (hplx_pdfact in FP32). */
int hplx_spdfact(hplx_ctx* ctx, int nbmin, int nbdiv, float* P, int lda,
                 int m, int jb, int* ipiv, int row_offset);

#ifdef __cplusplus
}
#endif

#endif /* HPLX_H */
