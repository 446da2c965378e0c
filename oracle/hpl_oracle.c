/*
 * hpl_oracle.c -- CPU restatement of rocHPL's P=Q=1 HPL_pdgesv path.
 *
 * TEST INFRASTRUCTURE ONLY.  Tests compare the HIP library in rochpl_amd/
 * against it: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in the product path links it.
 *
 * What it restates (reference files, paths relative to the rocHPL tree):
 *   - matgen:  src/matgen/HPL_pdrandmat_device.cpp:21-82,153-175 (LCG walk and
 *              value formula), src/matgen/HPL_xjumpm.cpp:81-91 (affine jump),
 *              include/hpl_pmatgen.hpp:35-39 (constants), HPL_ISEED=100 from
 *              include/hpl_ptest.hpp:57.
 *   - panel:   src/pfact/HPL_pdrpanrlT.cpp:93-160 (recursive right-looking
 *              split), src/pfact/HPL_pdpanrlT_device.cpp:147-319 (leaf: max|a|
 *              pivot, full-panel-width row swap, scale by division, rank-1
 *              update of the leaf columns); the left-looking and Crout
 *              variants: src/pfact/HPL_pdrpanllT.cpp:116-159 and
 *              HPL_pdrpancrT.cpp:116-175 (recursion), HPL_pdpanllT_device.cpp:
 *              7-376 and HPL_pdpancrT_device.cpp:7-354 (leaves), selected at
 *              src/HPL_pddriver.cpp:164-200.
 *   - swaps:   src/pgesv/HPL_pipid.cpp:19-164 + src/panel/HPL_pdpanel_swapids.cpp:22-41
 *              (composed permutation), src/pauxil/HPL_dlaswp00N_device.cpp:22-62.
 *   - update:  src/pgesv/HPL_pdupdateTT.cpp:95-144 (dtrsm L,U,T,Unit then
 *              dgemm N,T: A22 -= L21 * U12).
 *   - loop:    src/pgesv/HPL_pdgesv.cpp:168-245 (panel loop; the lookahead only
 *              reorders independent work, so the arithmetic per element is the
 *              loop below).
 *   - solve:   src/pgesv/HPL_pdtrsv_device.cpp:29-346 (U x = b').
 *   - check:   src/HPL_pdtest.cpp:303-432 (norms, r = b - A x, scaled
 *              residual), src/pauxil/HPL_pdlange_device.cpp:118-302 (norms;
 *              the NORM_I defect at :284 is reproduced only in the separate
 *              "r0 form"), src/auxil/HPL_dlamch.cpp:106-108 (eps = 2^-53).
 *
 * Deliberate deviation: an exactly-zero pivot is reported LAPACK-style
 * (info = 1-based global column of the first one) and its column is not
 * scaled; the reference divides by zero (HPL_pdpanrlT_device.cpp:232,317).
 *
 * Pinning (see tests/test_oracle.py, tests/test_oracle_variants.py): pivots
 * of every (pfact, rfact) variant equal LAPACK dgetrf's
 * (scipy.linalg.lu_factor), x equals scipy.linalg.lu_solve to rounding, the
 * matgen equals the exact-rational closed form and the O(log k) jump equals
 * sequential stepping, pipid equals brute-force application of the
 * interchanges, norms equal numpy's.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define HPLO_MULT 6364136223846793005ULL
#define HPLO_IADD 1ULL

typedef void (*hplo_dgemm_t)(const char*, const char*, const int*, const int*,
                             const int*, const double*, const double*,
                             const int*, const double*, const int*,
                             const double*, double*, const int*);
typedef void (*hplo_dtrsm_t)(const char*, const char*, const char*,
                             const char*, const int*, const int*,
                             const double*, const double*, const int*,
                             double*, const int*);

static hplo_dgemm_t g_dgemm = NULL;
static hplo_dtrsm_t g_dtrsm = NULL;

/* Optional host BLAS (Fortran ABI, 32-bit ints) used only for the timed CPU
 * baseline; NULL restores the built-in loops used by the parity tests. */
void hplo_set_blas(void* dgemm, void* dtrsm) {
  g_dgemm = (hplo_dgemm_t)dgemm;
  g_dtrsm = (hplo_dtrsm_t)dtrsm;
}

/* ------------------------------------------------------------------ */
/* matgen                                                              */
/* ------------------------------------------------------------------ */

/* Affine map of k LCG steps: s -> A*s + C (mod 2^64), by repeated squaring.
 * Same map HPL_xjumpm builds by a linear loop (HPL_xjumpm.cpp:81-91). */
void hplo_jump(uint64_t k, uint64_t mult, uint64_t iadd, uint64_t* A,
               uint64_t* C) {
  uint64_t ra = 1, rc = 0;       /* accumulated map */
  uint64_t pa = mult, pc = iadd; /* map of 2^b steps */
  while(k) {
    if(k & 1) {
      /* apply p after r: s -> pa*(ra*s+rc)+pc */
      rc = pa * rc + pc;
      ra = pa * ra;
    }
    /* square p */
    pc = pa * pc + pc;
    pa = pa * pa;
    k >>= 1;
  }
  *A = ra;
  *C = rc;
}

/* First state used by the matgen: one step from the seed
 * (HPL_pdrandmat_device.cpp:163, jump1 = 1). */
uint64_t hplo_start(uint64_t seed) { return HPLO_MULT * seed + HPLO_IADD; }

/* State of element index k = i + j*M. */
uint64_t hplo_state(uint64_t seed, uint64_t k) {
  uint64_t a, c;
  hplo_jump(k, HPLO_MULT, HPLO_IADD, &a, &c);
  return a * hplo_start(seed) + c;
}

/* Value formula of HPL_pdrandmat_device.cpp:65-73: 0.5 - s/2^64 with the
 * 64-bit state split in 16-bit limbs; every product is a power of two, so
 * this is RN(0.5 - RN(s*2^-32)*2^-32). */
double hplo_value(uint64_t r) {
  const double p1 = (double)(r & 0xFFFFULL);
  const double p2 = (double)((r >> 16) & 0xFFFFULL);
  const double p3 = (double)((r >> 32) & 0xFFFFULL);
  const double p4 = (double)((r >> 48) & 0xFFFFULL);
  const double lo = (p1 + p2 * 65536.0) / 2147483648.0 * 0.5;
  return 0.5 - ((lo + p3) + p4 * 65536.0) / 2147483648.0 * 0.5;
}

/* Generate rows [i0,i0+m) x cols [j0,j0+n) of the global M-row matrix into
 * A (column-major, lda).  diag_shift is added to A(i,i) for i < min(M, ncols
 * of the square part), i.e. to every element with i == j. */
void hplo_matgen_tile(int M, int i0, int j0, int m, int n, double* A, int lda,
                      int seed, double diag_shift) {
  uint64_t a1 = HPLO_MULT, c1 = HPLO_IADD;
  for(int j = 0; j < n; ++j) {
    uint64_t s = hplo_state((uint64_t)seed,
                            (uint64_t)(i0) + (uint64_t)(j0 + j) * (uint64_t)M);
    double* col = A + (size_t)j * (size_t)lda;
    for(int i = 0; i < m; ++i) {
      double v = hplo_value(s);
      if(diag_shift != 0.0 && (i0 + i) == (j0 + j)) v += diag_shift;
      col[i] = v;
      s      = a1 * s + c1;
    }
  }
}

/* Whole M x N matrix (HPL_pdrandmat(GRID, M, N, NB, A, LDA, ISEED)). */
void hplo_matgen(int M, int N, double* A, int lda, int seed,
                 double diag_shift) {
  hplo_matgen_tile(M, 0, 0, M, N, A, lda, seed, diag_shift);
}

/* ------------------------------------------------------------------ */
/* small dense kernels (column-major)                                  */
/* ------------------------------------------------------------------ */

#define AT(P, ld, i, j) ((P)[(size_t)(i) + (size_t)(j) * (size_t)(ld)])

/* C[m x n] -= A[m x k] * B[k x n] */
static void gemm_minus(int m, int n, int k, const double* A, int lda,
                       const double* B, int ldb, double* C, int ldc) {
  if(m <= 0 || n <= 0 || k <= 0) return;
  if(g_dgemm) {
    const double mone = -1.0, one = 1.0;
    g_dgemm("N", "N", &m, &n, &k, &mone, A, &lda, B, &ldb, &one, C, &ldc);
    return;
  }
  for(int j = 0; j < n; ++j) {
    double* c = &AT(C, ldc, 0, j);
    for(int p = 0; p < k; ++p) {
      const double b = AT(B, ldb, p, j);
      if(b == 0.0) continue;
      const double* a = &AT(A, lda, 0, p);
      for(int i = 0; i < m; ++i) c[i] -= a[i] * b;
    }
  }
}

/* B[k x n] = L^{-1} B, L unit lower triangular k x k (left, lower, no-trans,
 * unit; the same operator as the reference's dtrsm(L,U,T,Unit) on the
 * transposed L1, HPL_pdupdateTT.cpp:95-106). */
static void trsm_llnu(int k, int n, const double* L, int ldl, double* B,
                      int ldb) {
  if(k <= 0 || n <= 0) return;
  if(g_dtrsm) {
    const double one = 1.0;
    g_dtrsm("L", "L", "N", "U", &k, &n, &one, L, &ldl, B, &ldb);
    return;
  }
  for(int j = 0; j < n; ++j) {
    double* b = &AT(B, ldb, 0, j);
    for(int p = 0; p < k; ++p) {
      const double bp = b[p];
      if(bp == 0.0) continue;
      const double* l = &AT(L, ldl, 0, p);
      for(int i = p + 1; i < k; ++i) b[i] -= l[i] * bp;
    }
  }
}

static void swap_rows(double* P, int lda, int ncols, int r0, int r1) {
  if(r0 == r1) return;
  for(int c = 0; c < ncols; ++c) {
    double t        = AT(P, lda, r0, c);
    AT(P, lda, r0, c) = AT(P, lda, r1, c);
    AT(P, lda, r1, c) = t;
  }
}

/* ------------------------------------------------------------------ */
/* panel factorisation                                                 */
/* ------------------------------------------------------------------ */

/* Variant codes of HPL_T_FACT (include/hpl_pgesv.hpp). */
#define HPLO_LEFT 301
#define HPLO_CROUT 302
#define HPLO_RIGHT 303

/* Pivot step of column c: max |a| over rows c..m-1 (first index on ties),
 * interchange across the full panel width (the candidate rows are JB wide,
 * HPL_pdpanrlT_device.cpp:173-184), division of the column below the pivot
 * (:232).  Returns 0, or 1 when the pivot is an exact zero (the column is
 * then left unscaled). */
static int pivot_step(double* P, int lda, int m, int JB, int c, int* ipiv) {
  int    p    = c;
  double amax = fabs(AT(P, lda, c, c));
  for(int r = c + 1; r < m; ++r) {
    const double v = fabs(AT(P, lda, r, c));
    if(v > amax) {
      amax = v;
      p    = r;
    }
  }
  ipiv[c] = p;
  swap_rows(P, lda, JB, c, p);
  const double piv = AT(P, lda, c, c);
  if(piv == 0.0) return 1;
  for(int r = c + 1; r < m; ++r) AT(P, lda, r, c) /= piv;
  return 0;
}

/* Leaf on columns [c0, c0+w) of the panel P (m rows, JB columns, row 0 =
 * panel top); ipiv is panel relative.  Returns the 1-based panel column of
 * the first exact zero pivot, or 0.
 *   right (HPL_pdpanrlT_device.cpp:147-319): pivot, then a rank-1 update of
 *     the leaf columns right of the pivot column;
 *   left (HPL_pdpanllT_device.cpp:115-125,294-345): column c first receives every
 *     earlier leaf column's update (its rows above c become U by forward
 *     substitution, its rows from c on are reduced), then the pivot step;
 *   Crout (HPL_pdpancrT_device.cpp:131-140,311-325): column c's rows from c on receive the
 *     earlier columns' updates with the final U rows, the pivot step, then
 *     row c of U right of the pivot is finished (a_cj -= l_ck u_kj).
 * Every element receives the same products in the same order k = c0, c0+1,
 * ... in all three, so with these sequential loops the variants agree bit
 * for bit; only the moment of each update differs. */
static int leaf_v(int variant, double* P, int lda, int m, int JB, int c0,
                  int w, int* ipiv) {
  int info = 0;
  const int cend = c0 + w;
  for(int c = c0; c < cend; ++c) {
    if(variant == HPLO_LEFT) {
      for(int k = c0; k < c; ++k) {
        const double u = AT(P, lda, k, c);
        if(u == 0.0) continue;
        for(int r = k + 1; r < m; ++r) AT(P, lda, r, c) -= AT(P, lda, r, k) * u;
      }
    } else if(variant == HPLO_CROUT) {
      for(int k = c0; k < c; ++k) {
        const double u = AT(P, lda, k, c);
        if(u == 0.0) continue;
        for(int r = c; r < m; ++r) AT(P, lda, r, c) -= AT(P, lda, r, k) * u;
      }
    }
    if(pivot_step(P, lda, m, JB, c, ipiv)) {
      if(!info) info = c + 1;
      continue;
    }
    if(variant == HPLO_RIGHT) {
      for(int k = c + 1; k < cend; ++k) {
        const double u = AT(P, lda, c, k);
        if(u == 0.0) continue;
        for(int r = c + 1; r < m; ++r) AT(P, lda, r, k) -= AT(P, lda, r, c) * u;
      }
    } else if(variant == HPLO_CROUT) {
      for(int j = c + 1; j < cend; ++j)
        for(int k = c0; k < c; ++k) {
          const double u = AT(P, lda, k, j);
          if(u == 0.0) continue;
          AT(P, lda, c, j) -= AT(P, lda, c, k) * u;
        }
    }
  }
  return info;
}

/* Recursive split of the panel columns [c0, c0+n), rows [c0, m), in chunks
 * of nb = ceil(ceil(n / nbmin) / ndiv) * nbmin columns (HPL_pdrpanrlT.cpp:93),
 * each chunk factored by a recursive call (the leaf once n <= nbmin):
 *   right (HPL_pdrpanrlT.cpp:116-160): factor the chunk, then its U rows to
 *     the right (L11^{-1} A12) and the rows below (A22 -= L21 U12);
 *   left (HPL_pdrpanllT.cpp:116-159): first the chunk's rows above it
 *     (U = L^{-1} A over the jj columns left of it in this call) and the rows
 *     from its top down (A -= L U), then factor it;
 *   Crout (HPL_pdrpancrT.cpp:116-175): the chunk's rows from its top down
 *     (A -= L U with the final U rows above), factor it, then its own U rows
 *     right of it (A -= L U over the jj columns, then L11^{-1}). */
static int rpanel_v(int pfact, int rfact, double* P, int lda, int m, int JB,
                    int c0, int n, int nbmin, int ndiv, int* ipiv) {
  if(n <= nbmin) return leaf_v(pfact, P, lda, m, JB, c0, n, ipiv);
  int       info = 0;
  const int nb   = (((n + nbmin - 1) / nbmin + ndiv - 1) / ndiv) * nbmin;
  int       jj = 0, jb = nb, nrem = n;
  do {
    nrem -= jb;
    const int top = c0 + jj; /* first row and column of this chunk */
    if(jj > 0 && rfact == HPLO_LEFT)
      trsm_llnu(jj, jb, &AT(P, lda, c0, c0), lda, &AT(P, lda, c0, top), lda);
    if(jj > 0 && rfact != HPLO_RIGHT)
      gemm_minus(m - top, jb, jj, &AT(P, lda, top, c0), lda,
                 &AT(P, lda, c0, top), lda, &AT(P, lda, top, top), lda);
    const int i1 =
        rpanel_v(pfact, rfact, P, lda, m, JB, top, jb, nbmin, ndiv, ipiv);
    if(i1 && !info) info = i1;
    if(nrem > 0 && rfact == HPLO_RIGHT) {
      trsm_llnu(jb, nrem, &AT(P, lda, top, top), lda,
                &AT(P, lda, top, top + jb), lda);
      gemm_minus(m - top - jb, nrem, jb, &AT(P, lda, top + jb, top), lda,
                 &AT(P, lda, top, top + jb), lda,
                 &AT(P, lda, top + jb, top + jb), lda);
    } else if(nrem > 0 && rfact == HPLO_CROUT) {
      gemm_minus(jb, nrem, jj, &AT(P, lda, top, c0), lda,
                 &AT(P, lda, c0, top + jb), lda, &AT(P, lda, top, top + jb),
                 lda);
      trsm_llnu(jb, nrem, &AT(P, lda, top, top), lda,
                &AT(P, lda, top, top + jb), lda);
    }
    jj += jb;
    jb = nrem < nb ? nrem : nb;
  } while(nrem > 0);
  return info;
}

static int valid_fact(int f) {
  return f == HPLO_LEFT || f == HPLO_CROUT || f == HPLO_RIGHT;
}

/* HPL_pdfact at P=Q=1 with the panel (pfact) and recursion (rfact)
 * variants: factor the m x jb panel P in place; ipiv[0:jb] gets
 * panel-relative pivot rows.  Returns info, or -1 for an unknown variant. */
int hplo_pdfact_v(double* P, int lda, int m, int jb, int nbmin, int ndiv,
                  int pfact, int rfact, int* ipiv) {
  if(!valid_fact(pfact) || !valid_fact(rfact)) return -1;
  if(nbmin < 1) nbmin = 1;
  if(ndiv < 2) ndiv = 2;
  return rpanel_v(pfact, rfact, P, lda, m, jb, 0, jb, nbmin, ndiv, ipiv);
}

/* The right-looking panel and recursion (the reference's default). */
int hplo_pdfact(double* P, int lda, int m, int jb, int nbmin, int ndiv,
                int* ipiv) {
  return hplo_pdfact_v(P, lda, m, jb, nbmin, ndiv, HPLO_RIGHT, HPLO_RIGHT,
                       ipiv);
}

/* HPL_pipid + HPL_pdpanel_swapids (P=1 branch): from jb sequential
 * interchanges (row i <-> ipiv[i], panel relative, ipiv[i] >= i) build
 * permU[0:jb] = source row of top row i, permU[jb+i] = destination of the
 * displaced top row i, or -1.  Restated as an explicit simulation of the
 * interchanges on row labels; the reference's pair-list search produces the
 * same mapping. */
void hplo_pipid(int jb, const int* ipiv, int* permU) {
  /* labels for positions < jb, plus a map for positions >= jb */
  int* lab   = (int*)malloc(sizeof(int) * (size_t)(2 * jb + 1));
  int* ppos  = lab + jb; /* distinct positions >= jb, in first-touch order */
  int* plab  = (int*)malloc(sizeof(int) * (size_t)(jb + 1));
  int  np    = 0;
  for(int i = 0; i < jb; ++i) lab[i] = i;
  for(int i = 0; i < jb; ++i) {
    const int p = ipiv[i];
    if(p < jb) {
      int t  = lab[i];
      lab[i] = lab[p];
      lab[p] = t;
    } else {
      int s = -1;
      for(int q = 0; q < np; ++q)
        if(ppos[q] == p) {
          s = q;
          break;
        }
      if(s < 0) {
        s       = np++;
        ppos[s] = p;
        plab[s] = p;
      }
      int t   = lab[i];
      lab[i]  = plab[s];
      plab[s] = t;
    }
  }
  for(int i = 0; i < jb; ++i) {
    permU[i]      = lab[i];
    permU[jb + i] = -1;
  }
  for(int q = 0; q < np; ++q) {
    /* label plab[q] (< jb, see the interchange structure) ends at ppos[q] */
    permU[jb + plab[q]] = ppos[q];
  }
  free(plab);
  free(lab);
}

/* Apply k interchanges (row i <-> ipiv[i]) in order to ncols columns. */
void hplo_laswp(double* A, int lda, int ncols, int k, const int* ipiv) {
  for(int i = 0; i < k; ++i) swap_rows(A, lda, ncols, i, ipiv[i]);
}

/* dlaswp00N with the permU of hplo_pipid: rows [0,jb) gather, displaced top
 * rows scatter (HPL_dlaswp00N_device.cpp:22-62). */
void hplo_dlaswp00N(int jb, int ncols, double* A, int lda, const int* permU) {
  double* top = (double*)malloc(sizeof(double) * (size_t)(jb > 0 ? jb : 1));
  double* nw  = (double*)malloc(sizeof(double) * (size_t)(jb > 0 ? jb : 1));
  for(int c = 0; c < ncols; ++c) {
    double* col = A + (size_t)c * (size_t)lda;
    for(int i = 0; i < jb; ++i) top[i] = col[i];
    for(int i = 0; i < jb; ++i) {
      const int s = permU[i];
      nw[i]       = s < jb ? top[s] : col[s];
    }
    for(int i = 0; i < jb; ++i) col[i] = nw[i];
    for(int i = 0; i < jb; ++i)
      if(permU[jb + i] > -1) col[permU[jb + i]] = top[i];
  }
  free(nw);
  free(top);
}

/* U x = b, U upper triangular non-unit n x n (overwrites x). */
void hplo_trsv_upper(int n, const double* U, int ldu, double* x) {
  for(int j = n - 1; j >= 0; --j) {
    x[j] /= AT(U, ldu, j, j);
    const double xj = x[j];
    const double* u = &AT(U, ldu, 0, j);
    for(int i = 0; i < j; ++i) x[i] -= u[i] * xj;
  }
}

/* HPL_pdgesv at P=Q=1 on the augmented [A|b] (n x n+1, b = column n).
 * ipiv[0:n] gets global 0-based pivot rows (row i <-> ipiv[i]); x[0:n] gets
 * the solution.  Returns info (0, or 1-based column of the first exact zero
 * pivot). */
int hplo_pdgesv_v(int n, int nb, int nbmin, int ndiv, int pfact, int rfact,
                  double* A, int lda, int* ipiv, double* x) {
  if(!valid_fact(pfact) || !valid_fact(rfact)) return -1;
  int info = 0;
  int* permU = (int*)malloc(sizeof(int) * (size_t)(2 * (nb > 0 ? nb : 1)));
  for(int j = 0; j < n; j += nb) {
    const int jb = (n - j) < nb ? (n - j) : nb;
    const int m  = n - j;
    double*   P  = &AT(A, lda, j, j);
    const int i1 =
        hplo_pdfact_v(P, lda, m, jb, nbmin, ndiv, pfact, rfact, ipiv + j);
    if(i1 && !info) info = j + i1;
    /* trailing columns j+jb .. n (b included) */
    const int nt = n + 1 - (j + jb);
    hplo_pipid(jb, ipiv + j, permU);
    hplo_dlaswp00N(jb, nt, &AT(A, lda, j, j + jb), lda, permU);
    trsm_llnu(jb, nt, P, lda, &AT(A, lda, j, j + jb), lda);
    gemm_minus(m - jb, nt, jb, &AT(A, lda, j + jb, j), lda,
               &AT(A, lda, j, j + jb), lda, &AT(A, lda, j + jb, j + jb), lda);
    for(int i = 0; i < jb; ++i) ipiv[j + i] += j;
  }
  free(permU);
  for(int i = 0; i < n; ++i) x[i] = AT(A, lda, i, n);
  if(!info) hplo_trsv_upper(n, A, lda, x);
  return info;
}

int hplo_pdgesv(int n, int nb, int nbmin, int ndiv, double* A, int lda,
                int* ipiv, double* x) {
  return hplo_pdgesv_v(n, nb, nbmin, ndiv, HPLO_RIGHT, HPLO_RIGHT, A, lda,
                       ipiv, x);
}

/* ------------------------------------------------------------------ */
/* residual check (HPL_pdtest.cpp:303-432)                              */
/* ------------------------------------------------------------------ */

/* out[0] = scaled residual with the true inf-norm of r
 * out[1] = ||r||_oo, out[2] = the reference's printed form, which uses |r_0|
 *          in place of ||r||_oo (HPL_pdlange_device.cpp:284), scaled alike
 * out[3] = ||A||_oo, out[4] = ||A||_1, out[5] = ||x||_oo, out[6] = ||x||_1,
 * out[7] = ||b||_oo, out[8] = |r_0|.
 * [A|b] is regenerated column by column from the seed. */
void hplo_residual(int n, int seed, double diag_shift, const double* x,
                   double* out) {
  double* r    = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  double* rsum = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  double* col  = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double  anorm1 = 0.0;
  for(int j = 0; j < n; ++j) {
    hplo_matgen_tile(n, 0, j, n, 1, col, n, seed, diag_shift);
    double cs = 0.0;
    const double xj = x[j];
    for(int i = 0; i < n; ++i) {
      r[i] -= col[i] * xj;
      rsum[i] += fabs(col[i]);
      cs += fabs(col[i]);
    }
    if(cs > anorm1) anorm1 = cs;
  }
  hplo_matgen_tile(n, 0, n, n, 1, col, n, seed, diag_shift); /* b */
  double bnorm = 0.0, anormi = 0.0, rnorm = 0.0, xnormi = 0.0, xnorm1 = 0.0;
  for(int i = 0; i < n; ++i) {
    r[i] += col[i];
    if(fabs(col[i]) > bnorm) bnorm = fabs(col[i]);
    if(rsum[i] > anormi) anormi = rsum[i];
    if(fabs(r[i]) > rnorm) rnorm = fabs(r[i]);
    if(fabs(x[i]) > xnormi) xnormi = fabs(x[i]);
    xnorm1 += fabs(x[i]);
  }
  const double eps   = ldexp(1.0, -53);
  const double denom = eps * (anormi * xnormi + bnorm) * (double)n;
  const double r0    = n > 0 ? fabs(r[0]) : 0.0;
  out[0]             = n > 0 ? rnorm / denom : 0.0;
  out[1]             = rnorm;
  out[2]             = n > 0 ? r0 / denom : 0.0;
  out[3]             = anormi;
  out[4]             = anorm1;
  out[5]             = xnormi;
  out[6]             = xnorm1;
  out[7]             = bnorm;
  out[8]             = r0;
  free(col);
  free(rsum);
  free(r);
}

/* HPL flop counts: the reference's (2/3 N^3 + 3/2 N^2), HPL_pdtest.cpp:161-165,
 * and the 2/3 N^3 + 2 N^2 form used for the headline metric. */
double hplo_flops_ref(int n) {
  const double N = (double)n;
  return (2.0 / 3.0) * N * N * N + 1.5 * N * N;
}
double hplo_flops_hpl(int n) {
  const double N = (double)n;
  return (2.0 / 3.0) * N * N * N + 2.0 * N * N;
}
