/*
 * sanitize_main.c -- TEST INFRASTRUCTURE ONLY: runs every entry point of the
 * CPU oracle (hpl_oracle.c) on small inputs, for a build under
 * AddressSanitizer and UndefinedBehaviorSanitizer (`make -C oracle
 * sanitize`).  Exit status 0 when every call returned what it must; the
 * sanitizers abort on the first memory or undefined-behaviour error.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

/* the oracle's exported functions (hpl_oracle.c) */
void     hplo_jump(uint64_t k, uint64_t mult, uint64_t iadd, uint64_t* A,
               uint64_t* C);
uint64_t hplo_start(uint64_t seed);
uint64_t hplo_state(uint64_t seed, uint64_t k);
double   hplo_value(uint64_t r);
void     hplo_matgen_tile(int M, int i0, int j0, int m, int n, double* A,
                      int lda, int seed, double diag_shift);
void     hplo_matgen(int M, int N, double* A, int lda, int seed,
                 double diag_shift);
int      hplo_pdfact_v(double* P, int lda, int m, int jb, int nbmin, int ndiv,
                  int pfact, int rfact, int* ipiv);
void     hplo_pipid(int jb, const int* ipiv, int* permU);
void     hplo_laswp(double* A, int lda, int ncols, int k, const int* ipiv);
void     hplo_dlaswp00N(int jb, int ncols, double* A, int lda,
                    const int* permU);
void     hplo_trsv_upper(int n, const double* U, int ldu, double* x);
int      hplo_pdgesv_v(int n, int nb, int nbmin, int ndiv, int pfact,
                  int rfact, double* A, int lda, int* ipiv, double* x);
void     hplo_residual(int n, int seed, double diag_shift, const double* x,
                   double* out);
double   hplo_flops_ref(int n);
double   hplo_flops_hpl(int n);

static int fails = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if(!(c)) {                                                     \
      fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      ++fails;                                                     \
    }                                                              \
  } while(0)

static int hpl_ld(int n) { return ((n > 0 ? n : 0) + 95) / 128 * 128 + 32; }

int main(void) {
  /* LCG: the jump equals sequential stepping */
  uint64_t A = 0, C = 0;
  hplo_jump(5, 6364136223846793005ULL, 1ULL, &A, &C);
  uint64_t s = hplo_start(100), t = s;
  for(int k = 0; k < 5; ++k) t = 6364136223846793005ULL * t + 1ULL;
  CHECK(A * s + C == t);
  CHECK(hplo_state(100, 5) == t);
  CHECK(fabs(hplo_value(t)) <= 0.5);

  /* ragged sizes, every panel variant, a full solve each */
  const int sizes[][3] = {{1, 1, 1}, {7, 3, 2}, {65, 16, 4}, {130, 32, 8}};
  for(size_t q = 0; q < sizeof(sizes) / sizeof(sizes[0]); ++q) {
    const int n = sizes[q][0], nb = sizes[q][1], nbmin = sizes[q][2];
    const int lda = hpl_ld(n);
    double*   M   = malloc(sizeof(double) * (size_t)lda * (size_t)(n + 1));
    double*   x   = malloc(sizeof(double) * (size_t)n);
    int*      piv = malloc(sizeof(int) * (size_t)n);
    for(int pf = 301; pf <= 303; ++pf)
      for(int rf = 301; rf <= 303; ++rf) {
        hplo_matgen(n, n + 1, M, lda, 100, 0.0);
        CHECK(hplo_pdgesv_v(n, nb, nbmin, 2, pf, rf, M, lda, piv, x) == 0);
        double out[9];
        hplo_residual(n, 100, 0.0, x, out);
        CHECK(out[0] < 16.0);
      }
    CHECK(hplo_pdgesv_v(n, nb, nbmin, 2, 300, 303, M, lda, piv, x) == -1);
    free(piv);
    free(x);
    free(M);
  }

  /* a panel, its interchanges composed, applied two ways */
  {
    const int m = 50, jb = 12, nc = 9;
    double    P[50 * 12], B1[50 * 9], B2[50 * 9];
    int       ipiv[12], permU[24];
    hplo_matgen_tile(m, 0, 0, m, jb, P, m, 7, 0.0);
    CHECK(hplo_pdfact_v(P, m, m, jb, 4, 2, 301, 302, ipiv) == 0);
    hplo_pipid(jb, ipiv, permU);
    hplo_matgen_tile(m, 0, jb, m, nc, B1, m, 7, 0.0);
    for(int i = 0; i < m * nc; ++i) B2[i] = B1[i];
    hplo_laswp(B1, m, nc, jb, ipiv);
    hplo_dlaswp00N(jb, nc, B2, m, permU);
    for(int i = 0; i < m * nc; ++i) CHECK(B1[i] == B2[i]);
    double U[9] = {2, 0, 0, 1, 4, 0, 1, 1, 8}, y[3] = {4, 5, 8};
    hplo_trsv_upper(3, U, 3, y);
    CHECK(y[2] == 1.0 && y[1] == 1.0 && y[0] == 1.0);
  }
  CHECK(hplo_flops_ref(1000) > 0.0 && hplo_flops_hpl(1000) > 0.0);
  if(fails) {
    fprintf(stderr, "%d checks failed\n", fails);
    return 1;
  }
  puts("oracle sanitizer run: all checks passed");
  return 0;
}
