// xhplx -- the rocHPL driver surface (rochpl executable) over the hplx C ABI.
//
// Mirrors src/HPL_pddriver.cpp (test loops, summary), src/HPL_pdinfo.cpp
// (command line -P -Q -p -q -N -NB -f -it -i --version -h, the 31-line
// HPL.dat, parameter echo) and src/HPL_pdtest.cpp (matgen, timed
// HPL_pdgesv, W-line, residual check).  One process drives one GPU, so only
// the P=Q=1 grid runs; other grids are counted as skipped with a message, as
// the reference counts tests it cannot run.  Every PFACT and RFACT runs.
//
// Extra flags (not in the reference): --parse-only echoes the parameters and
// exits without touching the GPU; --dump DIR writes each solved test's x
// (N doubles) and pivot vector (N int32, 0-based global rows) to
// DIR/xhplx_N<N>_NB<NB>_PF<L|C|R>_NBMIN<nbmin>_NDIV<ndiv>_RF<L|C|R>_it<it>
// .{x,ipiv}, so that a checker can recompute the residual and compare the
// pivots.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include <sys/time.h>

#include "../../include/hplx.h"

namespace {

constexpr int kMaxParam = 20;  // HPL_MAX_PARAM, include/hpl_ptest.hpp
constexpr int kISeed    = 100; // HPL_ISEED

struct Params {
  std::string      outname = "HPL.out";
  int              devout  = 7;
  std::vector<int> N, NB, P, Q, PF, NBM, NDV, RF, TP, DH;
  int              pmap   = 1;
  double           thrsh  = 16.0;
  int              fswap  = HPLX_SWAP01;
  int              fsthr  = 64;
  int              L1notran = 0;
  int              Unotran  = 0;
  int              equil    = 0;
  int              align    = 8;
  double           frac     = 0.3;
  int              its      = 1;
  std::string      dump_dir;  // --dump: where x and the pivots go
};

struct Counts {
  int kpass = 0, kfail = 0, kskip = 0;
};

FILE* g_out = stderr;

void warn(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void warn(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::fprintf(stderr, "xhplx: ");
  std::vfprintf(stderr, fmt, ap);
  std::fprintf(stderr, "\n");
  va_end(ap);
}

int fact_of(int j) {
  return j == 0 ? HPLX_LEFT_LOOKING : j == 1 ? HPLX_CROUT : HPLX_RIGHT_LOOKING;
}

int top_of(int j) {
  static const int t[] = {HPLX_1RING, HPLX_1RING_M, HPLX_2RING,
                          HPLX_2RING_M, HPLX_BLONG, HPLX_BLONG_M};
  return (j >= 0 && j <= 4) ? t[j] : HPLX_BLONG_M;
}

// ---------------------------------------------------------------------------
// HPL.dat (src/HPL_pdinfo.cpp:585-1017)
// ---------------------------------------------------------------------------
struct LineReader {
  FILE* f;
  char  line[1024];
  bool  next() { return std::fgets(line, sizeof(line), f) != nullptr; }
  // first token
  std::string tok() {
    char b[256] = {0};
    std::sscanf(line, "%255s", b);
    return b;
  }
  std::vector<std::string> toks(int n) {
    std::vector<std::string> v;
    const char*               p = line;
    for(int i = 0; i < n; ++i) {
      char b[256] = {0};
      int  used   = 0;
      if(std::sscanf(p, "%255s%n", b, &used) != 1) break;
      v.push_back(b);
      p += used;
    }
    return v;
  }
};

bool read_count(LineReader& r, int& n, const char* what) {
  if(!r.next()) return false;
  n = std::atoi(r.tok().c_str());
  if(n < 1 || n > kMaxParam) {
    warn("Number of values of %s is less than 1 or greater than %d", what,
         kMaxParam);
    return false;
  }
  return true;
}

bool read_list(LineReader& r, int n, std::vector<int>& out) {
  if(!r.next()) return false;
  auto t = r.toks(n);
  if(static_cast<int>(t.size()) < n) return false;
  out.clear();
  for(auto& s : t) out.push_back(std::atoi(s.c_str()));
  return true;
}

int parse_hpldat(const char* path, Params& p) {
  FILE* f = std::fopen(path, "r");
  if(!f) {
    warn("cannot open file %s", path);
    return 1;
  }
  LineReader r{f, {0}};
  int        err = 0, n = 0;
  std::vector<int> tmp;
  auto fail = [&](const char* m) {
    if(m) warn("%s", m);
    err = 1;
  };
  do {
    if(!r.next() || !r.next()) { fail("HPL.dat too short"); break; }
    if(!r.next()) { fail("HPL.dat too short"); break; }
    p.outname = r.tok();
    if(!r.next()) { fail("HPL.dat too short"); break; }
    p.devout = std::atoi(r.tok().c_str());
    if(!read_count(r, n, "N") || !read_list(r, n, p.N)) { fail(nullptr); break; }
    for(int v : p.N)
      if(v < 0) { fail("Value of N less than 0"); break; }
    if(err) break;
    if(!read_count(r, n, "NB") || !read_list(r, n, p.NB)) { fail(nullptr); break; }
    for(int v : p.NB)
      if(v < 1) { fail("Value of NB less than 1"); break; }
    if(err) break;
    if(!r.next()) { fail("HPL.dat too short"); break; }
    p.pmap = std::atoi(r.tok().c_str()) == 0 ? 0 : 1;
    if(!read_count(r, n, "grids") || !read_list(r, n, p.P) ||
       !read_list(r, n, p.Q)) { fail(nullptr); break; }
    for(size_t i = 0; i < p.P.size(); ++i)
      if(p.P[i] < 1 || p.Q[i] < 1) { fail("Value of P or Q less than 1"); break; }
    if(err) break;
    if(!r.next()) { fail("HPL.dat too short"); break; }
    p.thrsh = std::atof(r.tok().c_str());
    if(!read_count(r, n, "PFACT") || !read_list(r, n, tmp)) { fail(nullptr); break; }
    p.PF.clear();
    for(int v : tmp) p.PF.push_back(fact_of(v));
    if(!read_count(r, n, "NBMIN") || !read_list(r, n, p.NBM)) { fail(nullptr); break; }
    for(int v : p.NBM)
      if(v < 1) { fail("Value of NBMIN less than 1"); break; }
    if(err) break;
    if(!read_count(r, n, "NDIV") || !read_list(r, n, p.NDV)) { fail(nullptr); break; }
    for(int v : p.NDV)
      if(v < 2) { fail("Value of NDIV less than 2"); break; }
    if(err) break;
    if(!read_count(r, n, "RFACT") || !read_list(r, n, tmp)) { fail(nullptr); break; }
    p.RF.clear();
    for(int v : tmp) p.RF.push_back(fact_of(v));
    if(!read_count(r, n, "BCAST") || !read_list(r, n, tmp)) { fail(nullptr); break; }
    p.TP.clear();
    for(int v : tmp) p.TP.push_back(top_of(v));
    if(!read_count(r, n, "DEPTH") || !read_list(r, n, p.DH)) { fail(nullptr); break; }
    for(int v : p.DH)
      if(v != 1) { fail("Lookahead depth must be 1"); break; }
    if(err) break;
    if(!r.next()) { fail("HPL.dat too short"); break; }
    {
      const int j = std::atoi(r.tok().c_str());
      p.fswap = j == 0 ? HPLX_SWAP00 : j == 1 ? HPLX_SWAP01
                                              : j == 2 ? HPLX_SW_MIX : HPLX_SWAP01;
      if(p.fswap != HPLX_SWAP01) { fail("Only long swapping (1) is implemented"); break; }
    }
    if(!r.next()) { fail("HPL.dat too short"); break; }
    p.fsthr = std::atoi(r.tok().c_str());
    if(p.fsthr <= 0) p.fsthr = 0;
    if(!r.next()) { fail("HPL.dat too short"); break; }
    p.L1notran = std::atoi(r.tok().c_str());
    if(p.L1notran != 0 && p.L1notran != 1) p.L1notran = 0;
    if(!r.next()) { fail("HPL.dat too short"); break; }
    p.Unotran = std::atoi(r.tok().c_str());
    if(p.Unotran != 0 && p.Unotran != 1) p.Unotran = 0;
    if(p.Unotran != 0) { fail("U  in no-transposed form unsupported"); break; }
    if(!r.next()) { fail("HPL.dat too short"); break; }
    p.equil = std::atoi(r.tok().c_str());
    if(p.equil != 0 && p.equil != 1) p.equil = 1;
    if(p.equil != 0) { fail("Equilibration currently unsupported"); break; }
    if(!r.next()) { fail("HPL.dat too short"); break; }
    p.align = std::atoi(r.tok().c_str());
    if(p.align <= 0) p.align = 4;
  } while(false);
  std::fclose(f);
  if(err) warn("Illegal input in file %s. Exiting ...", path);
  return err;
}

// ---------------------------------------------------------------------------
// echo (src/HPL_pdinfo.cpp:1242-1578)
// ---------------------------------------------------------------------------
void echo_ints(const char* name, const std::vector<int>& v) {
  std::fprintf(g_out, "\n%s:", name);
  for(size_t i = 0; i < v.size(); ++i) {
    if(i > 0 && i % 8 == 0) std::fprintf(g_out, "\n        ");
    std::fprintf(g_out, "%8d ", v[i]);
  }
}

const char* fact_name(int f) {
  return f == HPLX_LEFT_LOOKING ? "    Left " : f == HPLX_CROUT ? "   Crout "
                                                              : "   Right ";
}

void echo(const Params& p) {
  const char* bar = "================================================================================";
  std::fprintf(g_out, "%s\n", bar);
  std::fprintf(g_out, "HPLinpack benchmark driver -- rochpl_amd (MI355X, one GPU, P=Q=1 HPL_pdgesv path)\n");
  std::fprintf(g_out, "%s\n", bar);
  std::fprintf(g_out, "\nAn explanation of the input/output parameters follows:\n");
  std::fprintf(g_out, "T/V    : Wall time / encoded variant.\n");
  std::fprintf(g_out, "N      : The order of the coefficient matrix A.\n");
  std::fprintf(g_out, "NB     : The partitioning blocking factor.\n");
  std::fprintf(g_out, "P      : The number of process rows.\n");
  std::fprintf(g_out, "Q      : The number of process columns.\n");
  std::fprintf(g_out, "Time   : Time in seconds to solve the linear system.\n");
  std::fprintf(g_out, "Gflops : Rate of execution for solving the linear system.\n\n");
  std::fprintf(g_out, "The following parameter values will be used:\n");
  echo_ints("N      ", p.N);
  echo_ints("NB     ", p.NB);
  std::fprintf(g_out, "\nPMAP   : %s", p.pmap == 0 ? "Row-major process mapping"
                                                  : "Column-major process mapping");
  echo_ints("P      ", p.P);
  echo_ints("Q      ", p.Q);
  std::fprintf(g_out, "\nPFACT  :");
  for(int f : p.PF) std::fprintf(g_out, "%s", fact_name(f));
  echo_ints("NBMIN  ", p.NBM);
  echo_ints("NDIV   ", p.NDV);
  std::fprintf(g_out, "\nRFACT  :");
  for(int f : p.RF) std::fprintf(g_out, "%s", fact_name(f));
  std::fprintf(g_out, "\nBCAST  :");
  for(int t : p.TP) {
    static const char* nm[] = {"  1ring ", " 1ringM ", "  2ring ",
                               " 2ringM ", "  Blong ", " BlongM "};
    std::fprintf(g_out, "%s", nm[t - HPLX_1RING]);
  }
  echo_ints("DEPTH  ", p.DH);
  std::fprintf(g_out, "\nSWAP   : Spread-roll (long)");
  std::fprintf(g_out, "\nL1     : %s form", p.L1notran ? "no-transposed" : "transposed");
  std::fprintf(g_out, "\nU      : transposed form");
  std::fprintf(g_out, "\nEQUIL  : no");
  std::fprintf(g_out, "\nALIGN  : %d double precision words", p.align);
  std::fprintf(g_out, "\nFRAC   : %g (update split; single-GPU runs do not split)\n\n",
               p.frac);
  std::fprintf(g_out, "----------------------------------------------------------------------------\n\n");
  std::fprintf(g_out, "- The matrix A is randomly generated for each test.\n");
  std::fprintf(g_out, "- The following scaled residual check will be computed:\n");
  std::fprintf(g_out, "      ||Ax-b||_oo / ( eps * ( || x ||_oo * || A ||_oo + || b ||_oo ) * N )\n");
  std::fprintf(g_out, "- The relative machine precision (eps) is taken to be     %18.6e\n",
               std::ldexp(1.0, -53));
  std::fprintf(g_out, "- Computational tests pass if scaled residuals are less than %14.1f\n\n",
               p.thrsh);
}

double wall() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + 1e-6 * tv.tv_usec;
}

// --dump: x and the pivot vector of one solve, raw little-endian arrays
char fact_letter(int f) {
  return f == HPLX_LEFT_LOOKING ? 'L' : f == HPLX_CROUT ? 'C' : 'R';
}

bool dump_solution(hplx_ctx* ctx, const Params& p, const hplx_pmat& mat, int N,
                   int NB, int pf, int nbm, int ndv, int rf, int it) {
  std::vector<double> x(N);
  std::vector<int>    piv(N);
  if(hplx_memcpy_d2h(ctx, x.data(), mat.X, sizeof(double) * N) != HPLX_SUCCESS ||
     hplx_memcpy_d2h(ctx, piv.data(), mat.ipiv, sizeof(int) * N) != HPLX_SUCCESS)
    return false;
  char base[512];
  std::snprintf(base, sizeof(base),
                "%s/xhplx_N%d_NB%d_PF%c_NBMIN%d_NDIV%d_RF%c_it%d",
                p.dump_dir.c_str(), N, NB, fact_letter(pf), nbm, ndv,
                fact_letter(rf), it);
  const std::string names[2] = {std::string(base) + ".x",
                                std::string(base) + ".ipiv"};
  const void*  data[2]  = {x.data(), piv.data()};
  const size_t bytes[2] = {sizeof(double) * N, sizeof(int) * N};
  for(int f = 0; f < 2; ++f) {
    FILE* fp = std::fopen(names[f].c_str(), "wb");
    if(!fp) return false;
    const bool ok = std::fwrite(data[f], 1, bytes[f], fp) == bytes[f];
    if(std::fclose(fp) != 0 || !ok) return false;
  }
  return true;
}

// ---------------------------------------------------------------------------
// one test (src/HPL_pdtest.cpp:21-495)
// ---------------------------------------------------------------------------
void run_test(hplx_ctx* ctx, const Params& p, int N, int NB, int pf, int nbm,
              int ndv, int rf, int tp, int dh, Counts& cnt, bool& first) {
  hplx_palg a;
  hplx_palg_default(&a);
  a.btopo = tp;
  a.depth = dh;
  a.nbdiv = ndv;
  a.nbmin = nbm;
  a.pfact = pf;
  a.rfact = rf;
  a.fswap = p.fswap;
  a.fsthr = p.fsthr;
  a.equil = p.equil;
  a.L1notran = p.L1notran;
  a.Unotran  = p.Unotran;
  a.align    = p.align;
  a.frac     = p.frac;
  a.its      = p.its;
  hplx_pmat mat;
  if(hplx_pmat_alloc(ctx, N, NB, &mat) != HPLX_SUCCESS) {
    warn("[%d,%d] memory allocation failed for A, x and b: %s -- skip", N, NB,
         hplx_last_error(ctx));
    cnt.kskip++;
    return;
  }
  for(int it = 0; it < p.its; ++it) {
    hplx_pdrandmat(ctx, N, N + 1, NB, mat.A, mat.ld, kISeed, 0.0);
    hplx_synchronize(ctx);
    time_t      ts = std::time(nullptr);
    const double t0 = wall();
    const int    rc = hplx_pdgesv(ctx, &a, &mat);
    const double t  = wall() - t0;
    time_t       te = std::time(nullptr);
    if(rc < 0) {
      warn("HPL_pdgesv failed: %s", hplx_last_error(ctx));
      cnt.kfail++;
      continue;
    }
    if(first) {
      std::fprintf(g_out, "================================================================================\n");
      std::fprintf(g_out, "T/V                N    NB     P     Q               Time                 Gflops\n");
      std::fprintf(g_out, "--------------------------------------------------------------------------------\n");
      if(p.thrsh <= 0.0) first = false;
    }
    const double Nd = static_cast<double>(N);
    const double gflops = (Nd / 1.0e9) * (Nd / t) * ((2.0 / 3.0) * Nd + 1.5);
    const char   rfc = fact_letter(rf), pfc = fact_letter(pf);
    const char   ctop = static_cast<char>('0' + (tp - HPLX_1RING));
    std::fprintf(g_out, "W%c%1d%c%c%1d%c%1d%12d %5d %5d %5d %18.2f     %18.3e\n",
                 p.pmap == 0 ? 'R' : 'C', dh, ctop, rfc, ndv, pfc, nbm, N, NB,
                 1, 1, t, gflops);
    std::fprintf(g_out, "HPL_pdgesv() start time %s\n", std::ctime(&ts));
    std::fprintf(g_out, "HPL_pdgesv() end time   %s\n", std::ctime(&te));
    if(!p.dump_dir.empty() &&
       !dump_solution(ctx, p, mat, N, NB, pf, nbm, ndv, rf, it))
      warn("could not write the --dump files of N=%d NB=%d", N, NB);
    if(rc > 0) {
      warn("exact zero pivot at column %d", rc);
      cnt.kfail++;
      continue;
    }
    if(p.thrsh <= 0.0) {
      cnt.kpass++;
      continue;
    }
    double out[9];
    if(hplx_pdtest_check(ctx, &mat, kISeed, 0.0, out) != HPLX_SUCCESS) {
      warn("residual check failed to run: %s", hplx_last_error(ctx));
      cnt.kfail++;
      continue;
    }
    const double resid = out[0];
    if(resid < p.thrsh)
      cnt.kpass++;
    else
      cnt.kfail++;
    std::fprintf(g_out, "--------------------------------------------------------------------------------\n");
    std::fprintf(g_out, "||Ax-b||_oo/(eps*(||A||_oo*||x||_oo+||b||_oo)*N)= %16.7f ...... %s\n",
                 resid, resid < p.thrsh ? "PASSED" : "FAILED");
    if(resid >= p.thrsh) {
      std::fprintf(g_out, "||Ax-b||_oo  . . . . . . . . . . . . . . . . . = %18.6f\n", out[1]);
      std::fprintf(g_out, "||A||_oo . . . . . . . . . . . . . . . . . . . = %18.6f\n", out[3]);
      std::fprintf(g_out, "||A||_1  . . . . . . . . . . . . . . . . . . . = %18.6f\n", out[4]);
      std::fprintf(g_out, "||x||_oo . . . . . . . . . . . . . . . . . . . = %18.6f\n", out[5]);
      std::fprintf(g_out, "||x||_1  . . . . . . . . . . . . . . . . . . . = %18.6f\n", out[6]);
      std::fprintf(g_out, "||b||_oo . . . . . . . . . . . . . . . . . . . = %18.6f\n", out[7]);
    }
    std::fprintf(g_out, "(reference |r_0| form of the residual: %16.7f)\n", out[2]);
    std::fflush(g_out);
  }
  hplx_pmat_free(ctx, &mat);
}

void usage() {
  std::printf(
      "xhplx (rochpl_amd) command line options:\n"
      "-P  [ --ranksP ] arg (=1)          process grid rows (must be 1)\n"
      "-Q  [ --ranksQ ] arg (=1)          process grid columns (must be 1)\n"
      "-N  [ --sizeN ]  arg (=45312)      matrix order\n"
      "-NB [ --sizeNB ] arg (=384)        panel width\n"
      "-f  [ --frac ] arg (=0.3)          update split fraction\n"
      "-it arg (=1)                       iterations of each test\n"
      "-i  [ --input ]  arg (=HPL.dat)    input file; other options are then ignored\n"
      "-h  [ --help ]                     this message\n"
      "--version                          version\n"
      "--parse-only                       echo the parameters and exit\n"
      "--dump DIR                         write each test's x and pivots to DIR\n");
}

}  // namespace

int main(int argc, char** argv) {
  Params      p;
  int         P = 1, Q = 1, n = 45312, nb = 384, it = 1;
  double      frac = 0.3;
  bool        cmdline = false, inputfile = false, parse_only = false;
  std::string infile = "HPL.dat", dump_dir;
  for(int i = 1; i < argc; ++i) {
    const char* a   = argv[i];
    auto        val = [&](const char* flag) -> const char* {
      if(i + 1 >= argc) {
        warn("missing value for %s", flag);
        std::exit(1);
      }
      return argv[++i];
    };
    if(!std::strcmp(a, "-h") || !std::strcmp(a, "--help")) {
      usage();
      return 0;
    } else if(!std::strcmp(a, "--version")) {
      std::printf("rochpl_amd xhplx version: 0.1.0 (rocHPL 7.0.1 driver surface)\n");
      return 0;
    } else if(!std::strcmp(a, "--parse-only")) {
      parse_only = true;
    } else if(!std::strcmp(a, "--dump")) {
      dump_dir = val(a);
    } else if(!std::strcmp(a, "-P") || !std::strcmp(a, "--ranksP")) {
      P = std::atoi(val(a));
      cmdline = true;
      if(P < 1) { warn("Illegal value for P. Exiting ..."); return 1; }
    } else if(!std::strcmp(a, "-Q") || !std::strcmp(a, "--ranksQ")) {
      Q = std::atoi(val(a));
      cmdline = true;
      if(Q < 1) { warn("Illegal value for Q. Exiting ..."); return 1; }
    } else if(!std::strcmp(a, "-p") || !std::strcmp(a, "-q")) {
      (void)val(a);  // node-local grid shape: one process, nothing to map
      cmdline = true;
    } else if(!std::strcmp(a, "-N") || !std::strcmp(a, "--sizeN")) {
      n = std::atoi(val(a));
      cmdline = true;
      if(n < 1) { warn("Illegal value for N. Exiting ..."); return 1; }
    } else if(!std::strcmp(a, "-NB") || !std::strcmp(a, "--sizeNB")) {
      nb = std::atoi(val(a));
      cmdline = true;
      if(nb < 1) { warn("Illegal value for NB. Exiting ..."); return 1; }
    } else if(!std::strcmp(a, "-f") || !std::strcmp(a, "--frac")) {
      frac = std::atof(val(a));
    } else if(!std::strcmp(a, "-it")) {
      it = std::atoi(val(a));
    } else if(!std::strcmp(a, "-i") || !std::strcmp(a, "--input")) {
      infile    = val(a);
      inputfile = true;
    }
  }
  if(P * Q > 1) {
    warn("Need at least %d processes for these tests; xhplx runs one "
         "process on one GPU (P=Q=1)", P * Q);
    return 1;
  }
  if(!inputfile && cmdline) {
    p.N  = {n};
    p.NB = {nb};
    p.P  = {P};
    p.Q  = {Q};
    p.PF = {HPLX_RIGHT_LOOKING};
    p.NBM = {32};
    p.NDV = {2};
    p.RF  = {HPLX_RIGHT_LOOKING};
    p.TP  = {HPLX_1RING};
    p.DH  = {1};
    p.outname = "HPL.out";
    p.devout  = 8;  // a file
  } else if(parse_hpldat(infile.c_str(), p)) {
    return 1;
  }
  p.frac     = frac;
  p.its      = it > 0 ? it : 1;
  p.dump_dir = dump_dir;
  if(p.devout == 6)
    g_out = stdout;
  else if(p.devout == 7)
    g_out = stderr;
  else if(!(g_out = std::fopen(p.outname.c_str(), "w"))) {
    warn("cannot open file %s.", p.outname.c_str());
    return 1;
  }
  echo(p);
  if(parse_only) {
    if(g_out != stdout && g_out != stderr) std::fclose(g_out);
    return 0;
  }

  hplx_ctx* ctx = nullptr;
  if(hplx_ctx_create(0, &ctx) != HPLX_SUCCESS) {
    warn("no usable GPU (hplx_ctx_create failed)");
    return 1;
  }
  Counts cnt;
  bool   first = true;
  for(size_t g = 0; g < p.P.size(); ++g) {
    const int ntests = static_cast<int>(p.N.size() * p.NB.size() * p.PF.size() *
                                        p.NBM.size() * p.NDV.size() *
                                        p.RF.size() * p.TP.size() * p.DH.size());
    if(p.P[g] * p.Q[g] != 1) {
      warn("grid %dx%d needs %d processes; skipped (one GPU, P=Q=1)", p.P[g],
           p.Q[g], p.P[g] * p.Q[g]);
      cnt.kskip += ntests;
      continue;
    }
    // the reference's nesting (src/HPL_pddriver.cpp:141-157)
    for(int N : p.N)
      for(int NB : p.NB)
        for(int dh : p.DH)
          for(int tp : p.TP)
            for(int rf : p.RF)
              for(int pf : p.PF)
                for(int nbm : p.NBM)
                  for(int ndv : p.NDV)
                    run_test(ctx, p, N, NB, pf, nbm, ndv, rf, tp, dh, cnt, first);
  }
  hplx_ctx_destroy(ctx);

  const int ktest = cnt.kpass + cnt.kfail + cnt.kskip;
  std::fprintf(g_out, "================================================================================\n");
  std::fprintf(g_out, "\nFinished %6d tests with the following results:\n", ktest);
  if(p.thrsh > 0.0) {
    std::fprintf(g_out, "         %6d tests completed and passed residual checks,\n", cnt.kpass);
    std::fprintf(g_out, "         %6d tests completed and failed residual checks,\n", cnt.kfail);
    std::fprintf(g_out, "         %6d tests skipped because of illegal input values.\n", cnt.kskip);
  } else {
    std::fprintf(g_out, "         %6d tests completed without checking,\n", cnt.kpass);
    std::fprintf(g_out, "         %6d tests skipped because of illegal input values.\n", cnt.kskip);
  }
  std::fprintf(g_out, "--------------------------------------------------------------------------------\n");
  std::fprintf(g_out, "\nEnd of Tests.\n");
  std::fprintf(g_out, "================================================================================\n");
  if(g_out != stdout && g_out != stderr) std::fclose(g_out);
  return cnt.kfail ? 2 : 0;
}
