// C ABI of rochpl_amd (include/hplx.h): context, allocation, the P=Q=1
// HPL_pdgesv panel loop and the residual check.
//
// The panel loop restates src/pgesv/HPL_pdgesv.cpp:168-245 at P=Q=1 with the
// host out of the loop: every step is queued on the device (an update
// stream and, with the depth-1 look-ahead, a panel stream linked by events)
// and no host synchronisation happens until the back-solve.  Per panel
// (width jb, at global row/column j):
//   1. panel factorisation in place (HPL_pdfact)             -> ipiv[j:j+jb)
//   2. pivot composition on the device (HPL_pdpanel_swapids) -> permU
//   3. explicit inverse of the unit lower L11                -> Linv
//   4. row swaps of the trailing columns; the new top rows go to W
//      (HPL_pdlaswp_end / HPL_dlaswp00N)
//   5. U12 = Linv * W written in place over A12      (dtrsm of HPL_pdupdateTT)
//   6. A22 -= L21 * U12                              (dgemm of HPL_pdupdateTT)
// The look-ahead (here as in the reference) and the reference's split update
// only reorder these steps for overlap; each matrix element sees the same
// sequence of operations.  b is
// column n of the augmented matrix and is carried through 4-6, so after the
// loop the back-solve U x = b' finishes the solve (HPL_pdtrsv).
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <new>

#include "api_common.hpp"

namespace {

using hplx::Ctx;
using hplx::fail;
using hplx::use;
using hplx::align_up;

int hpl_ld(int n) { return ((n > 0 ? n : 0) + 95) / 128 * 128 + 32; }

constexpr int kCensusGrid = 16384;  // workgroups of the CU census

// The CU sets the look-ahead hands to a panel factorisation.  A census
// kernel records the __smid() of the CU each of kCensusGrid short
// workgroups ran on; vacate set r (r = 1 .. 32) then marks the first r CUs
// of every XCD, taken round-robin over its shader engines.  If the census
// did not see the same number of CUs on every XCD (or fewer than r), sets
// that cannot be formed stay empty and c.cus_per_xcd records how many can.
int cu_sets(Ctx& c, int* census_dev) {
  std::vector<int> ids(kCensusGrid);
  if(hplx::launch_cu_census(c.own_stream, census_dev, kCensusGrid) != hipSuccess ||
     hipMemcpyAsync(ids.data(), census_dev, sizeof(int) * kCensusGrid,
                    hipMemcpyDeviceToHost, c.own_stream) != hipSuccess ||
     hipStreamSynchronize(c.own_stream) != hipSuccess)
    return HPLX_ERR_HIP;
  bool seen[512] = {false};
  for(int v : ids)
    if(v >= 0 && v < 512) seen[v] = true;
  // per XCD: CU ids grouped by shader engine, then interleaved over engines
  std::vector<int> per_xcd[8];
  int              cnt[8] = {0};
  for(int x = 0; x < 8; ++x) {
    std::vector<int> se[4];
    for(int e = 0; e < 4; ++e)
      for(int u = 0; u < 16; ++u)
        if(seen[(x << 6) | (e << 4) | u]) se[e].push_back((x << 6) | (e << 4) | u);
    for(size_t k = 0; k < 16; ++k)
      for(int e = 0; e < 4; ++e)
        if(k < se[e].size()) per_xcd[x].push_back(se[e][k]);
    cnt[x] = static_cast<int>(per_xcd[x].size());
  }
  int r_max = cnt[0];
  for(int x = 1; x < 8; ++x) r_max = cnt[x] < r_max ? cnt[x] : r_max;
  if(r_max > hplx::kVacateSets) r_max = hplx::kVacateSets;
  c.cus_per_xcd = r_max;
  std::vector<unsigned long long> masks(hplx::kVacateSets * 8, 0ull);
  for(int r = 1; r <= r_max; ++r)
    for(int x = 0; x < 8; ++x)
      for(int k = 0; k < r; ++k) {
        const int id = per_xcd[x][k];
        masks[(r - 1) * 8 + (id >> 6)] |= 1ull << (id & 63);
      }
  if(hipMemcpy(c.vacate, masks.data(), masks.size() * 8,
               hipMemcpyHostToDevice) != hipSuccess)
    return HPLX_ERR_HIP;
  return HPLX_SUCCESS;
}

}  // namespace

namespace hplx {

// The solve with the factors of a pmat (hplx_pdgetrs), queued on s with the
// given scratch: permU (2 kMaxNB ints) and the split-K partial products of
// the transposed solve (work_doubles doubles; fewer only means fewer
// splits).  No synchronisation, no allocation.
hipError_t getrs_enqueue(hipStream_t s, const double* A, int N, int nb,
                         int lda, const int* ipiv, int trans, int nrhs,
                         double* B, int ldb, int* permU, double* work,
                         size_t work_doubles) {
#define HCHK(expr)                    \
  do {                                \
    const hipError_t e_ = (expr);     \
    if(e_ != hipSuccess) return e_;   \
  } while(0)
  const int    np = (N + nb - 1) / nb;
  // C (m x nrhs, ldb) -= Ablk (m x jb, lda) * Y (jb x nrhs, ldb) on FP64 MFMA
  auto sub_product = [&](int m, int jb, const double* Ablk, const double* Y,
                         double* Cblk) -> hipError_t {
    return hplx::launch_nn_sub(s, jb, m, nrhs, Ablk, lda, Y, ldb, Cblk, ldb);
  };
  if(!trans) {
    // L solve with the panel-by-panel interchanges, as b is carried through
    // the factorisation (HPL_pdgesv), then U x = y (HPL_pdtrsv)
    for(int p = 0; p < np; ++p) {
      const int j = p * nb, jb = (N - j) < nb ? (N - j) : nb;
      HCHK(hplx::launch_pipid(s, jb, ipiv + j, j, permU));
      HCHK(hplx::launch_swap_inplace(s, jb, nrhs, B + j, ldb, permU));
      HCHK(hplx::launch_tri_solve(s, true, false, true, jb, nrhs,
                                  A + j + static_cast<size_t>(j) * lda,
                                  lda, B + j, ldb));
      if(N - j - jb > 0)
        HCHK(sub_product(N - j - jb, jb,
                              A + j + jb + static_cast<size_t>(j) * lda, B + j,
                              B + j + jb));
    }
    for(int p = np - 1; p >= 0; --p) {
      const int j = p * nb, jb = (N - j) < nb ? (N - j) : nb;
      HCHK(hplx::launch_tri_solve(s, false, false, false, jb, nrhs,
                                  A + j + static_cast<size_t>(j) * lda,
                                  lda, B + j, ldb));
      if(j > 0)
        HCHK(sub_product(j, jb, A + static_cast<size_t>(j) * lda, B + j,
                              B));
    }
  } else {
    // A = P0^T L0 P1^T L1 ... U  =>  A^T z = g  is solved as
    // w = U^{-T} g, then for panels from the last: z_p -= L21^T z_below,
    // z_p = L11^{-T} z_p, z = P_p^T z.
    for(int p = 0; p < np; ++p) {
      const int j = p * nb, jb = (N - j) < nb ? (N - j) : nb, e = j + jb;
      HCHK(hplx::launch_tri_solve(s, true, true, false, jb, nrhs,
                                  A + j + static_cast<size_t>(j) * lda,
                                  lda, B + j, ldb));
      // z_below -= U12^T z_p (U12 = rows j..e of the columns right of the
      // panel)
      if(e < N)
        HCHK(hplx::launch_tn_sub(s, jb, N - e, nrhs,
                                 A + j + static_cast<size_t>(e) * lda,
                                 lda, B + j, ldb, B + e, ldb, work,
                                 work_doubles));
    }
    // z_p -= L21^T z_below reduces over all N - e rows below the panel for
    // a jb x nrhs result: split over K, the partials in the work buffer,
    // while they fit
    for(int p = np - 1; p >= 0; --p) {
      const int j = p * nb, jb = (N - j) < nb ? (N - j) : nb, e = j + jb;
      if(e < N)
        HCHK(hplx::launch_tn_sub(s, N - e, jb, nrhs,
                                 A + e + static_cast<size_t>(j) * lda,
                                 lda, B + e, ldb, B + j, ldb, work,
                                 work_doubles));
      HCHK(hplx::launch_tri_solve(s, false, true, true, jb, nrhs,
                                  A + j + static_cast<size_t>(j) * lda,
                                  lda, B + j, ldb));
      HCHK(hplx::launch_pipid(s, jb, ipiv + j, j, permU));
      HCHK(hplx::launch_swap_inverse(s, jb, nrhs, B + j, ldb, permU));
    }
  }
#undef HCHK
  return hipSuccess;
}

}  // namespace hplx

extern "C" {

void hplx_palg_default(hplx_palg* p) {
  if(!p) return;
  p->btopo    = HPLX_1RING;
  p->depth    = 1;
  p->nbdiv    = 2;
  p->nbmin    = 32;
  p->pfact    = HPLX_RIGHT_LOOKING;
  p->rfact    = HPLX_RIGHT_LOOKING;
  p->fswap    = HPLX_SWAP01;
  p->fsthr    = 64;
  p->equil    = 0;
  p->L1notran = 0;
  p->Unotran  = 0;
  p->align    = 8;
  p->frac     = 0.3;
  p->its      = 1;
}

int hplx_palg_check(const hplx_palg* palg) {
  char msg[256];
  return hplx::palg_status(palg, msg, sizeof(msg));
}

int hplx_ctx_create(int device, hplx_ctx** out) {
  if(!out) return HPLX_ERR_INVALID;
  *out          = nullptr;
  hplx_ctx* ctx = new(std::nothrow) hplx_ctx();
  if(!ctx) return HPLX_ERR_ALLOC;
  Ctx& c   = ctx->c;
  c.device = device;
  int ndev = 0;
  if(hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    delete ctx;
    return HPLX_ERR_HIP;
  }
  if(hipSetDevice(device) != hipSuccess ||
     hipDeviceGetAttribute(&c.num_cus, hipDeviceAttributeMultiprocessorCount,
                           device) != hipSuccess ||
     hipStreamCreateWithFlags(&c.own_stream, hipStreamNonBlocking) !=
         hipSuccess) {
    delete ctx;
    return HPLX_ERR_HIP;
  }
  c.stream = c.own_stream;
  {
    int per_cu = 0;
    if(hplx::leaf_occupancy(&per_cu) != hipSuccess || per_cu < 1) {
      (void)hipStreamDestroy(c.own_stream);
      delete ctx;
      return HPLX_ERR_HIP;
    }
    c.leaf_resident = per_cu * c.num_cus;
    // the persistent panel kernel is optional: no residency, no use
    int pk = 0;
    if(hplx::panel_kernel_occupancy(&pk) != hipSuccess) pk = 0;
    c.pk_resident = pk * c.num_cus;
  }
  if(const char* pe = std::getenv("HPLX_PROGRESS")) c.progress_every = std::atoi(pe);
  // small workspace
  const size_t nLinv  = static_cast<size_t>(hplx::kMaxNB) * hplx::kMaxNB * 8;
  const size_t nperm  = 2 * hplx::kMaxNB * sizeof(int);
  const size_t nslots = 2 * static_cast<size_t>(hplx::kMaxGrid + 1) *
                        hplx::kSlotStride * sizeof(double);
  size_t off_perm   = align_up(nLinv, 256);
  size_t off_slots  = align_up(off_perm + nperm, 256);
  size_t off_flags  = align_up(off_slots + nslots, 256);
  size_t off_dout   = off_flags + 256;
  size_t off_linv2  = align_up(off_dout + 256, 256);
  size_t off_perm2  = align_up(off_linv2 + nLinv, 256);
  size_t off_claim  = align_up(off_perm2 + nperm, 256);
  size_t off_stop   = align_up(off_claim + 2 * hplx::kClaimInts * sizeof(int), 256);
  size_t off_vacate = off_stop + 256;
  size_t off_census = align_up(off_vacate + hplx::kVacateSets * 8 * 8, 256);
  size_t off_clog   = align_up(off_census + kCensusGrid * sizeof(int), 256);
  size_t off_pops   = align_up(off_clog + 8 * hplx::kMaxClaimLog * sizeof(int), 256);
  size_t off_pbar   = align_up(off_pops + 2 * hplx::kMaxPanelOps * sizeof(hplx::PanelOp), 256);
  size_t total      = off_pbar + 256;
  auto undo = [&](int code) {
    if(c.small_base) (void)hipFree(c.small_base);
    if(c.ustream) (void)hipStreamDestroy(c.ustream);
    if(c.pstream) (void)hipStreamDestroy(c.pstream);
    (void)hipStreamDestroy(c.own_stream);
    delete ctx;
    return code;
  };
  if(hipStreamCreateWithFlags(&c.ustream, hipStreamNonBlocking) != hipSuccess ||
     hipStreamCreateWithFlags(&c.pstream, hipStreamNonBlocking) != hipSuccess)
    return undo(HPLX_ERR_HIP);
  if(hipMalloc(&c.small_base, total) != hipSuccess) return undo(HPLX_ERR_ALLOC);
  char* b   = static_cast<char*>(c.small_base);
  c.Linv    = reinterpret_cast<double*>(b);
  c.permU   = reinterpret_cast<int*>(b + off_perm);
  c.slots   = reinterpret_cast<double*>(b + off_slots);
  c.flags   = reinterpret_cast<int*>(b + off_flags);
  c.dout    = reinterpret_cast<double*>(b + off_dout);
  c.Linv2   = reinterpret_cast<double*>(b + off_linv2);
  c.permU2  = reinterpret_cast<int*>(b + off_perm2);
  c.claim   = reinterpret_cast<int*>(b + off_claim);
  c.stop    = reinterpret_cast<int*>(b + off_stop);
  c.vacate  = reinterpret_cast<unsigned long long*>(b + off_vacate);
  int* census = reinterpret_cast<int*>(b + off_census);
  c.claim_log = reinterpret_cast<int*>(b + off_clog);
  c.pops      = reinterpret_cast<hplx::PanelOp*>(b + off_pops);
  c.pbar      = reinterpret_cast<unsigned*>(b + off_pbar);
  if(hipMemset(c.small_base, 0, total) != hipSuccess ||
     hipDeviceSynchronize() != hipSuccess)
    return undo(HPLX_ERR_HIP);
  if(const char* la = std::getenv("HPLX_LOOKAHEAD")) c.lookahead = std::atoi(la) != 0;
  if(const char* pk = std::getenv("HPLX_PANEL_KERNEL")) {
    const int v = std::atoi(pk);
    c.panel_kernel = v >= 0 && v <= 2 ? v : c.panel_kernel;
  }
  if(cu_sets(c, census) != HPLX_SUCCESS) return undo(HPLX_ERR_HIP);
  *out = ctx;
  return HPLX_SUCCESS;
}

int hplx_ctx_destroy(hplx_ctx* ctx) {
  if(!ctx) return HPLX_SUCCESS;
  Ctx& c = ctx->c;
  (void)hipSetDevice(c.device);
  (void)hipStreamSynchronize(c.stream);
  if(c.ustream) (void)hipStreamSynchronize(c.ustream);
  if(c.pstream) (void)hipStreamSynchronize(c.pstream);
  for(hipEvent_t e : c.ev) (void)hipEventDestroy(e);
  for(hipEvent_t e : c.ev_prog) (void)hipEventDestroy(e);
  for(hipEvent_t e : c.ev_sync) (void)hipEventDestroy(e);
  for(hipEvent_t e : c.ev_swap) (void)hipEventDestroy(e);
  for(hipEvent_t e : c.ev_panel) (void)hipEventDestroy(e);
  if(c.small_base) (void)hipFree(c.small_base);
  if(c.ustream) (void)hipStreamDestroy(c.ustream);
  if(c.pstream) (void)hipStreamDestroy(c.pstream);
  if(c.own_stream) (void)hipStreamDestroy(c.own_stream);
  delete ctx;
  return HPLX_SUCCESS;
}

const char* hplx_last_error(const hplx_ctx* ctx) {
  return ctx ? ctx->c.last_error.c_str() : "null context";
}

int hplx_set_stream(hplx_ctx* ctx, void* stream) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  HIPCHK(c, hipStreamSynchronize(c.stream));
  c.stream = stream ? static_cast<hipStream_t>(stream) : c.own_stream;
  return HPLX_SUCCESS;
}

int hplx_synchronize(hplx_ctx* ctx) {
  int r = use(ctx);
  if(r) return r;
  HIPCHK(ctx->c, hipStreamSynchronize(ctx->c.stream));
  return HPLX_SUCCESS;
}

int hplx_num_cus(hplx_ctx* ctx) { return ctx ? ctx->c.num_cus : 0; }

int hplx_set_option(hplx_ctx* ctx, int option, int value) {
  int r = use(ctx);
  if(r) return r;
  if((option != HPLX_OPT_LOOKAHEAD && option != HPLX_OPT_PANEL_KERNEL) ||
     value < 0 || value > (option == HPLX_OPT_PANEL_KERNEL ? 2 : 1))
    return fail(ctx->c, HPLX_ERR_INVALID, "set_option: option %d value %d",
                option, value);
  (option == HPLX_OPT_LOOKAHEAD ? ctx->c.lookahead : ctx->c.panel_kernel) = value;
  return HPLX_SUCCESS;
}

int hplx_get_option(hplx_ctx* ctx, int option, int* value) {
  if(!ctx || !value) return HPLX_ERR_INVALID;
  if(option != HPLX_OPT_LOOKAHEAD && option != HPLX_OPT_PANEL_KERNEL)
    return fail(ctx->c, HPLX_ERR_INVALID, "get_option: option %d", option);
  *value = option == HPLX_OPT_LOOKAHEAD ? ctx->c.lookahead : ctx->c.panel_kernel;
  return HPLX_SUCCESS;
}

int hplx_panel_steps(const hplx_palg* palg, int jb, int* steps,
                     int max_steps) {
  if(!palg) return HPLX_ERR_INVALID;
  return hplx::panel_steps(hplx::panel_algo(*palg), jb, steps, max_steps);
}

int hplx_panel_kernel_runs(hplx_ctx* ctx) {
  if(!ctx) return HPLX_ERR_INVALID;
  return ctx->c.panel_kernel_runs;
}

int hplx_malloc(hplx_ctx* ctx, size_t bytes, void** dptr) {
  int r = use(ctx);
  if(r) return r;
  if(!dptr) return HPLX_ERR_INVALID;
  if(hipMalloc(dptr, bytes ? bytes : 1) != hipSuccess)
    return fail(ctx->c, HPLX_ERR_ALLOC, "hipMalloc(%zu) failed", bytes);
  return HPLX_SUCCESS;
}

int hplx_free(hplx_ctx* ctx, void* dptr) {
  int r = use(ctx);
  if(r) return r;
  HIPCHK(ctx->c, hipStreamSynchronize(ctx->c.stream));
  if(dptr) HIPCHK(ctx->c, hipFree(dptr));
  return HPLX_SUCCESS;
}

int hplx_memcpy_h2d(hplx_ctx* ctx, void* dst, const void* src, size_t bytes) {
  int r = use(ctx);
  if(r) return r;
  HIPCHK(ctx->c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice,
                                ctx->c.stream));
  HIPCHK(ctx->c, hipStreamSynchronize(ctx->c.stream));
  return HPLX_SUCCESS;
}

int hplx_memcpy_d2h(hplx_ctx* ctx, void* dst, const void* src, size_t bytes) {
  int r = use(ctx);
  if(r) return r;
  HIPCHK(ctx->c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost,
                                ctx->c.stream));
  HIPCHK(ctx->c, hipStreamSynchronize(ctx->c.stream));
  return HPLX_SUCCESS;
}

int hplx_pmat_alloc(hplx_ctx* ctx, int n, int nb, hplx_pmat* mat) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  if(!mat || n < 1 || nb < 1)
    return fail(c, HPLX_ERR_INVALID, "pmat_alloc: n=%d nb=%d", n, nb);
  if(nb > hplx::kMaxNB)
    return fail(c, HPLX_ERR_UNSUPPORTED, "NB=%d exceeds %d", nb, hplx::kMaxNB);
  std::memset(mat, 0, sizeof(*mat));
  mat->n  = n;
  mat->nb = nb;
  mat->ld = hpl_ld(n);
  mat->mp = n;
  mat->nq = n + 1;
  const size_t nA = static_cast<size_t>(mat->ld) * (n + 1);
  hplx::MatWs* ws = new(std::nothrow) hplx::MatWs();
  if(!ws) return HPLX_ERR_ALLOC;
  const size_t wl = static_cast<size_t>(nb) * (n + 1);
  const size_t cl = hplx::check_part_len(n);
  ws->W_len       = wl > cl ? wl : cl;
  if(hipMalloc(&mat->A, nA * sizeof(double)) != hipSuccess) {
    delete ws;
    return fail(c, HPLX_ERR_ALLOC, "hipMalloc of A|b (%zu bytes) failed",
                nA * sizeof(double));
  }
  if(hipMalloc(&mat->X, static_cast<size_t>(n) * sizeof(double)) !=
         hipSuccess ||
     hipMalloc(&mat->ipiv, static_cast<size_t>(n) * sizeof(int)) !=
         hipSuccess ||
     hipMalloc(&ws->W, ws->W_len * sizeof(double)) != hipSuccess) {
    if(mat->X) (void)hipFree(mat->X);
    if(mat->ipiv) (void)hipFree(mat->ipiv);
    (void)hipFree(mat->A);
    delete ws;
    std::memset(mat, 0, sizeof(*mat));
    return fail(c, HPLX_ERR_ALLOC, "hipMalloc of workspace failed");
  }
  mat->ws = ws;
  return HPLX_SUCCESS;
}

int hplx_pmat_free(hplx_ctx* ctx, hplx_pmat* mat) {
  int r = use(ctx);
  if(r) return r;
  if(!mat) return HPLX_ERR_INVALID;
  (void)hipStreamSynchronize(ctx->c.stream);
  hplx::MatWs* ws = static_cast<hplx::MatWs*>(mat->ws);
  if(ws) {
    if(ws->W) (void)hipFree(ws->W);
    delete ws;
  }
  if(mat->ipiv) (void)hipFree(mat->ipiv);
  if(mat->X) (void)hipFree(mat->X);
  if(mat->A) (void)hipFree(mat->A);
  std::memset(mat, 0, sizeof(*mat));
  return HPLX_SUCCESS;
}

int hplx_pdrandmat(hplx_ctx* ctx, int m, int n, int nb, double* A, int lda,
                   int seed, double diag_shift) {
  int r = use(ctx);
  if(r) return r;
  (void)nb;
  if(m < 0 || n < 0 || (m > 0 && (!A || lda < m)) || seed < 0)
    return fail(ctx->c, HPLX_ERR_INVALID, "pdrandmat: m=%d n=%d lda=%d", m, n,
                lda);
  HIPCHK(ctx->c, hplx::launch_matgen(ctx->c.stream, m, n, A, lda, seed,
                                     diag_shift));
  return HPLX_SUCCESS;
}

// Steps of the panel loop on the FP64 matrix (shared by the serial and the
// look-ahead schedules).
struct Fp64Loop {
  Ctx&         c;
  hplx_pmat*   mat;
  hplx::MatWs* ws;
  double*      A;
  int          N, nb, lda;
  hplx::PanelAlgo algo;

  double* Linv(int k) const { return (k & 1) ? c.Linv2 : c.Linv; }
  int*    permU(int k) const { return (k & 1) ? c.permU2 : c.permU; }
  int     width(int k) const { return (N - k * nb) < nb ? (N - k * nb) : nb; }

  // Panel k factored on stream s (its wide products use L11^{-1} buffer k as
  // scratch), then its composed pivots and L11^{-1} built into buffers k.
  int panel(int k, hipStream_t s) {
    const int j = k * nb, jb = width(k);
    double*   P = A + j + static_cast<size_t>(j) * lda;
    c.panel_stream_  = s;
    c.panel_scratch_ = Linv(k);
    // the factorisation, timed by its own event pair
    while(static_cast<int>(c.ev_panel.size()) < 2 * (c.panel_pairs + 1)) {
      hipEvent_t ev;
      HIPCHK(c, hipEventCreate(&ev));
      c.ev_panel.push_back(ev);
    }
    HIPCHK(c, hipEventRecord(c.ev_panel[2 * c.panel_pairs], s));
    const hipError_t e = hplx::panel_factor(c, algo, P, lda, N - j, jb,
                                            mat->ipiv + j, j);
    c.panel_stream_  = nullptr;
    c.panel_scratch_ = nullptr;
    HIPCHK(c, e);
    HIPCHK(c, hipEventRecord(c.ev_panel[2 * c.panel_pairs + 1], s));
    ++c.panel_pairs;
    c.panel_bytes += 16.0 * (N - j) * static_cast<double>(jb);
    HIPCHK(c, hplx::launch_pipid(s, jb, mat->ipiv + j, j, permU(k)));
    HIPCHK(c, hplx::launch_linv(s, jb, P, lda, Linv(k), nb));
    return HPLX_SUCCESS;
  }

  // Columns [c0, c0 + n) right of panel k (c0 relative to the first column
  // after the panel): row swaps into W, U12 = L11^{-1} W over A12, and (when
  // the rows below exist and the streamed update takes the shape) U12^T
  // into W.  Returns in *fast whether the streamed update applies.
  int prepare(int k, int c0, int n, hipStream_t s, bool* fast) {
    const int j = k * nb, jb = width(k), m = N - j;
    double*   P   = A + j + static_cast<size_t>(j) * lda;
    double*   A12 = A + j + static_cast<size_t>(j + jb + c0) * lda;
    // the interchanges, timed by their own event pair
    while(static_cast<int>(c.ev_swap.size()) < 2 * (c.swap_pairs + 1)) {
      hipEvent_t e;
      HIPCHK(c, hipEventCreate(&e));
      c.ev_swap.push_back(e);
    }
    HIPCHK(c, hipEventRecord(c.ev_swap[2 * c.swap_pairs], s));
    HIPCHK(c, hplx::launch_swap_gather(s, jb, n, A12, lda, permU(k), ws->W, nb));
    HIPCHK(c, hipEventRecord(c.ev_swap[2 * c.swap_pairs + 1], s));
    ++c.swap_pairs;
    c.swap_bytes += 4.0 * 8.0 * jb * static_cast<double>(n);
    HIPCHK(c, hplx::launch_dgemm(s, jb, n, jb, 1.0, Linv(k), nb, ws->W, nb,
                                 0.0, A12, lda, true));
    *fast = false;
    if(m - jb > 0) {
      const int ldu = hplx::ut_ld(n);
      *fast = hplx::update_fast_ok(m - jb, n, jb, P + jb, lda, ws->W, ldu,
                                   lda - j - jb, ldu);
      if(*fast) HIPCHK(c, hplx::launch_transpose(s, jb, n, A12, lda, ws->W, ldu));
    }
    return HPLX_SUCCESS;
  }

  // A22 -= L21 U12 on those columns in one launch, timed by an event pair.
  int gemm(int k, int c0, int n, hipStream_t s, bool fast) {
    const int j = k * nb, jb = width(k), m = N - j;
    if(m - jb <= 0) return HPLX_SUCCESS;
    double* P   = A + j + static_cast<size_t>(j) * lda;
    double* A12 = A + j + static_cast<size_t>(j + jb + c0) * lda;
    hipEvent_t e0 = c.ev[2 * c.ev_pairs], e1 = c.ev[2 * c.ev_pairs + 1];
    HIPCHK(c, hipEventRecord(e0, s));
    if(fast)
      HIPCHK(c, hplx::launch_update(s, m - jb, n, jb, P + jb, lda, ws->W,
                                    hplx::ut_ld(n), A12 + jb, lda,
                                    lda - j - jb, hplx::ut_ld(n)));
    else
      HIPCHK(c, hplx::launch_dgemm(s, m - jb, n, jb, -1.0, P + jb, lda, A12,
                                   lda, 1.0, A12 + jb, lda, false, 1));
    HIPCHK(c, hipEventRecord(e1, s));
    c.ev_kind[c.ev_pairs]  = 1;
    c.ev_flops[c.ev_pairs] = 2.0 * (m - jb) * static_cast<double>(n) * jb;
    c.gemm_flops += c.ev_flops[c.ev_pairs];
    c.ev_pairs++;
    return HPLX_SUCCESS;
  }
};

// Event from the context's pool (timing disabled), for cross-stream order.
using hplx::sync_event;
using hplx::kMinOverlapCols;

// The panel loop over the first `ncols` columns of mat->A (n+1 with b for
// hplx_pdgesv, n for hplx_pdgetrf).  Returns 0, an error code, or the
// 1-based column of the first exact zero pivot.
//
// Serial schedule (c.lookahead == 0): every step on c.stream in the order
// panel k, swaps, U12, update of all trailing columns.
//
// Look-ahead schedule (depth 1, src/pgesv/HPL_pdgesv.cpp:168-245 with
// nu0 = JB, src/panel/HPL_pdpanel_init.cpp:176-184): on the update stream U,
// step k first updates the columns of panel k+1 (swaps, U12, GEMM), then
// prepares the remaining columns (swaps, U12, U12^T) and queues their update
// as a persistent launch that leaves the CUs of a vacate set R free; the
// panel stream P factors panel k+1 on R meanwhile (after the preparation,
// so that nothing else is queued beside its co-resident leaf grids), sets
// the launch's stop flag and records an event; U then finishes the update's
// remaining tiles on every CU.  Each matrix element sees the same
// operations in the same order as in the serial schedule, so pivots and
// results are bit-identical.  Steps whose next panel needs more CUs than a
// vacate set holds, or whose remaining columns are few, run serially.
static int factor_loop(Ctx& c, const hplx_palg* palg, hplx_pmat* mat,
                       int ncols) {
  int r;
  if((r = hplx::check_palg(c, palg))) return r;
  if(!mat || !mat->A || !mat->ws) return fail(c, HPLX_ERR_INVALID, "bad pmat");
  const int N   = mat->n;
  const int nb  = mat->nb;
  const int lda = mat->ld;
  if(N > hplx::max_leaf_rows(c))
    return fail(c, HPLX_ERR_UNSUPPORTED, "N=%d exceeds the panel kernel's %d rows",
                N, hplx::max_leaf_rows(c));
  Fp64Loop L{c, mat, static_cast<hplx::MatWs*>(mat->ws), mat->A, N, nb, lda,
             hplx::panel_algo(*palg)};
  const int npanels = (N + nb - 1) / nb;
  while(static_cast<int>(c.ev.size()) < 6 * npanels) {
    hipEvent_t e;
    HIPCHK(c, hipEventCreate(&e));
    c.ev.push_back(e);
  }
  c.ev_pairs   = 0;
  c.gemm_flops = 0.0;
  c.swap_pairs = 0;
  c.swap_bytes = 0.0;
  c.panel_pairs = 0;
  c.panel_bytes = 0.0;
  c.ev_kind.assign(3 * npanels, 0);
  c.ev_flops.assign(3 * npanels, 0.0);
  c.splits.clear();
  c.claim_host.clear();
  c.panel_kernel_runs = 0;
  if(c.panel_kernel) {  // the recorded steps of both panel widths, up front
    int pr = hplx::panel_kernel_prepare(c, L.algo, N < nb ? N : nb);
    if(!pr && N > nb && N % nb) pr = hplx::panel_kernel_prepare(c, L.algo, N % nb);
    if(pr) return fail(c, pr, "persistent panel kernel: recording failed");
  }
  hplx::Progress prog(c, npanels, nb, N);
  HIPCHK(c, prog.init());

  const bool  la = c.lookahead && c.cus_per_xcd > 0;
  hipStream_t U  = la ? c.ustream : c.stream;
  hipStream_t P  = c.pstream;
  size_t      nev = 0;
  hipEvent_t  ev;
  if(la) {  // U and P start after the caller's stream
    if((r = sync_event(c, nev++, &ev))) return r;
    HIPCHK(c, hipEventRecord(ev, c.stream));
    HIPCHK(c, hipStreamWaitEvent(U, ev, 0));
    HIPCHK(c, hipStreamWaitEvent(P, ev, 0));
  }
  c.panel_stream_ = U;
  const hipError_t er = hplx::reset_panel_state(c);
  c.panel_stream_ = nullptr;
  HIPCHK(c, er);
  if((r = L.panel(0, U))) return r;
  for(int k = 0; k < npanels; ++k) {
    const int j = k * nb, jb = L.width(k);
    const int nt = ncols - j - jb;  // trailing columns (with b)
    const int jb1 = k + 1 < npanels ? L.width(k + 1) : 0;
    bool      fast;
    if(nt > 0) {
      // columns of panel k+1 first (all of them if there is no next panel)
      const int la_cols = jb1 > 0 ? jb1 : nt;
      const int rest    = nt - la_cols;
      if((r = L.prepare(k, 0, la_cols, U, &fast)) ||
         (r = L.gemm(k, 0, la_cols, U, fast)))
        return r;
      // the CUs the next panel's leaves need
      const int rR = hplx::vacate_set_size(N - j - jb);
      const bool overlap = la && jb1 > 0 && rest >= kMinOverlapCols &&
                           rR <= c.cus_per_xcd && rR < 28;
      if(overlap) {
        if((r = L.prepare(k, la_cols, rest, U, &fast))) return r;
        int* claim = c.claim + (k & 1) * hplx::kClaimInts;
        int* stop  = c.stop + (k & 1) * 32;
        HIPCHK(c, hplx::launch_zero_ints(U, claim, hplx::kClaimInts));
        HIPCHK(c, hplx::launch_zero_ints(U, stop, 1));
        hipEvent_t ev_prep, ev_pan;
        if((r = sync_event(c, nev++, &ev_prep)) ||
           (r = sync_event(c, nev++, &ev_pan)))
          return r;
        HIPCHK(c, hipEventRecord(ev_prep, U));
        // panel k+1 on P, on the CUs the update leaves free
        HIPCHK(c, hipStreamWaitEvent(P, ev_prep, 0));
        c.panel_cus = 8 * rR;
        r           = L.panel(k + 1, P);
        c.panel_cus = 0;
        if(r) return r;
        HIPCHK(c, hplx::launch_set_flag(P, stop));
        HIPCHK(c, hipEventRecord(ev_pan, P));
        const int m = N - j;
        if(m - jb > 0) {
          double* Pk  = L.A + j + static_cast<size_t>(j) * lda;
          double* A12 = L.A + j + static_cast<size_t>(j + jb + la_cols) * lda;
          const int    ldu   = hplx::ut_ld(rest);
          const double flops = 2.0 * (m - jb) * static_cast<double>(rest) * jb;
          auto pair = [&](int kind) {
            c.ev_kind[c.ev_pairs] = kind;
            return c.ev_pairs++;
          };
          if(fast) {
            // the claimed part beside the panel, then the resume part; each
            // timed by its own event pair
            const int pc = pair(2);
            HIPCHK(c, hipEventRecord(c.ev[2 * pc], U));
            HIPCHK(c, hplx::launch_update_claimed(
                          U, m - jb, rest, jb, Pk + jb, lda, L.ws->W, ldu,
                          A12 + jb, lda, lda - j - jb, ldu, claim,
                          c.vacate + 8 * (rR - 1), stop, 2 * c.num_cus));
            HIPCHK(c, hipEventRecord(c.ev[2 * pc + 1], U));
            const int ns = static_cast<int>(c.splits.size());
            if(ns < hplx::kMaxClaimLog)
              HIPCHK(c, hplx::launch_claim_log(U, claim, m - jb, rest,
                                               c.claim_log + 8 * ns));
            HIPCHK(c, hipStreamWaitEvent(U, ev_pan, 0));
            const int pr = pair(3);
            HIPCHK(c, hipEventRecord(c.ev[2 * pr], U));
            HIPCHK(c, hplx::launch_update_resume(
                          U, m - jb, rest, jb, Pk + jb, lda, L.ws->W, ldu,
                          A12 + jb, lda, lda - j - jb, ldu, claim));
            HIPCHK(c, hipEventRecord(c.ev[2 * pr + 1], U));
            c.splits.push_back({m - jb, rest, jb, pc, pr});
          } else {
            HIPCHK(c, hipStreamWaitEvent(U, ev_pan, 0));
            const int p1 = pair(1);
            HIPCHK(c, hipEventRecord(c.ev[2 * p1], U));
            HIPCHK(c, hplx::launch_dgemm(U, m - jb, rest, jb, -1.0, Pk + jb,
                                         lda, A12, lda, 1.0, A12 + jb, lda,
                                         false, 1));
            HIPCHK(c, hipEventRecord(c.ev[2 * p1 + 1], U));
            c.ev_flops[p1] = flops;
          }
          c.gemm_flops += flops;
        } else {
          HIPCHK(c, hipStreamWaitEvent(U, ev_pan, 0));
        }
      } else {
        if(jb1 > 0 && (r = L.panel(k + 1, U))) return r;
        if(rest > 0 && ((r = L.prepare(k, la_cols, rest, U, &fast)) ||
                        (r = L.gemm(k, la_cols, rest, U, fast))))
          return r;
      }
    } else if(jb1 > 0) {
      if((r = L.panel(k + 1, U))) return r;
    }
    HIPCHK(c, prog.panel_queued(k, U));
  }
  if(la) {  // the caller's stream continues after U and P
    hipEvent_t eu, ep;
    if((r = sync_event(c, nev++, &eu)) || (r = sync_event(c, nev++, &ep))) return r;
    HIPCHK(c, hipEventRecord(eu, U));
    HIPCHK(c, hipEventRecord(ep, P));
    HIPCHK(c, hipStreamWaitEvent(c.stream, eu, 0));
    HIPCHK(c, hipStreamWaitEvent(c.stream, ep, 0));
  }
  // zero pivot or barrier failure?  (one synchronisation; the claim log of
  // the split updates comes back with it)
  int flags[2] = {0, 0};
  HIPCHK(c, hipMemcpyAsync(flags, c.flags, sizeof(flags), hipMemcpyDeviceToHost,
                           c.stream));
  const size_t nlog = std::min<size_t>(c.splits.size(), hplx::kMaxClaimLog);
  c.claim_host.assign(8 * nlog, 0);
  if(nlog)
    HIPCHK(c, hipMemcpyAsync(c.claim_host.data(), c.claim_log,
                             sizeof(int) * 8 * nlog, hipMemcpyDeviceToHost,
                             c.stream));
  HIPCHK(c, hipStreamSynchronize(c.stream));
  if(flags[1])
    return fail(c, HPLX_ERR_HIP,
                "panel kernel wait timed out (workgroups not co-resident)");
  mat->info = flags[0];
  return flags[0];
}

int hplx_pdgesv(hplx_ctx* ctx, const hplx_palg* palg, hplx_pmat* mat) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  r      = factor_loop(c, palg, mat, mat ? mat->n + 1 : 0);
  if(r) return r;
  const int N = mat->n;
  hipStream_t s = c.stream;
  HIPCHK(c, hipMemcpyAsync(mat->X, mat->A + static_cast<size_t>(N) * mat->ld,
                           static_cast<size_t>(N) * sizeof(double),
                           hipMemcpyDeviceToDevice, s));
  HIPCHK(c, hplx::launch_trsv(s, N, mat->nb, mat->A, mat->ld, mat->X));
  HIPCHK(c, hipStreamSynchronize(s));
  return HPLX_SUCCESS;
}

int hplx_pdgetrf(hplx_ctx* ctx, const hplx_palg* palg, hplx_pmat* mat) {
  int r = use(ctx);
  if(r) return r;
  return factor_loop(ctx->c, palg, mat, mat ? mat->n : 0);
}

int hplx_pdgetrs(hplx_ctx* ctx, const hplx_pmat* mat, int trans, int nrhs,
                 double* B, int ldb) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  if(!mat || !mat->A || nrhs < 0 || (nrhs > 0 && (!B || ldb < mat->n)) ||
     (trans != 0 && trans != 1))
    return fail(c, HPLX_ERR_INVALID, "pdgetrs: bad arguments");
  if(nrhs == 0) return HPLX_SUCCESS;
  HIPCHK(c, hplx::getrs_enqueue(c.stream, mat->A, mat->n, mat->nb, mat->ld,
                                mat->ipiv, trans, nrhs, B, ldb, c.permU, c.Linv,
                                static_cast<size_t>(hplx::kMaxNB) * hplx::kMaxNB));
  HIPCHK(c, hipStreamSynchronize(c.stream));
  return HPLX_SUCCESS;
}

int hplx_pdgesv_adjoint(hplx_ctx* ctx, const hplx_pmat* mat, int nrhs,
                        const double* X, int ldx, const double* G, int ldg,
                        double* gradB, int ldgb, double* gradA, int ldga) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  if(!mat || !mat->A || nrhs < 1 || !X || !G || !gradB || ldx < mat->n ||
     ldg < mat->n || ldgb < mat->n || (gradA && ldga < mat->n))
    return fail(c, HPLX_ERR_INVALID, "pdgesv_adjoint: bad arguments");
  const int N = mat->n;
  HIPCHK(c, hipMemcpy2DAsync(gradB, static_cast<size_t>(ldgb) * sizeof(double),
                             G, static_cast<size_t>(ldg) * sizeof(double),
                             static_cast<size_t>(N) * sizeof(double), nrhs,
                             hipMemcpyDeviceToDevice, c.stream));
  if((r = hplx_pdgetrs(ctx, mat, 1, nrhs, gradB, ldgb))) return r;
  if(gradA) {
    // grad_A = -grad_B X^T: N x N from rank nrhs (HBM-bound store)
    HIPCHK(c, hplx::launch_outer_neg(c.stream, nrhs, N, N, gradB, ldgb, X,
                                     ldx, gradA, ldga));
    HIPCHK(c, hipStreamSynchronize(c.stream));
  }
  return HPLX_SUCCESS;
}

int hplx_pdtest_check(hplx_ctx* ctx, const hplx_pmat* mat, int seed,
                      double diag_shift, double* out) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  if(!mat || !mat->X || !mat->ws || !out)
    return fail(c, HPLX_ERR_INVALID, "pdtest_check: bad arguments");
  hplx::MatWs* ws = static_cast<hplx::MatWs*>(mat->ws);
  HIPCHK(c, hplx::launch_check(c.stream, mat->n, seed, diag_shift, mat->X,
                               ws->W, ws->W_len, c.dout));
  HIPCHK(c, hipMemcpyAsync(out, c.dout, 9 * sizeof(double),
                           hipMemcpyDeviceToHost, c.stream));
  HIPCHK(c, hipStreamSynchronize(c.stream));
  return HPLX_SUCCESS;
}

static int pdfact_algo(hplx_ctx* ctx, const hplx::PanelAlgo& algo, double* P,
                       int lda, int m, int jb, int* ipiv, int row_offset) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  if(!P || !ipiv || m < 1 || jb < 1 || jb > m || lda < m)
    return fail(c, HPLX_ERR_INVALID, "pdfact: m=%d jb=%d lda=%d", m, jb, lda);
  if(jb > hplx::kMaxNB || m > hplx::max_leaf_rows(c))
    return fail(c, HPLX_ERR_UNSUPPORTED, "pdfact: m=%d jb=%d too large", m, jb);
  c.panel_kernel_runs = 0;
  if(c.panel_kernel && (r = hplx::panel_kernel_prepare(c, algo, jb)))
    return fail(c, r, "persistent panel kernel: recording failed");
  HIPCHK(c, hplx::reset_panel_state(c));
  HIPCHK(c, hplx::panel_factor(c, algo, P, lda, m, jb, ipiv, row_offset));
  int flags[2] = {0, 0};
  HIPCHK(c, hipMemcpyAsync(flags, c.flags, sizeof(flags), hipMemcpyDeviceToHost,
                           c.stream));
  HIPCHK(c, hipStreamSynchronize(c.stream));
  if(flags[1])
    return fail(c, HPLX_ERR_HIP, "panel kernel grid barrier timed out");
  return flags[0] ? flags[0] - row_offset : HPLX_SUCCESS;
}

int hplx_pdfact(hplx_ctx* ctx, int nbmin, int nbdiv, double* P, int lda, int m,
                int jb, int* ipiv, int row_offset) {
  return pdfact_algo(ctx, hplx::panel_algo(nbmin, nbdiv), P, lda, m, jb, ipiv,
                     row_offset);
}

int hplx_pdfact_palg(hplx_ctx* ctx, const hplx_palg* palg, double* P, int lda,
                     int m, int jb, int* ipiv, int row_offset) {
  if(!ctx) return HPLX_ERR_INVALID;
  if(!palg) return fail(ctx->c, HPLX_ERR_INVALID, "pdfact_palg: null palg");
  if(!hplx::valid_fact(palg->pfact) || !hplx::valid_fact(palg->rfact))
    return fail(ctx->c, HPLX_ERR_INVALID, "pdfact_palg: pfact=%d rfact=%d",
                palg->pfact, palg->rfact);
  return pdfact_algo(ctx, hplx::panel_algo(*palg), P, lda, m, jb, ipiv,
                     row_offset);
}

int hplx_pipid(hplx_ctx* ctx, int jb, const int* ipiv, int* permU) {
  int r = use(ctx);
  if(r) return r;
  if(jb < 0 || jb > hplx::kMaxNB || (jb > 0 && (!ipiv || !permU)))
    return fail(ctx->c, HPLX_ERR_INVALID, "pipid: jb=%d", jb);
  HIPCHK(ctx->c, hplx::launch_pipid(ctx->c.stream, jb, ipiv, 0, permU));
  HIPCHK(ctx->c, hipStreamSynchronize(ctx->c.stream));
  return HPLX_SUCCESS;
}

int hplx_dlaswp00N(hplx_ctx* ctx, int jb, int n, double* A, int lda,
                   const int* permU) {
  int r = use(ctx);
  if(r) return r;
  if(jb < 0 || jb > hplx::kMaxNB || n < 0 || lda < jb)
    return fail(ctx->c, HPLX_ERR_INVALID, "dlaswp00N: jb=%d n=%d", jb, n);
  HIPCHK(ctx->c, hplx::launch_swap_inplace(ctx->c.stream, jb, n, A, lda, permU));
  HIPCHK(ctx->c, hipStreamSynchronize(ctx->c.stream));
  return HPLX_SUCCESS;
}

int hplx_pdupdate(hplx_ctx* ctx, double* A, int lda, int m, int jb, int n,
                  const int* permU, double* W) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  if(!A || !permU || !W || jb < 1 || jb > hplx::kMaxNB || m < jb || n < 0 ||
     lda < m)
    return fail(c, HPLX_ERR_INVALID, "pdupdate: m=%d jb=%d n=%d", m, jb, n);
  if(n == 0) return HPLX_SUCCESS;
  hipStream_t s   = c.stream;
  double*     A12 = A + static_cast<size_t>(jb) * lda;
  HIPCHK(c, hplx::launch_linv(s, jb, A, lda, c.Linv, jb));
  HIPCHK(c, hplx::launch_swap_gather(s, jb, n, A12, lda, permU, W, jb));
  HIPCHK(c, hplx::launch_dgemm(s, jb, n, jb, 1.0, c.Linv, jb, W, jb, 0.0, A12,
                               lda, true));
  if(m > jb) {
    // the same two paths as the panel loop: U12^T into W and the streamed
    // update, or the general GEMM
    const int ldu = hplx::ut_ld(n);
    if(hplx::update_fast_ok(m - jb, n, jb, A + jb, lda, W, ldu, lda - jb, ldu)) {
      HIPCHK(c, hplx::launch_transpose(s, jb, n, A12, lda, W, ldu));
      HIPCHK(c, hplx::launch_update(s, m - jb, n, jb, A + jb, lda, W, ldu,
                                    A12 + jb, lda, lda - jb, ldu));
    } else {
      HIPCHK(c, hplx::launch_dgemm(s, m - jb, n, jb, -1.0, A + jb, lda, A12,
                                   lda, 1.0, A12 + jb, lda, false));
    }
  }
  HIPCHK(c, hipStreamSynchronize(s));
  return HPLX_SUCCESS;
}

int hplx_dgemm(hplx_ctx* ctx, int m, int n, int k, double alpha,
               const double* A, int lda, const double* B, int ldb, double beta,
               double* C, int ldc) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  if(m < 0 || n < 0 || k < 0 || lda < (m > 1 ? m : 1) ||
     ldb < (k > 1 ? k : 1) || ldc < (m > 1 ? m : 1))
    return fail(c, HPLX_ERR_INVALID, "dgemm: bad sizes");
  if(!(alpha == 1.0 || alpha == -1.0) || !(beta == 0.0 || beta == 1.0))
    return fail(c, HPLX_ERR_UNSUPPORTED, "dgemm: alpha must be +-1, beta 0 or 1");
  HIPCHK(c, hplx::launch_dgemm(c.stream, m, n, k, alpha, A, lda, B, ldb, beta,
                               C, ldc, false));
  HIPCHK(c, hipStreamSynchronize(c.stream));
  return HPLX_SUCCESS;
}

int hplx_pdtrsv(hplx_ctx* ctx, int n, int nb, const double* A, int lda,
                double* x) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  if(n < 0 || nb < 1 || nb > hplx::kMaxNB || lda < n || (n > 0 && (!A || !x)))
    return fail(c, HPLX_ERR_INVALID, "pdtrsv: n=%d nb=%d", n, nb);
  if(n == 0) return HPLX_SUCCESS;
  HIPCHK(c, hipMemcpyAsync(x, A + static_cast<size_t>(n) * lda,
                           static_cast<size_t>(n) * sizeof(double),
                           hipMemcpyDeviceToDevice, c.stream));
  HIPCHK(c, hplx::launch_trsv(c.stream, n, nb, A, lda, x));
  HIPCHK(c, hipStreamSynchronize(c.stream));
  return HPLX_SUCCESS;
}

int hplx_update_timing(hplx_ctx* ctx, int kind, double* ms, int* launches,
                       double* flops) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  if(kind < 0 || kind > 3)
    return fail(c, HPLX_ERR_INVALID, "update_timing: kind %d", kind);
  HIPCHK(c, hipStreamSynchronize(c.stream));
  // flops of the split updates' two parts, from the claim log (a split
  // beyond the log counts wholly as resumed)
  std::vector<double> fl(c.ev_flops.begin(), c.ev_flops.end());
  for(size_t i = 0; i < c.splits.size(); ++i) {
    const auto&  sp    = c.splits[i];
    const double total = 2.0 * sp.m * static_cast<double>(sp.n) * sp.k;
    const double took  = 8 * i < c.claim_host.size()
                             ? hplx::claimed_flops(sp.m, sp.n, sp.k,
                                                   &c.claim_host[8 * i], sp.bm,
                                                   sp.bn)
                             : 0.0;
    fl[sp.pair_claimed] = took;
    fl[sp.pair_resume]  = total - took;
  }
  double tot = 0.0, f = 0.0;
  int    cnt = 0;
  for(int i = 0; i < c.ev_pairs; ++i) {
    if(kind && c.ev_kind[i] != kind) continue;
    float t = 0.f;
    HIPCHK(c, hipEventElapsedTime(&t, c.ev[2 * i], c.ev[2 * i + 1]));
    tot += t;
    f += fl[i];
    ++cnt;
  }
  if(ms) *ms = tot;
  if(launches) *launches = cnt;
  if(flops) *flops = f;
  return HPLX_SUCCESS;
}

int hplx_swap_timing(hplx_ctx* ctx, double* ms, int* launches,
                     double* bytes) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  HIPCHK(c, hipStreamSynchronize(c.stream));
  double tot = 0.0;
  for(int i = 0; i < c.swap_pairs; ++i) {
    float t = 0.f;
    HIPCHK(c, hipEventElapsedTime(&t, c.ev_swap[2 * i], c.ev_swap[2 * i + 1]));
    tot += t;
  }
  if(ms) *ms = tot;
  if(launches) *launches = c.swap_pairs;
  if(bytes) *bytes = c.swap_bytes;
  return HPLX_SUCCESS;
}

int hplx_panel_timing(hplx_ctx* ctx, double* ms, int* panels,
                      double* bytes) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  HIPCHK(c, hipStreamSynchronize(c.stream));
  double tot = 0.0;
  for(int i = 0; i < c.panel_pairs; ++i) {
    float t = 0.f;
    HIPCHK(c, hipEventElapsedTime(&t, c.ev_panel[2 * i], c.ev_panel[2 * i + 1]));
    tot += t;
  }
  if(ms) *ms = tot;
  if(panels) *panels = c.panel_pairs;
  if(bytes) *bytes = c.panel_bytes;
  return HPLX_SUCCESS;
}

int hplx_update_stats(hplx_ctx* ctx, double* gemm_ms, int* launches,
                      double* gemm_flops) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  HIPCHK(c, hipStreamSynchronize(c.stream));
  double tot = 0.0;
  for(int i = 0; i < c.ev_pairs; ++i) {
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c.ev[2 * i], c.ev[2 * i + 1]));
    tot += ms;
  }
  if(gemm_ms) *gemm_ms = tot;
  if(launches) *launches = c.ev_pairs;
  if(gemm_flops) *gemm_flops = c.gemm_flops;
  return HPLX_SUCCESS;
}

}  // extern "C"
