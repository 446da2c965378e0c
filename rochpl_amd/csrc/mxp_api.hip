// C ABI of the mixed-precision factorisation with FP64 iterative refinement
// (include/hplx.h, "Mixed precision").
//
// The panel loop is the FP64 loop of hplx_api.hip (src/pgesv/HPL_pdgesv.cpp
// :168-245 at P=Q=1) with the factors in FP32 or FP16 storage S:
//   1. panel factorisation in FP32 (in place for FP32 storage; FP16 panels
//      are copied to an FP32 buffer, factored, and written back)
//   2. pivot composition on the device                        -> permU
//   3. L11^{-1} in FP32
//   4. row swaps of the trailing columns; new top rows -> W (FP32)
//   5. U12 = L11^{-1} W in FP32, stored back in S
//   6. A22 -= L21 U12 in BF16 / FP8 (or FP32) on MFMA, FP32 accumulation
// b is not carried through the loop.  After the loop the diagonal blocks of
// L and U are inverted in FP32 (FP64 when the caller asks for an FP64
// correction solve).  The solves run on the device in that precision with
// the stored factors: x = U^{-1} L^{-1} P b with b generated in FP64 from
// the seed and rounded to FP32, then refinement in FP64 -- r = b - A x (A
// regenerated in FP64 tile by tile) and the residual norm -- with the
// correction d = LU \ r solved in FP32 and x += d accumulated in FP64.
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "api_common.hpp"
#include "mxp.hpp"

namespace {

using hplx::Ctx;
using hplx::fail;
using hplx::use;

struct MxpWs {
  float*  P32  = nullptr;  // FP16 storage: FP32 panel copy, ldp x nb
  int     ldp  = 0;
  float*  W    = nullptr;  // gathered top rows, nb x n (ld nb)
  float*  U    = nullptr;  // FP16 storage: U12 in FP32, nb x n (ld nb)
  float*  linv = nullptr;  // nb x nb
  void*   pa   = nullptr;  // packed L21
  void*   pb   = nullptr;  // packed U12
  int*    ea   = nullptr;
  int*    eb   = nullptr;
  double* r    = nullptr;  // refinement residual / correction, n
  double* part = nullptr;  // ir_part_len(n)
  double* out  = nullptr;  // 16 doubles
  // each panel's row interchanges as net (dst, src) moves of the solve's
  // vector: pairs of panel p at swp[2 swp_off[p] .. 2 swp_off[p+1])
  int*             swp = nullptr;  // device, up to 2n pairs
  std::vector<int> swp_off;
  bool             swp_ready = false;
  // inverses of the diagonal blocks of L (unit lower) and U, block J at
  // J nb nb (leading dimension = the block's size), and a block-sized
  // vector: FP32 values, or FP64 with solve_fp64 (room for FP64 either way)
  void* dinv_l = nullptr;
  void* dinv_u = nullptr;
  void* ytmp   = nullptr;
  float* x32   = nullptr;  // the FP32 vector of an FP32 solve, n
  bool  solve_fp64 = false;  // precision of the solves (set by the factorisation)
};

int hpl_ld(int n) { return ((n > 0 ? n : 0) + 95) / 128 * 128 + 32; }

size_t elem_size(int storage) { return storage == HPLX_STORE_FP16 ? 2 : 4; }

int check_opts(Ctx& c, const hplx_mxp_opts* o, int storage) {
  if(!o) return fail(c, HPLX_ERR_INVALID, "null mxp opts");
  if(o->update == HPLX_UPDATE_BF16) {
    if(o->terms < 1 || o->terms > 3)
      return fail(c, HPLX_ERR_INVALID, "BF16 terms must be 1, 2 or 3 (got %d)",
                  o->terms);
  } else if(o->update == HPLX_UPDATE_FP32) {
    if(storage != HPLX_STORE_FP32)
      return fail(c, HPLX_ERR_UNSUPPORTED, "the FP32 update needs FP32 storage");
  } else if(o->update != HPLX_UPDATE_FP8) {
    return fail(c, HPLX_ERR_INVALID, "unknown update %d", o->update);
  }
  if(o->max_its < 0 || o->max_its > 64)
    return fail(c, HPLX_ERR_INVALID, "max_its %d outside 0..64", o->max_its);
  if(o->solve_fp64 != 0 && o->solve_fp64 != 1)
    return fail(c, HPLX_ERR_INVALID, "solve_fp64 %d (0 or 1)", o->solve_fp64);
  if(!(o->threshold > 0.0)) return fail(c, HPLX_ERR_INVALID, "threshold <= 0");
  return HPLX_SUCCESS;
}

int check_mat(Ctx& c, const hplx_mxp_pmat* m) {
  if(!m || !m->A || !m->ws || m->n < 1 || m->nb < 1 ||
     (m->storage != HPLX_STORE_FP32 && m->storage != HPLX_STORE_FP16))
    return fail(c, HPLX_ERR_INVALID, "bad mxp pmat");
  return HPLX_SUCCESS;
}

int build_swaps(Ctx& c, const hplx_mxp_pmat* mat);

// Inverses of the diagonal blocks of the stored factors (all final once the
// factorisation is done: later interchanges do not touch earlier columns).
template <typename S, typename V>
int build_inverses_t(Ctx& c, const hplx_mxp_pmat* mat) {
  MxpWs*    ws = static_cast<MxpWs*>(mat->ws);
  const int N = mat->n, nb = mat->nb, lda = mat->ld;
  const S*  A     = static_cast<const S*>(mat->A);
  const int nfull = N / nb, rest = N - nfull * nb;
  const size_t tail = static_cast<size_t>(nfull) * nb * nb;
  for(const bool lower : {true, false}) {
    V* inv = static_cast<V*>(lower ? ws->dinv_l : ws->dinv_u);
    HIPCHK(c, hplx::launch_diag_inverse(c.stream, lower, nfull, nb, A, lda, inv));
    if(rest > 0)
      HIPCHK(c, hplx::launch_diag_inverse(
                    c.stream, lower, 1, rest,
                    A + static_cast<size_t>(nfull) * nb * (1 + static_cast<size_t>(lda)),
                    lda, inv + tail));
  }
  return HPLX_SUCCESS;
}

template <typename S>
int build_inverses(Ctx& c, const hplx_mxp_pmat* mat) {
  return static_cast<MxpWs*>(mat->ws)->solve_fp64
             ? build_inverses_t<S, double>(c, mat)
             : build_inverses_t<S, float>(c, mat);
}

// Steps of the mixed-precision panel loop (shared by the serial and the
// look-ahead schedules).
template <typename S>
struct MxpLoop {
  static constexpr bool kF16 = sizeof(S) == 2;
  Ctx&                 c;
  const hplx_palg*     palg;
  const hplx_mxp_opts* o;
  hplx_mxp_pmat*       mat;
  MxpWs*               ws;
  S*                   A;
  int                  N, nb, lda;

  int width(int k) const { return (N - k * nb) < nb ? (N - k * nb) : nb; }
  // the FP32 panel of step k (FP16 storage: its staged copy)
  float* panel_ptr(int k) const {
    const int j = k * nb;
    if constexpr(kF16)
      return ws->P32;
    else
      return A + j + static_cast<size_t>(j) * lda;
  }
  int panel_ld() const {
    if constexpr(kF16)
      return ws->ldp;
    else
      return lda;
  }

  // Panel k factored in FP32 on stream s (FP16 storage: staged in P32 and
  // written back).
  int panel(int k, hipStream_t s) {
    const int j = k * nb, jb = width(k), m = N - j;
    S*        Ap = A + j + static_cast<size_t>(j) * lda;
    if constexpr(kF16) HIPCHK(c, hplx::launch_copy_block(s, m, jb, Ap, lda, ws->P32, ws->ldp));
    c.panel_stream_ = s;
    const hipError_t e = hplx::panel_factor(c, hplx::panel_algo(*palg), panel_ptr(k),
                                            panel_ld(), m, jb, mat->ipiv + j, j);
    c.panel_stream_ = nullptr;
    HIPCHK(c, e);
    if constexpr(kF16) HIPCHK(c, hplx::launch_copy_block(s, m, jb, ws->P32, ws->ldp, Ap, lda));
    return HPLX_SUCCESS;
  }

  // Step k up to its update, on stream s: composed pivots, L11^{-1}, the
  // interchanges of the nt trailing columns (new top rows -> W), U12 in FP32
  // (stored back in S), and the operands packed for the low-precision GEMM.
  int prepare(int k, hipStream_t s, hplx::MxpPacked* pk, const float** U,
              int* ldu) {
    const int j = k * nb, jb = width(k), m = N - j, nt = N - j - jb;
    S*        A12 = A + j + static_cast<size_t>(j + jb) * lda;
    float*    P = panel_ptr(k);
    const int ldP = panel_ld();
    HIPCHK(c, hplx::launch_pipid(s, jb, mat->ipiv + j, j, c.permU));
    HIPCHK(c, hplx::launch_linv(s, jb, P, ldP, ws->linv, nb));
    HIPCHK(c, hplx::launch_swap_gather(s, jb, nt, A12, lda, c.permU, ws->W, nb));
    if constexpr(kF16) {
      HIPCHK(c, hplx::launch_sgemm(s, jb, nt, jb, 1.0, ws->linv, nb, ws->W, nb,
                                   0.0, ws->U, nb, true));
      HIPCHK(c, hplx::launch_copy_block(s, jb, nt, ws->U, nb, A12, lda));
      *U   = ws->U;
      *ldu = nb;
    } else {
      HIPCHK(c, hplx::launch_sgemm(s, jb, nt, jb, 1.0, ws->linv, nb, ws->W, nb,
                                   0.0, A12, lda, true));
      *U   = A12;
      *ldu = lda;
    }
    pk->a  = ws->pa;
    pk->b  = ws->pb;
    pk->ea = ws->ea;
    pk->eb = ws->eb;
    if(o->update != HPLX_UPDATE_FP32 && m - jb > 0)
      HIPCHK(c, hplx::launch_mxp_pack(s, o->update, o->terms, m - jb, jb, P + jb,
                                      ldP, nt, *U, *ldu, *pk));
    return HPLX_SUCCESS;
  }

  // C22 columns [c0, c0 + n) of step k -= L21 U12, on stream s, timed by an
  // event pair of the given kind (1 whole launch, 2 claimed, 3 resume)
  // carrying `flops`.
  int gemm(int k, int c0, int n, hipStream_t s, const hplx::MxpPacked& pk,
           const float* U, int ldu, const hplx::MxpLaunch& L, int kind,
           double flops) {
    const int j = k * nb, jb = width(k), m = N - j;
    if(m - jb <= 0 || n <= 0) return HPLX_SUCCESS;
    S*        C = A + j + jb + static_cast<size_t>(j + jb + c0) * lda;
    const int pi = c.ev_pairs++;
    c.ev_kind[pi]  = kind;
    c.ev_flops[pi] = flops;
    HIPCHK(c, hipEventRecord(c.ev[2 * pi], s));
    if(o->update == HPLX_UPDATE_FP32) {
      if constexpr(!kF16)
        HIPCHK(c, hplx::launch_sgemm(s, m - jb, n, jb, -1.0, panel_ptr(k) + jb,
                                     panel_ld(), U + static_cast<size_t>(c0) * ldu,
                                     ldu, 1.0, C, lda, false));
    } else {
      const hplx::MxpPacked q = c0 ? hplx::mxp_cols(pk, o->update, c0) : pk;
      HIPCHK(c, hplx::launch_mxp_gemm(s, o->update, o->terms, m - jb, n, jb, q,
                                      C, lda, L));
    }
    HIPCHK(c, hipEventRecord(c.ev[2 * pi + 1], s));
    return HPLX_SUCCESS;
  }
};

// The panel loop (src/pgesv/HPL_pdgesv.cpp:168-245 at P=Q=1) with depth-1
// look-ahead as in the FP64 loop: per step, the update of panel k+1's
// columns first, then panel k+1 factored on the panel stream while the rest
// of update k runs beside it (a claimed launch on the CUs the panel leaves,
// resumed on every CU once the panel is done).  Each C tile is computed by
// one launch with the same operands and K order in both schedules, so the
// factors are the serial schedule's bit for bit.
template <typename S>
int factor_loop(Ctx& c, const hplx_palg* palg, const hplx_mxp_opts* o,
                hplx_mxp_pmat* mat) {
  const int N = mat->n, nb = mat->nb, lda = mat->ld;
  if(N > hplx::max_leaf_rows(c))
    return fail(c, HPLX_ERR_UNSUPPORTED, "N=%d exceeds the panel kernel's %d rows",
                N, hplx::max_leaf_rows(c));
  MxpWs*     ws = static_cast<MxpWs*>(mat->ws);
  MxpLoop<S> L{c, palg, o, mat, ws, static_cast<S*>(mat->A), N, nb, lda};

  const int npanels = (N + nb - 1) / nb;
  while(static_cast<int>(c.ev.size()) < 6 * npanels) {
    hipEvent_t e;
    HIPCHK(c, hipEventCreate(&e));
    c.ev.push_back(e);
  }
  c.ev_pairs   = 0;
  c.gemm_flops = 0.0;
  c.ev_kind.assign(3 * npanels, 1);
  c.ev_flops.assign(3 * npanels, 0.0);
  c.splits.clear();
  c.claim_host.clear();
  hplx::Progress prog(c, npanels, nb, N);
  HIPCHK(c, prog.init());

  // the look-ahead needs the claimed GEMM (BF16 / FP8) and column splits on
  // whole 128-column tiles
  const bool la = c.lookahead && c.cus_per_xcd > 0 &&
                  o->update != HPLX_UPDATE_FP32 && nb % 128 == 0;
  hipStream_t U  = la ? c.ustream : c.stream;
  hipStream_t P  = c.pstream;
  size_t      nev = 0;
  hipEvent_t  ev;
  if(la) {  // U and P start after the caller's stream
    if(const int r = hplx::sync_event(c, nev++, &ev)) return r;
    HIPCHK(c, hipEventRecord(ev, c.stream));
    HIPCHK(c, hipStreamWaitEvent(U, ev, 0));
    HIPCHK(c, hipStreamWaitEvent(P, ev, 0));
  }
  ws->swp_ready = false;
  c.panel_stream_ = U;
  const hipError_t er = hplx::reset_panel_state(c);
  c.panel_stream_ = nullptr;
  HIPCHK(c, er);
  int r;
  if((r = L.panel(0, U))) return r;
  for(int k = 0; k < npanels; ++k) {
    const int j = k * nb, jb = L.width(k), m = N - j;
    const int nt  = N - j - jb;
    const int jb1 = k + 1 < npanels ? L.width(k + 1) : 0;
    if(nt > 0) {
      hplx::MxpPacked pk;
      const float*    Uf  = nullptr;
      int             ldu = 0;
      if((r = L.prepare(k, U, &pk, &Uf, &ldu))) return r;
      const double fl = 2.0 * (m - jb) * static_cast<double>(jb);
      c.gemm_flops += fl * nt;
      const int  rest = nt - jb1;
      const int  rR   = hplx::vacate_set_size(N - j - jb);
      const bool overlap = la && jb1 > 0 && rest >= hplx::kMinOverlapCols &&
                           rR <= c.cus_per_xcd && rR < 28 && m - jb > 0;
      if(overlap) {
        // panel k+1's columns, then that panel beside the rest of update k
        if((r = L.gemm(k, 0, jb1, U, pk, Uf, ldu, hplx::MxpLaunch(), 1, fl * jb1)))
          return r;
        int* claim = c.claim + (k & 1) * hplx::kClaimInts;
        int* stop  = c.stop + (k & 1) * 32;
        HIPCHK(c, hplx::launch_zero_ints(U, claim, hplx::kClaimInts));
        HIPCHK(c, hplx::launch_zero_ints(U, stop, 1));
        hipEvent_t ev_prep, ev_pan;
        if((r = hplx::sync_event(c, nev++, &ev_prep)) ||
           (r = hplx::sync_event(c, nev++, &ev_pan)))
          return r;
        HIPCHK(c, hipEventRecord(ev_prep, U));
        HIPCHK(c, hipStreamWaitEvent(P, ev_prep, 0));
        if((r = L.panel(k + 1, P))) return r;
        HIPCHK(c, hplx::launch_set_flag(P, stop));
        HIPCHK(c, hipEventRecord(ev_pan, P));
        hplx::MxpLaunch lc;
        lc.mode   = 1;
        lc.claim  = claim;
        lc.vacate = c.vacate + 8 * (rR - 1);
        lc.stop   = stop;
        const int pc = c.ev_pairs;
        if((r = L.gemm(k, jb1, rest, U, pk, Uf, ldu, lc, 2, 0.0))) return r;
        int bm = 0, bn = 0;
        hplx::mxp_tile(o->update, o->terms, &bm, &bn);
        const int ns = static_cast<int>(c.splits.size());
        if(ns < hplx::kMaxClaimLog)
          HIPCHK(c, hplx::launch_claim_log(U, claim, m - jb, rest,
                                           c.claim_log + 8 * ns, bm, bn));
        HIPCHK(c, hipStreamWaitEvent(U, ev_pan, 0));
        hplx::MxpLaunch lr;
        lr.mode  = 2;
        lr.claim = claim;
        const int pr = c.ev_pairs;
        if((r = L.gemm(k, jb1, rest, U, pk, Uf, ldu, lr, 3, fl * rest))) return r;
        Ctx::Split sp{m - jb, rest, jb, pc, pr};
        sp.bm = bm;
        sp.bn = bn;
        c.splits.push_back(sp);
      } else {
        if((r = L.gemm(k, 0, nt, U, pk, Uf, ldu, hplx::MxpLaunch(), 1, fl * nt)))
          return r;
        if(jb1 > 0 && (r = L.panel(k + 1, U))) return r;
      }
    } else if(jb1 > 0) {
      if((r = L.panel(k + 1, U))) return r;
    }
    HIPCHK(c, prog.panel_queued(k, U));
  }
  if(la) {  // the caller's stream continues after U and P
    hipEvent_t eu, ep;
    if((r = hplx::sync_event(c, nev++, &eu)) || (r = hplx::sync_event(c, nev++, &ep)))
      return r;
    HIPCHK(c, hipEventRecord(eu, U));
    HIPCHK(c, hipEventRecord(ep, P));
    HIPCHK(c, hipStreamWaitEvent(c.stream, eu, 0));
    HIPCHK(c, hipStreamWaitEvent(c.stream, ep, 0));
  }
  hipStream_t s = c.stream;
  int flags[2] = {0, 0};
  HIPCHK(c, hipMemcpyAsync(flags, c.flags, sizeof(flags), hipMemcpyDeviceToHost,
                           s));
  const size_t nlog = std::min<size_t>(c.splits.size(), hplx::kMaxClaimLog);
  c.claim_host.assign(8 * nlog, 0);
  if(nlog)
    HIPCHK(c, hipMemcpyAsync(c.claim_host.data(), c.claim_log,
                             sizeof(int) * 8 * nlog, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  if(flags[1])
    return fail(c, HPLX_ERR_HIP,
                "panel kernel wait timed out (workgroups not co-resident)");
  mat->info = flags[0];
  if(const int r2 = build_inverses<S>(c, mat)) return r2;
  if(const int r2 = build_swaps(c, mat)) return r2;
  return flags[0];
}

// The interchanges of panel p (y[i] <-> y[ipiv[i]] for its rows i in order)
// composed into the moves y[dst] = y_old[src] over the rows they touch, so
// that the solve applies them in one parallel step per panel.
int build_swaps(Ctx& c, const hplx_mxp_pmat* mat) {
  MxpWs*    ws = static_cast<MxpWs*>(mat->ws);
  const int N = mat->n, nb = mat->nb;
  std::vector<int> piv(N), cur(N), pairs;
  pairs.reserve(4 * static_cast<size_t>(N));
  HIPCHK(c, hipMemcpyAsync(piv.data(), mat->ipiv, sizeof(int) * N,
                           hipMemcpyDeviceToHost, c.stream));
  HIPCHK(c, hipStreamSynchronize(c.stream));
  for(int i = 0; i < N; ++i) cur[i] = i;
  ws->swp_off.assign(1, 0);
  for(int j = 0; j < N; j += nb) {
    const int jb = std::min(nb, N - j);
    for(int i = j; i < j + jb; ++i) {
      const int p = piv[i];
      if(p < i || p >= N)
        return fail(c, HPLX_ERR_INVALID, "pivot %d of row %d out of range", p, i);
      std::swap(cur[i], cur[p]);
    }
    // record every moved row once, restoring cur as it goes
    for(int i = j; i < j + jb; ++i)
      for(const int d : {i, piv[i]})
        if(cur[d] != d) {
          pairs.push_back(d);
          pairs.push_back(cur[d]);
          cur[d] = d;
        }
    ws->swp_off.push_back(static_cast<int>(pairs.size() / 2));
  }
  if(!pairs.empty())
    HIPCHK(c, hipMemcpy(ws->swp, pairs.data(), sizeof(int) * pairs.size(),
                        hipMemcpyHostToDevice));
  ws->swp_ready = true;
  return HPLX_SUCCESS;
}

// x = U^{-1} L^{-1} P x with the stored factors, in the vectors' precision
// V (FP32, or FP64 with solve_fp64).  Each diagonal block is applied as its
// inverse (one bandwidth-bound product) into ytmp, then the block's columns
// below (L) or above (U) are applied from ytmp, the same launch copying
// ytmp back into x.
template <typename S, typename V>
int solve(Ctx& c, const hplx_mxp_pmat* mat, V* x) {
  const int   N = mat->n, nb = mat->nb, lda = mat->ld;
  const S*    A = static_cast<const S*>(mat->A);
  hipStream_t s = c.stream;
  MxpWs* ws = static_cast<MxpWs*>(mat->ws);
  if(!ws->swp_ready) {
    const int r = build_swaps(c, mat);
    if(r) return r;
  }
  V*       y    = static_cast<V*>(ws->ytmp);
  const V* invl = static_cast<const V*>(ws->dinv_l);
  const V* invu = static_cast<const V*>(ws->dinv_u);
  // L y = P b, panel by panel (each panel's interchanges, then its L11 and
  // L21; L is left unpivoted by later panels)
  for(int j = 0; j < N; j += nb) {
    const int jb = (N - j) < nb ? (N - j) : nb;
    const int p0 = ws->swp_off[j / nb], p1 = ws->swp_off[j / nb + 1];
    HIPCHK(c, hplx::launch_apply_swaps(s, p1 - p0, ws->swp + 2 * p0, x));
    HIPCHK(c, hplx::launch_inv_gemv(s, jb, invl + static_cast<size_t>(j) * nb,
                                    x + j, y));
    if(N - j - jb > 0)
      HIPCHK(c, hplx::launch_gemv_sub(s, N - j - jb, jb,
                                      A + j + jb + static_cast<size_t>(j) * lda,
                                      lda, y, x + j + jb, x + j));
    else
      HIPCHK(c, hipMemcpyAsync(x + j, y, sizeof(V) * jb,
                               hipMemcpyDeviceToDevice, s));
  }
  // U x = y, blocks from the bottom
  const int nblk = (N + nb - 1) / nb;
  for(int J = nblk - 1; J >= 0; --J) {
    const int s0 = J * nb;
    const int kb = (N - s0) < nb ? (N - s0) : nb;
    HIPCHK(c, hplx::launch_inv_gemv(s, kb, invu + static_cast<size_t>(s0) * nb,
                                    x + s0, y));
    if(s0 > 0)
      HIPCHK(c, hplx::launch_gemv_sub(s, s0, kb, A + static_cast<size_t>(s0) * lda,
                                      lda, y, x, x + s0));
    else
      HIPCHK(c, hipMemcpyAsync(x + s0, y, sizeof(V) * kb,
                               hipMemcpyDeviceToDevice, s));
  }
  return HPLX_SUCCESS;
}

// x (FP64, in and out) = (LU)^{-1} P x: in FP64, or rounded to FP32, solved
// in FP32 and widened back.
template <typename S>
int solve_vec(Ctx& c, const hplx_mxp_pmat* mat, double* x) {
  MxpWs* ws = static_cast<MxpWs*>(mat->ws);
  if(ws->solve_fp64) return solve<S, double>(c, mat, x);
  HIPCHK(c, hplx::launch_to_f32(c.stream, mat->n, x, ws->x32));
  if(const int r = solve<S, float>(c, mat, ws->x32)) return r;
  HIPCHK(c, hplx::launch_to_f64(c.stream, mat->n, ws->x32, x));
  return HPLX_SUCCESS;
}

int solve_any(Ctx& c, const hplx_mxp_pmat* mat, double* x) {
  return mat->storage == HPLX_STORE_FP16 ? solve_vec<__half>(c, mat, x)
                                         : solve_vec<float>(c, mat, x);
}

int factor_any(Ctx& c, const hplx_palg* palg, const hplx_mxp_opts* o,
               hplx_mxp_pmat* mat) {
  int r;
  if((r = hplx::check_palg(c, palg))) return r;
  if((r = check_mat(c, mat))) return r;
  if((r = check_opts(c, o, mat->storage))) return r;
  static_cast<MxpWs*>(mat->ws)->solve_fp64 = o->solve_fp64 != 0;
  return mat->storage == HPLX_STORE_FP16 ? factor_loop<__half>(c, palg, o, mat)
                                         : factor_loop<float>(c, palg, o, mat);
}

double gemm_ms(Ctx& c) {
  double t = 0.0;
  for(int i = 0; i < c.ev_pairs; ++i) {
    float ms = 0.f;
    if(hipEventElapsedTime(&ms, c.ev[2 * i], c.ev[2 * i + 1]) == hipSuccess) t += ms;
  }
  return t;
}

}  // namespace

extern "C" {

void hplx_mxp_opts_default(hplx_mxp_opts* o) {
  if(!o) return;
  o->update    = HPLX_UPDATE_BF16;
  o->terms     = 3;
  o->max_its   = 50;
  o->solve_fp64 = 0;
  o->threshold = 16.0;
}

int hplx_mxp_alloc(hplx_ctx* ctx, int n, int nb, int storage,
                   hplx_mxp_pmat* mat) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  if(!mat || n < 1 || nb < 1 || nb > hplx::kMaxNB ||
     (storage != HPLX_STORE_FP32 && storage != HPLX_STORE_FP16))
    return fail(c, HPLX_ERR_INVALID, "mxp_alloc: n=%d nb=%d storage=%d", n, nb,
                storage);
  std::memset(mat, 0, sizeof(*mat));
  MxpWs* ws = new(std::nothrow) MxpWs();
  if(!ws) return HPLX_ERR_ALLOC;
  mat->n       = n;
  mat->nb      = nb;
  mat->ld      = hpl_ld(n);
  mat->storage = storage;
  mat->ws      = ws;
  const size_t es   = elem_size(storage);
  const size_t pack = std::max(hplx::mxp_pack_bytes(HPLX_UPDATE_BF16, 3, n, nb),
                               hplx::mxp_pack_bytes(HPLX_UPDATE_FP8, 1, n, nb));
  const size_t nsc  = hplx::mxp_scale_count(n, nb);
  // diagonal-block inverses: whole blocks of nb x nb, the last one smaller
  const size_t dinv = static_cast<size_t>((n + nb - 1) / nb) * nb * nb;
  bool ok = hipMalloc(&mat->A, es * mat->ld * static_cast<size_t>(n)) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&mat->X), 8 * static_cast<size_t>(n)) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&mat->ipiv), 4 * static_cast<size_t>(n)) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&ws->W), 4 * static_cast<size_t>(nb) * n) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&ws->linv), 4 * static_cast<size_t>(nb) * nb) == hipSuccess &&
            hipMalloc(&ws->pa, pack) == hipSuccess && hipMalloc(&ws->pb, pack) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&ws->ea), 4 * nsc) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&ws->eb), 4 * nsc) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&ws->r), 8 * static_cast<size_t>(n)) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&ws->part), 8 * hplx::ir_part_len(n)) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&ws->out), 16 * 8) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&ws->swp), 16 * static_cast<size_t>(n)) == hipSuccess &&
            hipMalloc(&ws->dinv_l, 8 * dinv) == hipSuccess &&
            hipMalloc(&ws->dinv_u, 8 * dinv) == hipSuccess &&
            hipMalloc(&ws->ytmp, 8 * static_cast<size_t>(nb)) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&ws->x32), 4 * static_cast<size_t>(n)) == hipSuccess;
  if(ok && storage == HPLX_STORE_FP16) {
    ws->ldp = static_cast<int>(hplx::align_up(n, 64));
    ok = hipMalloc(reinterpret_cast<void**>(&ws->P32), 4 * static_cast<size_t>(ws->ldp) * nb) == hipSuccess &&
         hipMalloc(reinterpret_cast<void**>(&ws->U), 4 * static_cast<size_t>(nb) * n) == hipSuccess;
  }
  if(!ok) {
    hplx_mxp_free(ctx, mat);
    return fail(c, HPLX_ERR_ALLOC, "mxp_alloc: out of device memory (n=%d)", n);
  }
  return HPLX_SUCCESS;
}

int hplx_mxp_free(hplx_ctx* ctx, hplx_mxp_pmat* mat) {
  int r = use(ctx);
  if(r) return r;
  if(!mat) return HPLX_SUCCESS;
  (void)hipStreamSynchronize(ctx->c.stream);
  if(MxpWs* ws = static_cast<MxpWs*>(mat->ws)) {
    for(void* p : {static_cast<void*>(ws->P32), static_cast<void*>(ws->W),
                   static_cast<void*>(ws->U), static_cast<void*>(ws->linv), ws->pa,
                   ws->pb, static_cast<void*>(ws->ea), static_cast<void*>(ws->eb),
                   static_cast<void*>(ws->r), static_cast<void*>(ws->part),
                   static_cast<void*>(ws->out), static_cast<void*>(ws->swp),
                   ws->dinv_l, ws->dinv_u, ws->ytmp,
                   static_cast<void*>(ws->x32)})
      if(p) (void)hipFree(p);
    delete ws;
  }
  if(mat->A) (void)hipFree(mat->A);
  if(mat->X) (void)hipFree(mat->X);
  if(mat->ipiv) (void)hipFree(mat->ipiv);
  std::memset(mat, 0, sizeof(*mat));
  return HPLX_SUCCESS;
}

int hplx_mxp_matgen(hplx_ctx* ctx, hplx_mxp_pmat* mat, int seed,
                    double diag_shift) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  if((r = check_mat(c, mat))) return r;
  if(seed < 0) return fail(c, HPLX_ERR_INVALID, "seed %d", seed);
  mat->seed       = seed;
  mat->diag_shift = diag_shift;
  if(mat->storage == HPLX_STORE_FP16)
    HIPCHK(c, hplx::launch_matgen(c.stream, mat->n, mat->n,
                                  static_cast<__half*>(mat->A), mat->ld, seed,
                                  diag_shift));
  else
    HIPCHK(c, hplx::launch_matgen(c.stream, mat->n, mat->n,
                                  static_cast<float*>(mat->A), mat->ld, seed,
                                  diag_shift));
  HIPCHK(c, hipStreamSynchronize(c.stream));
  return HPLX_SUCCESS;
}

int hplx_mxp_pdgetrf(hplx_ctx* ctx, const hplx_palg* palg,
                     const hplx_mxp_opts* opts, hplx_mxp_pmat* mat) {
  int r = use(ctx);
  if(r) return r;
  return factor_any(ctx->c, palg, opts, mat);
}

int hplx_mxp_pdgetrs(hplx_ctx* ctx, const hplx_mxp_pmat* mat, double* x) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  if((r = check_mat(c, mat))) return r;
  if(!x) return fail(c, HPLX_ERR_INVALID, "null x");
  if((r = solve_any(c, mat, x))) return r;
  HIPCHK(c, hipStreamSynchronize(c.stream));
  return HPLX_SUCCESS;
}

int hplx_mxp_residual(hplx_ctx* ctx, const hplx_mxp_pmat* mat,
                      const double* x, double* out) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  if((r = check_mat(c, mat))) return r;
  if(!x || !out) return fail(c, HPLX_ERR_INVALID, "null x/out");
  MxpWs* ws = static_cast<MxpWs*>(mat->ws);
  HIPCHK(c, hplx::launch_ir_residual(c.stream, mat->n, mat->seed, mat->diag_shift,
                                     x, ws->r, ws->part, ws->out));
  HIPCHK(c, hipMemcpyAsync(out, ws->out, 5 * sizeof(double),
                           hipMemcpyDeviceToHost, c.stream));
  HIPCHK(c, hipStreamSynchronize(c.stream));
  return HPLX_SUCCESS;
}

int hplx_mxp_pdgesv(hplx_ctx* ctx, const hplx_palg* palg,
                    const hplx_mxp_opts* opts, hplx_mxp_pmat* mat,
                    hplx_mxp_result* res) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  hplx_mxp_result local;
  hplx_mxp_result& R = res ? *res : local;
  std::memset(&R, 0, sizeof(R));
  const auto t0 = std::chrono::steady_clock::now();
  r = factor_any(c, palg, opts, mat);
  if(r) return r;
  const auto t1 = std::chrono::steady_clock::now();
  R.factor_ms       = std::chrono::duration<double, std::milli>(t1 - t0).count();
  R.update_ms       = gemm_ms(c);
  R.update_flops    = c.gemm_flops;
  R.update_launches = c.ev_pairs;

  MxpWs*      ws = static_cast<MxpWs*>(mat->ws);
  hipStream_t s  = c.stream;
  const int   N  = mat->n;
  HIPCHK(c, hplx::launch_gen_b(s, N, mat->seed, mat->X));
  if((r = solve_any(c, mat, mat->X))) return r;
  for(int it = 0;; ++it) {
    HIPCHK(c, hplx::launch_ir_residual(s, N, mat->seed, mat->diag_shift, mat->X,
                                       ws->r, ws->part, ws->out));
    double resid = 0.0;
    HIPCHK(c, hipMemcpyAsync(&resid, ws->out, sizeof(double),
                             hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    R.history[it] = resid;
    R.resid       = resid;
    R.its         = it;
    if(resid < opts->threshold) {
      R.converged = 1;
      break;
    }
    if(it >= opts->max_its || !(resid == resid)) break;
    if((r = solve_any(c, mat, ws->r))) return r;
    HIPCHK(c, hplx::launch_axpy(s, N, ws->r, mat->X));
  }
  R.solve_ms = std::chrono::duration<double, std::milli>(
                   std::chrono::steady_clock::now() - t1)
                   .count();
  return R.converged ? HPLX_SUCCESS : 1 + N;
}

int hplx_mxp_update(hplx_ctx* ctx, int update, int terms, int storage, int m,
                    int n, int k, const float* L, int ldl, const float* U,
                    int ldu, void* C, int ldc) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  if(m < 0 || n < 0 || k < 0 || k > hplx::kMaxNB || !L || !U || !C ||
     ldl < m || ldu < k || ldc < m ||
     (storage != HPLX_STORE_FP32 && storage != HPLX_STORE_FP16))
    return fail(c, HPLX_ERR_INVALID, "mxp_update: bad arguments");
  if(update == HPLX_UPDATE_BF16 && (terms < 1 || terms > 3))
    return fail(c, HPLX_ERR_INVALID, "terms %d", terms);
  if(m == 0 || n == 0 || k == 0) return HPLX_SUCCESS;
  hipStream_t s = c.stream;
  if(update == HPLX_UPDATE_FP32) {
    if(storage != HPLX_STORE_FP32)
      return fail(c, HPLX_ERR_UNSUPPORTED, "FP32 update needs FP32 storage");
    HIPCHK(c, hplx::launch_sgemm(s, m, n, k, -1.0, L, ldl, U, ldu, 1.0,
                                 static_cast<float*>(C), ldc, false));
    HIPCHK(c, hipStreamSynchronize(s));
    return HPLX_SUCCESS;
  }
  if(update != HPLX_UPDATE_BF16 && update != HPLX_UPDATE_FP8)
    return fail(c, HPLX_ERR_INVALID, "update %d", update);
  hplx::MxpPacked pk;
  const size_t ba = hplx::mxp_pack_bytes(update, terms, m, k);
  const size_t bb = hplx::mxp_pack_bytes(update, terms, n, k);
  const size_t sa = hplx::mxp_scale_count(m, k), sb = hplx::mxp_scale_count(n, k);
  if(hipMalloc(&pk.a, ba) != hipSuccess || hipMalloc(&pk.b, bb) != hipSuccess ||
     hipMalloc(reinterpret_cast<void**>(&pk.ea), 4 * sa) != hipSuccess ||
     hipMalloc(reinterpret_cast<void**>(&pk.eb), 4 * sb) != hipSuccess) {
    for(void* p : {pk.a, pk.b, static_cast<void*>(pk.ea), static_cast<void*>(pk.eb)})
      if(p) (void)hipFree(p);
    return fail(c, HPLX_ERR_ALLOC, "mxp_update: scratch allocation");
  }
  hipError_t e = hplx::launch_mxp_pack(s, update, terms, m, k, L, ldl, n, U,
                                       ldu, pk);
  if(e == hipSuccess)
    e = storage == HPLX_STORE_FP16
            ? hplx::launch_mxp_gemm(s, update, terms, m, n, k, pk,
                                    static_cast<__half*>(C), ldc)
            : hplx::launch_mxp_gemm(s, update, terms, m, n, k, pk,
                                    static_cast<float*>(C), ldc);
  const hipError_t es = hipStreamSynchronize(s);
  for(void* p : {pk.a, pk.b, static_cast<void*>(pk.ea), static_cast<void*>(pk.eb)})
    (void)hipFree(p);
  HIPCHK(c, e);
  HIPCHK(c, es);
  return HPLX_SUCCESS;
}

int hplx_spdfact(hplx_ctx* ctx, int nbmin, int nbdiv, float* P, int lda,
                 int m, int jb, int* ipiv, int row_offset) {
  int r = use(ctx);
  if(r) return r;
  Ctx& c = ctx->c;
  if(!P || !ipiv || m < 1 || jb < 1 || jb > m || lda < m || row_offset < 0)
    return fail(c, HPLX_ERR_INVALID, "spdfact: m=%d jb=%d lda=%d", m, jb, lda);
  if(m > hplx::max_leaf_rows(c))
    return fail(c, HPLX_ERR_UNSUPPORTED, "spdfact: m=%d too tall", m);
  HIPCHK(c, hplx::reset_panel_state(c));
  HIPCHK(c, hplx::panel_factor(c, hplx::panel_algo(nbmin, nbdiv), P, lda, m,
                               jb, ipiv, row_offset));
  int flags[2] = {0, 0};
  HIPCHK(c, hipMemcpyAsync(flags, c.flags, sizeof(flags), hipMemcpyDeviceToHost,
                           c.stream));
  HIPCHK(c, hipStreamSynchronize(c.stream));
  if(flags[1]) return fail(c, HPLX_ERR_HIP, "panel kernel wait timed out");
  return flags[0];
}

}  // extern "C"
