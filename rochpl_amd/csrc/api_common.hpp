// Pieces shared by the C-ABI translation units (hplx_api.hip, mxp_api.hip):
// the context object, error reporting, palg validation and the panel-loop
// progress lines.
#pragma once

#include <chrono>
#include <cstdarg>
#include <cstdio>

#include "hplx_internal.hpp"

struct hplx_ctx {
  hplx::Ctx c;
};

namespace hplx {

inline int fail(Ctx& c, int code, const char* fmt, ...) {
  char    buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  c.last_error = buf;
  return code;
}

#define HIPCHK(C, expr)                                                    \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if(e_ != hipSuccess)                                                   \
      return hplx::fail((C), HPLX_ERR_HIP, "%s failed: %s (%s:%d)", #expr, \
                        hipGetErrorString(e_), __FILE__, __LINE__);        \
  } while(0)

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Timing-disabled event i of the context's cross-stream pool (created on
// first use).
inline int sync_event(Ctx& c, size_t i, hipEvent_t* e) {
  while(c.ev_sync.size() <= i) {
    hipEvent_t x;
    const hipError_t r = hipEventCreateWithFlags(&x, hipEventDisableTiming);
    if(r != hipSuccess)
      return fail(c, HPLX_ERR_HIP, "hipEventCreateWithFlags: %s",
                  hipGetErrorString(r));
    c.ev_sync.push_back(x);
  }
  *e = c.ev_sync[i];
  return HPLX_SUCCESS;
}

inline int use(hplx_ctx* ctx) {
  if(!ctx) return HPLX_ERR_INVALID;
  hipError_t e = hipSetDevice(ctx->c.device);
  if(e != hipSuccess)
    return fail(ctx->c, HPLX_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
  return HPLX_SUCCESS;
}

// The P=Q=1 configurations this library builds (HPL_T_palg values):
// HPLX_SUCCESS, or an error code with its message in msg.  No device needed.
inline int palg_status(const hplx_palg* p, char* msg, size_t len) {
  if(!p) {
    std::snprintf(msg, len, "null palg");
    return HPLX_ERR_INVALID;
  }
  if(!valid_fact(p->pfact) || !valid_fact(p->rfact)) {
    std::snprintf(msg, len, "unknown panel factorisation (pfact=%d rfact=%d)",
                  p->pfact, p->rfact);
    return HPLX_ERR_INVALID;
  }
  // L1notran picks how the reference stores its copy of the panel's top
  // block (src/HPL_pddriver.cpp:166-199); the factors here stay in A, so
  // both forms run the same kernels.  U in non-transposed form is refused,
  // as the reference refuses it (src/HPL_pdinfo.cpp:979-986).
  if(p->L1notran != 0 && p->L1notran != 1) {
    std::snprintf(msg, len, "L1notran=%d (0 or 1)", p->L1notran);
    return HPLX_ERR_INVALID;
  }
  if(p->Unotran != 0) {
    std::snprintf(msg, len, "U in non-transposed form is not supported");
    return HPLX_ERR_UNSUPPORTED;
  }
  if(p->equil != 0) {
    std::snprintf(msg, len, "equilibration is not supported");
    return HPLX_ERR_UNSUPPORTED;
  }
  if(p->nbmin < 1 || p->nbdiv < 2) {
    std::snprintf(msg, len, "nbmin=%d nbdiv=%d", p->nbmin, p->nbdiv);
    return HPLX_ERR_INVALID;
  }
  // the reference refuses both in HPL_pdinfo (src/HPL_pdinfo.cpp:930-935
  // "Value of DEPTH must be 1", :950-955 "Value of SWAP must be 1")
  if(p->depth != 1) {
    std::snprintf(msg, len, "look-ahead depth must be 1 (got %d)", p->depth);
    return HPLX_ERR_UNSUPPORTED;
  }
  if(p->fswap != HPLX_SWAP01) {
    std::snprintf(msg, len,
                  "only long swapping (HPLX_SWAP01) is implemented (got %d)",
                  p->fswap);
    return HPLX_ERR_UNSUPPORTED;
  }
  return HPLX_SUCCESS;
}

inline int check_palg(Ctx& c, const hplx_palg* p) {
  char      msg[256];
  const int r = palg_status(p, msg, sizeof(msg));
  return r ? fail(c, r, "%s", msg) : HPLX_SUCCESS;
}

// Progress lines of a panel loop (HPLX_PROGRESS=<panels>, the counterpart of
// the reference's HPL_PROGRESS_REPORT build option): reports panels the
// device has finished, at most kLag panels behind what is queued, so the
// host never runs further ahead than that while it reports.
class Progress {
 public:
  static constexpr int kLag = 8;
  Progress(Ctx& c, int npanels, int nb, int n)
      : c_(c), npanels_(npanels), nb_(nb), n_(n),
        t0_(std::chrono::steady_clock::now()) {}

  hipError_t init() {
    if(c_.progress_every <= 0) return hipSuccess;
    while(static_cast<int>(c_.ev_prog.size()) < 2 * kLag) {
      hipEvent_t e;
      hipError_t r = hipEventCreateWithFlags(&e, hipEventDisableTiming);
      if(r != hipSuccess) return r;
      c_.ev_prog.push_back(e);
    }
    return hipSuccess;
  }

  // panel k's work has been queued; s: the stream its last step is on
  hipError_t panel_queued(int k, hipStream_t s) {
    if(c_.progress_every <= 0) return hipSuccess;
    hipError_t r = hipEventRecord(c_.ev_prog[k % (2 * kLag)], s);
    if(r != hipSuccess) return r;
    const int done = k - kLag;
    if(done < 0) return hipSuccess;
    r = hipEventSynchronize(c_.ev_prog[done % (2 * kLag)]);
    if(r != hipSuccess) return r;
    if(done % c_.progress_every == 0) {
      const double el =
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_)
              .count();
      std::fprintf(stderr, "[hplx] panel %d/%d done (column %d of %d) %.2f s\n",
                   done + 1, npanels_, (done + 1) * nb_, n_, el);
      std::fflush(stderr);
    }
    return hipSuccess;
  }

 private:
  Ctx&                                  c_;
  int                                   npanels_, nb_, n_;
  std::chrono::steady_clock::time_point t0_;
};

}  // namespace hplx
