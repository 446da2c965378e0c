// Panel factorisation, pivot composition, row swaps and small triangular
// solves.
//
// Panel (HPL_pdfact -> HPL_pdrpanrlT -> HPL_pdpanrlT at P=Q=1,
// src/pfact/HPL_pdfact.cpp:20-110, src/pfact/HPL_pdrpanrlT.cpp:93-160,
// src/pfact/HPL_pdpanrlT_device.cpp:23-468).
//   The recursion is the reference's: split n columns at
//   ((ceil(n/nbmin)+ndiv-1)/ndiv)*nbmin, factor the left part, solve the top
//   rows of the right part with the unit lower L11, update the rows below
//   with one FP64 MFMA GEMM.  Leaves of <= 32 columns run as ONE cooperative
//   kernel per leaf:
//     - every workgroup keeps its slice of the leaf (up to 1024 rows x 32
//       columns) in registers for the whole leaf, so the leaf reads and writes
//       HBM once;
//     - the pivot search is a wave64 shuffle max-loc, then a workgroup
//       max-loc through LDS, then one grid barrier per column after which every
//       workgroup reduces the published per-workgroup candidates itself;
//     - no host round-trip: the reference parks its leaf on a pinned-memory
//       flag once per column and lets a host thread run HPL_pdmxswp
//       (src/pfact/HPL_pdpanrlT_device.cpp:94-117,455-463);
//     - the row interchanges of the leaf are applied to the other panel
//       columns in the kernel's epilogue as one composed permutation.
//   Pivot rule: max |a|, ties to the lower row index (LAPACK idamax), so the
//   pivot vector equals dgetrf's on data without exact ties.  Multipliers are
//   a / pivot (IEEE division), as in the reference (:232,317).  An exact zero
//   pivot leaves its column unscaled and is reported LAPACK-style.
//
// Pivot composition (HPL_pipid + HPL_pdpanel_swapids P=1 branch,
// src/pgesv/HPL_pipid.cpp:19-164, src/panel/HPL_pdpanel_swapids.cpp:22-41)
// runs on the device, so no host copy or stream synchronisation sits in the
// panel loop.
//
// Row swaps (HPL_dlaswp00N, src/pauxil/HPL_dlaswp00N_device.cpp:22-62):
// one workgroup per column; the gather variant writes the new top rows to a
// separate buffer that feeds the U12 = L11^{-1} A12 GEMM.
#include <climits>

#include "leaf.hpp"

namespace hplx {
namespace {

__global__ void zero_kernel(double* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) p[i] = 0.0;
}

__global__ void reset_kernel(int* info, int* err) {
  *info = 0;
  *err  = 0;
}

// Device HPL_pipid: compose jb interchanges (row i <-> ipiv[i]-base) into
// permU[0:jb] (source of each top row) and permU[jb:2jb] (destination of
// each displaced top row, or -1), without a serial pass:
//   position i < jb is last written at step i, with what position p_i held
//   just before step i; a position q is written before step i only by the
//   latest earlier step i' with p_i' = q, which moved there what position i'
//   held just before step i'.  With g(i) = content of position i just before
//   step i:
//     g(i)       = g(last i' < i with p_i' = i), or i if none
//     top[i]     = g(i) if p_i == i, else g(last i' < i with p_i' = p_i), or
//                  p_i if none
//     a position q >= jb ends with g(last i with p_i = q).
//   Chains of g only go to smaller indices and are followed per thread.
__global__ __launch_bounds__(1024) void pipid_kernel(
    int jb, const int* __restrict__ ipiv, int base, int* __restrict__ permU) {
  __shared__ int sp[kMaxNB];
  __shared__ int sprev[kMaxNB];   // last i' < i with p_i' == p_i, or -1
  __shared__ int sto[kMaxNB];     // last i' < i with p_i' == i, or -1
  __shared__ int slast[kMaxNB];   // 1 if no later step targets p_i
  __shared__ int sg[kMaxNB];
  __shared__ int sout[2 * kMaxNB];
  const int t = threadIdx.x;
  // an interchange must target a row at or below its own; anything else is
  // treated as no interchange so that every LDS index stays in range
  for(int i = t; i < jb; i += blockDim.x) {
    const int p = ipiv[i] - base;
    sp[i]       = p < i ? i : p;
  }
  __syncthreads();
  for(int i = t; i < jb; i += blockDim.x) {
    int prev = -1, to = -1, last = 1;
    const int p = sp[i];
    for(int k2 = 0; k2 < jb; ++k2) {
      const int q = sp[k2];
      if(k2 < i) {
        if(q == p) prev = k2;
        if(q == i) to = k2;
      } else if(k2 > i && q == p) {
        last = 0;
      }
    }
    sprev[i] = prev;
    sto[i]   = to;
    slast[i] = last;
  }
  __syncthreads();
  for(int i = t; i < jb; i += blockDim.x) {
    int j = i;
    while(sto[j] >= 0) j = sto[j];  // strictly decreasing chain
    sg[i] = j;
  }
  __syncthreads();
  for(int i = t; i < jb; i += blockDim.x) {
    const int p = sp[i];
    int       top;
    if(p == i)
      top = sg[i];
    else
      top = sprev[i] >= 0 ? sg[sprev[i]] : p;
    sout[i]      = top;
    sout[jb + i] = -1;
  }
  __syncthreads();
  for(int i = t; i < jb; i += blockDim.x)
    if(sp[i] >= jb && slast[i]) sout[jb + sg[i]] = sp[i];
  __syncthreads();
  for(int i = t; i < 2 * jb; i += blockDim.x) permU[i] = sout[i];
}

// New top rows -> W (ld ldw); displaced top rows -> their destinations.
// S: storage type of A; WT: type of W (the FP16-stored factorisation gathers
// into FP32).
template <typename S, typename WT>
__global__ __launch_bounds__(256) void swap_gather_kernel(
    int jb, S* __restrict__ A, int lda, const int* __restrict__ permU,
    WT* __restrict__ W, int ldw) {
  constexpr int Q   = kMaxNB / 256;
  const int     col = blockIdx.x;
  S*            ac  = A + static_cast<size_t>(col) * lda;
  WT*           wc  = W + static_cast<size_t>(col) * ldw;
  S             keep[Q];
  int           dst[Q];
#pragma unroll
  for(int q = 0; q < Q; ++q) {
    const int i = threadIdx.x + 256 * q;
    dst[q]      = -1;
    keep[q]     = S(0);
    if(i < jb) {
      wc[i]  = static_cast<WT>(ac[permU[i]]);  // the store waits for the load
      dst[q] = permU[jb + i];
      if(dst[q] >= 0) keep[q] = ac[i];
    }
  }
  __syncthreads();  // every gather has been read before any scatter
#pragma unroll
  for(int q = 0; q < Q; ++q)
    if(dst[q] >= 0) ac[dst[q]] = keep[q];
}

// In-place HPL_dlaswp00N.
__global__ __launch_bounds__(256) void swap_inplace_kernel(
    int jb, double* __restrict__ A, int lda, const int* __restrict__ permU) {
  constexpr int Q = kMaxNB / 256;
  __shared__ double sg[kMaxNB];
  const int col = blockIdx.x;
  double*   ac  = A + static_cast<size_t>(col) * lda;
  double    keep[Q];
  int       dst[Q];
#pragma unroll
  for(int q = 0; q < Q; ++q) {
    const int i = threadIdx.x + 256 * q;
    dst[q]      = -1;
    keep[q]     = 0.0;
    if(i < jb) {
      sg[i]  = ac[permU[i]];
      dst[q] = permU[jb + i];
      if(dst[q] >= 0) keep[q] = ac[i];
    }
  }
  __syncthreads();
#pragma unroll
  for(int q = 0; q < Q; ++q) {
    const int i = threadIdx.x + 256 * q;
    if(i < jb) ac[i] = sg[i];
    if(dst[q] >= 0) ac[dst[q]] = keep[q];
  }
}

// Inverse of the composed interchanges (the transpose P^T of the row
// permutation, applied by the adjoint solve): what the forward swap moved
// from row src(i) to top row i goes back, and each displaced top row t
// returns from dst(t).
__global__ __launch_bounds__(256) void swap_inverse_kernel(
    int jb, double* __restrict__ A, int lda, const int* __restrict__ permU) {
  constexpr int Q = kMaxNB / 256;
  __shared__ double sv[kMaxNB];
  __shared__ double sd[kMaxNB];
  const int col = blockIdx.x;
  double*   ac  = A + static_cast<size_t>(col) * lda;
#pragma unroll
  for(int q = 0; q < Q; ++q) {
    const int i = threadIdx.x + 256 * q;
    if(i < jb) {
      sv[i]       = ac[i];
      const int d = permU[jb + i];
      sd[i]       = d >= 0 ? ac[d] : 0.0;
    }
  }
  __syncthreads();  // every value is read before any is written
#pragma unroll
  for(int q = 0; q < Q; ++q) {
    const int i = threadIdx.x + 256 * q;
    if(i < jb) {
      ac[permU[i]] = sv[i];
      if(permU[jb + i] >= 0) ac[i] = sd[i];
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

template <typename V>
const void* leaf_entry() {
  if constexpr(sizeof(V) == 8)
    return leaf_kernel_ptr<LeafKind::kF64>();
  else
    return leaf_kernel_ptr<LeafKind::kF32>();
}

template <typename V>
hipError_t launch_leaf(Ctx& c, V* P, int lda, int m, int JB, int c0, int w,
                       int* ipiv, int row_offset, int info_base) {
  constexpr int CH = kLeafThreads * kLeafRPT;
  const int     R  = m - c0;
  const int     G  = (R + CH - 1) / CH;
  if(G > kMaxLeafGrid || G < 1 || w < 1 || w > kLeafMax)
    return hipErrorInvalidValue;
  if(G > c.leaf_resident) return hipErrorInvalidValue;
  // sequence numbers: launch << 6 | column; slots restart when they wrap
  if(c.leaf_launches >= (1u << 26) - 1) {
    const int n = 2 * (kMaxGrid + 1) * kSlotStride;
    hipLaunchKernelGGL(zero_kernel, dim3((n + 255) / 256), dim3(256), 0,
                       panel_stream(c), c.slots, n);
    hipError_t e = hipGetLastError();
    if(e != hipSuccess) return e;
    c.leaf_launches = 0;
  }
  ++c.leaf_launches;
  leaf::LeafArgs la;
  la.P           = P;
  la.lda         = lda;
  la.m           = m;
  la.JB          = JB;
  la.c0          = c0;
  la.w           = w;
  la.ipiv        = ipiv;
  la.row_offset  = row_offset;
  la.info        = c.flags;
  la.info_base   = info_base;
  la.slots       = c.slots;
  la.seq_base    = c.leaf_launches << 6;
  la.err         = c.flags + 1;
  la.rows_per_wg = CH;
  la.grid        = G;
  // A plain launch: every workgroup is resident at once because the grid is
  // capped by the occupancy of this kernel (one workgroup per CU) and the
  // solve queues nothing else concurrently; the kernel's bounded waits turn
  // any violation into an error instead of a hang.  (The cooperative-launch
  // API adds nothing here and refuses to launch while another process holds
  // the device.)
  void* args[] = {&la};
  return hipLaunchKernel(leaf_entry<V>(), dim3(G), dim3(kLeafThreads), args,
                         0, panel_stream(c));
}

// C (m x n) -= A (m x k) B (k x n) inside the panel recursion
// (rocblas_dgemm at src/pfact/HPL_pdrpanrlT.cpp:142-154).  FP64 products at
// least one 128-column tile wide go through the streamed trailing-update
// kernel with B transposed into the context's L11^{-1} buffer, which is idle
// while a panel is factored; the others, and FP32, through the generic GEMM.
inline hipError_t panel_gemm(Ctx& c, int m, int n, int k, const double* A,
                             int lda, const double* B, int ldb, double* C,
                             int ldc) {
  const int   ldt = ut_ld(n);
  double*     Bt  = panel_scratch(c);
  hipStream_t s   = panel_stream(c);
  // rows the update may read below A: m rounded up to even (its row pairs
  // must keep their alignment).  Row m of A's k columns is readable: each
  // of them is followed by a further column of the panel (B's first
  // column), so the extra element lies inside the panel's storage.
  const int a_rows = (m + 1) & ~1;
  if(n >= 128 && static_cast<size_t>(k) * ldt <= static_cast<size_t>(kMaxNB) * kMaxNB &&
     update_fast_ok(m, n, k, A, lda, Bt, ldt, a_rows, ldt)) {
    hipError_t e = launch_transpose(s, k, n, B, ldb, Bt, ldt);
    if(e != hipSuccess) return e;
    return launch_update(s, m, n, k, A, lda, Bt, ldt, C, ldc, a_rows, ldt,
                         true);
  }
  return launch_dgemm(s, m, n, k, -1.0, A, lda, B, ldb, 1.0, C, ldc, false);
}
inline hipError_t panel_gemm(Ctx& c, int m, int n, int k, const float* A,
                             int lda, const float* B, int ldb, float* C,
                             int ldc) {
  return launch_sgemm(panel_stream(c), m, n, k, -1.0, A, lda, B, ldb, 1.0, C,
                      ldc, false);
}

// One level of the blocked panel loop: columns [c0, c0 + n) of the panel,
// rows [c0, m), in chunks of nb columns, each factored by factor(top, jb)
// (top = the chunk's first row and column), in the variant's order:
//   right (src/pfact/HPL_pdrpanrlT.cpp:116-160): factor the chunk, then its
//     U rows right of it (L11^{-1} A12) and the rows below (A22 -= L21 U12);
//   left (src/pfact/HPL_pdrpanllT.cpp:116-159): first the chunk's rows above
//     it (U = L^{-1} A over the jj columns left of it in this level) and its
//     rows from its top down (A -= L U), then factor it;
//   Crout (src/pfact/HPL_pdrpancrT.cpp:116-175): the chunk's rows from its
//     top down (A -= L U with the final U rows above), factor it, then its
//     own U rows right of it (A -= L U over the jj columns, then L11^{-1}).
// Row interchanges inside factor() span the whole panel width, so columns
// left and right of the chunk are permuted before they are used.
template <typename V, typename F>
hipError_t chunk_loop(Ctx& c, int variant, V* P, int lda, int m, int c0,
                      int n, int nb, F&& factor) {
  auto at = [&](int r, int col) {
    return P + r + static_cast<size_t>(col) * lda;
  };
  hipStream_t s = panel_stream(c);
  // X (k x nx at xr, xc) = L^{-1} X for the unit lower L at (tr, tc)
  auto trsm = [&](int k, int nx, int tr, int tc, int xr, int xc) {
    if(c.panel_rec) {
      PanelOp o{};
      o.kind = kPanelTrsm;
      o.k = k, o.n = nx, o.ar = tr, o.ac = tc, o.br = xr, o.bc = xc;
      c.panel_rec->push_back(o);
      return hipSuccess;
    }
    return launch_trsm_colsweep(s, k, nx, at(tr, tc), lda, at(xr, xc), lda);
  };
  // C -= A B with rows = m_fix (>= 0) or m - m_sub, n columns, depth k
  auto gemm = [&](int m_fix, int m_sub, int nc, int k, int ar, int ac, int br,
                  int bc, int cr, int cc) {
    if(c.panel_rec) {
      PanelOp o{};
      o.kind = kPanelGemm;
      o.m_fix = m_fix, o.m_sub = m_sub, o.n = nc, o.k = k;
      o.ar = ar, o.ac = ac, o.br = br, o.bc = bc, o.cr = cr, o.cc = cc;
      c.panel_rec->push_back(o);
      return hipSuccess;
    }
    return panel_gemm(c, m_fix >= 0 ? m_fix : m - m_sub, nc, k, at(ar, ac),
                      lda, at(br, bc), lda, at(cr, cc), lda);
  };
  int jj = 0, jb = nb < n ? nb : n;
  while(jj < n) {
    const int  top  = c0 + jj;
    const int  nrem = n - jj - jb;
    hipError_t e;
    if(jj > 0 && variant == HPLX_LEFT_LOOKING) {
      e = trsm(jj, jb, c0, c0, c0, top);
      if(e != hipSuccess) return e;
    }
    if(jj > 0 && variant != HPLX_RIGHT_LOOKING) {
      e = gemm(-1, top, jb, jj, top, c0, c0, top, top, top);
      if(e != hipSuccess) return e;
    }
    e = factor(top, jb);
    if(e != hipSuccess) return e;
    if(nrem > 0 && variant == HPLX_RIGHT_LOOKING) {
      e = trsm(jb, nrem, top, top, top, top + jb);
      if(e != hipSuccess) return e;
      e = gemm(-1, top + jb, nrem, jb, top + jb, top, top, top + jb, top + jb,
               top + jb);
      if(e != hipSuccess) return e;
    } else if(nrem > 0 && variant == HPLX_CROUT) {
      if(jj > 0) {
        e = gemm(jb, 0, nrem, jj, top, c0, c0, top + jb, top, top + jb);
        if(e != hipSuccess) return e;
      }
      e = trsm(jb, nrem, top, top, top, top + jb);
      if(e != hipSuccess) return e;
    }
    jj += jb;
    jb = (n - jj) < nb ? (n - jj) : nb;
  }
  return hipSuccess;
}

// Columns [c0, c0+n) of the panel, rows [c0, m), n <= nbmin: the panel leaf
// (HPL_pdpanrlT / HPL_pdpanllT / HPL_pdpancrT).  The register leaf factors up
// to kLeafMax columns in one kernel; it runs the right-looking schedule for
// all three pfact variants: within a leaf they give every element the same
// products in the same order (k ascending) and differ only in when each is
// applied, so the right-looking leaf computes the left-looking and Crout
// leaves' results exactly.  Wider leaves (nbmin > kLeafMax) are blocked into
// kLeafMax-column leaves in the pfact variant's order.
template <typename V>
hipError_t leaf_or_chunks(Ctx& c, int pfact, V* P, int lda, int m, int JB,
                          int c0, int n, int* ipiv, int row_offset,
                          int info_base) {
  auto leaf = [&](int top, int w) {
    if(c.panel_rec) {
      PanelOp o{};
      o.kind = kPanelLeaf;
      o.ac = top, o.n = w;
      c.panel_rec->push_back(o);
      return hipSuccess;
    }
    return launch_leaf(c, P, lda, m, JB, top, w, ipiv, row_offset, info_base);
  };
  if(n <= kLeafMax) return leaf(c0, n);
  return chunk_loop(c, pfact, P, lda, m, c0, n, kLeafMax, leaf);
}

// The recursive panel factorisation (HPL_pdrpan{rl,ll,cr}T): chunks of
// nb = ceil(ceil(n / nbmin) / nbdiv) * nbmin columns (HPL_pdrpanrlT.cpp:93),
// each factored recursively, in the rfact variant's order.
template <typename V>
hipError_t rpanel(Ctx& c, const PanelAlgo& a, V* P, int lda, int m, int JB,
                  int c0, int n, int* ipiv, int row_offset, int info_base) {
  if(n <= a.nbmin)
    return leaf_or_chunks(c, a.pfact, P, lda, m, JB, c0, n, ipiv, row_offset,
                          info_base);
  const int nb = (((n + a.nbmin - 1) / a.nbmin + a.nbdiv - 1) / a.nbdiv) * a.nbmin;
  return chunk_loop(c, a.rfact, P, lda, m, c0, n, nb, [&](int top, int jb) {
    return rpanel(c, a, P, lda, m, JB, top, jb, ipiv, row_offset, info_base);
  });
}

PanelAlgo normalized(PanelAlgo a) {
  if(a.nbmin < 1) a.nbmin = 1;
  if(a.nbdiv < 2) a.nbdiv = 2;
  return a;
}

// The device slot holding the recorded steps of (jb, a), or -1.
int recorded_slot(const Ctx& c, const PanelAlgo& a, int jb) {
  const int key[5] = {jb, a.nbmin, a.nbdiv, a.pfact, a.rfact};
  for(int s = 0; s < 2; ++s) {
    if(!c.pop_count[s]) continue;
    bool same = true;
    for(int i = 0; i < 5; ++i) same = same && c.pop_key[s][i] == key[i];
    if(same) return s;
  }
  return -1;
}

template <typename V>
hipError_t panel_factor_t(Ctx& c, PanelAlgo a, V* P, int lda, int m, int jb,
                          int* ipiv, int row_offset) {
  a = normalized(a);
  if(!valid_fact(a.pfact) || !valid_fact(a.rfact)) return hipErrorInvalidValue;
  if(jb > kMaxNB || jb > m) return hipErrorInvalidValue;
  if(m > max_leaf_rows(c)) return hipErrorInvalidValue;
  if constexpr(sizeof(V) == 8) {
    if(!c.panel_rec && (c.panel_kernel == 1 ||
                        (c.panel_kernel == 2 && m >= kPanelKernelMinRows))) {
      const int slot = recorded_slot(c, a, jb);
      if(slot >= 0) {
        bool       ran = false;
        hipError_t e   = launch_panel_kernel(c, slot, P, lda, m, jb, ipiv,
                                             row_offset, &ran);
        if(e != hipSuccess) return e;
        if(ran) {
          ++c.panel_kernel_runs;
          return hipSuccess;
        }
      }
    }
  }
  return rpanel(c, a, P, lda, m, jb, 0, jb, ipiv, row_offset, row_offset);
}

}  // namespace

// The steps of a panel of width jb under a (host only: in recording mode the
// recursion queues nothing).
int record_panel_steps(Ctx& c, const PanelAlgo& a, int jb,
                       std::vector<PanelOp>* ops) {
  c.panel_rec  = ops;
  hipError_t e = rpanel(c, a, static_cast<double*>(nullptr), 0, jb, jb, 0, jb,
                        nullptr, 0, 0);
  c.panel_rec  = nullptr;
  return e == hipSuccess ? HPLX_SUCCESS : HPLX_ERR_HIP;
}

int panel_kernel_prepare(Ctx& c, const PanelAlgo& a0, int jb) {
  PanelAlgo a = normalized(a0);
  if(!valid_fact(a.pfact) || !valid_fact(a.rfact) || jb < 1 || jb > kMaxNB)
    return HPLX_ERR_INVALID;
  // a list in use stays: the next recording replaces the other slot
  const int have = recorded_slot(c, a, jb);
  if(have >= 0) {
    c.pop_next = have ^ 1;
    return HPLX_SUCCESS;
  }
  std::vector<PanelOp> ops;
  if(int r = record_panel_steps(c, a, jb, &ops)) return r;
  if(ops.size() > static_cast<size_t>(kMaxPanelOps)) return HPLX_SUCCESS;
  const int slot = c.pop_next;
  c.pop_next ^= 1;
  c.pop_count[slot] = 0;
  if(hipMemcpy(c.pops + static_cast<size_t>(slot) * kMaxPanelOps, ops.data(),
               ops.size() * sizeof(PanelOp), hipMemcpyHostToDevice) != hipSuccess)
    return HPLX_ERR_HIP;
  int nleaf = 0;
  for(const PanelOp& o : ops) nleaf += o.kind == kPanelLeaf;
  const int key[5] = {jb, a.nbmin, a.nbdiv, a.pfact, a.rfact};
  for(int i = 0; i < 5; ++i) c.pop_key[slot][i] = key[i];
  c.pop_leaves[slot] = nleaf;
  c.pop_count[slot]  = static_cast<int>(ops.size());
  c.pop_host[slot]   = std::move(ops);
  return HPLX_SUCCESS;
}

int panel_steps(const PanelAlgo& a0, int jb, int* out, int max_steps) {
  const PanelAlgo a = normalized(a0);
  if(!valid_fact(a.pfact) || !valid_fact(a.rfact) || jb < 1 || jb > kMaxNB ||
     max_steps < 0 || (max_steps > 0 && !out))
    return HPLX_ERR_INVALID;
  Ctx                  c;  // recording only: no device state is touched
  std::vector<PanelOp> ops;
  if(int r = record_panel_steps(c, a, jb, &ops)) return r;
  const int n = static_cast<int>(ops.size());
  for(int i = 0; i < n && i < max_steps; ++i) {
    const PanelOp& o = ops[i];
    const int      v[kPanelStepInts] = {o.kind, o.m_fix, o.m_sub, o.n, o.k, o.ar,
                                        o.ac, o.br, o.bc, o.cr, o.cc};
    for(int f = 0; f < kPanelStepInts; ++f) out[i * kPanelStepInts + f] = v[f];
  }
  return n;
}

hipError_t leaf_occupancy(int* blocks_per_cu) {
  // the smaller of the FP64 and FP32 leaves' occupancies
  int        b64 = 0, b32 = 0;
  hipError_t e   = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &b64, leaf_kernel_ptr<LeafKind::kF64>(), kLeafThreads, 0);
  if(e != hipSuccess) return e;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &b32, leaf_kernel_ptr<LeafKind::kF32>(), kLeafThreads, 0);
  if(e != hipSuccess) return e;
  *blocks_per_cu = b64 < b32 ? b64 : b32;
  return hipSuccess;
}

hipError_t reset_panel_state(Ctx& c) {
  hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(1), 0, panel_stream(c), c.flags,
                     c.flags + 1);
  return hipGetLastError();
}

hipError_t panel_factor(Ctx& c, const PanelAlgo& a, double* P, int lda,
                        int m, int jb, int* ipiv, int row_offset) {
  return panel_factor_t(c, a, P, lda, m, jb, ipiv, row_offset);
}

hipError_t panel_factor(Ctx& c, const PanelAlgo& a, float* P, int lda,
                        int m, int jb, int* ipiv, int row_offset) {
  return panel_factor_t(c, a, P, lda, m, jb, ipiv, row_offset);
}

hipError_t launch_pipid(hipStream_t s, int jb, const int* ipiv, int ipiv_base,
                        int* permU) {
  if(jb <= 0) return hipSuccess;
  if(jb > kMaxNB) return hipErrorInvalidValue;
  hipLaunchKernelGGL(pipid_kernel, dim3(1), dim3(1024), 0, s, jb, ipiv,
                     ipiv_base, permU);
  return hipGetLastError();
}

template <typename S, typename WT>
hipError_t swap_gather_t(hipStream_t s, int jb, int n, S* A, int lda,
                         const int* permU, WT* W, int ldw) {
  if(jb <= 0 || n <= 0) return hipSuccess;
  if(jb > kMaxNB) return hipErrorInvalidValue;
  hipLaunchKernelGGL((swap_gather_kernel<S, WT>), dim3(n), dim3(256), 0, s, jb,
                     A, lda, permU, W, ldw);
  return hipGetLastError();
}

hipError_t launch_swap_gather(hipStream_t s, int jb, int n, double* A, int lda,
                              const int* permU, double* W, int ldw) {
  return swap_gather_t(s, jb, n, A, lda, permU, W, ldw);
}

hipError_t launch_swap_gather(hipStream_t s, int jb, int n, float* A, int lda,
                              const int* permU, float* W, int ldw) {
  return swap_gather_t(s, jb, n, A, lda, permU, W, ldw);
}

hipError_t launch_swap_gather(hipStream_t s, int jb, int n, __half* A,
                              int lda, const int* permU, float* W, int ldw) {
  return swap_gather_t(s, jb, n, A, lda, permU, W, ldw);
}

hipError_t launch_swap_inverse(hipStream_t s, int jb, int n, double* A,
                               int lda, const int* permU) {
  if(jb <= 0 || n <= 0) return hipSuccess;
  if(jb > kMaxNB) return hipErrorInvalidValue;
  hipLaunchKernelGGL(swap_inverse_kernel, dim3(n), dim3(256), 0, s, jb, A, lda,
                     permU);
  return hipGetLastError();
}

hipError_t launch_swap_inplace(hipStream_t s, int jb, int n, double* A,
                               int lda, const int* permU) {
  if(jb <= 0 || n <= 0) return hipSuccess;
  if(jb > kMaxNB) return hipErrorInvalidValue;
  hipLaunchKernelGGL(swap_inplace_kernel, dim3(n), dim3(256), 0, s, jb, A, lda,
                     permU);
  return hipGetLastError();
}

}  // namespace hplx
