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

__global__ void reset_kernel(uint32_t* counter, int* info, int* err) {
  *counter = 0;
  *info    = 0;
  *err     = 0;
}

// Solve L X = B in place; L unit lower k x k.  One workgroup per column of
// B; with `identity` B starts as the identity (explicit inverse of L).
__global__ __launch_bounds__(256) void trsm_colsweep_kernel(
    int k, const double* __restrict__ L, int ldl, double* __restrict__ B,
    int ldb, int identity) {
  __shared__ double sx[kMaxNB];
  const int col = blockIdx.x;
  double*   bc  = B + static_cast<size_t>(col) * ldb;
  for(int i = threadIdx.x; i < k; i += blockDim.x)
    sx[i] = identity ? (i == col ? 1.0 : 0.0) : bc[i];
  __syncthreads();
  const int p0 = identity ? col : 0;
  for(int p = p0; p < k - 1; ++p) {
    const double xp = sx[p];
    if(xp != 0.0) {
      const double* lc = L + static_cast<size_t>(p) * ldl;
      for(int i = p + 1 + threadIdx.x; i < k; i += blockDim.x)
        sx[i] = fma(-lc[i], xp, sx[i]);
    }
    __syncthreads();
  }
  for(int i = threadIdx.x; i < k; i += blockDim.x) bc[i] = sx[i];
}

// Device HPL_pipid: compose jb interchanges (row i <-> ipiv[i]-base) into
// permU[0:jb] (source of each top row) and permU[jb:2jb] (destination of
// each displaced top row, or -1).
__global__ __launch_bounds__(1024) void pipid_kernel(
    int jb, const int* __restrict__ ipiv, int base, int* __restrict__ permU) {
  __shared__ int sp[kMaxNB];
  __shared__ int sfirst[kMaxNB];
  __shared__ int slab[kMaxNB];
  __shared__ int splab[kMaxNB];
  __shared__ int sout[2 * kMaxNB];
  const int t = threadIdx.x;
  // an interchange must target a row at or below its own; anything else is
  // treated as no interchange so that every LDS index stays in range
  for(int i = t; i < jb; i += blockDim.x) {
    const int p = ipiv[i] - base;
    sp[i]       = p < i ? i : p;
  }
  __syncthreads();
  // slot of an outside position = index of its first occurrence
  for(int i = t; i < jb; i += blockDim.x) {
    int f = -1;
    if(sp[i] >= jb) {
      f = i;
      for(int k = 0; k < i; ++k)
        if(sp[k] == sp[i]) {
          f = k;
          break;
        }
    }
    sfirst[i] = f;
  }
  __syncthreads();
  if(t == 0) {
    for(int i = 0; i < jb; ++i) {
      slab[i] = i;
      if(sfirst[i] == i) splab[i] = sp[i];
    }
    for(int i = 0; i < jb; ++i) {
      const int p = sp[i];
      if(p < jb) {
        const int x = slab[i];
        slab[i]     = slab[p];
        slab[p]     = x;
      } else {
        const int s = sfirst[i];
        const int x = slab[i];
        slab[i]     = splab[s];
        splab[s]    = x;
      }
    }
  }
  __syncthreads();
  for(int i = t; i < jb; i += blockDim.x) {
    sout[i]      = slab[i];
    sout[jb + i] = -1;
  }
  __syncthreads();
  // positions >= jb only ever receive rows that started in the top jb
  for(int i = t; i < jb; i += blockDim.x)
    if(sfirst[i] == i) sout[jb + splab[i]] = sp[i];
  __syncthreads();
  for(int i = t; i < 2 * jb; i += blockDim.x) permU[i] = sout[i];
}

// New top rows -> W (ld ldw); displaced top rows -> their destinations.
__global__ __launch_bounds__(256) void swap_gather_kernel(
    int jb, double* __restrict__ A, int lda, const int* __restrict__ permU,
    double* __restrict__ W, int ldw) {
  constexpr int Q   = kMaxNB / 256;
  const int     col = blockIdx.x;
  double*       ac  = A + static_cast<size_t>(col) * lda;
  double*       wc  = W + static_cast<size_t>(col) * ldw;
  double        keep[Q];
  int           dst[Q];
#pragma unroll
  for(int q = 0; q < Q; ++q) {
    const int i = threadIdx.x + 256 * q;
    dst[q]      = -1;
    keep[q]     = 0.0;
    if(i < jb) {
      wc[i]  = ac[permU[i]];  // the store waits for the gathered value
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

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

hipError_t launch_leaf(Ctx& c, double* P, int lda, int m,
                       int JB, int c0, int w, int* ipiv, int row_offset,
                       int info_base) {
  const int R    = m - c0;
  const int Gmax = c.num_cus < kMaxGrid ? c.num_cus : kMaxGrid;
  int       rpt  = 1;
  while(rpt < kLeafMaxRPT &&
        (R + kLeafThreads * rpt - 1) / (kLeafThreads * rpt) > Gmax)
    rpt *= 2;
  const int G = (R + kLeafThreads * rpt - 1) / (kLeafThreads * rpt);
  if(G > Gmax || G < 1 || w < 1 || w > kLeafMax) return hipErrorInvalidValue;
  leaf::LeafArgs la;
  la.P          = P;
  la.lda        = lda;
  la.m          = m;
  la.JB         = JB;
  la.c0         = c0;
  la.w          = w;
  la.ipiv       = ipiv;
  la.row_offset = row_offset;
  la.info       = c.flags;
  la.info_base  = info_base;
  la.slots      = c.slots;
  la.counter    = c.counter;
  la.base       = static_cast<uint32_t>(c.barrier_epoch);
  la.err        = c.flags + 1;
  void*       args[] = {&la};
  const void* fn     = rpt == 1   ? leaf_kernel_ptr<1>()
                       : rpt == 2 ? leaf_kernel_ptr<2>()
                                  : leaf_kernel_ptr<4>();
  hipError_t e = hipLaunchCooperativeKernel(fn, dim3(G), dim3(kLeafThreads),
                                            args, 0, c.stream);
  if(e == hipSuccess)
    c.barrier_epoch += static_cast<uint64_t>(w) * static_cast<uint64_t>(G);
  return e;
}

// Columns [c0, c0+n) of the panel, rows [c0, m).
hipError_t leaf_or_chunks(Ctx& c, double* P, int lda, int m,
                          int JB, int c0, int n, int* ipiv, int row_offset,
                          int info_base) {
  if(n <= kLeafMax)
    return launch_leaf(c, P, lda, m, JB, c0, n, ipiv, row_offset,
                       info_base);
  // nbmin wider than the register leaf: right-looking over 32-column leaves
  for(int k = 0; k < n; k += kLeafMax) {
    const int w = (n - k) < kLeafMax ? (n - k) : kLeafMax;
    hipError_t e = launch_leaf(c, P, lda, m, JB, c0 + k, w, ipiv,
                               row_offset, info_base);
    if(e != hipSuccess) return e;
    const int top  = c0 + k;
    const int rest = n - k - w;
    if(rest > 0) {
      double* L11 = P + top + static_cast<size_t>(top) * lda;
      double* A12 = P + top + static_cast<size_t>(top + w) * lda;
      e = launch_trsm_colsweep(c.stream, w, rest, L11, lda, A12, lda);
      if(e != hipSuccess) return e;
      e = launch_dgemm(c.stream, m - top - w, rest, w, -1.0, L11 + w, lda, A12,
                       lda, 1.0, A12 + w, lda, false);
      if(e != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

hipError_t rpanel(Ctx& c, int nbmin, int nbdiv, double* P,
                  int lda, int m, int JB, int c0, int n, int* ipiv,
                  int row_offset, int info_base) {
  if(n <= nbmin)
    return leaf_or_chunks(c, P, lda, m, JB, c0, n, ipiv, row_offset,
                          info_base);
  const int nb = (((n + nbmin - 1) / nbmin + nbdiv - 1) / nbdiv) * nbmin;
  int       jj = 0, jb = nb, nrem = n;
  do {
    nrem -= jb;
    hipError_t e = rpanel(c, nbmin, nbdiv, P, lda, m, JB, c0 + jj, jb,
                          ipiv, row_offset, info_base);
    if(e != hipSuccess) return e;
    const int top = c0 + jj;
    if(nrem > 0) {
      double* L11 = P + top + static_cast<size_t>(top) * lda;
      double* A12 = P + top + static_cast<size_t>(top + jb) * lda;
      e = launch_trsm_colsweep(c.stream, jb, nrem, L11, lda, A12, lda);
      if(e != hipSuccess) return e;
      e = launch_dgemm(c.stream, m - top - jb, nrem, jb, -1.0, L11 + jb, lda,
                       A12, lda, 1.0, A12 + jb, lda, false);
      if(e != hipSuccess) return e;
    }
    jj += jb;
    jb = nrem < nb ? nrem : nb;
  } while(nrem > 0);
  return hipSuccess;
}

}  // namespace

hipError_t reset_panel_state(Ctx& c) {
  hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(1), 0, c.stream, c.counter,
                     c.flags, c.flags + 1);
  c.barrier_epoch = 0;
  return hipGetLastError();
}

hipError_t panel_factor(Ctx& c, int nbmin, int nbdiv, double* P, int lda,
                        int m, int jb, int* ipiv, int row_offset) {
  if(nbmin < 1) nbmin = 1;
  if(nbdiv < 2) nbdiv = 2;
  if(jb > kMaxNB || jb > m) return hipErrorInvalidValue;
  if(m > c.num_cus * kLeafThreads * kLeafMaxRPT) return hipErrorInvalidValue;
  return rpanel(c, nbmin, nbdiv, P, lda, m, jb, 0, jb, ipiv, row_offset,
                row_offset);
}

hipError_t launch_pipid(hipStream_t s, int jb, const int* ipiv, int ipiv_base,
                        int* permU) {
  if(jb <= 0) return hipSuccess;
  if(jb > kMaxNB) return hipErrorInvalidValue;
  hipLaunchKernelGGL(pipid_kernel, dim3(1), dim3(1024), 0, s, jb, ipiv,
                     ipiv_base, permU);
  return hipGetLastError();
}

hipError_t launch_swap_gather(hipStream_t s, int jb, int n, double* A, int lda,
                              const int* permU, double* W, int ldw) {
  if(jb <= 0 || n <= 0) return hipSuccess;
  if(jb > kMaxNB) return hipErrorInvalidValue;
  hipLaunchKernelGGL(swap_gather_kernel, dim3(n), dim3(256), 0, s, jb, A, lda,
                     permU, W, ldw);
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

hipError_t launch_linv(hipStream_t s, int jb, const double* L, int ldl,
                       double* Linv, int ldi) {
  if(jb <= 0) return hipSuccess;
  if(jb > kMaxNB) return hipErrorInvalidValue;
  hipLaunchKernelGGL(trsm_colsweep_kernel, dim3(jb), dim3(256), 0, s, jb, L,
                     ldl, Linv, ldi, 1);
  return hipGetLastError();
}

hipError_t launch_trsm_colsweep(hipStream_t s, int k, int n, const double* L,
                                int ldl, double* B, int ldb) {
  if(k <= 0 || n <= 0) return hipSuccess;
  if(k > kMaxNB) return hipErrorInvalidValue;
  hipLaunchKernelGGL(trsm_colsweep_kernel, dim3(n), dim3(256), 0, s, k, L, ldl,
                     B, ldb, 0);
  return hipGetLastError();
}

}  // namespace hplx
