// Register-resident cooperative leaf of the panel factorisation
// (HPL_pdpanrlT, src/pfact/HPL_pdpanrlT_device.cpp:23-468): one launch per
// leaf of <= 32 columns, one grid barrier per column, pivot search by wave64
// shuffles and LDS, no host round-trip.  Included by k_leaf{1,2,4}.hip, each
// of which instantiates one rows-per-thread variant (the fully unrolled
// column steps make each instance a large compile unit).
#pragma once

#include <climits>

#include "hplx_internal.hpp"

namespace hplx {
namespace leaf {

constexpr uint32_t kSpinLimit = 1u << 22;

__device__ inline bool better(double v, int r, double bv, int br) {
  return v > bv || (v == bv && r < br);
}

__device__ inline void st_agent(double* p, double v) {
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(p),
                     static_cast<unsigned long long>(__double_as_longlong(v)),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline double ld_agent(const double* p) {
  return __longlong_as_double(static_cast<long long>(__hip_atomic_load(
      reinterpret_cast<unsigned long long*>(const_cast<double*>(p)),
      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
}

// Grid-wide barrier over a monotonically increasing arrival counter.  Every
// wait is bounded: after kSpinLimit polls the workgroup raises *err and all
// workgroups leave the kernel, so a grid that is not co-resident cannot hang
// the device.  Returns false once *err is set.
__device__ inline bool grid_sync(uint32_t* counter, uint32_t target, int* err,
                          int* s_flag) {
  __threadfence();
  __syncthreads();
  if(threadIdx.x == 0) {
    __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_AGENT);
    int      ok    = 1;
    uint32_t spins = 0;
    while(static_cast<int32_t>(__hip_atomic_load(counter, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT) -
                               target) < 0) {
      ++spins;
      if((spins & 1023u) == 0 &&
         __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        ok = 0;
        break;
      }
      if(spins > kSpinLimit) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    *s_flag = ok;
  }
  __syncthreads();
  return *s_flag != 0;
}

struct LeafArgs {
  double*   P;          // panel top-left (panel row 0)
  int       lda;
  int       m;          // panel rows
  int       JB;         // panel width (row interchanges span it)
  int       c0;         // first leaf column (panel relative)
  int       w;          // leaf width, <= 32
  int*      ipiv;       // panel-relative pivots (+ row_offset) written here
  int       row_offset;
  int*      info;       // first exact zero pivot, 1-based global column
  int       info_base;  // global column of panel column 0
  double*   slots;      // 2 x (kMaxGrid+1) x kSlotStride
  uint32_t* counter;
  uint32_t  base;       // counter value when this launch starts
  int*      err;
};


// Shared-memory scratch of the leaf kernel.
struct LeafShared {
  double u[kLeafMax];
  double top[kLeafMax];
  double wv[kLeafThreads / 64];
  int    wr[kLeafThreads / 64];
  int    piv[kLeafMax];
  int    flag;
  int    lab[kLeafMax];
  int    ppos[kLeafMax];
  int    plab[kLeafMax];
  int    dst[2 * kLeafMax];
  int    src[2 * kLeafMax];
  int    np;
};

// One pivot step of the leaf for compile-time column JJ, so that every
// register of the row tile is addressed statically (no scratch spill).
template <int RPT, int JJ>
__device__ __forceinline__ void leaf_step(const LeafArgs& a,
                                          double (&v)[RPT][kLeafMax],
                                          const int (&rows)[RPT],
                                          LeafShared& sh, bool& alive) {
  constexpr int CH   = kLeafThreads * RPT;
  const int     t    = threadIdx.x;
  const int     lane = t & 63;
  const int     wave = t >> 6;
  const int     b    = blockIdx.x;
  const int     G    = gridDim.x;
  const int     gc   = a.c0 + JJ;  // current top row == current column
  // ---- workgroup candidate: max |a(r, gc)| over rows r >= gc
  double bv = -1.0;
  int    br = INT_MAX;
#pragma unroll
  for(int q = 0; q < RPT; ++q) {
    if(rows[q] >= gc && rows[q] < a.m) {
      const double x = fabs(v[q][JJ]);
      if(better(x, rows[q], bv, br)) {
        bv = x;
        br = rows[q];
      }
    }
  }
#pragma unroll
  for(int o = 32; o > 0; o >>= 1) {
    const double ov  = __shfl_xor(bv, o);
    const int    orr = __shfl_xor(br, o);
    if(better(ov, orr, bv, br)) {
      bv = ov;
      br = orr;
    }
  }
  if(lane == 0) {
    sh.wv[wave] = bv;
    sh.wr[wave] = br;
  }
  __syncthreads();
  bv = sh.wv[0];
  br = sh.wr[0];
#pragma unroll
  for(int k = 1; k < kLeafThreads / 64; ++k)
    if(better(sh.wv[k], sh.wr[k], bv, br)) {
      bv = sh.wv[k];
      br = sh.wr[k];
    }
  // ---- publish candidate (value, row, leaf segment) and the top row
  const size_t par   = static_cast<size_t>(JJ & 1) * (kMaxGrid + 1);
  double*      slot  = a.slots + (par + b) * kSlotStride;
  double*      tslot = a.slots + (par + kMaxGrid) * kSlotStride;
  if(t == 0) {
    st_agent(slot + 0, bv);
    st_agent(slot + 1, static_cast<double>(br));
  }
#pragma unroll
  for(int q = 0; q < RPT; ++q) {
    if(rows[q] == br) {
#pragma unroll
      for(int c = 0; c < kLeafMax; ++c) st_agent(slot + 2 + c, v[q][c]);
    }
    if(rows[q] == gc) {
#pragma unroll
      for(int c = 0; c < kLeafMax; ++c) st_agent(tslot + 2 + c, v[q][c]);
    }
  }
  alive = grid_sync(a.counter,
                    a.base + static_cast<uint32_t>(JJ + 1) *
                                 static_cast<uint32_t>(G),
                    a.err, &sh.flag);
  if(!alive) return;
  // ---- every workgroup reduces all candidates itself
  double cv = -1.0;
  int    cr = INT_MAX;
  for(int i = t; i < G; i += kLeafThreads) {
    const double* si = a.slots + (par + i) * kSlotStride;
    const double  x  = ld_agent(si);
    const int     r  = static_cast<int>(ld_agent(si + 1));
    if(better(x, r, cv, cr)) {
      cv = x;
      cr = r;
    }
  }
#pragma unroll
  for(int o = 32; o > 0; o >>= 1) {
    const double ov  = __shfl_xor(cv, o);
    const int    orr = __shfl_xor(cr, o);
    if(better(ov, orr, cv, cr)) {
      cv = ov;
      cr = orr;
    }
  }
  if(lane == 0) {
    sh.wv[wave] = cv;
    sh.wr[wave] = cr;
  }
  __syncthreads();
  cv = sh.wv[0];
  cr = sh.wr[0];
#pragma unroll
  for(int k = 1; k < kLeafThreads / 64; ++k)
    if(better(sh.wv[k], sh.wr[k], cv, cr)) {
      cv = sh.wv[k];
      cr = sh.wr[k];
    }
  const int     p  = cr;
  const int     wb = (p - a.c0) / CH;
  const double* ws = a.slots + (par + wb) * kSlotStride;
  if(t < kLeafMax)
    sh.u[t] = ld_agent(ws + 2 + t);
  else if(t < 2 * kLeafMax)
    sh.top[t - kLeafMax] = ld_agent(tslot + 2 + (t - kLeafMax));
  if(t == 0) sh.piv[JJ] = p;
  __syncthreads();
  const double piv = sh.u[JJ];
  if(b == 0 && t == 0) {
    a.ipiv[gc] = p + a.row_offset;
    if(piv == 0.0 && *a.info == 0) *a.info = a.info_base + gc + 1;
  }
  // ---- interchange rows gc and p (leaf columns; registers)
#pragma unroll
  for(int q = 0; q < RPT; ++q) {
    if(rows[q] == gc) {
#pragma unroll
      for(int c = 0; c < kLeafMax; ++c) v[q][c] = sh.u[c];
    } else if(rows[q] == p) {
#pragma unroll
      for(int c = 0; c < kLeafMax; ++c) v[q][c] = sh.top[c];
    }
  }
  // ---- scale column gc and rank-1 update the rest of the leaf
#pragma unroll
  for(int q = 0; q < RPT; ++q) {
    if(rows[q] > gc && rows[q] < a.m) {
      double l = v[q][JJ];
      if(piv != 0.0) l = l / piv;
      v[q][JJ] = l;
#pragma unroll
      for(int c = JJ + 1; c < kLeafMax; ++c) v[q][c] = fma(-l, sh.u[c], v[q][c]);
    }
  }
  __syncthreads();  // sh.u / sh.top / sh.wv are rewritten next column
}

template <int RPT, int JJ>
struct LeafSteps {
  __device__ __forceinline__ static void run(const LeafArgs& a,
                                             double (&v)[RPT][kLeafMax],
                                             const int (&rows)[RPT],
                                             LeafShared& sh, bool& alive) {
    if(JJ < a.w && alive) leaf_step<RPT, JJ>(a, v, rows, sh, alive);
    LeafSteps<RPT, JJ + 1>::run(a, v, rows, sh, alive);
  }
};

template <int RPT>
struct LeafSteps<RPT, kLeafMax> {
  __device__ __forceinline__ static void run(const LeafArgs&,
                                             double (&)[RPT][kLeafMax],
                                             const int (&)[RPT], LeafShared&,
                                             bool&) {}
};

template <int RPT>
__global__ __launch_bounds__(kLeafThreads) void leaf_kernel(LeafArgs a) {
  constexpr int CH = kLeafThreads * RPT;  // rows per workgroup
  const int     t    = threadIdx.x;
  const int     lane = t & 63;
  const int     wave = t >> 6;
  const int     b    = blockIdx.x;
  const int     G    = gridDim.x;

  __shared__ LeafShared sh;

  if(__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    return;

  int    rows[RPT];
  double v[RPT][kLeafMax];
#pragma unroll
  for(int q = 0; q < RPT; ++q) rows[q] = a.c0 + b * CH + q * kLeafThreads + t;
#pragma unroll
  for(int c = 0; c < kLeafMax; ++c) {
    const double* col = a.P + static_cast<size_t>(a.c0 + c) * a.lda;
#pragma unroll
    for(int q = 0; q < RPT; ++q)
      v[q][c] = (c < a.w && rows[q] < a.m) ? col[rows[q]] : 0.0;
  }

  bool alive = true;
  LeafSteps<RPT, 0>::run(a, v, rows, sh, alive);
  if(!alive) return;

  // ---- write the leaf back
#pragma unroll
  for(int c = 0; c < kLeafMax; ++c) {
    if(c < a.w) {
      double* col = a.P + static_cast<size_t>(a.c0 + c) * a.lda;
#pragma unroll
      for(int q = 0; q < RPT; ++q)
        if(rows[q] < a.m) col[rows[q]] = v[q][c];
    }
  }

  // ---- compose the w interchanges and apply them to the other panel columns
  if(t == 0) {
    const int top_end = a.c0 + a.w;
    int       nout    = 0;
    for(int i = 0; i < a.w; ++i) sh.lab[i] = a.c0 + i;
    for(int jj = 0; jj < a.w; ++jj) {
      const int p = sh.piv[jj];
      if(p < top_end) {
        const int x         = sh.lab[jj];
        sh.lab[jj]           = sh.lab[p - a.c0];
        sh.lab[p - a.c0]     = x;
      } else {
        int s = -1;
        for(int k = 0; k < nout; ++k)
          if(sh.ppos[k] == p) {
            s = k;
            break;
          }
        if(s < 0) {
          s         = nout++;
          sh.ppos[s] = p;
          sh.plab[s] = p;
        }
        const int x = sh.lab[jj];
        sh.lab[jj]   = sh.plab[s];
        sh.plab[s]   = x;
      }
    }
    int np = 0;
    for(int i = 0; i < a.w; ++i)
      if(sh.lab[i] != a.c0 + i) {
        sh.dst[np] = a.c0 + i;
        sh.src[np] = sh.lab[i];
        ++np;
      }
    for(int k = 0; k < nout; ++k)
      if(sh.plab[k] != sh.ppos[k]) {
        sh.dst[np] = sh.ppos[k];
        sh.src[np] = sh.plab[k];
        ++np;
      }
    sh.np = np;
  }
  __syncthreads();
  const int np     = sh.np;
  const int others = a.JB - a.w;
  if(np > 0) {
    for(int idx = b * (kLeafThreads / 64) + wave; idx < others;
        idx += G * (kLeafThreads / 64)) {
      const int c   = idx < a.c0 ? idx : idx + a.w;
      double*   col = a.P + static_cast<size_t>(c) * a.lda;
      if(lane < np) {
        // one instruction loads every source before any lane stores
        const double x = col[sh.src[lane]];
        col[sh.dst[lane]] = x;
      }
    }
  }
}

}  // namespace leaf

// Kernel entry of one rows-per-thread variant (defined in k_leaf<RPT>.hip).
template <int RPT>
const void* leaf_kernel_ptr();

}  // namespace hplx
