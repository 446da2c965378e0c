// Register-resident cooperative leaf of the panel factorisation
// (HPL_pdpanrlT, src/pfact/HPL_pdpanrlT_device.cpp:23-468): one launch per
// leaf of <= 32 columns.  Instantiated in k_leaf*.hip for each (register
// type T, storage type S) pair: FP64 for the HPL solve, FP32 for the panels
// of the mixed-precision factorisations.
//
// Every workgroup keeps its slice of the leaf (up to 512*RPT rows x 32
// columns) in registers for the whole leaf.  Per column:
//   1. each workgroup finds its candidate max |a| (wave64 shuffles, LDS);
//   2. it publishes the candidate's value, row index and 32-wide row segment
//      in its slot, then a sequence number with release ordering; workgroup
//      0, which holds the current top row, publishes that row likewise;
//   3. one wave of every workgroup polls all slots' sequence numbers (each
//      lane its slots at once) and takes one acquire fence; then the workgroup reads all candidates
//      (values and row segments) in one pass into LDS, reduces them itself
//      and takes the winner's row from LDS;
//   4. registers are updated: row interchange, division by the pivot,
//      rank-1 update of the remaining leaf columns.
// Slots are double-buffered by column parity, so a workgroup that runs ahead
// never overwrites a slot another workgroup still reads.  A grid of one
// workgroup skips the slots.  The reference instead parks its leaf on a
// pinned host flag once per column and lets a host thread run HPL_pdmxswp
// (:94-117,455-463).
//
// Steps 1 and 4 touch the register tile and are instantiated per column
// (static register indices); steps 2-3 are one out-of-line function.
//
// Every wait is bounded: after kSpinLimit polls a workgroup raises *err and
// all workgroups leave the kernel, so a grid that is not co-resident cannot
// hang the device.
#pragma once

#include <climits>

#include "hplx_internal.hpp"

namespace hplx {
namespace leaf {

constexpr uint32_t kSpinLimit = 1u << 22;
constexpr int      kWaves     = kLeafThreads / 64;
// slot layout (doubles): [0] |value|, [1] row, [2] sequence, [4..35] row
constexpr int kSlotVal = 0, kSlotRow = 1, kSlotSeq = 2, kSlotData = 4;

__device__ inline bool better(double v, int r, double bv, int br) {
  return v > bv || (v == bv && r < br);
}

__device__ inline unsigned long long ld_seq(const double* p) {
  return __hip_atomic_load(
      reinterpret_cast<unsigned long long*>(const_cast<double*>(p)),
      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void st_seq(double* p, unsigned long long v) {
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), v,
                     __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

struct LeafArgs {
  void*    P;          // panel top-left (panel row 0), storage type S
  int      lda;
  int      m;          // panel rows
  int      JB;         // panel width (row interchanges span it)
  int      c0;         // first leaf column (panel relative)
  int      w;          // leaf width, <= 32
  int*     ipiv;       // panel-relative pivots (+ row_offset) written here
  int      row_offset;
  int*     info;       // first exact zero pivot, 1-based global column
  int      info_base;  // global column of panel column 0
  double*  slots;      // 2 x (kMaxGrid+1) x kSlotStride
  uint32_t seq_base;   // unique per launch; column j publishes seq_base+j+1
  int*     err;
  int      rows_per_wg;  // kLeafThreads * RPT
};

// Shared-memory scratch of the leaf kernel (the exchange works in double;
// FP32 rows convert exactly).
struct LeafShared {
  double cand[kMaxLeafGrid][kLeafMax];  // every workgroup's candidate row
  double cv[kMaxLeafGrid];
  int    cr[kMaxLeafGrid];
  double mine[kLeafMax];  // this workgroup's candidate row
  double u[kLeafMax];     // pivot row
  double top[kLeafMax];   // current top row
  double wv[kWaves];
  int    wr[kWaves];
  int    piv[kLeafMax];
  int    lab[kLeafMax];
  int    ppos[kLeafMax];
  int    plab[kLeafMax];
  int    dst[2 * kLeafMax];
  int    src[2 * kLeafMax];
  int    np;
  int    p;
};

template <typename T, typename S>
__device__ inline T load_s(const S* p) {
  return static_cast<T>(*p);
}

template <typename T, typename S>
__device__ inline void store_s(S* p, T v) {
  *p = static_cast<S>(v);
}

// Workgroup max-loc of (v, r); every thread gets the result.
__device__ inline void block_maxloc(double& v, int& r, LeafShared& sh) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for(int o = 32; o > 0; o >>= 1) {
    const double ov  = __shfl_xor(v, o);
    const int    orr = __shfl_xor(r, o);
    if(better(ov, orr, v, r)) {
      v = ov;
      r = orr;
    }
  }
  if(lane == 0) {
    sh.wv[wave] = v;
    sh.wr[wave] = r;
  }
  __syncthreads();
  v = sh.wv[0];
  r = sh.wr[0];
#pragma unroll
  for(int k = 1; k < kWaves; ++k)
    if(better(sh.wv[k], sh.wr[k], v, r)) {
      v = sh.wv[k];
      r = sh.wr[k];
    }
  __syncthreads();  // sh.wv may be rewritten by the next call
}

// Lane l of one wave waits for slots l, l+64, ... (< G) of `slots`, and for
// `extra` if not null, to show sequence `want`.  All of a lane's slots are
// read at once on every round, so the wait costs one memory latency once
// they are published.  False on timeout or error.
__device__ inline bool poll_slots(const double* slots, int G,
                                  const double* extra, unsigned long long want,
                                  int* err) {
  constexpr int kPer = (kMaxGrid + 63) / 64;
  const int     lane = threadIdx.x & 63;
  unsigned      pending = extra ? 1u << kPer : 0u;
#pragma unroll
  for(int q = 0; q < kPer; ++q)
    if(lane + 64 * q < G) pending |= 1u << q;
  uint32_t spins = 0;
  while(pending) {
    unsigned long long sv[kPer + 1];
#pragma unroll
    for(int q = 0; q < kPer; ++q)
      if(pending & (1u << q))
        sv[q] = ld_seq(slots + (lane + 64 * q) * kSlotStride + kSlotSeq);
    if(pending & (1u << kPer)) sv[kPer] = ld_seq(extra + kSlotSeq);
#pragma unroll
    for(int q = 0; q <= kPer; ++q)
      if((pending & (1u << q)) && sv[q] == want) pending &= ~(1u << q);
    if(!pending) break;
    ++spins;
    if((spins & 255u) == 0 &&
       __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      return false;
    if(spins > kSpinLimit) {
      __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

// Steps 2-3 of a column: sh.mine holds this workgroup's candidate row (value
// bv, row br) and, in workgroup 0, sh.top holds the top row.  On return
// sh.p is the pivot row, sh.u its segment and sh.top the top row in every
// workgroup.  Returns false if the grid has to stop.
__device__ __attribute__((noinline)) bool exchange(const LeafArgs* ap, int jj,
                                                   double bv, int br,
                                                   LeafShared* shp) {
  const LeafArgs& a  = *ap;
  LeafShared&     sh = *shp;
  const int       t  = threadIdx.x;
  const int       b  = blockIdx.x;
  const int       G  = gridDim.x;
  if(G == 1) {
    if(t < kLeafMax) sh.u[t] = sh.mine[t];
    if(t == 0) sh.p = br;
    __syncthreads();
    return true;
  }
  // ---- publish (value, row, segment), then the sequence number
  const unsigned long long seq =
      static_cast<unsigned long long>(a.seq_base) + jj + 1;
  const size_t par   = static_cast<size_t>(jj & 1) * (kMaxGrid + 1);
  double*      slot  = a.slots + (par + b) * kSlotStride;
  double*      tslot = a.slots + (par + kMaxGrid) * kSlotStride;
  if(t < kLeafMax) {
    slot[kSlotData + t] = sh.mine[t];
  } else if(t == kLeafMax) {
    slot[kSlotVal] = bv;
    slot[kSlotRow] = static_cast<double>(br);
  } else if(b == 0 && t >= 64 && t < 64 + kLeafMax) {
    // the top row c0 + jj < c0 + 32 always lives in workgroup 0
    tslot[kSlotData + (t - 64)] = sh.top[t - 64];
  }
  // The workgroup barrier orders every wave's slot stores before thread 0's
  // release store of the sequence number, so one release per slot (one L2
  // write-back) publishes them all; no per-wave fence is needed.
  __syncthreads();
  if(t == 0) {
    st_seq(slot + kSlotSeq, seq);
    if(b == 0) st_seq(tslot + kSlotSeq, seq);
  }
  // ---- wave 0 waits for every candidate and the top row (lane l polls
  // slots l, l+64, ...; lane 0 also the top row), then takes one acquire
  // fence; the barrier extends it to the other waves, which then read the
  // candidates.
  int ok = 1;
  if(t < 64) {
    ok = poll_slots(a.slots + par * kSlotStride, G, t == 0 ? tslot : nullptr,
                    seq, a.err);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  if(!__syncthreads_and(ok)) return false;
  for(int i = t; i < G * kLeafMax; i += kLeafThreads) {
    const int g = i / kLeafMax, c = i % kLeafMax;
    sh.cand[g][c] = a.slots[(par + g) * kSlotStride + kSlotData + c];
  }
  if(t < G) {
    const double* s = a.slots + (par + t) * kSlotStride;
    sh.cv[t]        = s[kSlotVal];
    sh.cr[t]        = static_cast<int>(s[kSlotRow]);
  } else if(t >= kLeafThreads - kLeafMax && b != 0) {
    const int c = t - (kLeafThreads - kLeafMax);
    sh.top[c]   = tslot[kSlotData + c];
  }
  __syncthreads();
  double cv = -1.0;
  int    cr = INT_MAX;
  if(t < G) {
    cv = sh.cv[t];
    cr = sh.cr[t];
  }
  block_maxloc(cv, cr, sh);
  const int wb = (cr - a.c0) / a.rows_per_wg;
  if(t < kLeafMax) sh.u[t] = sh.cand[wb][t];
  if(t == 0) sh.p = cr;
  __syncthreads();
  return true;
}

// One pivot step for compile-time column JJ: every register of the row tile
// is addressed statically (no scratch spill).
template <typename T, int RPT, int JJ>
__device__ __forceinline__ void leaf_step(const LeafArgs& a,
                                          T (&v)[RPT][kLeafMax],
                                          const int (&rows)[RPT],
                                          LeafShared& sh, bool& alive) {
  const int t  = threadIdx.x;
  const int b  = blockIdx.x;
  const int gc = a.c0 + JJ;  // current top row == current column
  // ---- step 1: workgroup candidate max |a(r, gc)| over rows r >= gc
  double bv = -1.0;
  int    br = INT_MAX;
#pragma unroll
  for(int q = 0; q < RPT; ++q) {
    if(rows[q] >= gc && rows[q] < a.m) {
      const double x = fabs(static_cast<double>(v[q][JJ]));
      if(better(x, rows[q], bv, br)) {
        bv = x;
        br = rows[q];
      }
    }
  }
  block_maxloc(bv, br, sh);
#pragma unroll
  for(int q = 0; q < RPT; ++q) {
    if(rows[q] == br) {
#pragma unroll
      for(int c = 0; c < kLeafMax; ++c) sh.mine[c] = static_cast<double>(v[q][c]);
    }
    if(rows[q] == gc) {
#pragma unroll
      for(int c = 0; c < kLeafMax; ++c) sh.top[c] = static_cast<double>(v[q][c]);
    }
  }
  __syncthreads();
  // ---- steps 2-3
  if(!exchange(&a, JJ, bv, br, &sh)) {
    alive = false;
    return;
  }
  const int p = sh.p;
  if(t == 0) sh.piv[JJ] = p;
  const T piv = static_cast<T>(sh.u[JJ]);
  if(b == 0 && t == 0) {
    a.ipiv[gc] = p + a.row_offset;
    if(piv == T(0) && *a.info == 0) *a.info = a.info_base + gc + 1;
  }
  // ---- step 4: interchange rows gc and p, scale, rank-1 update
#pragma unroll
  for(int q = 0; q < RPT; ++q) {
    if(rows[q] == gc) {
#pragma unroll
      for(int c = 0; c < kLeafMax; ++c) v[q][c] = static_cast<T>(sh.u[c]);
    } else if(rows[q] == p) {
#pragma unroll
      for(int c = 0; c < kLeafMax; ++c) v[q][c] = static_cast<T>(sh.top[c]);
    }
  }
#pragma unroll
  for(int q = 0; q < RPT; ++q) {
    if(rows[q] > gc && rows[q] < a.m) {
      T l = v[q][JJ];
      if(piv != T(0)) l = l / piv;
      v[q][JJ] = l;
#pragma unroll
      for(int c = JJ + 1; c < kLeafMax; ++c)
        v[q][c] = fma(-l, static_cast<T>(sh.u[c]), v[q][c]);
    }
  }
  __syncthreads();  // sh.u / sh.top / sh.mine are rewritten next column
}

template <typename T, int RPT, int JJ>
struct LeafSteps {
  __device__ __forceinline__ static void run(const LeafArgs& a,
                                             T (&v)[RPT][kLeafMax],
                                             const int (&rows)[RPT],
                                             LeafShared& sh, bool& alive) {
    if(JJ < a.w && alive) leaf_step<T, RPT, JJ>(a, v, rows, sh, alive);
    LeafSteps<T, RPT, JJ + 1>::run(a, v, rows, sh, alive);
  }
};

template <typename T, int RPT>
struct LeafSteps<T, RPT, kLeafMax> {
  __device__ __forceinline__ static void run(const LeafArgs&,
                                             T (&)[RPT][kLeafMax],
                                             const int (&)[RPT], LeafShared&,
                                             bool&) {}
};

template <typename T, typename S, int RPT>
__global__ __launch_bounds__(kLeafThreads) void leaf_kernel(LeafArgs a) {
  constexpr int CH   = kLeafThreads * RPT;  // rows per workgroup
  const int     t    = threadIdx.x;
  const int     lane = t & 63;
  const int     wave = t >> 6;
  const int     b    = blockIdx.x;
  const int     G    = gridDim.x;
  S*            P    = static_cast<S*>(a.P);

  __shared__ LeafShared sh;

  if(__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    return;

  int rows[RPT];
  T   v[RPT][kLeafMax];
#pragma unroll
  for(int q = 0; q < RPT; ++q) rows[q] = a.c0 + b * CH + q * kLeafThreads + t;
#pragma unroll
  for(int c = 0; c < kLeafMax; ++c) {
    const S* col = P + static_cast<size_t>(a.c0 + c) * a.lda;
#pragma unroll
    for(int q = 0; q < RPT; ++q)
      v[q][c] = (c < a.w && rows[q] < a.m) ? load_s<T, S>(col + rows[q]) : T(0);
  }

  bool alive = true;
  LeafSteps<T, RPT, 0>::run(a, v, rows, sh, alive);
  if(!alive) return;

  // ---- write the leaf back
#pragma unroll
  for(int c = 0; c < kLeafMax; ++c) {
    if(c < a.w) {
      S* col = P + static_cast<size_t>(a.c0 + c) * a.lda;
#pragma unroll
      for(int q = 0; q < RPT; ++q)
        if(rows[q] < a.m) store_s<T, S>(col + rows[q], v[q][c]);
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
        const int x      = sh.lab[jj];
        sh.lab[jj]       = sh.lab[p - a.c0];
        sh.lab[p - a.c0] = x;
      } else {
        int s = -1;
        for(int k = 0; k < nout; ++k)
          if(sh.ppos[k] == p) {
            s = k;
            break;
          }
        if(s < 0) {
          s          = nout++;
          sh.ppos[s] = p;
          sh.plab[s] = p;
        }
        const int x = sh.lab[jj];
        sh.lab[jj]  = sh.plab[s];
        sh.plab[s]  = x;
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
  if(np > 0 && lane < np) {
    // a wave takes columns idx, idx + stride, ...; each column's sources are
    // all loaded (one instruction) before any of its destinations is stored,
    // and kBatch columns are in flight at once
    constexpr int kBatch = 8;
    const int     stride = G * kWaves;
    const int     src = sh.src[lane], dst = sh.dst[lane];
    for(int idx0 = b * kWaves + wave; idx0 < others; idx0 += kBatch * stride) {
      S x[kBatch];
#pragma unroll
      for(int u = 0; u < kBatch; ++u) {
        const int idx = idx0 + u * stride;
        const int c   = idx < a.c0 ? idx : idx + a.w;
        if(idx < others) x[u] = P[src + static_cast<size_t>(c) * a.lda];
      }
#pragma unroll
      for(int u = 0; u < kBatch; ++u) {
        const int idx = idx0 + u * stride;
        const int c   = idx < a.c0 ? idx : idx + a.w;
        if(idx < others) P[dst + static_cast<size_t>(c) * a.lda] = x[u];
      }
    }
  }
}

}  // namespace leaf

// Kernel entry of one variant (defined in k_leaf*.hip): FP64, or FP32.
enum class LeafKind { kF64, kF32 };
template <LeafKind K>
const void* leaf_kernel_ptr();

}  // namespace hplx
