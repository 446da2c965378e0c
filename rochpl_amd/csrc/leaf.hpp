// Register-resident cooperative leaf of the panel factorisation
// (HPL_pdpanrlT, src/pfact/HPL_pdpanrlT_device.cpp:23-468): one launch per
// leaf of <= 32 columns.  Instantiated in k_leaf*.hip for each (register
// type T, storage type S) pair: FP64 for the HPL solve, FP32 for the panels
// of the mixed-precision factorisations.
//
// Every workgroup keeps its slice of the leaf (up to 512*RPT rows x 32
// columns) in registers for the whole leaf.  Per column:
//   1. each workgroup finds its candidate max |a| (DPP lane exchanges within
//      a wave, then the 8 waves through LDS);
//   2. the thread holding the candidate row stores it straight from its
//      registers into the workgroup's slot (value, row index, 32-wide row);
//      in workgroup 0 the thread holding the current top row stores it into
//      the same slot; after a workgroup barrier one thread publishes a
//      sequence number with release ordering;
//   3. one wave of every workgroup polls all slots' sequence numbers (each
//      lane its slots at once) and takes one acquire fence; then every thread
//      loads a share of all candidate rows (16-byte loads, all in flight
//      before the first LDS store) and the top row into LDS, and each wave
//      reduces the candidates itself (no further barrier);
//   4. registers are updated: row interchange, division by the pivot,
//      rank-1 update of the remaining leaf columns.
// Slots are double-buffered by column parity, so a workgroup that runs ahead
// never overwrites a slot another workgroup still reads.  A grid of one
// workgroup skips the slots.  The reference instead parks its leaf on a
// pinned host flag once per column and lets a host thread run HPL_pdmxswp
// (:94-117,455-463).
//
// Each column's steps are instantiated per column (static register indices).
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
// slot layout (doubles): [0] |value|, [1] row, [2] sequence, [4..35] the
// candidate row, [36..67] the top row (workgroup 0's slot only)
constexpr int kSlotVal = 0, kSlotRow = 1, kSlotSeq = 2, kSlotData = 4,
              kSlotTop = 36;
static_assert(kSlotTop + kLeafMax <= kSlotStride, "slot too small");

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
  int      grid;         // workgroups taking part: 0 .. grid-1 (<= gridDim.x)
};

// Shared-memory scratch of the leaf kernel (the exchange works in double;
// FP32 rows convert exactly).
struct LeafShared {
  alignas(16) double cand[kMaxLeafGrid][kLeafMax];  // every workgroup's candidate row
  double cv[kMaxLeafGrid];
  int    cr[kMaxLeafGrid];
  double top[kLeafMax];  // current top row
  double wv[kWaves];
  int    wr[kWaves];
  int    ok;
  int    piv[kLeafMax];
  int    lab[kLeafMax];
  int    ppos[kLeafMax];
  int    plab[kLeafMax];
  int    dst[2 * kLeafMax];
  int    src[2 * kLeafMax];
  int    np;
};

template <typename T, typename S>
__device__ inline T load_s(const S* p) {
  return static_cast<T>(*p);
}

template <typename T, typename S>
__device__ inline void store_s(S* p, T v) {
  *p = static_cast<S>(v);
}

// One step of a wave max-loc: combine (v, r) with the lane that the DPP
// control CTRL selects (rows outside ROW_MASK keep their value).
template <int CTRL, int ROW_MASK>
__device__ inline void maxloc_dpp(double& v, int& r) {
  const long long b  = __double_as_longlong(v);
  const int       lo = static_cast<int>(b), hi = static_cast<int>(b >> 32);
  const int ol = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW_MASK, 0xF, false);
  const int oh = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW_MASK, 0xF, false);
  const int orr = __builtin_amdgcn_update_dpp(r, r, CTRL, ROW_MASK, 0xF, false);
  const double ov = __longlong_as_double(
      (static_cast<long long>(oh) << 32) | static_cast<unsigned int>(ol));
  if(better(ov, orr, v, r)) {
    v = ov;
    r = orr;
  }
}

// Wave max-loc of (v, r); every lane gets the result.  Within a row of 16
// lanes: xor 1, xor 2 (quad permutes), half-row and row mirrors; across
// rows: broadcasts of lane 15 (into rows 1, 3) and lane 31 (into rows 2,
// 3), leaving the result in lane 63.
__device__ inline void wave_maxloc(double& v, int& r) {
  maxloc_dpp<0xB1, 0xF>(v, r);   // quad_perm [1,0,3,2]
  maxloc_dpp<0x4E, 0xF>(v, r);   // quad_perm [2,3,0,1]
  maxloc_dpp<0x141, 0xF>(v, r);  // row_half_mirror
  maxloc_dpp<0x140, 0xF>(v, r);  // row_mirror
  maxloc_dpp<0x142, 0xA>(v, r);  // row_bcast:15
  maxloc_dpp<0x143, 0xC>(v, r);  // row_bcast:31
  const long long b  = __double_as_longlong(v);
  const int       lo = __builtin_amdgcn_readlane(static_cast<int>(b), 63);
  const int       hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), 63);
  r                  = __builtin_amdgcn_readlane(r, 63);
  v = __longlong_as_double((static_cast<long long>(hi) << 32) |
                           static_cast<unsigned int>(lo));
}

// Workgroup max-loc of (v, r); every thread gets the result.  sh.wv is next
// written after further barriers of the same column, so none is needed here
// after the read.
__device__ inline void block_maxloc(double& v, int& r, LeafShared& sh) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  wave_maxloc(v, r);
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

// Steps 2-3 of a column after the candidate row (and in workgroup 0 the
// top row) has been stored into this workgroup's slot: publish, wait for
// every workgroup, load all candidates and the top row, reduce.  Returns
// the pivot row (sh.cand[its workgroup] holds its segment, sh.top the top
// row), or -1 if the grid has to stop.
__device__ __forceinline__ int exchange(const LeafArgs* ap, int jj,
                                                  LeafShared* shp) {
  const LeafArgs& a    = *ap;
  LeafShared&     sh   = *shp;
  const int       t    = threadIdx.x;
  const int       lane = t & 63;
  const int       b    = blockIdx.x;
  const int       G    = a.grid;
  const unsigned long long seq =
      static_cast<unsigned long long>(a.seq_base) + jj + 1;
  const double* slots = a.slots + static_cast<size_t>(jj & 1) * (kMaxGrid + 1) * kSlotStride;
  // The workgroup barrier orders every wave's slot stores before thread 0's
  // release store of the sequence number, so one release per slot (one L2
  // write-back) publishes them all.
  __syncthreads();
  if(t == 0) st_seq(const_cast<double*>(slots) + b * kSlotStride + kSlotSeq, seq);
  if(t < 64) {
    const int ok = poll_slots(slots, G, nullptr, seq, a.err);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if(t == 0) sh.ok = ok;
  }
  __syncthreads();
  if(!sh.ok) return -1;
  if(t < G) {
    sh.cv[t] = slots[t * kSlotStride + kSlotVal];
    sh.cr[t] = static_cast<int>(slots[t * kSlotStride + kSlotRow]);
  } else if(t >= kLeafThreads - kLeafMax) {
    const int c = t - (kLeafThreads - kLeafMax);
    sh.top[c]   = slots[kSlotTop + c];
  }
  // all candidate rows as 16-byte loads, issued four per thread before
  // their LDS stores (one memory latency per 128 workgroups)
  typedef double d2 __attribute__((ext_vector_type(2)));
  constexpr int kHalf = kLeafMax / 2;
  for(int i0 = t; i0 < G * kHalf; i0 += 4 * kLeafThreads) {
    d2 rv[4];
#pragma unroll
    for(int k = 0; k < 4; ++k) {
      const int i = i0 + k * kLeafThreads;
      if(i < G * kHalf)
        rv[k] = *reinterpret_cast<const d2*>(
            &slots[(i / kHalf) * kSlotStride + kSlotData + 2 * (i % kHalf)]);
    }
#pragma unroll
    for(int k = 0; k < 4; ++k) {
      const int i = i0 + k * kLeafThreads;
      if(i < G * kHalf)
        *reinterpret_cast<d2*>(&sh.cand[i / kHalf][2 * (i % kHalf)]) = rv[k];
    }
  }
  __syncthreads();
  // every wave reduces the candidates itself
  double cv = -1.0;
  int    cr = INT_MAX;
  for(int g = lane; g < G; g += 64)
    if(better(sh.cv[g], sh.cr[g], cv, cr)) {
      cv = sh.cv[g];
      cr = sh.cr[g];
    }
  wave_maxloc(cv, cr);
  return cr;
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
  // ---- steps 2-3
  int           p;
  const double* u;
  if(a.grid == 1) {
#pragma unroll
    for(int q = 0; q < RPT; ++q)
      if(rows[q] == br) {
#pragma unroll
        for(int c = 0; c < kLeafMax; ++c) sh.cand[0][c] = static_cast<double>(v[q][c]);
      }
    if(t == JJ) {  // the top row c0 + JJ: workgroup 0, first row set
#pragma unroll
      for(int c = 0; c < kLeafMax; ++c) sh.top[c] = static_cast<double>(v[0][c]);
    }
    __syncthreads();
    p = br;
    u = sh.cand[0];
  } else {
    double* slot = a.slots + (static_cast<size_t>(JJ & 1) * (kMaxGrid + 1) + b) * kSlotStride;
#pragma unroll
    for(int q = 0; q < RPT; ++q)
      if(rows[q] == br) {
#pragma unroll
        for(int c = 0; c < kLeafMax; ++c) slot[kSlotData + c] = static_cast<double>(v[q][c]);
      }
    if(t == 0) {
      slot[kSlotVal] = bv;
      slot[kSlotRow] = static_cast<double>(br);
    }
    if(b == 0 && t == JJ) {  // the top row c0 + JJ
#pragma unroll
      for(int c = 0; c < kLeafMax; ++c) slot[kSlotTop + c] = static_cast<double>(v[0][c]);
    }
    p = exchange(&a, JJ, &sh);
    if(p < 0) {
      alive = false;
      return;
    }
    u = sh.cand[(p - a.c0) / a.rows_per_wg];
  }
  if(t == 0) sh.piv[JJ] = p;
  const T piv = static_cast<T>(u[JJ]);
  if(b == 0 && t == 0) {
    a.ipiv[gc] = p + a.row_offset;
    if(piv == T(0) && *a.info == 0) *a.info = a.info_base + gc + 1;
  }
  // ---- step 4: interchange rows gc and p, scale, rank-1 update
#pragma unroll
  for(int q = 0; q < RPT; ++q) {
    if(rows[q] == gc) {
#pragma unroll
      for(int c = 0; c < kLeafMax; ++c) v[q][c] = static_cast<T>(u[c]);
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
        v[q][c] = fma(-l, static_cast<T>(u[c]), v[q][c]);
    }
  }
  // sh.cand / sh.top are rewritten next column only after the barrier in
  // its block_maxloc, which every wave reaches after this update
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

// One leaf on workgroups 0 .. a.grid-1: the body of leaf_kernel, also run
// by the persistent panel kernel (k_panel_kernel.hip) once per leaf.
// Workgroups whose rows all lie at or below a.m take part in every exchange
// with an empty candidate.
template <typename T, typename S, int RPT>
__device__ __forceinline__ void leaf_run(const LeafArgs& a, LeafShared& sh) {
  constexpr int CH   = kLeafThreads * RPT;  // rows per workgroup
  const int     t    = threadIdx.x;
  const int     lane = t & 63;
  const int     wave = t >> 6;
  const int     b    = blockIdx.x;
  const int     G    = a.grid;
  S*            P    = static_cast<S*>(a.P);

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

template <typename T, typename S, int RPT>
__global__ __launch_bounds__(kLeafThreads) void leaf_kernel(LeafArgs a) {
  __shared__ LeafShared sh;
  if(__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    return;
  leaf_run<T, S, RPT>(a, sh);
}

}  // namespace leaf

// Kernel entry of one variant (defined in k_leaf*.hip): FP64, or FP32.
enum class LeafKind { kF64, kF32 };
template <LeafKind K>
const void* leaf_kernel_ptr();

}  // namespace hplx
