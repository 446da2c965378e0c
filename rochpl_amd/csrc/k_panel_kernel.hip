// The persistent panel kernel: one launch factors a whole FP64 panel
// (HPL_pdfact -> HPL_pdrpan{rl,ll,cr}T -> HPL_pdpan{rl,ll,cr}T,
// src/pfact/HPL_pdrpanrlT.cpp:19-161), recursion included, instead of the
// sequence of leaf, triangular-solve and product launches k_panel.hip queues.
//
// The host records the recursion's steps once per (panel width, algorithm)
// in panel-relative coordinates (PanelOp, k_panel.hip's chunk_loop in
// recording mode) and keeps them in device memory; the kernel walks the list
// with its whole grid and a grid barrier after every step:
//   - a leaf is the cooperative register leaf (leaf.hpp, leaf_run), its
//     pivot exchange through the same candidate slots, one sequence number
//     per leaf;
//   - a triangular solve gives each workgroup right-hand-side columns
//     b, b + G, ... (tri_solve.hpp, the same per-column routine as
//     tri_solve_kernel);
//   - a product first transposes its U block into the scratch buffer (32 x 32
//     tiles through LDS), passes a barrier, then runs the streamed update's
//     128 x 128 tiles b, b + G, ... (update_tile.hpp, the routine of the
//     trailing update).
// The grid is one workgroup of 512 threads on every CU the panel has, all
// resident at once: with the look-ahead the CUs the claimed update leaves
// free, else the whole device.  A leaf runs on the workgroups that hold its
// rows (ceil((m - c0) / 1536), as its own launch would), the others wait at
// the barrier; the products and solves use the whole grid.
//
// Grid barrier: every wave waits for its own stores, the workgroup meets at
// a barrier, then thread 0 adds 1 to the counter with release ordering at
// device scope (writing back the XCD's L2), polls it until all G
// workgroups of this barrier have arrived and takes an acquire fence
// (invalidating the stale lines of its CU and L2).  The counter is zeroed
// before each launch, so barrier i waits for (i + 1) G arrivals.  Waits are
// bounded: past the limit the workgroup raises the error flag and every
// workgroup leaves, as the leaf's own waits do.
#include "hplx_internal.hpp"
#include "leaf.hpp"
#include "tri_solve.hpp"
#include "update_tile.hpp"

namespace hplx {
namespace {

using namespace upd;

struct PkArgs {
  double*        P;     // panel top-left
  int            lda;
  int            m;     // panel rows
  int            JB;    // panel width
  int            nops;
  const PanelOp* ops;
  int*           ipiv;
  int            row_offset;
  int*           info;
  int            info_base;
  double*        slots;
  uint32_t       leaf_base;  // launch number of the first leaf
  int*           err;
  unsigned*      bar;        // grid barrier counter, zero at launch
  double*        Bt;         // kMaxNB^2 doubles: the products' U^T
};

constexpr uint32_t kBarSpinLimit = 1u << 22;

__device__ bool grid_sync(const PkArgs& a, unsigned target, int* ok_sh) {
  wait_vm<0>();  // this wave's global stores have reached its L2
  __syncthreads();
  if(threadIdx.x == 0) {
    __hip_atomic_fetch_add(a.bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    int      ok    = 1;
    uint32_t spins = 0;
    while(__hip_atomic_load(a.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
          target) {
      ++spins;
      if((spins & 255u) == 0 &&
         __hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        ok = 0;
        break;
      }
      if(spins > kBarSpinLimit) {
        __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    *ok_sh = ok;
  }
  __syncthreads();
  return *ok_sh != 0;
}

__global__ __launch_bounds__(kLeafThreads) void panel_kernel(PkArgs a) {
  __shared__ leaf::LeafShared sh;
  __shared__ double           tt[32][33];  // transpose tile
  __shared__ int              ok_sh;
  HPLX_USLOT(0) HPLX_USLOT(1) HPLX_USLOT(2) HPLX_USLOT(3)

  if(__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    return;
  const int G = gridDim.x, b = blockIdx.x, t = threadIdx.x;
  unsigned  nbar   = 0;
  uint32_t  leaf_i = 0;
  for(int i = 0; i < a.nops; ++i) {
    // the step's fields as wave-uniform scalars (SGPRs), not VGPRs
    PanelOp o;
    {
      const int* src = reinterpret_cast<const int*>(a.ops + i);
      int*       dst = reinterpret_cast<int*>(&o);
#pragma unroll
      for(int f = 0; f < static_cast<int>(sizeof(PanelOp) / sizeof(int)); ++f)
        dst[f] = __builtin_amdgcn_readfirstlane(src[f]);
    }
    if(o.kind == kPanelLeaf) {
      // the leaf's own grid, as its launch would have: the workgroups that
      // hold rows [ac, m); the others wait at the barrier
      constexpr int CH = kLeafThreads * kLeafRPT;
      const int     gl = (a.m - o.ac + CH - 1) / CH;
      leaf::LeafArgs la;
      la.P           = a.P;
      la.lda         = a.lda;
      la.m           = a.m;
      la.JB          = a.JB;
      la.c0          = o.ac;
      la.w           = o.n;
      la.ipiv        = a.ipiv;
      la.row_offset  = a.row_offset;
      la.info        = a.info;
      la.info_base   = a.info_base;
      la.slots       = a.slots;
      la.seq_base    = (a.leaf_base + leaf_i++) << 6;
      la.err         = a.err;
      la.rows_per_wg = CH;
      la.grid        = gl;
      if(b < gl) leaf::leaf_run<double, double, kLeafRPT>(la, sh);
    } else if(o.kind == kPanelTrsm) {
      const double* T = a.P + o.ar + static_cast<size_t>(o.ac) * a.lda;
      for(int col = b; col < o.n; col += G)
        trs::tri_solve_col<double, double, true, false, true, false>(
            o.k, T, a.lda, a.P + o.br + static_cast<size_t>(o.bc + col) * a.lda,
            col);
    } else {
      const int rows = o.m_fix >= 0 ? o.m_fix : a.m - o.m_sub;
      if(rows <= 0) continue;  // uniform over the grid: no barrier owed
      const int ldt = ut_ld(o.n);
      // U (k x n at br, bc) -> Bt (row kk holds U(kk, 0 .. n-1))
      const double* B  = a.P + o.br + static_cast<size_t>(o.bc) * a.lda;
      const int     gx = (o.k + 31) / 32, gy = (o.n + 31) / 32;
      const int     tx = t & 31, ty = t >> 5;  // 32 x 16
      for(int tile = b; tile < gx * gy; tile += G) {
        const int r0 = (tile % gx) * 32, c0 = (tile / gx) * 32;
#pragma unroll
        for(int q = 0; q < 2; ++q) {
          const int c = c0 + ty + 16 * q, r = r0 + tx;
          if(r < o.k && c < o.n) tt[ty + 16 * q][tx] = B[r + static_cast<size_t>(c) * a.lda];
        }
        __syncthreads();
#pragma unroll
        for(int q = 0; q < 2; ++q) {
          const int r = r0 + ty + 16 * q, c = c0 + tx;
          if(r < o.k && c < o.n) a.Bt[c + static_cast<size_t>(r) * ldt] = tt[tx][ty + 16 * q];
        }
        __syncthreads();
      }
      if(!grid_sync(a, ++nbar * G, &ok_sh)) return;
      const double* A  = a.P + o.ar + static_cast<size_t>(o.ac) * a.lda;
      double*       C  = a.P + o.cr + static_cast<size_t>(o.cc) * a.lda;
      const int     tm = (rows + 127) / 128, tn = (o.n + 127) / 128;
      for(int tile = b; tile < tm * tn; tile += G) {
        if(tile > b) __syncthreads();  // the LDS ring is free again
        update_tile(tile, rows, o.n, o.k, A, a.lda, a.Bt, ldt, C, a.lda, tn, tm,
                    (rows + 1) & ~1, ldt, sA0, sB0, sA1, sB1, sA2, sB2, sA3,
                    sB3);
      }
    }
    if(!grid_sync(a, ++nbar * G, &ok_sh)) return;
  }
}

}  // namespace

hipError_t panel_kernel_occupancy(int* blocks_per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(
      blocks_per_cu, reinterpret_cast<const void*>(panel_kernel), kLeafThreads,
      0);
}

hipError_t launch_panel_kernel(Ctx& c, int slot, double* P, int lda, int m,
                               int jb, int* ipiv, int row_offset, bool* ran) {
  *ran = false;
  const std::vector<PanelOp>& ops = c.pop_host[slot];
  constexpr int CH = kLeafThreads * kLeafRPT;
  const int     gl = (m + CH - 1) / CH;  // the first leaf's workgroups
  if(gl < 1 || gl > kMaxLeafGrid || gl > c.pk_resident) return hipSuccess;
  // the grid: every CU the panel has (the vacated CUs beside a claimed
  // update, else the whole device), so that the products run on all of them
  int G = c.panel_cus > 0 ? c.panel_cus : c.pk_resident;
  G     = G < c.pk_resident ? G : c.pk_resident;
  G     = G > gl ? G : gl;
  // every product must take the streamed tile (the launch sequence falls
  // back to the generic GEMM for the others)
  double* Bt = panel_scratch(c);
  for(const PanelOp& o : ops) {
    if(o.kind != kPanelGemm) continue;
    const int rows = o.m_fix >= 0 ? o.m_fix : m - o.m_sub;
    if(rows <= 0) continue;
    const int ldt = ut_ld(o.n);
    if(static_cast<size_t>(o.k) * ldt > static_cast<size_t>(kMaxNB) * kMaxNB ||
       !update_fast_ok(rows, o.n, o.k, P + o.ar + static_cast<size_t>(o.ac) * lda,
                       lda, Bt, ldt, (rows + 1) & ~1, ldt))
      return hipSuccess;
  }
  hipStream_t s = panel_stream(c);
  // leaf sequence numbers: launch << 6 | column; slots restart when they wrap
  const uint32_t nleaf = static_cast<uint32_t>(c.pop_leaves[slot]);
  if(c.leaf_launches + nleaf >= (1u << 26) - 1) {
    hipError_t e = launch_zero_ints(s, reinterpret_cast<int*>(c.slots),
                                    4 * (kMaxGrid + 1) * kSlotStride);
    if(e != hipSuccess) return e;
    c.leaf_launches = 0;
  }
  PkArgs a;
  a.P          = P;
  a.lda        = lda;
  a.m          = m;
  a.JB         = jb;
  a.nops       = static_cast<int>(ops.size());
  a.ops        = c.pops + static_cast<size_t>(slot) * kMaxPanelOps;
  a.ipiv       = ipiv;
  a.row_offset = row_offset;
  a.info       = c.flags;
  a.info_base  = row_offset;
  a.slots      = c.slots;
  a.leaf_base  = c.leaf_launches + 1;
  a.err        = c.flags + 1;
  a.bar        = c.pbar;
  a.Bt         = Bt;
  c.leaf_launches += nleaf;
  hipError_t e = launch_zero_ints(s, reinterpret_cast<int*>(c.pbar), 1);
  if(e != hipSuccess) return e;
  hipLaunchKernelGGL(panel_kernel, dim3(G), dim3(kLeafThreads), 0, s, a);
  e = hipGetLastError();
  if(e == hipSuccess) *ran = true;
  return e;
}

}  // namespace hplx
