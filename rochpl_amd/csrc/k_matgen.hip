// Matrix generator and the on-the-fly residual check.
//
// Matgen restates HPL_pdrandmat (src/matgen/HPL_pdrandmat_device.cpp:21-201)
// at P=Q=1: element (i,j) of an M-row matrix is the value of LCG state
// i + j*M.  The reference walks one column per thread (uncoalesced stores);
// here each thread owns one row of a 256-row x 256-column tile and steps by
// the M-step map between columns, so every store instruction writes 256
// consecutive doubles of one column.
//
// The check restates HPL_pdtest's verification (src/HPL_pdtest.cpp:303-432)
// without re-reading A: [A|b] is regenerated tile by tile in registers while
// r = b - A x, the row sums (||A||_oo) and the column sums (||A||_1) are
// accumulated.  ||r||_oo is the true maximum over all rows; the reference's
// |r_0| form (src/pauxil/HPL_pdlange_device.cpp:284) is reported beside it.
#include "lcg.hpp"
#include "mxp.hpp"

namespace hplx {
namespace {

constexpr int kGenRows = 256;  // rows per block (one per thread)
constexpr int kGenCols = 256;  // columns per block
constexpr int kChkCols = 4096; // columns per partial-sum chunk (row pass)
constexpr int kChkRows = 4096; // rows per partial-sum chunk (column pass)

// T: stored type.  FP32 storage holds RN_fp32(v); FP16 storage holds
// RN_fp16(RN_fp32(v)) of the FP64 generator value v.
template <typename T>
__device__ inline T store_as(double v) {
  if constexpr(sizeof(T) == 2) {
    float f = static_cast<float>(v);
    // keep the two roundings: the compiler may not merge them into one
    // double -> half rounding (which differs on 1 in 2^13 values)
    asm volatile("" : "+v"(f));
    return __float2half(f);
  } else {
    return static_cast<T>(v);
  }
}

template <typename T>
__global__ __launch_bounds__(kGenRows) void matgen_kernel(
    int m, int n, T* __restrict__ A, int lda, uint64_t seed, double shift,
    uint64_t colA, uint64_t colC, int gx) {
  const int bx = blockIdx.x % gx;
  const int by = blockIdx.x / gx;
  const int i  = bx * kGenRows + threadIdx.x;
  const int j0 = by * kGenCols;
  if(i >= m) return;
  const int jend = min(n, j0 + kGenCols);
  uint64_t s =
      lcg_state(seed, static_cast<uint64_t>(i) +
                          static_cast<uint64_t>(j0) * static_cast<uint64_t>(m));
  T* col = A + static_cast<size_t>(j0) * lda + i;
  for(int j = j0; j < jend; ++j) {
    double v = lcg_value(s);
    if(i == j) v += shift;
    *col = store_as<T>(v);
    col += lda;
    s = colA * s + colC;
  }
}

// Row pass: part_r[c*n + i] = sum_{j in chunk c} a(i,j) x_j,
//           part_a[c*n + i] = sum_{j in chunk c} |a(i,j)|.
__global__ __launch_bounds__(256) void check_rows_kernel(
    int n, uint64_t seed, double shift, const double* __restrict__ X,
    double* __restrict__ part_r, double* __restrict__ part_a, uint64_t colA,
    uint64_t colC, int gx) {
  __shared__ double sx[kChkCols];
  const int bx = blockIdx.x % gx;
  const int c  = blockIdx.x / gx;
  const int j0 = c * kChkCols;
  const int jn = min(n, j0 + kChkCols) - j0;
  for(int t = threadIdx.x; t < jn; t += blockDim.x) sx[t] = X[j0 + t];
  __syncthreads();
  const int i = bx * 256 + threadIdx.x;
  if(i >= n) return;
  uint64_t s = lcg_state(seed, static_cast<uint64_t>(i) +
                                   static_cast<uint64_t>(j0) *
                                       static_cast<uint64_t>(n));
  double acc = 0.0, aabs = 0.0;
  for(int t = 0; t < jn; ++t) {
    double v = lcg_value(s);
    if(i == j0 + t) v += shift;
    acc  = fma(v, sx[t], acc);
    aabs += fabs(v);
    s = colA * s + colC;
  }
  part_r[static_cast<size_t>(c) * n + i] = acc;
  part_a[static_cast<size_t>(c) * n + i] = aabs;
}

// Column pass: part_c[rc*n + j] = sum_{i in row chunk rc} |a(i,j)|.
__global__ __launch_bounds__(256) void check_cols_kernel(
    int n, uint64_t seed, double shift, double* __restrict__ part_c, int gx) {
  const int bx = blockIdx.x % gx;
  const int rc = blockIdx.x / gx;
  const int j  = bx * 256 + threadIdx.x;
  if(j >= n) return;
  const int i0 = rc * kChkRows;
  const int in = min(n, i0 + kChkRows);
  uint64_t s = lcg_state(seed, static_cast<uint64_t>(i0) +
                                   static_cast<uint64_t>(j) *
                                       static_cast<uint64_t>(n));
  double aabs = 0.0;
  for(int i = i0; i < in; ++i) {
    double v = lcg_value(s);
    if(i == j) v += shift;
    aabs += fabs(v);
    s = kLcgMult * s + kLcgAdd;
  }
  part_c[static_cast<size_t>(rc) * n + j] = aabs;
}

__device__ inline double block_max(double v, double* sh) {
  for(int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  const int w = threadIdx.x >> 6;
  if((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double r = sh[0];
  for(int k = 1; k < (int)(blockDim.x >> 6); ++k) r = fmax(r, sh[k]);
  __syncthreads();
  return r;
}

__device__ inline double block_sum(double v, double* sh) {
  for(int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x >> 6;
  if((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double r = sh[0];
  for(int k = 1; k < (int)(blockDim.x >> 6); ++k) r += sh[k];
  __syncthreads();
  return r;
}

// Per-row/column finish.  Block b writes 6 partials to bp[b*8 + 0..5]:
// max|r_i|, max rowsum, max colsum, max|x_i|, sum|x_i|, max|b_i|; block 0
// also writes r_0 to bp[6].
__global__ __launch_bounds__(256) void check_finish_kernel(
    int n, int nch_c, int nch_r, uint64_t seed, const double* __restrict__ X,
    const double* __restrict__ part_r, const double* __restrict__ part_a,
    const double* __restrict__ part_c, double* __restrict__ bp) {
  __shared__ double sh[4];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double r = 0.0, rs = 0.0, cs = 0.0, xa = 0.0, ba = 0.0;
  if(i < n) {
    // b_i = value of state i + n*n
    const double b = lcg_value(lcg_state(
        seed, static_cast<uint64_t>(i) +
                  static_cast<uint64_t>(n) * static_cast<uint64_t>(n)));
    double ax = 0.0;
    for(int c = 0; c < nch_c; ++c) {
      ax += part_r[static_cast<size_t>(c) * n + i];
      rs += part_a[static_cast<size_t>(c) * n + i];
    }
    for(int c = 0; c < nch_r; ++c) cs += part_c[static_cast<size_t>(c) * n + i];
    r  = b - ax;
    xa = fabs(X[i]);
    ba = fabs(b);
    if(i == 0) bp[6] = fabs(r);
  }
  const double m_r  = block_max(fabs(r), sh);
  const double m_rs = block_max(rs, sh);
  const double m_cs = block_max(cs, sh);
  const double m_x  = block_max(xa, sh);
  const double s_x  = block_sum(xa, sh);
  const double m_b  = block_max(ba, sh);
  if(threadIdx.x == 0) {
    double* o = bp + 8 + static_cast<size_t>(blockIdx.x) * 8;
    o[0] = m_r;
    o[1] = m_rs;
    o[2] = m_cs;
    o[3] = m_x;
    o[4] = s_x;
    o[5] = m_b;
  }
}

__global__ __launch_bounds__(256) void check_reduce_kernel(
    int n, int nblk, const double* __restrict__ bp, double* __restrict__ out) {
  __shared__ double sh[4];
  double v[6] = {0, 0, 0, 0, 0, 0};
  // sums over blocks are taken in block order by one thread for a fixed
  // rounding order; the maxima are order-free.
  for(int b = threadIdx.x; b < nblk; b += blockDim.x) {
    const double* o = bp + 8 + static_cast<size_t>(b) * 8;
    v[0] = fmax(v[0], o[0]);
    v[1] = fmax(v[1], o[1]);
    v[2] = fmax(v[2], o[2]);
    v[3] = fmax(v[3], o[3]);
    v[5] = fmax(v[5], o[5]);
  }
  const double rnorm = block_max(v[0], sh);
  const double anormi = block_max(v[1], sh);
  const double anorm1 = block_max(v[2], sh);
  const double xnormi = block_max(v[3], sh);
  const double bnorm  = block_max(v[5], sh);
  if(threadIdx.x == 0) {
    double xs = 0.0;
    for(int b = 0; b < nblk; ++b) xs += bp[8 + static_cast<size_t>(b) * 8 + 4];
    const double eps   = 0x1p-53;
    const double denom = eps * (anormi * xnormi + bnorm) * static_cast<double>(n);
    const double r0    = bp[6];
    out[0] = rnorm / denom;
    out[1] = rnorm;
    out[2] = r0 / denom;
    out[3] = anormi;
    out[4] = anorm1;
    out[5] = xnormi;
    out[6] = xs;
    out[7] = bnorm;
    out[8] = r0;
  }
}

// Refinement residual, per row: r_i = b_i - sum_c part_r[c][i] (chunks in
// order), row sums of |A| likewise; block maxima of |r_i|, the row sums,
// |x_i| and |b_i| to bp[4*block + 0..3].
__global__ __launch_bounds__(256) void ir_finish_kernel(
    int n, int nch_c, uint64_t seed, const double* __restrict__ X,
    const double* __restrict__ part_r, const double* __restrict__ part_a,
    double* __restrict__ R, double* __restrict__ bp) {
  __shared__ double sh[4];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double r = 0.0, rs = 0.0, xa = 0.0, ba = 0.0;
  if(i < n) {
    const double b = lcg_value(lcg_state(
        seed, static_cast<uint64_t>(i) +
                  static_cast<uint64_t>(n) * static_cast<uint64_t>(n)));
    double ax = 0.0;
    for(int c = 0; c < nch_c; ++c) {
      ax += part_r[static_cast<size_t>(c) * n + i];
      rs += part_a[static_cast<size_t>(c) * n + i];
    }
    r    = b - ax;
    R[i] = r;
    xa   = fabs(X[i]);
    ba   = fabs(b);
  }
  const double m_r  = block_max(fabs(r), sh);
  const double m_rs = block_max(rs, sh);
  const double m_x  = block_max(xa, sh);
  const double m_b  = block_max(ba, sh);
  if(threadIdx.x == 0) {
    double* o = bp + static_cast<size_t>(blockIdx.x) * 4;
    o[0]      = m_r;
    o[1]      = m_rs;
    o[2]      = m_x;
    o[3]      = m_b;
  }
}

__global__ __launch_bounds__(256) void ir_reduce_kernel(
    int n, int nblk, const double* __restrict__ bp, double* __restrict__ out) {
  __shared__ double sh[4];
  double v[4] = {0, 0, 0, 0};
  for(int b = threadIdx.x; b < nblk; b += blockDim.x)
#pragma unroll
    for(int q = 0; q < 4; ++q) v[q] = fmax(v[q], bp[static_cast<size_t>(b) * 4 + q]);
  const double rnorm = block_max(v[0], sh);
  const double anorm = block_max(v[1], sh);
  const double xnorm = block_max(v[2], sh);
  const double bnorm = block_max(v[3], sh);
  if(threadIdx.x == 0) {
    const double eps = 0x1p-53;
    out[0] = rnorm / (eps * (anorm * xnorm + bnorm) * static_cast<double>(n));
    out[1] = rnorm;
    out[2] = anorm;
    out[3] = xnorm;
    out[4] = bnorm;
  }
}

}  // namespace

template <typename T>
hipError_t matgen_t(hipStream_t s, int m, int n, T* A, int lda, int seed,
                    double diag_shift) {
  if(m <= 0 || n <= 0) return hipSuccess;
  const Affine colj = lcg_jump(static_cast<uint64_t>(m));
  const int    gx   = (m + kGenRows - 1) / kGenRows;
  const int    gy   = (n + kGenCols - 1) / kGenCols;
  hipLaunchKernelGGL((matgen_kernel<T>), dim3(static_cast<unsigned>(gx) * gy),
                     dim3(kGenRows), 0, s, m, n, A, lda,
                     static_cast<uint64_t>(seed), diag_shift, colj.a, colj.c,
                     gx);
  return hipGetLastError();
}

hipError_t launch_matgen(hipStream_t s, int m, int n, double* A, int lda,
                         int seed, double diag_shift) {
  return matgen_t(s, m, n, A, lda, seed, diag_shift);
}
hipError_t launch_matgen(hipStream_t s, int m, int n, float* A, int lda,
                         int seed, double diag_shift) {
  return matgen_t(s, m, n, A, lda, seed, diag_shift);
}
hipError_t launch_matgen(hipStream_t s, int m, int n, __half* A, int lda,
                         int seed, double diag_shift) {
  return matgen_t(s, m, n, A, lda, seed, diag_shift);
}

size_t check_part_len(int n) {
  if(n <= 0) return 16;
  const size_t nch_c = (n + kChkCols - 1) / kChkCols;
  const size_t nch_r = (n + kChkRows - 1) / kChkRows;
  const size_t nblk  = (n + 255) / 256;
  return (2 * nch_c + nch_r) * static_cast<size_t>(n) + 8 + 8 * nblk + 16;
}

hipError_t launch_check(hipStream_t s, int n, int seed, double diag_shift,
                        const double* X, double* part, size_t part_len,
                        double* out_dev) {
  if(n <= 0) return hipErrorInvalidValue;
  if(part_len < check_part_len(n)) return hipErrorInvalidValue;
  const int nch_c = (n + kChkCols - 1) / kChkCols;
  const int nch_r = (n + kChkRows - 1) / kChkRows;
  const int nblk  = (n + 255) / 256;
  double* part_r = part;
  double* part_a = part_r + static_cast<size_t>(nch_c) * n;
  double* part_c = part_a + static_cast<size_t>(nch_c) * n;
  double* bp     = part_c + static_cast<size_t>(nch_r) * n;
  const Affine colj = lcg_jump(static_cast<uint64_t>(n));
  const uint64_t sd = static_cast<uint64_t>(seed);
  hipLaunchKernelGGL(check_rows_kernel,
                     dim3(static_cast<unsigned>(nblk) * nch_c), dim3(256), 0, s,
                     n, sd, diag_shift, X, part_r, part_a, colj.a, colj.c,
                     nblk);
  hipLaunchKernelGGL(check_cols_kernel,
                     dim3(static_cast<unsigned>(nblk) * nch_r), dim3(256), 0, s,
                     n, sd, diag_shift, part_c, nblk);
  hipLaunchKernelGGL(check_finish_kernel, dim3(nblk), dim3(256), 0, s, n,
                     nch_c, nch_r, sd, X, part_r, part_a, part_c, bp);
  hipLaunchKernelGGL(check_reduce_kernel, dim3(1), dim3(256), 0, s, n, nblk,
                     bp, out_dev);
  return hipGetLastError();
}

size_t ir_part_len(int n) {
  if(n <= 0) return 16;
  const size_t nch_c = (n + kChkCols - 1) / kChkCols;
  const size_t nblk  = (n + 255) / 256;
  return 2 * nch_c * static_cast<size_t>(n) + 4 * nblk + 16;
}

hipError_t launch_ir_residual(hipStream_t s, int n, int seed,
                              double diag_shift, const double* x, double* r,
                              double* part, double* out) {
  if(n <= 0) return hipErrorInvalidValue;
  const int nch_c = (n + kChkCols - 1) / kChkCols;
  const int nblk  = (n + 255) / 256;
  double* part_r = part;
  double* part_a = part_r + static_cast<size_t>(nch_c) * n;
  double* bp     = part_a + static_cast<size_t>(nch_c) * n;
  const Affine   colj = lcg_jump(static_cast<uint64_t>(n));
  const uint64_t sd   = static_cast<uint64_t>(seed);
  hipLaunchKernelGGL(check_rows_kernel,
                     dim3(static_cast<unsigned>(nblk) * nch_c), dim3(256), 0, s,
                     n, sd, diag_shift, x, part_r, part_a, colj.a, colj.c,
                     nblk);
  hipLaunchKernelGGL(ir_finish_kernel, dim3(nblk), dim3(256), 0, s, n, nch_c,
                     sd, x, part_r, part_a, r, bp);
  hipLaunchKernelGGL(ir_reduce_kernel, dim3(1), dim3(256), 0, s, n, nblk, bp,
                     out);
  return hipGetLastError();
}

}  // namespace hplx
