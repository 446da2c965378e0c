// Small kernels of the mixed-precision path: storage <-> FP32 block copies,
// vector row interchanges (a panel's interchanges as net moves), FP64 <->
// FP32 vector conversions, x += d, and the right-hand side b generated in
// FP64 from the matrix generator (column n of [A|b], state i + n*n).
#include "lcg.hpp"
#include "mxp.hpp"

namespace hplx {
namespace {

template <typename D, typename S>
__device__ inline D cvt(S x) {
  if constexpr(sizeof(D) == 2)
    return __float2half(static_cast<float>(x));
  else if constexpr(sizeof(S) == 2)
    return static_cast<D>(__half2float(x));
  else
    return static_cast<D>(x);
}

// B (m x n) = A (m x n), converted.  Block: 256 rows of one column; gx
// blocks per column.
template <typename S, typename D>
__global__ __launch_bounds__(256) void copy_block_kernel(
    int m, int gx, const S* __restrict__ A, int lda, D* __restrict__ B,
    int ldb) {
  const int j = blockIdx.x / gx;
  const int i = (blockIdx.x % gx) * 256 + threadIdx.x;
  if(i < m)
    B[i + static_cast<size_t>(j) * ldb] =
        cvt<D>(A[i + static_cast<size_t>(j) * lda]);
}

// y[dst_i] = y[src_i] for the cnt pairs (one workgroup; every source is read
// before any destination is written).
template <typename V>
__global__ __launch_bounds__(1024) void apply_swaps_kernel(
    int cnt, const int* __restrict__ pairs, V* __restrict__ y) {
  constexpr int kPer = (kMaxSwapPairs + 1023) / 1024;
  V             v[kPer];
#pragma unroll
  for(int q = 0; q < kPer; ++q) {
    const int i = threadIdx.x + 1024 * q;
    if(i < cnt) v[q] = y[pairs[2 * i + 1]];
  }
  __syncthreads();
#pragma unroll
  for(int q = 0; q < kPer; ++q) {
    const int i = threadIdx.x + 1024 * q;
    if(i < cnt) y[pairs[2 * i]] = v[q];
  }
}

template <typename D>
__global__ __launch_bounds__(256) void axpy_kernel(int n,
                                                   const D* __restrict__ d,
                                                   double* __restrict__ x) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if(i < n) x[i] += static_cast<double>(d[i]);
}

template <typename S, typename D>
__global__ __launch_bounds__(256) void convert_kernel(int n,
                                                      const S* __restrict__ x,
                                                      D* __restrict__ y) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if(i < n) y[i] = static_cast<D>(x[i]);
}

__global__ __launch_bounds__(256) void gen_b_kernel(int n, uint64_t seed,
                                                    double* __restrict__ y) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if(i < n)
    y[i] = lcg_value(lcg_state(seed, static_cast<uint64_t>(i) +
                                         static_cast<uint64_t>(n) *
                                             static_cast<uint64_t>(n)));
}

template <typename S, typename D>
hipError_t copy_t(hipStream_t s, int m, int n, const S* A, int lda, D* B,
                  int ldb) {
  if(m <= 0 || n <= 0) return hipSuccess;
  const long gx = (m + 255) / 256;
  if(gx * n > 0x7fffffffL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((copy_block_kernel<S, D>),
                     dim3(static_cast<unsigned>(gx * n)), dim3(256), 0, s, m,
                     static_cast<int>(gx), A, lda, B, ldb);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_copy_block(hipStream_t s, int m, int n, const __half* A,
                             int lda, float* B, int ldb) {
  return copy_t(s, m, n, A, lda, B, ldb);
}
hipError_t launch_copy_block(hipStream_t s, int m, int n, const float* A,
                             int lda, __half* B, int ldb) {
  return copy_t(s, m, n, A, lda, B, ldb);
}
hipError_t launch_copy_block(hipStream_t s, int m, int n, const double* A,
                             int lda, float* B, int ldb) {
  return copy_t(s, m, n, A, lda, B, ldb);
}
hipError_t launch_copy_block(hipStream_t s, int m, int n, const double* A,
                             int lda, __half* B, int ldb) {
  return copy_t(s, m, n, A, lda, B, ldb);
}

namespace {
template <typename V>
hipError_t apply_swaps_t(hipStream_t s, int cnt, const int* pairs, V* y) {
  if(cnt <= 0) return hipSuccess;
  if(cnt > kMaxSwapPairs) return hipErrorInvalidValue;
  hipLaunchKernelGGL(apply_swaps_kernel<V>, dim3(1), dim3(1024), 0, s, cnt,
                     pairs, y);
  return hipGetLastError();
}
}  // namespace

hipError_t launch_apply_swaps(hipStream_t s, int cnt, const int* pairs,
                              double* y) {
  return apply_swaps_t(s, cnt, pairs, y);
}

hipError_t launch_apply_swaps(hipStream_t s, int cnt, const int* pairs,
                              float* y) {
  return apply_swaps_t(s, cnt, pairs, y);
}

namespace {
template <typename D>
hipError_t axpy_t(hipStream_t s, int n, const D* d, double* x) {
  if(n <= 0) return hipSuccess;
  hipLaunchKernelGGL(axpy_kernel<D>, dim3((n + 255) / 256), dim3(256), 0, s, n,
                     d, x);
  return hipGetLastError();
}
}  // namespace

hipError_t launch_axpy(hipStream_t s, int n, const double* d, double* x) {
  return axpy_t(s, n, d, x);
}

hipError_t launch_axpy(hipStream_t s, int n, const float* d, double* x) {
  return axpy_t(s, n, d, x);
}

namespace {
template <typename S, typename D>
hipError_t convert_t(hipStream_t s, int n, const S* x, D* y) {
  if(n <= 0) return hipSuccess;
  hipLaunchKernelGGL((convert_kernel<S, D>), dim3((n + 255) / 256), dim3(256),
                     0, s, n, x, y);
  return hipGetLastError();
}
}  // namespace

hipError_t launch_to_f32(hipStream_t s, int n, const double* x, float* y) {
  return convert_t(s, n, x, y);
}

hipError_t launch_to_f64(hipStream_t s, int n, const float* y, double* x) {
  return convert_t(s, n, y, x);
}

hipError_t launch_gen_b(hipStream_t s, int n, int seed, double* y) {
  if(n <= 0) return hipSuccess;
  hipLaunchKernelGGL(gen_b_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n,
                     static_cast<uint64_t>(seed), y);
  return hipGetLastError();
}

}  // namespace hplx
