// hplx_pdgetrs_noalloc: the solve with stored factors (hplx_pdgetrs) for
// callers that cannot allocate device memory (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc).  It needs no context:
// its scratch is a file-scope device array, resolved once per device with
// hipGetSymbolAddress, and the work is queued on the caller's stream with
// no allocation and no synchronisation.
#include <mutex>

#include "api_common.hpp"

namespace {

// Scratch of every call on one device: the composed pivots of a panel
// (2 kMaxNB ints) and the split-K partial products of the transposed solve
// (kMaxNB^2 doubles, the size hplx_pdgetrs gives them, so both entry points
// split alike and return the same bits).  8 MiB + 8 KiB per device.
constexpr size_t kWorkDoubles = static_cast<size_t>(hplx::kMaxNB) * hplx::kMaxNB;
constexpr size_t kPermInts    = 2 * static_cast<size_t>(hplx::kMaxNB);
constexpr size_t kScratchBytes = kWorkDoubles * sizeof(double) + kPermInts * sizeof(int);

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex    g_lock;
void*         g_ptr[kMaxDevices] = {};  // the array's address on each device

// Device address of `scratch` on the current device (resolved on the first
// call there), or null with the HIP error in *err.
void* scratch_on_current_device(hipError_t* err) {
  int dev = 0;
  *err    = hipGetDevice(&dev);
  if(*err != hipSuccess) return nullptr;
  if(dev < 0 || dev >= kMaxDevices) {
    *err = hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> g(g_lock);
  if(!g_ptr[dev]) {
    void* p = nullptr;
    *err    = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if(*err != hipSuccess) return nullptr;
    g_ptr[dev] = p;
  }
  return g_ptr[dev];
}

}  // namespace

extern "C" {

int hplx_pdgetrs_noalloc(const hplx_pmat* mat, int trans, int nrhs, double* B,
                         int ldb, void* stream) {
  if(!mat || !mat->A || !mat->ipiv || mat->n < 0 || mat->nb < 1 ||
     mat->ld < mat->n || nrhs < 0 || (nrhs > 0 && (!B || ldb < mat->n)) ||
     (trans != 0 && trans != 1))
    return HPLX_ERR_INVALID;
  // the scratch holds the composed pivots of one panel of at most kMaxNB
  if(mat->nb > hplx::kMaxNB) return HPLX_ERR_UNSUPPORTED;
  if(nrhs == 0 || mat->n == 0) return HPLX_SUCCESS;
  hipError_t e   = hipSuccess;
  void*      ptr = scratch_on_current_device(&e);
  if(!ptr) return HPLX_ERR_HIP;
  double* work  = static_cast<double*>(ptr);
  int*    permU = reinterpret_cast<int*>(work + kWorkDoubles);
  e = hplx::getrs_enqueue(static_cast<hipStream_t>(stream), mat->A, mat->n,
                          mat->nb, mat->ld, mat->ipiv, trans, nrhs, B, ldb,
                          permU, work, kWorkDoubles);
  return e == hipSuccess ? HPLX_SUCCESS : HPLX_ERR_HIP;
}

}  // extern "C"
