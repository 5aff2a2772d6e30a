// gb_noalloc.hip -- an entry point that allocates no device memory.
//
// gb_pagerank_ref_noalloc runs the reference-form PageRank loop of
// gb_pagerank_ref (the same kernels, bit-identical results) with its scratch
// -- diff[V] and wb[V] (FP64) and the not-converged flag -- in a fixed
// file-scope device array instead of buffers allocated per call.  It is for
// callers that cannot allocate: serving loops, harnesses that forbid
// hipMalloc.  Its limit is the size of that array: V <= GB_NOALLOC_PR_MAX_V.
//
// The array's name is a host-side stand-in in host code; its device address
// is looked up once per device with hipGetSymbolAddress (at gb_init, else at
// the first call on a device) and only that address reaches the kernels.
// Every call on a device shares that device's array, which keeps the last
// call's contents; the loop initialises diff and wb and clears the flag
// itself.  Calls that overlap on one device race on the array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "gb_internal.hpp"

namespace {

constexpr int64_t kMaxV = GB_NOALLOC_PR_MAX_V;
// diff[kMaxV], wb[kMaxV] (double), then the flag (int) in its own 256 bytes
constexpr size_t kScratchBytes = 2 * sizeof(double) * (size_t)kMaxV + 256;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

// The current device's scratch address, looked up on first use.
gb_status scratch_for(int32_t device, void **out) {
  if (device < 0 || device >= kMaxDevices)
    return gb_fail(GB_ERR_UNSUPPORTED, "no-allocation scratch: device %d beyond %d", device,
                   kMaxDevices);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return gb_fail(GB_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                     hipGetErrorString(e));
    }
    g_scratch[device] = p;
  }
  *out = g_scratch[device];
  return GB_OK;
}

}  // namespace

// Called by gb_init with `device` current.
gb_status gb_noalloc_setup(int32_t device) {
  void *p = nullptr;
  return scratch_for(device, &p);
}

extern "C" gb_status gb_pagerank_ref_noalloc(int32_t V, const int32_t *d_irp,
                                             const int32_t *d_icol, const double *d_recip,
                                             int32_t max_iter, double *d_x, int32_t *iterations,
                                             void *stream) {
  if (!d_irp || !d_recip || !d_x)
    return gb_fail(GB_ERR_INVALID, "gb_pagerank_ref_noalloc: null argument");
  if (V <= 0) return gb_fail(GB_ERR_INVALID, "gb_pagerank_ref_noalloc: empty graph");
  if ((int64_t)V > kMaxV)
    return gb_fail(GB_ERR_UNSUPPORTED,
                   "gb_pagerank_ref_noalloc: V = %d exceeds the fixed scratch (at most %lld "
                   "vertices)",
                   V, (long long)kMaxV);
  int dev = 0;
  GB_HIP(hipGetDevice(&dev));
  void *base = nullptr;
  GB_TRY(scratch_for(dev, &base));
  double *diff = (double *)base;
  double *wb = diff + kMaxV;
  int *flag = (int *)(wb + kMaxV);
  return gb_pagerank_ref_core(V, d_irp, d_icol, d_recip, max_iter, d_x, iterations,
                              gb_stream(stream), diff, wb, flag, nullptr, nullptr);
}
