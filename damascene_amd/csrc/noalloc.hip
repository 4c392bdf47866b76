// noalloc.hip -- a k-means entry point that allocates no device memory.
//
// dmsc_kmeans_noalloc runs the same kernels as dmsc_kmeans (textons.hip),
// with their scratch (the k-means state and the per-pixel |x|^2) carved out
// of one file-scope device array instead of the context's grow-only arenas
// (8 MiB: about 2.09 million pixels, a 1920 x 1080 image fits).  Each
// GPU holds its own copy of the array from the time the library's code
// object is loaded until the process ends.  In host code `scratch` names a
// host-side stand-in, not device memory: the device address of each GPU's
// copy comes from hipGetSymbolAddress (called with that device current) and
// only that address is ever handed to a kernel or a copy.
#include <mutex>

#include "dmsc_internal.hpp"
#include "kernels.hpp"

namespace dmsc {

constexpr size_t kScratchBytes = (size_t)8 << 20;  // DMSC_KMEANS_NOALLOC_SCRATCH_BYTES

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

// Scratch bytes a call over n pixels needs: the state and |x|^2 per pixel
// (pitch_floats(n) floats); both precisions need the same.
size_t noalloc_need(long n) { return round256(kmeans_state_bytes()) + sizeof(float) * (size_t)pitch_floats(n); }

}  // namespace

// Looks up the current device's copy of `scratch` once per device and keeps
// its address.  The caller has made `device` current.
int noalloc_scratch_init(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return fail(DMSC_ERR_ARG, "device %d out of range for the fixed scratch", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    DMSC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(DMSC_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
    g_scratch_ptr[device] = p;
  }
  if (out) *out = g_scratch_ptr[device];
  return DMSC_OK;
}

}  // namespace dmsc

using namespace dmsc;

extern "C" {

size_t dmsc_kmeans_noalloc_max_pixels(void) {
  long lo = 0, hi = (1L << 28) - 1;  // largest n with noalloc_need(n) <= kScratchBytes
  while (lo < hi) {
    const long mid = (lo + hi + 1) / 2;
    if (noalloc_need(mid) <= kScratchBytes) lo = mid;
    else hi = mid - 1;
  }
  return (size_t)lo;
}

int dmsc_kmeans_noalloc(dmsc_ctx* ctx, int w, int h, int k, int nfilters, const float* resp, int* labels,
                        int max_iter, int conv_thresh, int precision, int* changes, int* iterations) {
  DMSC_ARG(ctx != nullptr, "null context");
  DMSC_ARG(w > 0 && h > 0 && resp && labels, "bad arguments");
  DMSC_ARG((k == 32 || k == 64) && nfilters == kNFilters, "kmeans supports k=32 or 64 over 34 filters");
  DMSC_ARG(max_iter >= 1 && max_iter <= 1000, "max_iter out of range");
  DMSC_ARG(conv_thresh == 0, "only conv_thresh=0 (the reference default) is supported");
  DMSC_ARG(precision == DMSC_PREC_FP32 || precision == DMSC_PREC_MIXED, "bad precision");
  const long n = (long)w * h;
  DMSC_ARG(n < (1L << 28) && noalloc_need(n) <= kScratchBytes,
           "%dx%d needs %zu bytes of k-means scratch; the fixed scratch holds %zu (at most %zu pixels)", w, h,
           noalloc_need(n < (1L << 28) ? n : 0), kScratchBytes, dmsc_kmeans_noalloc_max_pixels());
  DMSC_HIP(hipSetDevice(ctx->device));
  void* base = nullptr;
  DMSC_TRY(noalloc_scratch_init(ctx->device, &base));
  char* state = (char*)base;
  float* pdot = (float*)(state + round256(kmeans_state_bytes()));
  if (precision == DMSC_PREC_MIXED) {
    DMSC_TRY(launch_kmeans_mixed(ctx->stream, w, h, k, resp, labels, pdot, state, max_iter));
  } else {
    DMSC_TRY(launch_kmeans(ctx->stream, w, h, k, nfilters, resp, labels, pdot, state, max_iter));
  }
  if (!changes && !iterations) return DMSC_OK;  // enqueued only; no read-back, no synchronisation
  return kmeans_read_state(ctx->stream, state, changes, iterations);
}

}  // extern "C"
