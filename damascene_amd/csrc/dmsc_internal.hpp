// dmsc_internal.hpp -- shared declarations of the MI355X gPb library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "dmsc.h"
#include "tables.hpp"

namespace dmsc {

#define DMSC_HIP(call)                                                            \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return ::dmsc::fail(DMSC_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call, \
                          hipGetErrorString(e_));                                 \
  } while (0)

// ---------------------------------------------------------------- context
constexpr int kArenas = 6;

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  // grow-only device scratch arenas (reused across calls on this context):
  // 0 pipeline planes, 1 local-cue frames, 2 small solver state, 3 Krylov
  // basis, 4-5 batch/backward work
  void* arena[kArenas] = {};
  size_t arena_bytes[kArenas] = {};
  float* d_texton_coeffs = nullptr;
  int* h_pinned = nullptr;  // pinned host block (64 KiB) for asynchronous read-backs
  hipStream_t aux = nullptr;  // second stream (default priority) for work overlapped with `stream`
  hipEvent_t aux_ev[2] = {};
  std::vector<Ctx*> lanes;  // child contexts of batch calls, each with its own stream
  // dmsc_gpb runs the L/a/b local cues on `aux` beside the textons; batch
  // lanes do not (the lanes already overlap whole images, and a second
  // stream per lane makes them contend for the hardware queues)
  bool overlap_stages = true;
  // how the last eigensolve on this context stopped (DMSC_STOP_*) and the
  // largest residual |A u - lambda u| of its wanted pairs at that point
  int eig_stop = 0;
  double eig_residual = 0.0;
  ~Ctx();
  int pinned(int** out);
  int aux_stream(hipStream_t* s, hipEvent_t* ev0, hipEvent_t* ev1);
  int scratch(int slot, size_t bytes, void** ptr);
  int adopt(int slot, void* p, size_t bytes);  // replace arena `slot` by p
  int texton_coeffs(const float** out);
};

}  // namespace dmsc

struct dmsc_ctx : public dmsc::Ctx {};
