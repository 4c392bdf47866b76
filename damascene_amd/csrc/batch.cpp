// batch.cpp -- many images of uneven size through the gPb pipeline.
//
// The reference handles one image per process run (damascene.cu:130-207,
// 269-470).  A batch call here spreads its images over `lanes` child
// contexts of the caller's context, each with its own stream (default
// priority) and host thread.  A lane always takes the largest image left:
// long eigensolves start first (longest-processing-time order) and a lane's
// scratch arenas are sized by the first image it sees.  One image's Lanczos
// steps are dependent launches that leave most of the chip idle while each
// block reduces the previous step's partial sums; steps of other images on
// other streams fill those gaps.  Every image runs the kernels and launch
// shapes of dmsc_gpb, so its results are bit-identical to a single-image
// call.
#include <algorithm>
#include <atomic>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "dmsc_internal.hpp"
#include "kernels.hpp"

namespace dmsc {

int batch_lanes = 4;  // default lane count (tune key "batch_lanes")

namespace {

constexpr int kMaxLanes = 16;

int make_lanes(Ctx* ctx, int lanes) {
  DMSC_HIP(hipSetDevice(ctx->device));
  while ((int)ctx->lanes.size() < lanes) {
    hipStream_t s = nullptr;
    DMSC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    dmsc_ctx* c = new dmsc_ctx();
    c->device = ctx->device;
    c->stream = s;
    c->own_stream = true;
    c->overlap_stages = false;
    ctx->lanes.push_back(c);
  }
  return DMSC_OK;
}

}  // namespace
}  // namespace dmsc

using namespace dmsc;

extern "C" int dmsc_gpb_batch(dmsc_ctx* ctx, int count, const dmsc_batch_image* images,
                              const dmsc_gpb_params* params, int lanes) {
  DMSC_ARG(ctx != nullptr, "null context");
  DMSC_ARG(count >= 0, "negative image count");
  if (count == 0) return DMSC_OK;
  DMSC_ARG(images != nullptr, "null image list");
  for (int i = 0; i < count; i++) {
    const dmsc_batch_image& im = images[i];
    DMSC_ARG(im.w >= kMinSide && im.h >= kMinSide && (long)im.w * im.h < (1L << 28),
             "image %d: size %dx%d out of range", i, im.w, im.h);
    DMSC_ARG(im.d_rgb != nullptr && im.out != nullptr, "image %d: null pointer", i);
  }
  if (lanes <= 0) lanes = batch_lanes;
  lanes = std::max(1, std::min({lanes, kMaxLanes, count}));
  DMSC_TRY(make_lanes(ctx, lanes));

  // the lanes start after the work already on the caller's stream
  hipEvent_t start = nullptr;
  DMSC_HIP(hipEventCreateWithFlags(&start, hipEventDisableTiming));
  int rc = DMSC_OK;
  if (hipEventRecord(start, ctx->stream) != hipSuccess) rc = fail(DMSC_ERR_HIP, "hipEventRecord failed");
  for (int l = 0; l < lanes && rc == DMSC_OK; l++)
    if (hipStreamWaitEvent(ctx->lanes[l]->stream, start, 0) != hipSuccess)
      rc = fail(DMSC_ERR_HIP, "hipStreamWaitEvent failed");
  if (rc != DMSC_OK) {
    (void)hipEventDestroy(start);
    return rc;
  }

  std::vector<int> order(count);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (long)images[a].w * images[a].h > (long)images[b].w * images[b].h;
  });
  std::atomic<int> next{0};
  std::vector<int> lane_rc(lanes, DMSC_OK);
  std::vector<std::string> lane_msg(lanes);
  auto work = [&](int l) {
    if (hipSetDevice(ctx->device) != hipSuccess) {
      lane_rc[l] = DMSC_ERR_HIP;
      lane_msg[l] = "hipSetDevice failed on a batch lane";
      next.store(count);
      return;
    }
    dmsc_ctx* lane = static_cast<dmsc_ctx*>(ctx->lanes[l]);
    for (;;) {
      const int i = next.fetch_add(1);
      if (i >= count) break;
      const dmsc_batch_image& im = images[order[i]];
      const int r = dmsc_gpb(lane, im.w, im.h, im.d_rgb, params, im.out, 0);
      if (r != DMSC_OK) {
        lane_rc[l] = r;
        lane_msg[l] = std::string(dmsc_last_error_string()) + " (batch image " + std::to_string(order[i]) + ")";
        next.store(count);  // the other lanes stop after their current image
        break;
      }
    }
  };
  std::vector<std::thread> threads;
  threads.reserve(lanes - 1);
  for (int l = 1; l < lanes; l++) threads.emplace_back(work, l);
  work(0);
  for (auto& t : threads) t.join();

  // every image is finished when the call returns
  for (int l = 0; l < lanes; l++)
    if (hipStreamSynchronize(ctx->lanes[l]->stream) != hipSuccess && rc == DMSC_OK)
      rc = fail(DMSC_ERR_HIP, "batch lane %d: stream synchronisation failed", l);
  (void)hipEventDestroy(start);
  for (int l = 0; l < lanes; l++)
    if (lane_rc[l] != DMSC_OK) return fail(lane_rc[l], "%s", lane_msg[l].c_str());
  return rc;
}
