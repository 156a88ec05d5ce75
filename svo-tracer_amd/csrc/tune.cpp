// svo_autotune: picks the launch options of a (batched) frame by timing real frames.
//
// Coordinate descent over the work-group size (64..1024 lanes), the work chunk a wave claims
// per atomic (64..1024 cones: the queue chunk), the refill threshold (lanes idle before a
// wave takes new cones) and the depth of the LDS short stack (levels kept per lane; below it
// a pop walks down from the deepest kept level).  Every option gives bit-identical results, and re-running a frame
// leaves the same state (usage ticks are written with the same frame tick, request counts
// and outputs are overwritten), so tuning on the caller's own frame is safe.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "svo_internal.hpp"
#include "svo_runtime.hpp"

using namespace svo;

namespace {

FrameBuffers::B named(Context *ctx, int name) {
  FrameBuffers::B b;
  auto p = ctx->get(name);
  if (p) {
    b.ptr = p->ptr;
    b.bytes = p->bytes;
  }
  return b;
}

FrameBuffers context_buffers(Context *ctx) {
  FrameBuffers fb;
  fb.bases = named(ctx, SVO_BUF_BASES);
  fb.blocks = named(ctx, SVO_BUF_BLOCKS);
  fb.usages = named(ctx, SVO_BUF_USAGES);
  fb.requests = named(ctx, SVO_BUF_CHILD_REQUESTS);
  fb.request_id = named(ctx, SVO_BUF_CHILD_REQUEST_ID);
  fb.rgba = named(ctx, SVO_BUF_OUTPUT_IMAGE);
  fb.hit = named(ctx, SVO_BUF_HIT_INDICES);
  fb.visits = named(ctx, SVO_BUF_VISIT_COUNTS);
  fb.hashes = named(ctx, SVO_BUF_VISIT_HASHES);
  fb.status = named(ctx, SVO_BUF_RAY_STATUS);
  fb.locations = named(ctx, SVO_BUF_LOCATIONS);
  fb.depths = named(ctx, SVO_BUF_DEPTHS);
  fb.bg_queue = named(ctx, SVO_BUF_BACKGROUND_QUEUE);
  fb.mat_queue = named(ctx, SVO_BUF_MATERIAL_QUEUE);
  fb.bg_count = named(ctx, SVO_BUF_BACKGROUND_QUEUE_ID);
  fb.mat_count = named(ctx, SVO_BUF_MATERIAL_QUEUE_ID);
  return fb;
}

// Average milliseconds per frame of `frames` frames with options `o` (after one warm-up).
int time_frames(Context *ctx, const svo_trace_input *ins, const float *ap, uint32_t n, const svo_frame_options &o,
                uint32_t frames, float *ms) {
  const FrameBuffers fb = context_buffers(ctx);
  int rc = trace_batch_impl(ctx, ins, ap, n, o, fb);
  if (rc) return rc;
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return fail(SVO_ERR_HIP, "hipEventCreate");
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return fail(SVO_ERR_HIP, "hipEventCreate");
  }
  (void)hipEventRecord(a, ctx->stream);
  for (uint32_t f = 0; f < frames && !rc; f++) rc = trace_batch_impl(ctx, ins, ap, n, o, fb);
  (void)hipEventRecord(b, ctx->stream);
  hipError_t e = hipEventSynchronize(b);
  float t = 0.f;
  if (!rc && e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc) return rc;
  if (e != hipSuccess) return hip_fail(e, "autotune timing");
  *ms = t / (float)frames;
  return SVO_OK;
}

}  // namespace

extern "C" int svo_autotune(svo_context *ctx, const svo_trace_input *inputs, const float *apertures,
                            uint32_t n_cameras, uint32_t frames, const svo_frame_options *base,
                            svo_frame_options *best, float *best_ms) {
  if (!ctx || !inputs || !best) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_autotune: null argument");
  if (frames == 0) frames = 3;
  svo_frame_options cur;
  std::memset(&cur, 0, sizeof cur);
  if (base) cur = *base;
  float cur_ms = 0.f;
  int rc = time_frames(ctx, inputs, apertures, n_cameras, cur, frames, &cur_ms);
  if (rc) return rc;
  auto try_field = [&](uint32_t svo_frame_options::*field, const std::vector<uint32_t> &values) -> int {
    for (uint32_t v : values) {
      if (cur.*field == v) continue;
      svo_frame_options cand = cur;
      cand.*field = v;
      float ms = 0.f;
      int r = time_frames(ctx, inputs, apertures, n_cameras, cand, frames, &ms);
      if (r == SVO_ERR_INVALID_ARGUMENT) continue;  // e.g. a work-group too large for LDS
      if (r) return r;
      if (ms < cur_ms) {
        cur = cand;
        cur_ms = ms;
      }
    }
    return SVO_OK;
  };
  if ((rc = try_field(&svo_frame_options::workgroup_size, {64, 128, 256, 512, 1024}))) return rc;
  if ((rc = try_field(&svo_frame_options::work_chunk, {64, 128, 256, 512, 1024}))) return rc;
  if ((rc = try_field(&svo_frame_options::refill_lanes, {16, 32, 48, 64}))) return rc;
  if ((rc = try_field(&svo_frame_options::stack_levels, {4, 6, 8, 12, 16, 24}))) return rc;
  *best = cur;
  if (best_ms) *best_ms = cur_ms;
  return SVO_OK;
}
