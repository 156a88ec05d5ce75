// svo_trace_frame: one primary frame on the device (MainWindow.RunWavefront,
// MainWindow.cs:356-426, without any host round-trip per traversal iteration).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "svo_internal.hpp"
#include "svo_kernels.hpp"
#include "svo_runtime.hpp"

namespace svo {

#define HIP_TRY(expr, what)                          \
  do {                                               \
    hipError_t e_ = (expr);                          \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

int update_camera_table(Context *ctx, const svo_trace_input &in) {
  const uint32_t W = in.screen_size[0], H = in.screen_size[1];
  if (ctx->lut_valid && ctx->lut_w == W && ctx->lut_h == H && std::memcmp(&ctx->lut_fov0, &in.fov[0], 4) == 0 &&
      std::memcmp(&ctx->lut_fov1, &in.fov[1], 4) == 0)
    return SVO_OK;
  std::vector<float> lut(2 * ((size_t)W + H));
  camera_lut(in, lut.data(), lut.data() + W, lut.data() + 2 * W, lut.data() + 2 * W + H);
  int rc = ensure_scratch(ctx->scratch_lut, lut.size() * sizeof(float));
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(ctx->scratch_lut->ptr, lut.data(), lut.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream),
          "camera table upload");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "camera table sync");
  ctx->lut_w = W;
  ctx->lut_h = H;
  ctx->lut_fov0 = in.fov[0];
  ctx->lut_fov1 = in.fov[1];
  ctx->lut_valid = true;
  return SVO_OK;
}

namespace {

int check_size(const FrameBuffers::B &b, size_t need, const char *what) {
  if (b.ptr && b.bytes < need)
    return fail(SVO_ERR_BUFFER_TOO_SMALL, std::string("trace frame: buffer ") + what + " holds " +
                                              std::to_string(b.bytes) + " bytes, needs " + std::to_string(need));
  return SVO_OK;
}

int choose_launch(Context *ctx, uint32_t B, uint32_t n_bases, const svo_frame_options &o, bool *bases_in_lds,
                  uint32_t *stack_levels, LaunchShape *shape) {
  int wg = o.workgroup_size ? (int)o.workgroup_size : 256;
  if (wg != 64 && wg != 128 && wg != 256 && wg != 512)
    return fail(SVO_ERR_INVALID_ARGUMENT, "trace frame: workgroup_size must be 64, 128, 256 or 512");
  uint32_t max_depth = o.max_tree_depth ? o.max_tree_depth : (ctx->tree_max_depth ? ctx->tree_max_depth : 24u);
  if (max_depth > 60) return fail(SVO_ERR_INVALID_ARGUMENT, "trace frame: max_tree_depth must be <= 60");
  uint32_t levels = o.stack_levels ? o.stack_levels : (max_depth >= B + 2 ? max_depth - (B + 2) + 1 : 1u);
  levels = std::max(1u, std::min(levels, 32u));
  const size_t lds_limit = 64 * 1024;
  size_t bases_bytes = ((size_t)n_bases * 2 + 3) & ~(size_t)3;
  *bases_in_lds = bases_bytes <= 32 * 1024;
  if (!*bases_in_lds) bases_bytes = 0;
  while (levels > 1 && bases_bytes + (size_t)levels * wg * 4 > lds_limit) levels--;
  size_t lds = bases_bytes + (size_t)levels * wg * 4;
  if (lds > lds_limit) return fail(SVO_ERR_INVALID_ARGUMENT, "trace frame: workgroup too large for LDS");
  const int waves = o.waves_per_simd ? (int)o.waves_per_simd : 6;
  if (waves != 6)
    return fail(SVO_ERR_INVALID_ARGUMENT, "trace frame: waves_per_simd must be 0 (default) or 6");
  int per_cu = 0;
  if (trace_occupancy(wg, waves, lds, &per_cu) != 0 || per_cu <= 0)
    return fail(SVO_ERR_HIP, "trace frame: occupancy query failed");
  int grid = per_cu * std::max(1, ctx->cu_count);
  grid = std::max(8, (grid / 8) * 8);
  shape->workgroup = wg;
  shape->waves = waves;
  shape->grid = grid;
  shape->lds_bytes = lds;
  *stack_levels = levels;
  return SVO_OK;
}

}  // namespace

int trace_frame_impl(Context *ctx, const svo_trace_input &in, const svo_frame_options &o, const FrameBuffers &fb) {
  const uint32_t W = in.screen_size[0], H = in.screen_size[1], B = in.base_depth;
  if (W == 0 || H == 0 || (uint64_t)W * H > (1ull << 31))
    return fail(SVO_ERR_INVALID_ARGUMENT, "trace frame: screen size must be non-zero and at most 2^31 cones");
  if (B < 1 || B > 7) return fail(SVO_ERR_INVALID_ARGUMENT, "trace frame: base_depth must be 1..7");
  const size_t N = (size_t)W * H;
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  if (!fb.bases.ptr) return fail(SVO_ERR_UNKNOWN_BUFFER, "trace frame: buffer Bases not initialised");
  if (!fb.blocks.ptr) return fail(SVO_ERR_UNKNOWN_BUFFER, "trace frame: buffer Blocks not initialised");
  const uint32_t n_bases = pow_sum(B);
  int rc;
  if ((rc = check_size(fb.bases, (size_t)n_bases * 2, "Bases"))) return rc;
  if ((rc = check_size(fb.request_id, 4, "ChildRequestID"))) return rc;
  if ((rc = check_size(fb.rgba, 16 * N, "OutputImage"))) return rc;
  if ((rc = check_size(fb.hit, 4 * N, "HitIndices"))) return rc;
  if ((rc = check_size(fb.visits, 4 * N, "VisitCounts"))) return rc;
  if ((rc = check_size(fb.hashes, 8 * N, "VisitHashes"))) return rc;
  if ((rc = check_size(fb.status, N, "RayStatus"))) return rc;

  FrameParams p;
  std::memset(&p, 0, sizeof p);
  p.bases = (const uint16_t *)fb.bases.ptr;
  p.n_bases = n_bases;
  p.blocks = (const uint4 *)fb.blocks.ptr;
  p.n_blocks = (uint32_t)std::min<size_t>(fb.blocks.bytes / 16, 0xffffffffu);
  p.usage16 = (uint16_t *)fb.usages.ptr;
  p.n_usages = fb.usages.ptr ? (uint32_t)std::min<size_t>(fb.usages.bytes / 8, 0xffffffffu) : 0u;
  svo_child_request *requests = (svo_child_request *)fb.requests.ptr;
  uint32_t *request_id = (uint32_t *)fb.request_id.ptr;
  const uint32_t req_capacity =
      requests ? (uint32_t)std::min<size_t>(fb.requests.bytes / sizeof(svo_child_request), in.max_child_request_id) : 0u;
  p.rgba = (float4 *)fb.rgba.ptr;
  p.hit = (uint32_t *)fb.hit.ptr;
  p.visits = (uint32_t *)fb.visits.ptr;
  p.hashes = (uint64_t *)fb.hashes.ptr;
  p.status = (uint8_t *)fb.status.ptr;
  if (o.flags & SVO_FRAME_STATE) {
    if (!fb.locations.ptr || !fb.depths.ptr)
      return fail(SVO_ERR_UNKNOWN_BUFFER, "trace frame: SVO_FRAME_STATE needs Locations and Depths");
    if ((rc = check_size(fb.locations, 32 * N, "Locations"))) return rc;
    if ((rc = check_size(fb.depths, N, "Depths"))) return rc;
    p.loc_out = (uint64_t *)fb.locations.ptr;
    p.depth_out = (uint8_t *)fb.depths.ptr;
  }
  if (o.flags & SVO_FRAME_QUEUES) {
    if (!fb.bg_queue.ptr || !fb.mat_queue.ptr || !fb.bg_count.ptr || !fb.mat_count.ptr)
      return fail(SVO_ERR_UNKNOWN_BUFFER, "trace frame: SVO_FRAME_QUEUES needs Background/MaterialQueue and their IDs");
    if ((rc = check_size(fb.bg_queue, 4 * N, "BackgroundQueue"))) return rc;
    if ((rc = check_size(fb.mat_queue, 4 * N, "MaterialQueue"))) return rc;
    if ((rc = check_size(fb.bg_count, 4, "BackgroundQueueID"))) return rc;
    if ((rc = check_size(fb.mat_count, 4, "MaterialQueueID"))) return rc;
  }
  const bool queues = (o.flags & SVO_FRAME_QUEUES) != 0;
  const bool compaction = requests || request_id || queues;
  if (compaction && !p.status) {
    if ((rc = ensure_scratch(ctx->scratch_status, N))) return rc;
    p.status = (uint8_t *)ctx->scratch_status->ptr;
  }
  if (requests || request_id) {
    if ((rc = ensure_scratch(ctx->scratch_requests, N * sizeof(svo_child_request)))) return rc;
    p.req_sparse = (svo_child_request *)ctx->scratch_requests->ptr;
  }
  if ((rc = ensure_scratch(ctx->scratch_small, kSmallScratchBytes))) return rc;
  char *small = (char *)ctx->scratch_small->ptr;
  p.counters = (uint32_t *)(small + kCountersOffset);
  p.totals = (unsigned long long *)(small + kTotalsOffset);
  if ((rc = update_camera_table(ctx, in))) return rc;
  p.lut = (const float *)ctx->scratch_lut->ptr;

  for (int k = 0; k < 9; k++) p.facing[k] = in.facing[k];
  for (int k = 0; k < 3; k++) p.origin[k] = in.origin[k];
  p.dof[0] = in.dof[0];
  p.dof[1] = in.dof[1];
  p.cone_fov = in.fov[0] / (float)W;  // setupInitialTrace: FoV = FoV[0] / ScreenSize.x
  p.W = W;
  p.H = H;
  p.B = B;
  p.tick = in.tick;
  p.tiles_x = (W + 7) / 8;
  const uint32_t tiles = p.tiles_x * ((H + 7) / 8);
  for (int r = 0; r <= 8; r++) p.range_begin[r] = (uint32_t)(((uint64_t)tiles * r) / 8);
  p.refill = o.refill_lanes ? std::min<uint32_t>(o.refill_lanes, 64u) : 64u;
  // defaults: a wave takes one 8x8 tile per claim and refills only when all its lanes are
  // idle -- spatial coherence of the 64 cones matters more than lane utilisation
  p.chunk = o.work_chunk ? ((std::min(o.work_chunk, 1u << 20) + 63u) / 64u) * 64u : 64u;

  bool bases_in_lds = false;
  uint32_t levels = 1;
  LaunchShape shape;
  if ((rc = choose_launch(ctx, B, n_bases, o, &bases_in_lds, &levels, &shape))) return rc;
  p.bases_in_lds = bases_in_lds ? 1u : 0u;
  p.stack_levels = levels;

  if (launch_frame_prepare(p.counters, p.totals, ctx->stream) != 0)
    return hip_fail(hipGetLastError(), "trace frame: prepare launch");
  std::pair<hipEvent_t, hipEvent_t> *ev = nullptr;
  if (ctx->kernel_timing) {
    if (ctx->kernel_events_used == ctx->kernel_events.size()) {
      std::pair<hipEvent_t, hipEvent_t> pr;
      HIP_TRY(hipEventCreate(&pr.first), "hipEventCreate");
      HIP_TRY(hipEventCreate(&pr.second), "hipEventCreate");
      ctx->kernel_events.push_back(pr);
    }
    ev = &ctx->kernel_events[ctx->kernel_events_used++];
    HIP_TRY(hipEventRecord(ev->first, ctx->stream), "hipEventRecord");
  }
  if (launch_trace_frame(p, shape, ctx->stream) != 0) return hip_fail(hipGetLastError(), "trace frame: trace launch");
  if (ev) HIP_TRY(hipEventRecord(ev->second, ctx->stream), "hipEventRecord");

  if (compaction) {
    const uint32_t chunks = compaction_chunks((uint32_t)N);
    if ((rc = ensure_scratch(ctx->scratch_counts, (size_t)std::max(1u, chunks) * 16))) return rc;
    CompactParams c;
    std::memset(&c, 0, sizeof c);
    c.status = p.status;
    c.n = (uint32_t)N;
    c.counts = (uint32_t *)ctx->scratch_counts->ptr;
    if (queues) {
      c.bg_queue = (uint32_t *)fb.bg_queue.ptr;
      c.mat_queue = (uint32_t *)fb.mat_queue.ptr;
      c.bg_count = (uint32_t *)fb.bg_count.ptr;
      c.mat_count = (uint32_t *)fb.mat_count.ptr;
    }
    c.req_count = request_id;
    c.req_sparse = p.req_sparse;
    c.requests = requests;
    c.req_capacity = req_capacity;
    c.totals = p.totals;
    if (launch_compaction(c, ctx->stream) != 0) return hip_fail(hipGetLastError(), "trace frame: compaction launch");
  }
  ctx->last_rays = N;
  return SVO_OK;
}

}  // namespace svo

using namespace svo;

static FrameBuffers::B named(Context *ctx, int name) {
  FrameBuffers::B b;
  auto p = ctx->get(name);
  if (p) {
    b.ptr = p->ptr;
    b.bytes = p->bytes;
  }
  return b;
}

extern "C" int svo_trace_frame(svo_context *ctx, const svo_trace_input *input, const svo_frame_options *options) {
  if (!ctx || !input) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_trace_frame: null argument");
  svo_frame_options o;
  std::memset(&o, 0, sizeof o);
  if (options) o = *options;
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
  return trace_frame_impl(ctx, *input, o, fb);
}
