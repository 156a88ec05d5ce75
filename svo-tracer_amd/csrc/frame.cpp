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

// Concatenated camera tables of a batch (cached while the resolutions and FoVs repeat);
// offsets[c] = first float of camera c's table.
int update_batch_tables(Context *ctx, const svo_trace_input *ins, uint32_t n, std::vector<uint32_t> *offsets) {
  std::vector<float> key;
  key.reserve(4 * (size_t)n);
  size_t total = 0;
  offsets->resize(n);
  for (uint32_t c = 0; c < n; c++) {
    key.push_back((float)ins[c].screen_size[0]);
    key.push_back((float)ins[c].screen_size[1]);
    key.push_back(ins[c].fov[0]);
    key.push_back(ins[c].fov[1]);
    (*offsets)[c] = (uint32_t)total;
    total += 2 * ((size_t)ins[c].screen_size[0] + ins[c].screen_size[1]);
  }
  if (ctx->batch_lut_valid && ctx->batch_lut_key.size() == key.size() &&
      std::memcmp(ctx->batch_lut_key.data(), key.data(), key.size() * sizeof(float)) == 0)
    return SVO_OK;
  std::vector<float> lut(total);
  for (uint32_t c = 0; c < n; c++) {
    const uint32_t W = ins[c].screen_size[0], H = ins[c].screen_size[1];
    float *t = lut.data() + (*offsets)[c];
    camera_lut(ins[c], t, t + W, t + 2 * W, t + 2 * W + H);
  }
  int rc = ensure_scratch(ctx->scratch_batch_lut, lut.size() * sizeof(float));
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(ctx->scratch_batch_lut->ptr, lut.data(), lut.size() * sizeof(float), hipMemcpyHostToDevice,
                         ctx->stream),
          "camera tables upload");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "camera tables sync");
  ctx->batch_lut_key = key;
  ctx->batch_lut_valid = true;
  return SVO_OK;
}

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

int frame_scratch(FrameScratch &sc, std::shared_ptr<DeviceBuffer> &b, size_t bytes, void **out) {
  if (sc.arena) {
    const size_t at = (sc.arena_used + 255) & ~(size_t)255;
    if (at > sc.arena_bytes || bytes > sc.arena_bytes - at)
      return fail(SVO_ERR_CAPACITY, "trace frame: the frame needs more than the " + std::to_string(sc.arena_bytes) +
                                        " bytes of its scratch region (" + std::to_string(at + bytes) + " so far)");
    sc.arena_used = at + bytes;
    *out = sc.arena + at;
    return SVO_OK;
  }
  int rc = ensure_scratch(b, bytes);
  if (rc) return rc;
  *out = b->ptr;
  return SVO_OK;
}

namespace {

int check_size(const FrameBuffers::B &b, size_t need, const char *what) {
  if (b.ptr && b.bytes < need)
    return fail(SVO_ERR_BUFFER_TOO_SMALL, std::string("trace frame: buffer ") + what + " holds " +
                                              std::to_string(b.bytes) + " bytes, needs " + std::to_string(need));
  return SVO_OK;
}

int choose_launch(Context *ctx, uint32_t B, uint32_t n_bases, size_t extra_lds, const svo_frame_options &o,
                  bool *bases_in_lds, uint32_t *stack_levels, LaunchShape *shape) {
  int wg = o.workgroup_size ? (int)o.workgroup_size : 256;
  if (wg != 64 && wg != 128 && wg != 256 && wg != 512 && wg != 1024)
    return fail(SVO_ERR_INVALID_ARGUMENT, "trace frame: workgroup_size must be 64, 128, 256, 512 or 1024");
  uint32_t max_depth = o.max_tree_depth ? o.max_tree_depth : (ctx->tree_max_depth ? ctx->tree_max_depth : 24u);
  if (max_depth > 60) return fail(SVO_ERR_INVALID_ARGUMENT, "trace frame: max_tree_depth must be <= 60");
  uint32_t levels = o.stack_levels ? o.stack_levels : (max_depth >= B + 2 ? max_depth - (B + 2) + 1 : 1u);
  levels = std::max(1u, std::min(levels, 32u));
  const size_t lds_limit = 64 * 1024;
  size_t bases_bytes = ((size_t)n_bases * 2 + 3) & ~(size_t)3;
  *bases_in_lds = bases_bytes <= 32 * 1024;
  if (!*bases_in_lds) bases_bytes = 0;
  bases_bytes += extra_lds;  // camera records
  // the short stack: per level and lane a block address, its Child field (u32 each) and its
  // chunk (u16)
  const auto stack_bytes = [&](uint32_t n) { return ((size_t)n * wg * kStackEntryBytes + 15) & ~(size_t)15; };
  while (levels > 1 && bases_bytes + stack_bytes(levels) > lds_limit) levels--;
  if (bases_bytes + stack_bytes(levels) > lds_limit)
    return fail(SVO_ERR_INVALID_ARGUMENT, "trace frame: workgroup too large for LDS");
  // register budget: 5 waves per SIMD (96 VGPRs) keeps the step loop free of spills
  const int waves = o.waves_per_simd ? (int)o.waves_per_simd : 5;
  if (waves < 4 || waves > 6)
    return fail(SVO_ERR_INVALID_ARGUMENT, "trace frame: waves_per_simd must be 0 (default, 5), 4, 5 or 6");
  // the work-groups a CU holds by registers alone; unless the caller fixed the stack depth,
  // the stack and a ticked-usage table of at least 256 slots keep that many (a pop below the
  // stack walks down from its deepest level, so every depth is correct; a deep stack that
  // halves the CU's waves is not worth it -- a pool whose depth is not known would otherwise
  // get 19 levels and 2 work-groups -- and neither is one that squeezes the table: 9 levels
  // with 64 slots took 1.89 ms for a 4K streamed frame, 8 levels with 512 slots 1.77; the
  // 64-camera batch wants the deeper stack, 6 levels with 256 slots beating 5 with 1024)
  int by_regs = 0, per_cu = 0;
  if (trace_occupancy(wg, waves, bases_bytes + stack_bytes(1), &by_regs) != 0 || by_regs <= 0)
    return fail(SVO_ERR_HIP, "trace frame: occupancy query failed");
  const size_t table_bytes = o.stack_levels ? 0 : (size_t)kMinTickedSlots * 4;
  for (;;) {
    if (trace_occupancy(wg, waves, bases_bytes + stack_bytes(levels) + table_bytes, &per_cu) != 0 || per_cu <= 0)
      return fail(SVO_ERR_HIP, "trace frame: occupancy query failed");
    if (o.stack_levels || per_cu >= by_regs || levels == 1) break;
    levels--;
  }
  if (!o.stack_levels && trace_occupancy(wg, waves, bases_bytes + stack_bytes(levels), &per_cu) != 0)
    return fail(SVO_ERR_HIP, "trace frame: occupancy query failed");
  size_t lds = bases_bytes + stack_bytes(levels);
  // the largest table of ticked usage groups that keeps per_cu work-groups on a CU
  uint32_t slots = 0;
  for (uint32_t t = kTickedSlots; t >= 64u && !slots; t >>= 1) {
    int with = 0;
    if (lds + (size_t)t * 4 <= lds_limit && trace_occupancy(wg, waves, lds + (size_t)t * 4, &with) == 0 &&
        with >= per_cu)
      slots = t;
  }
  lds += (size_t)slots * 4;
  int grid = per_cu * std::max(1, ctx->cu_count);
  grid = std::max(8, (grid / 8) * 8);
  shape->workgroup = wg;
  shape->waves = waves;
  shape->grid = grid;
  shape->lds_bytes = lds;
  shape->ticked_slots = slots;
  *stack_levels = levels;
  return SVO_OK;
}

}  // namespace

int trace_frame_impl(Context *ctx, const svo_trace_input &in, const svo_frame_options &o, const FrameBuffers &fb) {
  return trace_batch_impl(ctx, &in, nullptr, 1, o, fb);
}

// Fills the per-frame fields of the camera records (pose, cone aperture, tick); the layout
// fields (sizes, tile and cone offsets, table offsets) are set by plan_batch.
void fill_camera_poses(const svo_trace_input *ins, const float *apertures, uint32_t n_cams, CamRec *cams) {
  for (uint32_t c = 0; c < n_cams; c++) {
    CamRec &r = cams[c];
    for (int k = 0; k < 9; k++) r.facing[k] = ins[c].facing[k];
    for (int k = 0; k < 3; k++) r.origin[k] = ins[c].origin[k];
    // setupInitialTrace: FoV = FoV[0] / ScreenSize.x; an aperture of k pixels scales it
    const float a = apertures ? apertures[c] : 1.0f;
    r.cone_fov = (a * ins[c].fov[0]) / (float)ins[c].screen_size[0];
    r.dof[0] = ins[c].dof[0];
    r.dof[1] = ins[c].dof[1];
    r.pad[0] = ins[0].tick;
  }
}

int validate_batch(const svo_trace_input *ins, const float *apertures, uint32_t n_cams, size_t *cones) {
  if (!ins || n_cams == 0) return fail(SVO_ERR_INVALID_ARGUMENT, "trace frame: no camera");
  if (n_cams > kMaxBatchCameras)
    return fail(SVO_ERR_INVALID_ARGUMENT, "trace frame: at most " + std::to_string(kMaxBatchCameras) + " cameras");
  uint64_t total = 0;
  for (uint32_t c = 0; c < n_cams; c++) {
    const uint32_t w = ins[c].screen_size[0], h = ins[c].screen_size[1];
    if (w == 0 || h == 0) return fail(SVO_ERR_INVALID_ARGUMENT, "trace frame: screen size must be non-zero");
    if (ins[c].base_depth != ins[0].base_depth || ins[c].tick != ins[0].tick)
      return fail(SVO_ERR_INVALID_ARGUMENT, "trace frame: the cameras of a batch share BaseDepth and Tick");
    if (apertures && !(apertures[c] > 0.0f && apertures[c] <= 64.0f))
      return fail(SVO_ERR_INVALID_ARGUMENT, "trace frame: aperture must be in (0, 64] pixels");
    total += (uint64_t)w * h;
  }
  if (total > (1ull << 31)) return fail(SVO_ERR_INVALID_ARGUMENT, "trace frame: at most 2^31 cones per frame");
  const uint32_t B = ins[0].base_depth;
  if (B < 1 || B > 7) return fail(SVO_ERR_INVALID_ARGUMENT, "trace frame: base_depth must be 1..7");
  *cones = (size_t)total;
  return SVO_OK;
}

// Everything of a frame but its submission: buffers checked, scratch sized, camera tables
// uploaded, launch shape chosen.  Camera records are written to `host_cams` and
// read by the kernel from `dev_cams`.
int plan_batch(Context *ctx, FrameScratch &sc, const svo_trace_input *ins, const float *apertures, uint32_t n_cams,
               const svo_frame_options &o, const FrameBuffers &fb, CamRec *host_cams, CamRec *dev_cams,
               const float *lut_dev, const uint32_t *lut_offsets, BatchPlan *plan) {
  size_t N = 0;
  // planning a direct frame may move the frame scratch: kept requests are valid again only
  // when this frame keeps its own
  if (&sc == &ctx->frame_scratch) ctx->kept.valid = false;
  int rc = validate_batch(ins, apertures, n_cams, &N);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  const svo_trace_input &in = ins[0];
  const uint32_t B = in.base_depth;
  if (!fb.bases.ptr) return fail(SVO_ERR_UNKNOWN_BUFFER, "trace frame: buffer Bases not initialised");
  if (!fb.blocks.ptr) return fail(SVO_ERR_UNKNOWN_BUFFER, "trace frame: buffer Blocks not initialised");
  const uint32_t n_bases = pow_sum(B);
  if ((rc = check_size(fb.bases, (size_t)n_bases * 2, "Bases"))) return rc;
  if ((rc = check_size(fb.request_id, 4, "ChildRequestID"))) return rc;
  if ((rc = check_size(fb.rgba, 16 * N, "OutputImage"))) return rc;
  if ((rc = check_size(fb.hit, 4 * N, "HitIndices"))) return rc;
  if ((rc = check_size(fb.visits, 4 * N, "VisitCounts"))) return rc;
  if ((rc = check_size(fb.hashes, 8 * N, "VisitHashes"))) return rc;
  if ((rc = check_size(fb.status, N, "RayStatus"))) return rc;

  FrameParams &p = plan->p;
  std::memset(&p, 0, sizeof p);
  plan->views = nullptr;
  p.bases = (const uint16_t *)fb.bases.ptr;
  p.n_bases = n_bases;
  p.blocks = (const uint4 *)fb.blocks.ptr;
  p.n_blocks = (uint32_t)std::min<size_t>(fb.blocks.bytes / 16, 0xffffffffu);
  p.usage16 = (uint16_t *)fb.usages.ptr;
  p.n_usages = fb.usages.ptr ? (uint32_t)std::min<size_t>(fb.usages.bytes / 8, 0xffffffffu) : 0u;
  svo_child_request *requests = (svo_child_request *)fb.requests.ptr;
  uint32_t *request_id = (uint32_t *)fb.request_id.ptr;
  // kept requests stay in the context's frame scratch for svo_stream_update; the capacity is
  // what ChildRequests would have stored
  const bool keep = (o.flags & SVO_FRAME_KEEP_REQUESTS) != 0;
  if (keep && (&sc != &ctx->frame_scratch || sc.arena))
    return fail(SVO_ERR_NOT_SUPPORTED,
                "trace frame: SVO_FRAME_KEEP_REQUESTS is for direct frames (svo_trace_frame, svo_trace_batch)");
  if (keep && !request_id)
    return fail(SVO_ERR_UNKNOWN_BUFFER, "trace frame: SVO_FRAME_KEEP_REQUESTS needs buffer ChildRequestID");
  const uint32_t req_capacity =
      requests ? (uint32_t)std::min<size_t>(fb.requests.bytes / sizeof(svo_child_request), in.max_child_request_id)
               : (keep ? in.max_child_request_id : 0u);
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
  const bool queues = (o.flags & SVO_FRAME_QUEUES) != 0;
  if (queues) {
    if (!fb.bg_queue.ptr || !fb.mat_queue.ptr || !fb.bg_count.ptr || !fb.mat_count.ptr)
      return fail(SVO_ERR_UNKNOWN_BUFFER, "trace frame: SVO_FRAME_QUEUES needs Background/MaterialQueue and their IDs");
    if ((rc = check_size(fb.bg_queue, 4 * N, "BackgroundQueue"))) return rc;
    if ((rc = check_size(fb.mat_queue, 4 * N, "MaterialQueue"))) return rc;
    if ((rc = check_size(fb.bg_count, 4, "BackgroundQueueID"))) return rc;
    if ((rc = check_size(fb.mat_count, 4, "MaterialQueueID"))) return rc;
  }
  plan->compaction = requests || request_id || queues;
  void *sp = nullptr;
  if (plan->compaction && !p.status) {
    if ((rc = frame_scratch(sc, sc.status, N, &sp))) return rc;
    p.status = (uint8_t *)sp;
  }
  if (requests || request_id) {
    if ((rc = frame_scratch(sc, sc.requests, N * sizeof(svo_child_request), &sp))) return rc;
    p.req_sparse = (svo_child_request *)sp;
  }
  uint32_t *req_cones = nullptr;
  const uint32_t kept_capacity = (uint32_t)std::min<size_t>(req_capacity, N);
  if (keep) {
    if ((rc = frame_scratch(sc, sc.req_keys, N * sizeof(uint2), &sp))) return rc;
    p.req_keys = (uint2 *)sp;
    if ((rc = frame_scratch(sc, sc.req_cones, (size_t)std::max(kept_capacity, 1u) * 4, &sp))) return rc;
    req_cones = (uint32_t *)sp;
  }
  if ((rc = frame_scratch(sc, sc.small, kSmallScratchBytes, &sp))) return rc;
  char *small = (char *)sp;
  p.counters = (uint32_t *)(small + kCountersOffset);
  p.totals = (unsigned long long *)(small + kTotalsOffset);

  uint32_t tiles = 0;
  {
    std::vector<uint32_t> offsets;
    if (lut_dev) {
      p.lut = lut_dev;
    } else {
      if ((rc = update_batch_tables(ctx, ins, n_cams, &offsets))) return rc;
      p.lut = (const float *)ctx->scratch_batch_lut->ptr;
      lut_offsets = offsets.data();
    }
    uint64_t ray0 = 0;
    for (uint32_t c = 0; c < n_cams; c++) {
      CamRec &r = host_cams[c];
      std::memset(&r, 0, sizeof r);
      r.W = ins[c].screen_size[0];
      r.H = ins[c].screen_size[1];
      r.tiles_x = (r.W + 7) / 8;
      r.tile_begin = tiles;
      r.ray_begin = (uint32_t)ray0;
      r.lut_offset = lut_offsets[c];
      tiles += r.tiles_x * ((r.H + 7) / 8);
      ray0 += (uint64_t)r.W * r.H;
    }
    fill_camera_poses(ins, apertures, n_cams, host_cams);
    p.cams = dev_cams;
    p.n_cams = n_cams;
  }
  p.B = B;
  // eight bands of tiles, one per XCD (work-group b runs on XCD b mod 8); without the split
  // every wave claims from one shared range
  const bool split = (o.flags & SVO_FRAME_NO_XCD_SPLIT) == 0;
  for (int r = 0; r <= 8; r++) p.range_begin[r] = split ? (uint32_t)(((uint64_t)tiles * r) / 8) : (r ? tiles : 0u);
  p.refill = o.refill_lanes ? std::min<uint32_t>(o.refill_lanes, 64u) : 64u;
  // defaults: a wave takes one 8x8 tile per claim and refills only when all its lanes are
  // idle -- spatial coherence of the 64 cones matters more than lane utilisation
  p.chunk = o.work_chunk ? ((std::min(o.work_chunk, 1u << 20) + 63u) / 64u) * 64u : 64u;
  if (!(o.flags & SVO_FRAME_RASTER_ORDER) && tiles > 0 && !sc.arena) {
    const DeviceBuffer *old_cost = sc.tile_cost.get(), *old_views = sc.tile_views.get();
    const uint32_t view_words = 8u * n_cams * kViewWords;
    if ((rc = ensure_scratch(sc.tile_cost, (size_t)tiles * 4))) return rc;
    if ((rc = ensure_scratch(sc.tile_order, (size_t)tiles * 4))) return rc;
    if ((rc = ensure_scratch(sc.tile_views, (size_t)view_words * 4))) return rc;
    // costs measured on another tiling (or a new buffer) start from zero, and views recorded
    // for another camera count (or none) from unrecorded: the first frame runs in raster order
    if (sc.tile_cost.get() != old_cost || std::memcmp(sc.sched_ranges, p.range_begin, sizeof p.range_begin) != 0) {
      if (launch_zero_u32((uint32_t *)sc.tile_cost->ptr, tiles, ctx->stream) != 0)
        return hip_fail(hipGetLastError(), "trace frame: schedule reset");
      std::memcpy(sc.sched_ranges, p.range_begin, sizeof p.range_begin);
    }
    if (sc.tile_views.get() != old_views || sc.sched_cams != n_cams) {
      if (launch_zero_u32((uint32_t *)sc.tile_views->ptr, view_words, ctx->stream) != 0)
        return hip_fail(hipGetLastError(), "trace frame: schedule reset");
      sc.sched_cams = n_cams;
    }
    p.tile_cost = (uint32_t *)sc.tile_cost->ptr;
    p.order = (const uint32_t *)sc.tile_order->ptr;
    plan->views = (float *)sc.tile_views->ptr;
  }

  bool bases_in_lds = false;
  uint32_t levels = 1;
  if ((rc = choose_launch(ctx, B, n_bases, (size_t)n_cams * sizeof(CamRec), o, &bases_in_lds, &levels,
                          &plan->shape)))
    return rc;
  p.bases_in_lds = bases_in_lds ? 1u : 0u;
  p.stack_levels = levels;
  p.ticked_slots = plan->shape.ticked_slots;

  CompactParams &c = plan->compact;
  std::memset(&c, 0, sizeof c);
  if (plan->compaction) {
    const uint32_t chunks = compaction_chunks((uint32_t)N);
    if ((rc = frame_scratch(sc, sc.counts, (size_t)std::max(1u, chunks) * 16, &sp))) return rc;
    c.status = p.status;
    c.n = (uint32_t)N;
    c.counts = (uint32_t *)sp;
    if (queues) {
      c.bg_queue = (uint32_t *)fb.bg_queue.ptr;
      c.mat_queue = (uint32_t *)fb.mat_queue.ptr;
      c.bg_count = (uint32_t *)fb.bg_count.ptr;
      c.mat_count = (uint32_t *)fb.mat_count.ptr;
    }
    c.req_count = request_id;
    c.req_sparse = p.req_sparse;
    c.requests = keep ? nullptr : requests;
    c.req_cones = req_cones;
    c.req_capacity = keep ? kept_capacity : req_capacity;
    c.totals = p.totals;
  }
  plan->cones = N;
  plan->n_cams = n_cams;
  // a direct frame replaces the frame scratch the kept requests of an earlier one live in
  if (&sc == &ctx->frame_scratch) {
    ctx->kept.valid = keep;
    ctx->kept.cones = req_cones;
    ctx->kept.keys = p.req_keys;
    ctx->kept.records = p.req_sparse;
    ctx->kept.capacity = kept_capacity;
  }
  return SVO_OK;
}

int launch_begin(const BatchPlan &plan, const CamRec *host_cams, uint32_t n_inline, hipStream_t stream) {
  FrameBegin f;
  std::memset(&f, 0, sizeof f);
  f.counters = plan.p.counters;
  f.totals = plan.p.totals;
  f.cams = const_cast<CamRec *>(plan.p.cams);
  f.n_cams = n_inline;
  for (uint32_t c = 0; c < n_inline; c++) f.cam[c] = host_cams[c];
  if (plan.p.order) {
    f.bands = 8;
    f.n_views = plan.n_cams;
    f.views = plan.views;
    f.cost = plan.p.tile_cost;
    f.order = const_cast<uint32_t *>(plan.p.order);
    std::memcpy(f.range_begin, plan.p.range_begin, sizeof f.range_begin);
  }
  return launch_frame_begin(f, stream);
}

// The stream work of one planned frame: camera records, counter reset and tile order (one
// launch), trace kernel, compaction.  No host synchronisation and no memset nodes
// (capturable).  eager: submitted now, not captured; then up to kInlineCams camera records
// ride in the first launch's argument (the staging records are not read after the call
// returns), more are copied and ctx->cam_copied follows the copy, and the trace kernel is
// timed when kernel timing is on.
int submit_batch(Context *ctx, const BatchPlan &plan, const CamRec *host_cams, bool eager) {
  const uint32_t n_inline = eager && plan.n_cams <= kInlineCams ? plan.n_cams : 0u;
  if (plan.n_cams && !n_inline) {
    HIP_TRY(hipMemcpyAsync((void *)plan.p.cams, host_cams, plan.n_cams * sizeof(CamRec), hipMemcpyHostToDevice,
                           ctx->stream),
            "camera upload");
    if (eager) HIP_TRY(hipEventRecord(ctx->cam_copied, ctx->stream), "hipEventRecord");
  }
  if (launch_begin(plan, host_cams, n_inline, ctx->stream) != 0)
    return hip_fail(hipGetLastError(), "trace frame: frame begin launch");
  std::pair<hipEvent_t, hipEvent_t> *ev = nullptr;
  if (eager && ctx->kernel_timing) {
    if (ctx->kernel_events_used == ctx->kernel_events.size()) {
      std::pair<hipEvent_t, hipEvent_t> pr;
      HIP_TRY(hipEventCreate(&pr.first), "hipEventCreate");
      HIP_TRY(hipEventCreate(&pr.second), "hipEventCreate");
      ctx->kernel_events.push_back(pr);
    }
    ev = &ctx->kernel_events[ctx->kernel_events_used++];
    HIP_TRY(hipEventRecord(ev->first, ctx->stream), "hipEventRecord");
  }
  if (launch_trace_frame(plan.p, plan.shape, ctx->stream) != 0)
    return hip_fail(hipGetLastError(), "trace frame: trace launch");
  if (ev) HIP_TRY(hipEventRecord(ev->second, ctx->stream), "hipEventRecord");
  if (plan.compaction && launch_compaction(plan.compact, ctx->stream) != 0)
    return hip_fail(hipGetLastError(), "trace frame: compaction launch");
  ctx->last_totals = plan.p.totals;
  return SVO_OK;
}

int trace_batch_impl(Context *ctx, const svo_trace_input *ins, const float *apertures, uint32_t n_cams,
                     const svo_frame_options &o, const FrameBuffers &fb) {
  if (!ctx->cam_host) {
    HIP_TRY(hipHostMalloc((void **)&ctx->cam_host, kMaxBatchCameras * sizeof(CamRec), hipHostMallocDefault),
            "hipHostMalloc");
    HIP_TRY(hipEventCreateWithFlags(&ctx->cam_copied, hipEventDisableTiming), "hipEventCreate");
  } else {
    // the previous frame's camera upload must have read the staging records
    HIP_TRY(hipEventSynchronize(ctx->cam_copied), "camera table reuse");
  }
  int rc = ensure_scratch(ctx->frame_scratch.cams, kMaxBatchCameras * sizeof(CamRec));
  if (rc) return rc;
  BatchPlan plan;
  if ((rc = plan_batch(ctx, ctx->frame_scratch, ins, apertures, n_cams, o, fb, ctx->cam_host,
                       (CamRec *)ctx->frame_scratch.cams->ptr, nullptr, nullptr, &plan)))
    return rc;
  if ((rc = submit_batch(ctx, plan, ctx->cam_host, true))) return rc;
  ctx->last_rays = plan.cones;
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

FrameBuffers svo::frame_buffers(Context *ctx) {
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

extern "C" int svo_trace_frame(svo_context *ctx, const svo_trace_input *input, const svo_frame_options *options) {
  if (!ctx || !input) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_trace_frame: null argument");
  svo_frame_options o;
  std::memset(&o, 0, sizeof o);
  if (options) o = *options;
  const FrameBuffers fb = frame_buffers(ctx);
  return trace_frame_impl(ctx, *input, o, fb);
}

extern "C" int svo_trace_batch(svo_context *ctx, const svo_trace_input *inputs, const float *apertures,
                               uint32_t n_cameras, const svo_frame_options *options) {
  if (!ctx || !inputs) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_trace_batch: null argument");
  svo_frame_options o;
  std::memset(&o, 0, sizeof o);
  if (options) o = *options;
  const FrameBuffers fb = frame_buffers(ctx);
  return trace_batch_impl(ctx, inputs, apertures, n_cameras, o, fb);
}
