// svo_enqueue dispatch: resolves the arguments bound with SetArg (by the parameter
// names of ComputeManager.cs:313-344) and launches the HIP kernels of the
// kernel-by-kernel path, appending to queues in work-item order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>

#include "svo_internal.hpp"
#include "svo_runtime.hpp"
#include "svo_wavefront.hpp"
#include "svo_material_kernels.hpp"
#include "svo_wavefront_kernels.hpp"

namespace svo {

namespace {

const char *kernel_label(int k) {
  static const char *names[] = {"prune", "graft", "trace", "test", "init", "runBaseTrace", "runBlockTrace",
                                "evaluateBackground", "evaluateMaterial", "resolveRemainders", "drawTrace"};
  return names[k];
}

struct Resolver {
  Context *ctx;
  int kernel;
  int rc = SVO_OK;

  const Arg *find(const char *param) {
    auto k = ctx->args.find(kernel);
    if (k == ctx->args.end()) return nullptr;
    auto a = k->second.find(param);
    return a == k->second.end() ? nullptr : &a->second;
  }
  // a bound buffer; `required` makes a missing binding an error
  FrameBuffers::B buf(const char *param, bool required = true) {
    FrameBuffers::B b;
    if (rc) return b;
    const Arg *a = find(param);
    if (!a || !a->buffer) {
      if (required)
        rc = fail(SVO_ERR_MISSING_ARGUMENT, std::string("EnqueueNDRangeKernel: ") + kernel_label(kernel) + " argument " +
                                                param + " is not set");
      return b;
    }
    b.ptr = a->buffer->ptr;
    b.bytes = a->buffer->bytes;
    return b;
  }
  bool input(svo_trace_input *out) {
    if (rc) return false;
    const Arg *a = find("input");
    if (!a || a->bytes.size() != sizeof(svo_trace_input)) {
      rc = fail(SVO_ERR_MISSING_ARGUMENT, std::string("EnqueueNDRangeKernel: ") + kernel_label(kernel) +
                                              " argument input (TraceInput) is not set");
      return false;
    }
    std::memcpy(out, a->bytes.data(), sizeof *out);
    return true;
  }
};

uint32_t count_of(const FrameBuffers::B &b, size_t elem) { return (uint32_t)std::min<size_t>(b.bytes / elem, 0xffffffffu); }

int scratch_for(Context *ctx, uint32_t n, bool staging) {
  int rc;
  if ((rc = ensure_scratch(ctx->wave_codes, std::max<size_t>(n, 1)))) return rc;
  if ((rc = ensure_scratch(ctx->wave_values, std::max<size_t>(n, 1) * 4))) return rc;
  if ((rc = ensure_scratch(ctx->wave_counts, (size_t)std::max(1u, append_chunks(n)) * kAppendTargets * 4))) return rc;
  if (staging && (rc = ensure_scratch(ctx->wave_staging, std::max<size_t>(n, 1) * sizeof(svo_child_request)))) return rc;
  return SVO_OK;
}

AppendParams append_base(Context *ctx, uint32_t n) {
  AppendParams a;
  std::memset(&a, 0, sizeof a);
  a.codes = (const uint8_t *)ctx->wave_codes->ptr;
  a.values = (const uint32_t *)ctx->wave_values->ptr;
  a.n = n;
  a.counts = (uint32_t *)ctx->wave_counts->ptr;
  return a;
}

void set_queue(AppendParams &a, int q, const FrameBuffers::B &queue, const FrameBuffers::B &counter) {
  a.queue[q] = (uint32_t *)queue.ptr;
  a.capacity[q] = count_of(queue, 4);
  a.counter[q] = (uint32_t *)counter.ptr;
}

WaveState wave_state(Resolver &r, bool need_parents, bool need_weightings) {
  WaveState s;
  std::memset(&s, 0, sizeof s);
  auto o = r.buf("Origins"), d = r.buf("Directions"), f = r.buf("FoVs"), l = r.buf("Locations"), dp = r.buf("Depths");
  auto w = r.buf("Weightings", need_weightings);
  auto pt = r.buf("ParentTraces", need_parents);
  s.origins = (float4 *)o.ptr;
  s.directions = (float4 *)d.ptr;
  s.fovs = (float *)f.ptr;
  s.locations = (uint64_t *)l.ptr;
  s.depths = (uint8_t *)dp.ptr;
  s.weightings = (uint8_t *)w.ptr;
  s.parents = (uint32_t *)pt.ptr;
  uint32_t n = std::min({count_of(o, 16), count_of(d, 16), count_of(f, 4), count_of(l, 32), count_of(dp, 1)});
  if (w.ptr) n = std::min(n, count_of(w, 1));
  if (pt.ptr) n = std::min(n, count_of(pt, 4));
  s.n_rays = n;
  return s;
}

int check_counter(Resolver &r, const FrameBuffers::B &c, const char *name) {
  if (r.rc) return r.rc;
  if (c.bytes < 4) return fail(SVO_ERR_BUFFER_TOO_SMALL, std::string("EnqueueNDRangeKernel: ") + name + " is smaller than 4 bytes");
  return SVO_OK;
}

int enqueue_init(Context *ctx, Resolver &r, uint32_t gsx, uint32_t gsy) {
  svo_trace_input in;
  if (!r.input(&in)) return r.rc;
  InitParams p;
  std::memset(&p, 0, sizeof p);
  p.s = wave_state(r, true, true);
  auto bq = r.buf("BaseTraceQueue"), bqid = r.buf("BaseTraceQueueID"), gq = r.buf("BackgroundQueue"),
       gqid = r.buf("BackgroundQueueID");
  auto cr = r.buf("ColourRs", false), cg = r.buf("ColourGs", false), cb = r.buf("ColourBs", false);
  auto fr = r.buf("FinalColourRs"), fg = r.buf("FinalColourGs"), fbb = r.buf("FinalColourBs"), fw = r.buf("FinalWeightings");
  if (r.rc) return r.rc;
  int rc;
  if ((rc = check_counter(r, bqid, "BaseTraceQueueID")) || (rc = check_counter(r, gqid, "BackgroundQueueID"))) return rc;
  if (in.screen_size[0] == 0 || in.screen_size[1] == 0) return fail(SVO_ERR_INVALID_ARGUMENT, "init: screen size is zero");
  uint32_t n = p.s.n_rays;
  n = std::min({n, count_of(fr, 4), count_of(fg, 4), count_of(fbb, 4), count_of(fw, 4)});
  if (cr.ptr) n = std::min(n, count_of(cr, 1));
  if (cg.ptr) n = std::min(n, count_of(cg, 1));
  if (cb.ptr) n = std::min(n, count_of(cb, 1));
  p.s.n_rays = n;
  if ((rc = update_camera_table(ctx, in))) return rc;
  const uint64_t items = (uint64_t)gsx * gsy;
  if ((rc = scratch_for(ctx, (uint32_t)items, false))) return rc;
  p.lut = (const float *)ctx->scratch_lut->ptr;
  for (int k = 0; k < 9; k++) p.facing[k] = in.facing[k];
  for (int k = 0; k < 3; k++) p.origin[k] = in.origin[k];
  p.cone_fov = in.fov[0] / (float)in.screen_size[0];
  p.W = in.screen_size[0];
  p.H = in.screen_size[1];
  p.gsx = gsx;
  p.gsy = gsy;
  p.colour_r = (uint8_t *)cr.ptr;
  p.colour_g = (uint8_t *)cg.ptr;
  p.colour_b = (uint8_t *)cb.ptr;
  p.final_r = (uint32_t *)fr.ptr;
  p.final_g = (uint32_t *)fg.ptr;
  p.final_b = (uint32_t *)fbb.ptr;
  p.final_w = (uint32_t *)fw.ptr;
  p.codes = (uint8_t *)ctx->wave_codes->ptr;
  p.values = (uint32_t *)ctx->wave_values->ptr;
  if (launch_init(p, ctx->stream)) return hip_fail(hipGetLastError(), "init launch");
  AppendParams a = append_base(ctx, (uint32_t)items);
  set_queue(a, kQueueBase, bq, bqid);
  set_queue(a, kQueueBackground, gq, gqid);
  if (launch_append(a, ctx->stream)) return hip_fail(hipGetLastError(), "init append");
  return SVO_OK;
}

int enqueue_step(Context *ctx, Resolver &r, uint32_t n, bool block) {
  svo_trace_input in;
  if (!r.input(&in)) return r.rc;
  if (in.base_depth < 1 || in.base_depth > 7) return fail(SVO_ERR_INVALID_ARGUMENT, "trace step: base_depth must be 1..7");
  StepParams p;
  std::memset(&p, 0, sizeof p);
  p.s = wave_state(r, false, true);
  auto inq = r.buf(block ? "BlockTraces" : "BaseTraces");
  auto bq = r.buf("BaseTraceQueue"), bqid = r.buf("BaseTraceQueueID"), kq = r.buf("BlockTraceQueue"),
       kqid = r.buf("BlockTraceQueueID"), gq = r.buf("BackgroundQueue"), gqid = r.buf("BackgroundQueueID");
  FrameBuffers::B bases, blocks, usages, reqs, reqid, mq, mqid;
  if (block) {
    blocks = r.buf("Blocks");
    usages = r.buf("Usages", false);
    reqs = r.buf("ChildRequests", false);
    reqid = r.buf("ChildRequestID", false);
    mq = r.buf("MaterialQueue");
    mqid = r.buf("MaterialQueueID");
  } else {
    bases = r.buf("Bases");
  }
  if (r.rc) return r.rc;
  int rc;
  if ((rc = check_counter(r, bqid, "BaseTraceQueueID")) || (rc = check_counter(r, kqid, "BlockTraceQueueID")) ||
      (rc = check_counter(r, gqid, "BackgroundQueueID")))
    return rc;
  if (block && (rc = check_counter(r, mqid, "MaterialQueueID"))) return rc;
  if (reqid.ptr && (rc = check_counter(r, reqid, "ChildRequestID"))) return rc;
  if (inq.bytes < (size_t)n * 4) return fail(SVO_ERR_BUFFER_TOO_SMALL, "trace step: work size exceeds the input queue");
  if (!block && bases.bytes < (size_t)pow_sum(in.base_depth) * 2)
    return fail(SVO_ERR_BUFFER_TOO_SMALL, "runBaseTrace: Bases is smaller than PowSum(BaseDepth)");
  if ((rc = scratch_for(ctx, n, block))) return rc;
  p.in_queue = (const uint32_t *)inq.ptr;
  p.n = n;
  p.bases = (const uint16_t *)bases.ptr;
  p.B = in.base_depth;
  p.blocks = (const uint4 *)blocks.ptr;
  p.n_blocks = count_of(blocks, 16);
  p.usage16 = (uint16_t *)usages.ptr;
  p.n_usages = count_of(usages, 8);
  p.tick = in.tick;
  p.dof0 = in.dof[0];
  p.dof1 = in.dof[1];
  p.codes = (uint8_t *)ctx->wave_codes->ptr;
  p.values = (uint32_t *)ctx->wave_values->ptr;
  p.req_staging = block ? (svo_child_request *)ctx->wave_staging->ptr : nullptr;
  if ((block ? launch_block_step(p, ctx->stream) : launch_base_step(p, ctx->stream)))
    return hip_fail(hipGetLastError(), "trace step launch");
  AppendParams a = append_base(ctx, n);
  set_queue(a, kQueueBase, bq, bqid);
  set_queue(a, kQueueBlock, kq, kqid);
  set_queue(a, kQueueBackground, gq, gqid);
  if (block) {
    set_queue(a, kQueueMaterial, mq, mqid);
    a.req_staging = p.req_staging;
    a.requests = (svo_child_request *)reqs.ptr;
    a.capacity[4] = reqs.ptr ? std::min<uint32_t>(count_of(reqs, sizeof(svo_child_request)), in.max_child_request_id) : 0u;
    a.counter[4] = (uint32_t *)reqid.ptr;
  }
  if (launch_append(a, ctx->stream)) return hip_fail(hipGetLastError(), "trace step append");
  return SVO_OK;
}

int enqueue_shade(Context *ctx, Resolver &r, uint32_t n, bool background) {
  svo_trace_input in;
  if (!r.input(&in)) return r.rc;
  ShadeParams p;
  std::memset(&p, 0, sizeof p);
  auto q = r.buf(background ? "BackgroundQueue" : "MaterialQueue");
  auto pt = r.buf("ParentTraces"), w = r.buf("Weightings");
  auto fr = r.buf("FinalColourRs"), fg = r.buf("FinalColourGs"), fbb = r.buf("FinalColourBs"), fw = r.buf("FinalWeightings");
  FrameBuffers::B dirs, cr, cg, cb, blocks, locs, depths;
  if (background) {
    dirs = r.buf("Directions");
    cr = r.buf("ColourRs");
    cg = r.buf("ColourGs");
    cb = r.buf("ColourBs");
  } else {
    blocks = r.buf("Blocks");
    locs = r.buf("Locations");
    depths = r.buf("Depths");
  }
  if (r.rc) return r.rc;
  if (q.bytes < (size_t)n * 4) return fail(SVO_ERR_BUFFER_TOO_SMALL, "shade: work size exceeds the queue");
  p.queue = (const uint32_t *)q.ptr;
  p.n = n;
  uint32_t nr = std::min(count_of(pt, 4), count_of(w, 1));
  if (background) nr = std::min({nr, count_of(dirs, 16), count_of(cr, 1), count_of(cg, 1), count_of(cb, 1)});
  else nr = std::min({nr, count_of(locs, 32), count_of(depths, 1)});
  p.n_rays = nr;
  p.directions = (const float4 *)dirs.ptr;
  p.parents = (const uint32_t *)pt.ptr;
  p.weightings = (const uint8_t *)w.ptr;
  p.colour_r = (const uint8_t *)cr.ptr;
  p.colour_g = (const uint8_t *)cg.ptr;
  p.colour_b = (const uint8_t *)cb.ptr;
  p.final_r = (uint32_t *)fr.ptr;
  p.final_g = (uint32_t *)fg.ptr;
  p.final_b = (uint32_t *)fbb.ptr;
  p.final_w = (uint32_t *)fw.ptr;
  p.n_pixels = std::min({count_of(fr, 4), count_of(fg, 4), count_of(fbb, 4), count_of(fw, 4)});
  p.ambient = in.ambient_light_level;
  p.blocks = (const uint4 *)blocks.ptr;
  p.n_blocks = count_of(blocks, 16);
  p.locations = (const uint64_t *)locs.ptr;
  p.depths = (const uint8_t *)depths.ptr;
  p.B = in.base_depth;
  if ((background ? launch_background(p, ctx->stream) : launch_remainders(p, ctx->stream)))
    return hip_fail(hipGetLastError(), "shade launch");
  return SVO_OK;
}

int enqueue_draw(Context *ctx, Resolver &r, uint32_t gsx, uint32_t gsy) {
  svo_trace_input in;
  if (!r.input(&in)) return r.rc;
  auto fr = r.buf("FinalColourRs"), fg = r.buf("FinalColourGs"), fbb = r.buf("FinalColourBs"), fw = r.buf("FinalWeightings");
  auto out = r.buf("outputImage");
  if (r.rc) return r.rc;
  DrawParams p;
  std::memset(&p, 0, sizeof p);
  p.final_r = (const uint32_t *)fr.ptr;
  p.final_g = (const uint32_t *)fg.ptr;
  p.final_b = (const uint32_t *)fbb.ptr;
  p.final_w = (const uint32_t *)fw.ptr;
  p.n_pixels = std::min({count_of(fr, 4), count_of(fg, 4), count_of(fbb, 4), count_of(fw, 4)});
  p.out = (float4 *)out.ptr;
  p.n_out = count_of(out, 16);
  p.W = in.screen_size[0];
  p.gsx = gsx;
  p.gsy = gsy;
  if (launch_draw(p, ctx->stream)) return hip_fail(hipGetLastError(), "drawTrace launch");
  return SVO_OK;
}

// evaluateMaterial, wavefront.cl:1568-1739 (MainWindow.EvaulateMaterial binds the Root*
// buffers after swapping them with the current trace arrays, MainWindow.cs:533-571).
int enqueue_material(Context *ctx, Resolver &r, uint32_t n) {
  svo_trace_input in;
  if (!r.input(&in)) return r.rc;
  auto blocks = r.buf("Blocks"), usages = r.buf("Usages", false), queue = r.buf("MaterialQueue");
  auto rd = r.buf("RootDirections"), rl = r.buf("RootLocations"), rdp = r.buf("RootDepths"),
       rw = r.buf("RootWeightings"), rp = r.buf("RootParentTraces");
  auto o = r.buf("Origins"), d = r.buf("Directions"), f = r.buf("FoVs"), l = r.buf("Locations"), dp = r.buf("Depths"),
       w = r.buf("Weightings"), cr = r.buf("ColourRs"), cg = r.buf("ColourGs"), cb = r.buf("ColourBs"),
       lum = r.buf("Luminosities"), pt = r.buf("ParentTraces"), fw = r.buf("FinalWeightings");
  auto acc = r.buf("AccumulatorID"), bgq = r.buf("BackgroundQueue"), bgc = r.buf("BackgroundQueueID");
  if (r.rc) return r.rc;
  int rc;
  if ((rc = check_counter(r, acc, "AccumulatorID")) || (rc = check_counter(r, bgc, "BackgroundQueueID"))) return rc;
  if (queue.bytes < (size_t)n * 4) return fail(SVO_ERR_BUFFER_TOO_SMALL, "evaluateMaterial: work size exceeds MaterialQueue");
  // the reference kernel returns at once for these parameters
  if (in.fov_constant <= 0.0f || in.fov_multiplier >= 1.0f || in.fov_multiplier <= -1.0f) return SVO_OK;
  MaterialParams p;
  std::memset(&p, 0, sizeof p);
  p.queue = (const uint32_t *)queue.ptr;
  p.n = n;
  p.root_directions = (const float4 *)rd.ptr;
  p.root_locations = (const uint64_t *)rl.ptr;
  p.root_depths = (const uint8_t *)rdp.ptr;
  p.root_weightings = (const uint8_t *)rw.ptr;
  p.root_parents = (const uint32_t *)rp.ptr;
  p.n_root = std::min({count_of(rd, 16), count_of(rl, 32), count_of(rdp, 1), count_of(rw, 1), count_of(rp, 4)});
  p.blocks = (const uint4 *)blocks.ptr;
  p.n_blocks = count_of(blocks, 16);
  p.usage16 = (uint16_t *)usages.ptr;
  p.n_usages = usages.ptr ? count_of(usages, 8) : 0;
  p.B = in.base_depth;
  p.tick = in.tick;
  p.fov_multiplier = in.fov_multiplier;
  p.fov_constant = in.fov_constant;
  p.weighting_multiplier = in.weighting_multiplier;
  p.weighting_constant = in.weighting_constant;
  p.acc_counter = (uint32_t *)acc.ptr;
  p.origins = (float4 *)o.ptr;
  p.directions = (float4 *)d.ptr;
  p.fovs = (float *)f.ptr;
  p.locations = (uint64_t *)l.ptr;
  p.depths = (uint8_t *)dp.ptr;
  p.weightings = (uint8_t *)w.ptr;
  p.colour_r = (uint8_t *)cr.ptr;
  p.colour_g = (uint8_t *)cg.ptr;
  p.colour_b = (uint8_t *)cb.ptr;
  p.luminosities = (uint8_t *)lum.ptr;
  p.parents = (uint32_t *)pt.ptr;
  p.n_out = std::min({count_of(o, 16), count_of(d, 16), count_of(f, 4), count_of(l, 32), count_of(dp, 1),
                      count_of(w, 1), count_of(cr, 1), count_of(cg, 1), count_of(cb, 1), count_of(lum, 1),
                      count_of(pt, 4)});
  p.final_w = (uint32_t *)fw.ptr;
  p.n_final = count_of(fw, 4);
  hipStream_t s = ctx->stream;
  auto &sc = ctx->material_scratch;
  const uint32_t chunks = material_chunks(n);
  if ((rc = ensure_scratch(sc[0], (size_t)n * 4)) || (rc = ensure_scratch(sc[1], (size_t)n * 4)) ||
      (rc = ensure_scratch(sc[2], ((size_t)chunks + 1) * 4)))
    return rc;
  uint32_t *counts = (uint32_t *)sc[0]->ptr, *first = (uint32_t *)sc[1]->ptr, *sums = (uint32_t *)sc[2]->ptr;
  if (launch_material_count(p, counts, s) || launch_material_scan(counts, n, first, sums, s))
    return hip_fail(hipGetLastError(), "evaluateMaterial: count");
  uint32_t total = 0;
  hipError_t e = hipMemcpyAsync(&total, sums + chunks, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "evaluateMaterial: total");
  const uint32_t fchunks = material_chunks(total);
  if ((rc = ensure_scratch(sc[3], (size_t)std::max(total, 1u) * 4)) || (rc = ensure_scratch(sc[4], ((size_t)fchunks + 1) * 4)))
    return rc;
  p.bg_flags = (uint32_t *)sc[3]->ptr;
  p.flag_capacity = total;
  uint32_t *fsums = (uint32_t *)sc[4]->ptr;
  if (launch_material_write(p, first, s) || launch_material_scan(p.bg_flags, total, nullptr, fsums, s) ||
      launch_material_append(p.bg_flags, total, fsums, p.acc_counter, (const uint32_t *)bgc.ptr, (uint32_t *)bgq.ptr,
                             count_of(bgq, 4), s) ||
      launch_material_counters(p.acc_counter, (uint32_t *)bgc.ptr, sums + chunks, fsums + fchunks, s))
    return hip_fail(hipGetLastError(), "evaluateMaterial launch");
  return SVO_OK;
}

// KernelName.Trace: the single-kernel form (kernel.cl:1531-1558) maps onto the fused frame.
int enqueue_trace(Context *ctx, Resolver &r, uint32_t gsx, uint32_t gsy) {
  svo_trace_input in;
  if (!r.input(&in)) return r.rc;
  if (gsx != in.screen_size[0] || gsy != in.screen_size[1])
    return fail(SVO_ERR_INVALID_ARGUMENT, "trace: work size must equal TraceInput.ScreenSize");
  FrameBuffers fb;
  fb.bases = r.buf("bases");
  fb.blocks = r.buf("blocks");
  fb.usages = r.buf("usage", false);
  fb.request_id = r.buf("childRequestId", false);
  fb.requests = r.buf("childRequests", false);
  fb.rgba = r.buf("outputImage");
  if (r.rc) return r.rc;
  svo_frame_options o;
  std::memset(&o, 0, sizeof o);
  return trace_frame_impl(ctx, in, o, fb);
}

}  // namespace

int enqueue_wavefront(Context *ctx, int kernel, uint32_t dims, const size_t *g) {
  const uint32_t gsx = (uint32_t)g[0], gsy = dims == 2 ? (uint32_t)g[1] : 1u;
  Resolver r{ctx, kernel};
  int rc = hipSetDevice(ctx->device) == hipSuccess ? SVO_OK : fail(SVO_ERR_HIP, "hipSetDevice");
  if (rc) return rc;
  switch (kernel) {
    case SVO_KERNEL_INIT:
      return enqueue_init(ctx, r, gsx, gsy);
    case SVO_KERNEL_RUN_BASE_TRACE:
      return enqueue_step(ctx, r, gsx * gsy, false);
    case SVO_KERNEL_RUN_BLOCK_TRACE:
      return enqueue_step(ctx, r, gsx * gsy, true);
    case SVO_KERNEL_EVALUATE_BACKGROUND:
      return enqueue_shade(ctx, r, gsx * gsy, true);
    case SVO_KERNEL_RESOLVE_REMAINDERS:
      return enqueue_shade(ctx, r, gsx * gsy, false);
    case SVO_KERNEL_DRAW_TRACE:
      return enqueue_draw(ctx, r, gsx, gsy);
    case SVO_KERNEL_EVALUATE_MATERIAL:
      return enqueue_material(ctx, r, gsx * gsy);
    case SVO_KERNEL_TRACE:
      return enqueue_trace(ctx, r, gsx, gsy);
    case SVO_KERNEL_PRUNE:
    case SVO_KERNEL_GRAFT:
      return enqueue_tree(ctx, kernel, gsx * gsy);
    case SVO_KERNEL_TEST:
      // the reference's argument table names a kernel "test" that its programs do not define
      return fail(SVO_ERR_UNKNOWN_KERNEL, "CreateKernel:Test: no kernel named test");
    default:
      return fail(SVO_ERR_NOT_SUPPORTED, std::string("svo_enqueue: ") + kernel_label(kernel) + " is not implemented");
  }
}

}  // namespace svo
