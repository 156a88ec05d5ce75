// Captured frame submission: the stream work of a batched frame (camera upload, counter
// reset, trace kernel, compaction) recorded once into a HIP graph and replayed per frame
// with new camera poses and tick.  The graph owns its scratch, camera tables and pinned
// camera staging, so direct frames on the same context never move memory a replay uses.
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "svo_internal.hpp"
#include "svo_runtime.hpp"

namespace svo {

namespace {

#define HIP_TRY(expr, what)                          \
  do {                                               \
    hipError_t e_ = (expr);                          \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

FrameBuffers::B named(Context *ctx, int name) {
  FrameBuffers::B b;
  auto p = ctx->get(name);
  if (p) {
    b.ptr = p->ptr;
    b.bytes = p->bytes;
  }
  return b;
}

}  // namespace

}  // namespace svo

using namespace svo;

struct svo_frame_graph {
  Context *ctx = nullptr;
  uint32_t n_cams = 0;
  bool per_camera = false;
  std::vector<svo_trace_input> layout;  // resolutions, FoVs, BaseDepth captured
  std::vector<float> apertures;
  std::vector<FrameScratch> scratch;    // one per submission (per camera in per-camera mode)
  std::shared_ptr<DeviceBuffer> lut, cams;
  CamRec *host = nullptr;               // pinned camera records, read by the graph's copy nodes
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t done = nullptr;
  std::vector<const void *> totals;

  ~svo_frame_graph() {
    if (done) (void)hipEventSynchronize(done);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (done) (void)hipEventDestroy(done);
    if (host) (void)hipHostFree(host);
  }
};

static FrameBuffers context_buffers(Context *ctx) {
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

static int capture(svo_frame_graph *g, const svo_trace_input *ins, const svo_frame_options &o0) {
  Context *ctx = g->ctx;
  const FrameBuffers fb = context_buffers(ctx);
  svo_frame_options o = o0;
  o.flags |= SVO_FRAME_CAMERA_TABLE;
  // camera tables of all cameras, owned by the graph
  std::vector<uint32_t> offsets(g->n_cams);
  size_t total = 0;
  for (uint32_t c = 0; c < g->n_cams; c++) {
    offsets[c] = (uint32_t)total;
    total += 2 * ((size_t)ins[c].screen_size[0] + ins[c].screen_size[1]);
  }
  std::vector<float> lut(total);
  for (uint32_t c = 0; c < g->n_cams; c++) {
    const uint32_t W = ins[c].screen_size[0], H = ins[c].screen_size[1];
    float *t = lut.data() + offsets[c];
    camera_lut(ins[c], t, t + W, t + 2 * W, t + 2 * W + H);
  }
  int rc = ensure_scratch(g->lut, lut.size() * sizeof(float));
  if (!rc) rc = ensure_scratch(g->cams, (size_t)g->n_cams * sizeof(CamRec));
  if (rc) return rc;
  HIP_TRY(hipMemcpy(g->lut->ptr, lut.data(), lut.size() * sizeof(float), hipMemcpyHostToDevice), "graph tables");
  HIP_TRY(hipHostMalloc((void **)&g->host, (size_t)g->n_cams * sizeof(CamRec), hipHostMallocDefault), "hipHostMalloc");
  HIP_TRY(hipEventCreateWithFlags(&g->done, hipEventDisableTiming), "hipEventCreate");

  const float *ap = g->apertures.empty() ? nullptr : g->apertures.data();
  const uint32_t subs = g->per_camera ? g->n_cams : 1u;
  g->scratch.resize(subs);
  std::vector<BatchPlan> plans(subs);
  CamRec *dev = (CamRec *)g->cams->ptr;
  for (uint32_t s = 0; s < subs; s++) {
    const uint32_t first = g->per_camera ? s : 0u, count = g->per_camera ? 1u : g->n_cams;
    if ((rc = plan_batch(ctx, g->scratch[s], ins + first, ap ? ap + first : nullptr, count, o, fb, g->host + first,
                         dev + first, (const float *)g->lut->ptr, offsets.data() + first, &plans[s])))
      return rc;
    g->totals.push_back(plans[s].p.totals);
  }
  HIP_TRY(hipStreamSynchronize(ctx->stream), "graph capture: sync");
  HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  for (uint32_t s = 0; s < subs && !rc; s++) {
    const uint32_t first = g->per_camera ? s : 0u;
    rc = submit_batch(ctx, plans[s], g->host + first, false);
  }
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  g->graph = graph;
  HIP_TRY(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  return SVO_OK;
}

extern "C" int svo_frame_graph_create(svo_context *ctx, const svo_trace_input *inputs, const float *apertures,
                                      uint32_t n_cameras, const svo_frame_options *options, uint32_t per_camera,
                                      svo_frame_graph **out) {
  if (!ctx || !inputs || !out) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_frame_graph_create: null argument");
  *out = nullptr;
  size_t cones = 0;
  int rc = validate_batch(inputs, apertures, n_cameras, &cones);
  if (rc) return rc;
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(SVO_ERR_HIP, "hipSetDevice");
  svo_frame_options o;
  std::memset(&o, 0, sizeof o);
  if (options) o = *options;
  std::unique_ptr<svo_frame_graph> g(new svo_frame_graph);
  g->ctx = ctx;
  g->n_cams = n_cameras;
  g->per_camera = per_camera != 0;
  g->layout.assign(inputs, inputs + n_cameras);
  if (apertures) g->apertures.assign(apertures, apertures + n_cameras);
  if ((rc = capture(g.get(), inputs, o))) return rc;
  *out = g.release();
  return SVO_OK;
}

extern "C" int svo_frame_graph_launch(svo_context *ctx, svo_frame_graph *g, const svo_trace_input *inputs) {
  if (!ctx || !g || !inputs || g->ctx != ctx) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_frame_graph_launch: bad argument");
  for (uint32_t c = 0; c < g->n_cams; c++) {
    const svo_trace_input &a = inputs[c], &b = g->layout[c];
    if (a.screen_size[0] != b.screen_size[0] || a.screen_size[1] != b.screen_size[1] ||
        std::memcmp(a.fov, b.fov, sizeof a.fov) != 0 || a.base_depth != b.base_depth ||
        a.max_child_request_id != b.max_child_request_id)
      return fail(SVO_ERR_INVALID_ARGUMENT,
                  "svo_frame_graph_launch: camera " + std::to_string(c) + " differs from the captured layout");
    if (a.tick != inputs[0].tick) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_frame_graph_launch: cameras share Tick");
  }
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(SVO_ERR_HIP, "hipSetDevice");
  // the previous replay's copy nodes must have read the pinned records
  if (hipEventSynchronize(g->done) != hipSuccess) return fail(SVO_ERR_HIP, "svo_frame_graph_launch: previous replay");
  const float *ap = g->apertures.empty() ? nullptr : g->apertures.data();
  if (g->per_camera) {
    for (uint32_t c = 0; c < g->n_cams; c++) fill_camera_poses(inputs + c, ap ? ap + c : nullptr, 1, g->host + c);
  } else {
    fill_camera_poses(inputs, ap, g->n_cams, g->host);
  }
  // the replayed frames write ChildRequests: earlier kept requests are no longer the last frame's
  ctx->kept.valid = false;
  hipError_t e = hipGraphLaunch(g->exec, ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(g->done, ctx->stream);
  if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
  size_t cones = 0;
  for (uint32_t c = 0; c < g->n_cams; c++) cones += (size_t)inputs[c].screen_size[0] * inputs[c].screen_size[1];
  ctx->last_rays = g->per_camera ? (size_t)inputs[g->n_cams - 1].screen_size[0] * inputs[g->n_cams - 1].screen_size[1]
                                 : cones;
  ctx->last_totals = g->totals.back();
  return SVO_OK;
}

extern "C" int svo_frame_graph_destroy(svo_context *ctx, svo_frame_graph *g) {
  if (!g) return SVO_OK;
  if (ctx) {
    (void)hipStreamSynchronize(ctx->stream);
    for (const void *t : g->totals)
      if (ctx->last_totals == t) ctx->last_totals = nullptr;
  }
  delete g;
  return SVO_OK;
}
