// Differentiable compositing frames (svo_composite) and the Adam step over the leaf
// parameters (svo_adam_step).  No reference counterpart: the reference stops every cone at
// its first hit; with every opacity 1 the composite reduces to that result.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

#include "svo_internal.hpp"
#include "svo_opt_kernels.hpp"
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

extern "C" int svo_composite(svo_context *ctx, const svo_trace_input *views, const float *apertures,
                             uint32_t n_views, const svo_composite_options *options, double *loss) {
  if (!ctx || !views) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_composite: null argument");
  svo_composite_options co;
  std::memset(&co, 0, sizeof co);
  if (options) co = *options;
  const uint32_t K = co.max_layers ? co.max_layers : kMaxLayers;
  if (K > kMaxLayers) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_composite: max_layers must be 1..4");
  const bool adjoint = (co.flags & SVO_COMPOSITE_ADJOINT) != 0;

  FrameBuffers fb;
  fb.bases = named(ctx, SVO_BUF_BASES);
  fb.blocks = named(ctx, SVO_BUF_BLOCKS);
  fb.rgba = named(ctx, SVO_BUF_OUTPUT_IMAGE);
  const FrameBuffers::B params = named(ctx, SVO_BUF_LEAF_PARAMS), target = named(ctx, SVO_BUF_TARGET_IMAGE),
                        grads = named(ctx, SVO_BUF_LEAF_GRADIENTS), layers = named(ctx, SVO_BUF_LAYERS);
  if (!fb.blocks.ptr) return fail(SVO_ERR_UNKNOWN_BUFFER, "svo_composite: buffer Blocks not initialised");
  const size_t n_blocks = fb.blocks.bytes / 16;
  if (!params.ptr || params.bytes < n_blocks * 16)
    return fail(SVO_ERR_UNKNOWN_BUFFER, "svo_composite: LeafParams must hold a float4 per block");
  if (adjoint && (!target.ptr || !grads.ptr || grads.bytes < n_blocks * 16))
    return fail(SVO_ERR_UNKNOWN_BUFFER, "svo_composite: the adjoint needs TargetImage and LeafGradients");

  if (!ctx->cam_host) {
    HIP_TRY(hipHostMalloc((void **)&ctx->cam_host, kMaxBatchCameras * sizeof(CamRec), hipHostMallocDefault),
            "hipHostMalloc");
    HIP_TRY(hipEventCreateWithFlags(&ctx->cam_copied, hipEventDisableTiming), "hipEventCreate");
  } else {
    HIP_TRY(hipEventSynchronize(ctx->cam_copied), "camera table reuse");
  }
  int rc = ensure_scratch(ctx->frame_scratch.cams, kMaxBatchCameras * sizeof(CamRec));
  if (rc) return rc;
  svo_frame_options o;
  std::memset(&o, 0, sizeof o);
  o.flags = SVO_FRAME_CAMERA_TABLE;
  o.workgroup_size = 256;
  BatchPlan plan;
  if ((rc = plan_batch(ctx, ctx->frame_scratch, views, apertures, n_views, o, fb, ctx->cam_host,
                       (CamRec *)ctx->frame_scratch.cams->ptr, nullptr, nullptr, &plan)))
    return rc;
  const size_t N = plan.cones;
  if (target.ptr && target.bytes < N * 16)
    return fail(SVO_ERR_BUFFER_TOO_SMALL, "svo_composite: TargetImage must hold a float4 per cone");
  if ((co.flags & SVO_COMPOSITE_LAYERS) && (!layers.ptr || layers.bytes < N * K * 4))
    return fail(SVO_ERR_BUFFER_TOO_SMALL, "svo_composite: Layers must hold max_layers u32 per cone");

  FrameParams &p = plan.p;
  p.leaf_params = (const float4 *)params.ptr;
  p.target = (const float4 *)target.ptr;
  p.grads = adjoint ? (float4 *)grads.ptr : nullptr;
  p.layers_out = (co.flags & SVO_COMPOSITE_LAYERS) ? (uint32_t *)layers.ptr : nullptr;
  p.loss = (double *)&p.totals[3];  // cleared with the frame totals
  p.max_layers = K;
  int per_cu = 0;
  if (composite_occupancy(plan.shape.lds_bytes, &per_cu) != 0 || per_cu <= 0)
    return fail(SVO_ERR_HIP, "svo_composite: occupancy query failed");
  plan.shape.grid = std::max(8, (per_cu * std::max(1, ctx->cu_count) / 8) * 8);

  const uint32_t n_inline = n_views <= kInlineCams ? n_views : 0u;
  if (!n_inline) {
    HIP_TRY(hipMemcpyAsync((void *)p.cams, ctx->cam_host, n_views * sizeof(CamRec), hipMemcpyHostToDevice,
                           ctx->stream),
            "camera upload");
    HIP_TRY(hipEventRecord(ctx->cam_copied, ctx->stream), "hipEventRecord");
  }
  if (launch_begin(plan, ctx->cam_host, n_inline, ctx->stream) != 0)
    return hip_fail(hipGetLastError(), "svo_composite: frame begin launch");
  if (launch_composite(p, plan.shape, adjoint ? kModeAdjoint : kModeComposite, ctx->stream) != 0)
    return hip_fail(hipGetLastError(), "svo_composite: launch");
  ctx->last_rays = N;
  ctx->last_totals = p.totals;
  if (loss) {
    HIP_TRY(hipMemcpyAsync(loss, p.loss, sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "svo_composite: loss");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "svo_composite: sync");
  }
  return SVO_OK;
}

extern "C" int svo_adam_step(svo_context *ctx, const svo_adam_options *options) {
  if (!ctx || !options) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_adam_step: null argument");
  const FrameBuffers::B params = named(ctx, SVO_BUF_LEAF_PARAMS), grads = named(ctx, SVO_BUF_LEAF_GRADIENTS),
                        m = named(ctx, SVO_BUF_ADAM_M), v = named(ctx, SVO_BUF_ADAM_V);
  if (!params.ptr || !grads.ptr || !m.ptr || !v.ptr)
    return fail(SVO_ERR_UNKNOWN_BUFFER, "svo_adam_step: LeafParams, LeafGradients, AdamM and AdamV are needed");
  const size_t n = params.bytes / 16;
  if (grads.bytes < n * 16 || m.bytes < n * 16 || v.bytes < n * 16)
    return fail(SVO_ERR_BUFFER_TOO_SMALL, "svo_adam_step: gradient and moment buffers must match LeafParams");
  if (options->step < 1) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_adam_step: step counts from 1");
  AdamParams a;
  a.lr = options->learning_rate;
  a.beta1 = options->beta1;
  a.beta2 = options->beta2;
  a.eps = options->epsilon;
  a.bias1 = (float)(1.0 - std::pow((double)a.beta1, (double)options->step));
  a.bias2 = (float)(1.0 - std::pow((double)a.beta2, (double)options->step));
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(SVO_ERR_HIP, "hipSetDevice");
  if (launch_adam((float4 *)params.ptr, (float4 *)grads.ptr, (float4 *)m.ptr, (float4 *)v.ptr,
                  (uint32_t)std::min<size_t>(n, 0xffffffffu), a, ctx->cu_count, ctx->stream) != 0)
    return hip_fail(hipGetLastError(), "svo_adam_step: launch");
  return SVO_OK;
}
