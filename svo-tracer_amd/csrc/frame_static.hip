// svo_trace_frame_static: the fused primary frame of svo_trace_frame without any allocation.
//
// Every piece of device scratch the frame uses -- the work counters and step totals, the
// camera record, the camera table, the per-cone status bytes, the staged child requests and
// the compaction counts -- is carved from one file-scope device array, `scratch`, of
// kStaticScratchBytes.  Each GPU the library runs this entry point on holds its own copy of
// the array from the first call until the process ends.  Every call on a GPU shares that
// GPU's copy: calls that overlap on one GPU (from several host threads or contexts) race.
// What must start at zero is cleared by the frame's first launch (counters and totals); the
// rest is fully written before it is read.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "svo_internal.hpp"
#include "svo_kernels.hpp"
#include "svo_runtime.hpp"

namespace svo {
namespace {

constexpr size_t kStaticScratchBytes = size_t(96) << 20;
constexpr int kMaxStaticDevices = 64;

// Host code never uses this name as an address (on the host it is only a stand-in for the
// device array); it uses the per-device address from hipGetSymbolAddress below.
static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

std::mutex g_scratch_mu;
uint8_t *g_scratch_ptr[kMaxStaticDevices] = {};

// The device address of `scratch` on `device`, which must be the current device; looked up
// once per device (the lookup loads this file's code object on that device).
int scratch_address(int device, uint8_t **out) {
  if (device < 0 || device >= kMaxStaticDevices)
    return fail(SVO_ERR_NOT_SUPPORTED, "svo_trace_frame_static: device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "svo_trace_frame_static: hipGetSymbolAddress");
    if (!ptr) return fail(SVO_ERR_HIP, "svo_trace_frame_static: no device address for the scratch array");
    g_scratch_ptr[device] = static_cast<uint8_t *>(ptr);
  }
  *out = g_scratch_ptr[device];
  return SVO_OK;
}

// Uploads the camera table (cos_u[W], sin_u[W], cos_v[H], sin_v[H]) in chunks staged on the
// host stack; each chunk's copy completes before the stack buffer is refilled.
int upload_camera_table(const svo_trace_input &in, float *dev, hipStream_t stream) {
  constexpr size_t kChunk = 2048;
  float host[kChunk];
  const size_t n = 2 * ((size_t)in.screen_size[0] + in.screen_size[1]);
  for (size_t k0 = 0; k0 < n; k0 += kChunk) {
    const size_t m = n - k0 < kChunk ? n - k0 : kChunk;
    for (size_t k = 0; k < m; k++) host[k] = camera_lut_at(in, k0 + k);
    hipError_t e = hipMemcpyAsync(dev + k0, host, m * sizeof(float), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "svo_trace_frame_static: camera table upload");
  }
  return SVO_OK;
}

}  // namespace
}  // namespace svo

using namespace svo;

extern "C" int svo_trace_frame_static(svo_context *ctx, const svo_trace_input *input,
                                      const svo_frame_options *options) {
  if (!ctx || !input) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_trace_frame_static: null argument");
  svo_frame_options o;
  std::memset(&o, 0, sizeof o);
  if (options) o = *options;
  size_t cones = 0;
  int rc = validate_batch(input, nullptr, 1, &cones);
  if (rc) return rc;
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(e, "svo_trace_frame_static: hipSetDevice");
  uint8_t *base = nullptr;
  if ((rc = scratch_address(ctx->device, &base))) return rc;

  FrameScratch sc;
  sc.arena = base;
  sc.arena_bytes = kStaticScratchBytes;
  void *cam_dev = nullptr, *lut_dev = nullptr;
  const size_t lut_bytes = 2 * ((size_t)input->screen_size[0] + input->screen_size[1]) * sizeof(float);
  if ((rc = frame_scratch(sc, sc.cams, sizeof(CamRec), &cam_dev)) ||
      (rc = frame_scratch(sc, sc.cams, lut_bytes, &lut_dev)))
    return fail(SVO_ERR_CAPACITY, std::string("svo_trace_frame_static: ") + svo_last_error());
  if ((rc = upload_camera_table(*input, (float *)lut_dev, ctx->stream))) return rc;

  CamRec host_cam;
  const uint32_t lut_offset = 0;
  BatchPlan plan;
  rc = plan_batch(ctx, sc, input, nullptr, 1, o, frame_buffers(ctx), &host_cam, (CamRec *)cam_dev,
                  (const float *)lut_dev, &lut_offset, &plan);
  if (rc == SVO_ERR_CAPACITY)
    return fail(SVO_ERR_CAPACITY, "svo_trace_frame_static: the frame's scratch (" + std::to_string(sc.arena_used) +
                                      " bytes and more) exceeds the static maximum of " +
                                      std::to_string(kStaticScratchBytes) + " bytes");
  if (rc) return rc;
  // this frame writes ChildRequests: earlier kept requests are no longer the last frame's
  ctx->kept.valid = false;
  // the camera record rides in the first launch's argument; nothing else is copied
  if (launch_begin(plan, &host_cam, 1, ctx->stream) != 0)
    return hip_fail(hipGetLastError(), "svo_trace_frame_static: frame begin launch");
  if (launch_trace_frame(plan.p, plan.shape, ctx->stream) != 0)
    return hip_fail(hipGetLastError(), "svo_trace_frame_static: trace launch");
  if (plan.compaction && launch_compaction(plan.compact, ctx->stream) != 0)
    return hip_fail(hipGetLastError(), "svo_trace_frame_static: compaction launch");
  ctx->last_totals = plan.p.totals;
  ctx->last_rays = plan.cones;
  return SVO_OK;
}
