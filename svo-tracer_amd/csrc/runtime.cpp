// Host runtime of libsvo: the C ABI of include/svo.h.
//
// Mirrors SvoTracer.Kernel.ComputeManager (ComputeManager.cs:10-367): a context
// owns one HIP stream (the single in-order queue) and device buffers keyed by
// BufferName; kernel arguments are bound by (KernelName, parameter name) and, as
// with clSetKernelArg, a buffer binding captures the allocation itself, so Swap
// only changes which allocation a *name* refers to.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "svo_internal.hpp"
#include "svo_kernels.hpp"
#include "svo_runtime.hpp"

namespace svo {

static thread_local std::string g_error;

int fail(int code, const std::string &msg) {
  g_error = msg;
  return code;
}

DeviceBuffer::~DeviceBuffer() {
  if (ptr) (void)hipFree(ptr);
}

int hip_fail(hipError_t e, const char *what) {
  return fail(e == hipErrorOutOfMemory ? SVO_ERR_OUT_OF_MEMORY : SVO_ERR_HIP,
              std::string(what) + ": " + hipGetErrorString(e));
}

int new_buffer(size_t bytes, std::shared_ptr<DeviceBuffer> *out) {
  auto b = std::make_shared<DeviceBuffer>();
  b->bytes = bytes;
  if (bytes) {
    hipError_t e = hipMalloc(&b->ptr, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  }
  *out = b;
  return SVO_OK;
}

// Grows a context-owned scratch allocation to at least `bytes`.
int ensure_scratch(std::shared_ptr<DeviceBuffer> &b, size_t bytes) {
  if (b && b->bytes >= bytes) return SVO_OK;
  if (b) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_fail(e, "hipDeviceSynchronize");
  }
  std::shared_ptr<DeviceBuffer> nb;
  int rc = new_buffer(std::max<size_t>(bytes, 256), &nb);
  if (rc) return rc;
  b = nb;
  return SVO_OK;
}

int context_scan_sums(Context *ctx, size_t bytes, uint32_t **out) {
  *out = nullptr;
  if (!bytes) return SVO_OK;
  int rc = ensure_scratch(ctx->scan_sums, bytes);
  if (rc) return rc;
  *out = (uint32_t *)ctx->scan_sums->ptr;
  return SVO_OK;
}

// Kernel parameter names in argument order, ComputeManager.cs:313-344.
const std::vector<std::string> &kernel_params(int kernel) {
  static const std::vector<std::vector<std::string>> tables = {
      {"bases", "blocks", "usage", "childRequestId", "childRequests", "parentSize", "parentResidency", "parents",
       "dereferenceQueue", "dereferenceRemaining", "semaphor", "pruning", "pruningBlockData", "pruningAddresses",
       "inputData"},
      {"blocks", "usage", "childRequestId", "childRequests", "parentSize", "parentResidency", "parents",
       "dereferenceQueue", "dereferenceRemaining", "semaphor", "grafting", "graftingBlocks", "graftingAddresses",
       "holdingAddresses", "addressPosition", "inputData"},
      {"bases", "blocks", "usage", "childRequestId", "childRequests", "outputImage", "input"},
      {"Origins", "Directions", "FoVs", "Locations", "Depths", "BaseTraceQueue", "BaseTraceQueueID",
       "BackgroundQueue", "BackgroundQueueID", "ParentTraces", "ColourRs", "ColourGs", "ColourBs", "Weightings",
       "FinalColourRs", "FinalColourGs", "FinalColourBs", "FinalWeightings", "input"},
      {"Origins", "Directions", "FoVs", "Locations", "Depths", "BaseTraceQueue", "BaseTraceQueueID",
       "BackgroundQueue", "BackgroundQueueID", "ParentTraces", "ColourRs", "ColourGs", "ColourBs", "Weightings",
       "FinalColourRs", "FinalColourGs", "FinalColourBs", "FinalWeightings", "input"},
      {"Bases", "Origins", "Directions", "FoVs", "Locations", "Weightings", "Depths", "BaseTraces",
       "BlockTraceQueue", "BlockTraceQueueID", "BaseTraceQueue", "BaseTraceQueueID", "BackgroundQueue",
       "BackgroundQueueID", "input"},
      {"Blocks", "Usages", "ChildRequestID", "ChildRequests", "Origins", "Directions", "FoVs", "Locations",
       "Weightings", "Depths", "BlockTraces", "BlockTraceQueue", "BlockTraceQueueID", "BaseTraceQueue",
       "BaseTraceQueueID", "BackgroundQueue", "BackgroundQueueID", "MaterialQueue", "MaterialQueueID",
       "ParentTraces", "input"},
      {"BackgroundQueue", "Directions", "ParentTraces", "ColourRs", "ColourGs", "ColourBs", "Weightings",
       "FinalColourRs", "FinalColourGs", "FinalColourBs", "FinalWeightings", "input"},
      {"Blocks", "Usages", "Origins", "Directions", "FoVs", "Locations", "Depths", "Weightings", "ColourRs",
       "ColourGs", "ColourBs", "RayLengths", "Luminosities", "MaterialQueue", "ParentTraces", "RootDirections",
       "RootLocations", "RootDepths", "RootWeightings", "RootParentTraces", "BaseTraceQueue", "BaseTraceQueueID",
       "FinalWeightings", "AccumulatorID", "BackgroundQueue", "BackgroundQueueID", "input"},
      {"Blocks", "Usages", "Locations", "Depths", "MaterialQueue", "FinalColourRs", "FinalColourGs",
       "FinalColourBs", "FinalWeightings", "Weightings", "ParentTraces", "input"},
      {"FinalColourRs", "FinalColourGs", "FinalColourBs", "FinalWeightings", "outputImage", "input"},
  };
  return tables[kernel];
}

bool valid_buffer_name(int name) {
  return (name >= 0 && name < SVO_BUF_REFERENCE_COUNT) || (name >= SVO_BUF_OUTPUT_IMAGE && name <= SVO_BUF_ADAM_V);
}

std::shared_ptr<DeviceBuffer> Context::get(int name) const {
  auto it = buffers.find(name);
  return it == buffers.end() ? nullptr : it->second;
}

}  // namespace svo

using namespace svo;

#define CHECK_CTX(ctx, fn) \
  if (!(ctx)) return fail(SVO_ERR_INVALID_ARGUMENT, fn ": null context")

#define HIP_TRY(expr, what)                  \
  do {                                       \
    hipError_t e_ = (expr);                  \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

extern "C" {

const char *svo_last_error(void) { return g_error.c_str(); }
int svo_abi_version(void) { return SVO_ABI_VERSION; }

int svo_device_count(int *count) {
  if (!count) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_device_count: null argument");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return SVO_OK;
}

int svo_create(int device_ordinal, svo_context **out) {
  if (!out) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_create: null argument");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(SVO_ERR_NO_DEVICE, "svo_create: no HIP device available");
  if (device_ordinal < 0 || device_ordinal >= n) return fail(SVO_ERR_NO_DEVICE, "svo_create: device ordinal out of range");
  HIP_TRY(hipSetDevice(device_ordinal), "hipSetDevice");
  std::unique_ptr<svo_context> ctx(new svo_context());
  ctx->device = device_ordinal;
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device_ordinal), "hipGetDeviceProperties");
  ctx->cu_count = prop.multiProcessorCount;
  ctx->arch = prop.gcnArchName;
  HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
  HIP_TRY(hipEventCreate(&ctx->timer_begin), "hipEventCreate");
  HIP_TRY(hipEventCreate(&ctx->timer_end), "hipEventCreate");
  *out = ctx.release();
  return SVO_OK;
}

int svo_destroy(svo_context *ctx) {
  if (!ctx) return SVO_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->buffers.clear();
  ctx->args.clear();
  ctx->frame_scratch = FrameScratch();
  ctx->scratch_lut.reset();
  for (auto &pr : ctx->kernel_events) {
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  (void)hipEventDestroy(ctx->timer_begin);
  (void)hipEventDestroy(ctx->timer_end);
  for (auto &b : ctx->tree_scratch) b.reset();
  if (ctx->tree_host) (void)hipHostFree(ctx->tree_host);
  if (ctx->stream_host) (void)hipHostFree(ctx->stream_host);
  if (ctx->id_host) (void)hipHostFree(ctx->id_host);
  if (ctx->cam_copied) (void)hipEventSynchronize(ctx->cam_copied);
  if (ctx->cam_host) (void)hipHostFree(ctx->cam_host);
  if (ctx->cam_copied) (void)hipEventDestroy(ctx->cam_copied);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return SVO_OK;
}

int svo_buffer_init(svo_context *ctx, int name, size_t bytes, const void *host_data, uint32_t flags) {
  CHECK_CTX(ctx, "svo_buffer_init");
  if (!valid_buffer_name(name)) return fail(SVO_ERR_UNKNOWN_BUFFER, "svo_buffer_init: unknown buffer name " + std::to_string(name));
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  std::shared_ptr<DeviceBuffer> b;
  int rc = new_buffer(bytes, &b);
  if (rc) return rc;
  b->flags = flags;
  if (bytes) {
    if (host_data) {
      HIP_TRY(hipMemcpyAsync(b->ptr, host_data, bytes, hipMemcpyHostToDevice, ctx->stream), "svo_buffer_init: copy");
      HIP_TRY(hipStreamSynchronize(ctx->stream), "svo_buffer_init: sync");
    } else {
      HIP_TRY(hipMemsetAsync(b->ptr, 0, bytes, ctx->stream), "svo_buffer_init: clear");
    }
  }
  ctx->buffers[name] = b;  // the previous allocation is released once nothing binds it
  return SVO_OK;
}

int svo_buffer_release(svo_context *ctx, int name) {
  CHECK_CTX(ctx, "svo_buffer_release");
  auto it = ctx->buffers.find(name);
  if (it == ctx->buffers.end()) return fail(SVO_ERR_UNKNOWN_BUFFER, "svo_buffer_release: buffer not initialised");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "svo_buffer_release: sync");
  ctx->buffers.erase(it);
  return SVO_OK;
}

int svo_buffer_size(svo_context *ctx, int name, size_t *bytes) {
  CHECK_CTX(ctx, "svo_buffer_size");
  auto b = ctx->get(name);
  if (!b) return fail(SVO_ERR_UNKNOWN_BUFFER, "svo_buffer_size: buffer " + std::to_string(name) + " not initialised");
  if (bytes) *bytes = b->bytes;
  return SVO_OK;
}

static int checked_range(svo_context *ctx, int name, size_t offset, size_t bytes, const char *fn,
                         std::shared_ptr<DeviceBuffer> *out) {
  auto b = ctx->get(name);
  if (!b) return fail(SVO_ERR_UNKNOWN_BUFFER, std::string(fn) + ": buffer " + std::to_string(name) + " not initialised");
  if (offset > b->bytes || bytes > b->bytes - offset)
    return fail(SVO_ERR_BUFFER_TOO_SMALL, std::string(fn) + ": range exceeds buffer " + std::to_string(name));
  *out = b;
  return SVO_OK;
}

int svo_buffer_write(svo_context *ctx, int name, size_t offset, const void *src, size_t bytes) {
  CHECK_CTX(ctx, "svo_buffer_write");
  if (!src && bytes) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_buffer_write: null source");
  std::shared_ptr<DeviceBuffer> b;
  int rc = checked_range(ctx, name, offset, bytes, "svo_buffer_write", &b);
  if (rc) return rc;
  if (!bytes) return SVO_OK;
  HIP_TRY(hipMemcpyAsync((char *)b->ptr + offset, src, bytes, hipMemcpyHostToDevice, ctx->stream), "svo_buffer_write");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "svo_buffer_write: sync");
  return SVO_OK;
}

int svo_buffer_read(svo_context *ctx, int name, size_t offset, void *dst, size_t bytes) {
  CHECK_CTX(ctx, "svo_buffer_read");
  if (!dst && bytes) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_buffer_read: null destination");
  std::shared_ptr<DeviceBuffer> b;
  int rc = checked_range(ctx, name, offset, bytes, "svo_buffer_read", &b);
  if (rc) return rc;
  if (!bytes) return SVO_OK;
  HIP_TRY(hipMemcpyAsync(dst, (const char *)b->ptr + offset, bytes, hipMemcpyDeviceToHost, ctx->stream), "svo_buffer_read");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "svo_buffer_read: sync");
  return SVO_OK;
}

int svo_buffer_fill(svo_context *ctx, int name, const void *pattern, size_t pattern_bytes, size_t offset, size_t bytes) {
  CHECK_CTX(ctx, "svo_buffer_fill");
  if (!pattern || pattern_bytes == 0 || pattern_bytes > 128 || bytes % pattern_bytes != 0)
    return fail(SVO_ERR_INVALID_ARGUMENT, "svo_buffer_fill: pattern must be 1..128 bytes and divide the range");
  std::shared_ptr<DeviceBuffer> b;
  int rc = checked_range(ctx, name, offset, bytes, "svo_buffer_fill", &b);
  if (rc) return rc;
  if (!bytes) return SVO_OK;
  std::vector<uint8_t> host(bytes);
  for (size_t i = 0; i < bytes; i += pattern_bytes) std::memcpy(&host[i], pattern, pattern_bytes);
  HIP_TRY(hipMemcpyAsync((char *)b->ptr + offset, host.data(), bytes, hipMemcpyHostToDevice, ctx->stream), "svo_buffer_fill");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "svo_buffer_fill: sync");
  return SVO_OK;
}

int svo_buffer_copy(svo_context *ctx, int dst_name, size_t dst_offset, int src_name, size_t src_offset, size_t bytes) {
  CHECK_CTX(ctx, "svo_buffer_copy");
  std::shared_ptr<DeviceBuffer> d, s;
  int rc = checked_range(ctx, dst_name, dst_offset, bytes, "svo_buffer_copy", &d);
  if (rc) return rc;
  rc = checked_range(ctx, src_name, src_offset, bytes, "svo_buffer_copy", &s);
  if (rc) return rc;
  if (!bytes) return SVO_OK;
  HIP_TRY(hipMemcpyAsync((char *)d->ptr + dst_offset, (const char *)s->ptr + src_offset, bytes, hipMemcpyDeviceToDevice,
                         ctx->stream),
          "svo_buffer_copy");
  return SVO_OK;
}

int svo_reset_id(svo_context *ctx, int name, uint32_t *count_out) {
  CHECK_CTX(ctx, "svo_reset_id");
  std::shared_ptr<DeviceBuffer> b;
  int rc = checked_range(ctx, name, 0, 4, "svo_reset_id", &b);
  if (rc) return rc;
  // read into a pinned word and clear on the device: one round trip per call
  if (!ctx->id_host) HIP_TRY(hipHostMalloc((void **)&ctx->id_host, 4, hipHostMallocDefault), "hipHostMalloc");
  HIP_TRY(hipMemcpyAsync(ctx->id_host, b->ptr, 4, hipMemcpyDeviceToHost, ctx->stream), "svo_reset_id: read");
  HIP_TRY(hipMemsetAsync(b->ptr, 0, 4, ctx->stream), "svo_reset_id: clear");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "svo_reset_id: sync");
  if (count_out) *count_out = *ctx->id_host;
  return SVO_OK;
}

int svo_zero_id(svo_context *ctx, int name) {
  CHECK_CTX(ctx, "svo_zero_id");
  std::shared_ptr<DeviceBuffer> b;
  int rc = checked_range(ctx, name, 0, 4, "svo_zero_id", &b);
  if (rc) return rc;
  HIP_TRY(hipMemsetAsync(b->ptr, 0, 4, ctx->stream), "svo_zero_id");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "svo_zero_id: sync");
  return SVO_OK;
}

int svo_swap(svo_context *ctx, int a, int b) {
  CHECK_CTX(ctx, "svo_swap");
  auto ia = ctx->buffers.find(a), ib = ctx->buffers.find(b);
  if (ia == ctx->buffers.end() || ib == ctx->buffers.end())
    return fail(SVO_ERR_UNKNOWN_BUFFER, "svo_swap: both buffers must be initialised");
  std::swap(ia->second, ib->second);
  return SVO_OK;
}

int svo_buffer_device_pointer(svo_context *ctx, int name, void **device_ptr) {
  CHECK_CTX(ctx, "svo_buffer_device_pointer");
  auto b = ctx->get(name);
  if (!b) return fail(SVO_ERR_UNKNOWN_BUFFER, "svo_buffer_device_pointer: buffer not initialised");
  if (device_ptr) *device_ptr = b->ptr;
  return SVO_OK;
}

int svo_set_arg_buffer(svo_context *ctx, int kernel, const char *param, int name) {
  CHECK_CTX(ctx, "svo_set_arg_buffer");
  if (kernel < 0 || kernel >= SVO_KERNEL_COUNT) return fail(SVO_ERR_UNKNOWN_KERNEL, "svo_set_arg_buffer: unknown kernel");
  if (!param) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_set_arg_buffer: null parameter name");
  const auto &names = kernel_params(kernel);
  if (std::find(names.begin(), names.end(), param) == names.end())
    return fail(SVO_ERR_UNKNOWN_PARAMETER, std::string("svo_set_arg_buffer: no parameter ") + param);
  auto b = ctx->get(name);
  if (!b) return fail(SVO_ERR_UNKNOWN_BUFFER, "svo_set_arg_buffer: buffer " + std::to_string(name) + " not initialised");
  Arg a;
  a.buffer = b;
  a.buffer_name = name;
  ctx->args[kernel][param] = a;
  return SVO_OK;
}

int svo_set_local_arg(svo_context *ctx, int kernel, const char *param, int name) {
  CHECK_CTX(ctx, "svo_set_local_arg");
  if (kernel < 0 || kernel >= SVO_KERNEL_COUNT) return fail(SVO_ERR_UNKNOWN_KERNEL, "svo_set_local_arg: unknown kernel");
  if (!param) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_set_local_arg: null parameter name");
  const auto &names = kernel_params(kernel);
  if (std::find(names.begin(), names.end(), param) == names.end())
    return fail(SVO_ERR_UNKNOWN_PARAMETER, std::string("svo_set_local_arg: no parameter ") + param);
  if (!ctx->get(name))
    return fail(SVO_ERR_UNKNOWN_BUFFER, "svo_set_local_arg: buffer " + std::to_string(name) + " not initialised");
  return fail(SVO_ERR_INVALID_ARGUMENT,
              std::string("svo_set_local_arg: ") + param + " is not a local-memory argument of this kernel");
}

int svo_set_arg_bytes(svo_context *ctx, int kernel, const char *param, const void *value, size_t bytes) {
  CHECK_CTX(ctx, "svo_set_arg_bytes");
  if (kernel < 0 || kernel >= SVO_KERNEL_COUNT) return fail(SVO_ERR_UNKNOWN_KERNEL, "svo_set_arg_bytes: unknown kernel");
  if (!param || (!value && bytes)) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_set_arg_bytes: null argument");
  const auto &names = kernel_params(kernel);
  if (std::find(names.begin(), names.end(), param) == names.end())
    return fail(SVO_ERR_UNKNOWN_PARAMETER, std::string("svo_set_arg_bytes: no parameter ") + param);
  std::string p(param);
  size_t want = 0;
  if (p == "input") want = sizeof(svo_trace_input);
  if (p == "inputData") want = sizeof(svo_update_input_data);
  if (want && bytes != want)
    return fail(SVO_ERR_INVALID_ARGUMENT, "svo_set_arg_bytes: " + p + " must be " + std::to_string(want) + " bytes");
  Arg a;
  a.bytes.assign((const uint8_t *)value, (const uint8_t *)value + bytes);
  ctx->args[kernel][p] = a;
  return SVO_OK;
}

int svo_flush(svo_context *ctx) {
  CHECK_CTX(ctx, "svo_flush");
  return SVO_OK;  // work is submitted to the stream as it is enqueued
}

int svo_finish(svo_context *ctx) {
  CHECK_CTX(ctx, "svo_finish");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "svo_finish");
  return SVO_OK;
}

int svo_timer_begin(svo_context *ctx) {
  CHECK_CTX(ctx, "svo_timer_begin");
  HIP_TRY(hipEventRecord(ctx->timer_begin, ctx->stream), "hipEventRecord");
  return SVO_OK;
}

int svo_timer_end(svo_context *ctx, float *elapsed_ms) {
  CHECK_CTX(ctx, "svo_timer_end");
  HIP_TRY(hipEventRecord(ctx->timer_end, ctx->stream), "hipEventRecord");
  HIP_TRY(hipEventSynchronize(ctx->timer_end), "hipEventSynchronize");
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, ctx->timer_begin, ctx->timer_end), "hipEventElapsedTime");
  if (elapsed_ms) *elapsed_ms = ms;
  return SVO_OK;
}

int svo_kernel_timing_enable(svo_context *ctx, int enable) {
  CHECK_CTX(ctx, "svo_kernel_timing_enable");
  ctx->kernel_timing = enable != 0;
  return SVO_OK;
}

int svo_kernel_timing_get(svo_context *ctx, double *total_ms, uint64_t *launches, int reset) {
  CHECK_CTX(ctx, "svo_kernel_timing_get");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "svo_kernel_timing_get: sync");
  double sum = 0;
  for (size_t i = 0; i < ctx->kernel_events_used; i++) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->kernel_events[i].first, ctx->kernel_events[i].second), "hipEventElapsedTime");
    sum += ms;
  }
  if (total_ms) *total_ms = sum;
  if (launches) *launches = ctx->kernel_events_used;
  if (reset) ctx->kernel_events_used = 0;
  return SVO_OK;
}

int svo_frame_stats_get(svo_context *ctx, svo_frame_stats *stats) {
  CHECK_CTX(ctx, "svo_frame_stats_get");
  if (!stats) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_frame_stats_get: null argument");
  if (!ctx->last_totals) return fail(SVO_ERR_NOT_SUPPORTED, "svo_frame_stats_get: no frame traced yet");
  unsigned long long t[4] = {0, 0, 0, 0};
  HIP_TRY(hipMemcpyAsync(t, ctx->last_totals, sizeof t, hipMemcpyDeviceToHost, ctx->stream),
          "svo_frame_stats_get");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "svo_frame_stats_get: sync");
  stats->base_steps = t[0];
  stats->block_steps = t[1];
  stats->child_requests = t[2];
  stats->rays = ctx->last_rays;
  return SVO_OK;
}

int svo_convert_probe(svo_context *ctx, int kind, const void *in, void *out, size_t n) {
  CHECK_CTX(ctx, "svo_convert_probe");
  if ((n && (!in || !out)) || (kind != 0 && kind != 1))
    return fail(SVO_ERR_INVALID_ARGUMENT, "svo_convert_probe: bad argument");
  if (!n) return SVO_OK;
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  const size_t in_bytes = n * (kind == 0 ? 4 : 8), out_bytes = n * (kind == 0 ? 8 : 4);
  std::shared_ptr<DeviceBuffer> a, b;
  int rc = new_buffer(in_bytes, &a);
  if (!rc) rc = new_buffer(out_bytes, &b);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(a->ptr, in, in_bytes, hipMemcpyHostToDevice, ctx->stream), "svo_convert_probe: upload");
  if (launch_convert((uint32_t)kind, a->ptr, b->ptr, n, ctx->stream) != 0)
    return hip_fail(hipGetLastError(), "svo_convert_probe: launch");
  HIP_TRY(hipMemcpyAsync(out, b->ptr, out_bytes, hipMemcpyDeviceToHost, ctx->stream), "svo_convert_probe: download");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "svo_convert_probe: sync");
  return SVO_OK;
}

int svo_context_load_tree(svo_context *ctx, const svo_tree *tree, int with_usages, uint16_t usage_tick) {
  CHECK_CTX(ctx, "svo_context_load_tree");
  if (!tree) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_context_load_tree: null tree");
  int rc = svo_buffer_init(ctx, SVO_BUF_BASES, tree->t.bases.size() * 2, tree->t.bases.data(), SVO_BUFFER_READ_ONLY);
  if (rc) return rc;
  rc = svo_buffer_init(ctx, SVO_BUF_BLOCKS, tree->t.blocks.size() * 16, tree->t.blocks.data(), SVO_BUFFER_READ_ONLY);
  if (rc) return rc;
  ctx->tree_max_depth = tree->t.max_depth;
  if (with_usages) {
    size_t groups = (tree->t.block_count + 7) / 8;
    std::vector<svo_usage> u(std::max<size_t>(groups, 1));
    rc = svo_tree_initial_usages(tree, usage_tick, u.data(), groups);
    if (rc) return rc;
    rc = svo_buffer_init(ctx, SVO_BUF_USAGES, groups * sizeof(svo_usage), u.data(), SVO_BUFFER_READ_WRITE);
    if (rc) return rc;
  }
  return SVO_OK;
}

}  // extern "C"
