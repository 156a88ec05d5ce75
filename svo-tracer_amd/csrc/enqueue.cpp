// svo_enqueue: Enqueue(KernelName, workSize), ComputeManager.cs:72-78.
#include <string>

#include "svo_internal.hpp"
#include "svo_runtime.hpp"
#include "svo_wavefront.hpp"

using namespace svo;

extern "C" int svo_enqueue(svo_context *ctx, int kernel, uint32_t dims, const size_t *global_size) {
  if (!ctx) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_enqueue: null context");
  if (kernel < 0 || kernel >= SVO_KERNEL_COUNT)
    return fail(SVO_ERR_UNKNOWN_KERNEL, "Kernel name " + std::to_string(kernel) + " not found");
  if (dims < 1 || dims > 2 || !global_size) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_enqueue: work size must have 1 or 2 dimensions");
  size_t items = global_size[0] * (dims == 2 ? global_size[1] : 1);
  if (items == 0) return SVO_OK;
  if (items > 0xffffffffu) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_enqueue: work size exceeds 2^32 items");
  // a frame traced kernel by kernel writes ChildRequests: earlier kept requests are no longer
  // the last frame's
  if (kernel == SVO_KERNEL_TRACE || kernel == SVO_KERNEL_INIT || kernel == SVO_KERNEL_RUN_BASE_TRACE ||
      kernel == SVO_KERNEL_RUN_BLOCK_TRACE)
    ctx->kept.valid = false;
  return enqueue_wavefront(ctx, kernel, dims, global_size);
}
