// Kernel-by-kernel wavefront path (placeholder until the kernels land).
#include <string>

#include "svo_internal.hpp"
#include "svo_wavefront.hpp"

namespace svo {

int enqueue_wavefront(Context *, int kernel, uint32_t, const size_t *) {
  return fail(SVO_ERR_NOT_SUPPORTED, "svo_enqueue: kernel " + std::to_string(kernel) + " has no implementation yet");
}

}  // namespace svo
