// Kernel-by-kernel path behind svo_enqueue (the reference's wavefront kernels).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "svo_runtime.hpp"

namespace svo {

// Resolves the arguments bound to `kernel` and launches its HIP implementation.
int enqueue_wavefront(Context *ctx, int kernel, uint32_t dims, const size_t *global_size);

}  // namespace svo
