// Launch interface of the leaf-parameter optimiser.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace svo {

struct AdamParams {
  float lr, beta1, beta2, eps;
  float bias1, bias2;  // 1 - beta^t of this step
};

int launch_adam(float4 *params, float4 *grads, float4 *m, float4 *v, uint32_t n, const AdamParams &a,
                int cu_count, hipStream_t stream);

}  // namespace svo
