// Adam update of the per-block leaf parameters (colour and opacity) from accumulated
// gradients, with the parameters kept in [0, 1]; the gradients are cleared for the next
// step.  Elementwise over the float4 of every block: HBM-bound, 64 B per block
// (params, gradient, two moments: read and write).
#include "svo_device.hpp"
#include "svo_opt_kernels.hpp"

namespace svo {
namespace {

__global__ __launch_bounds__(256) void adam_kernel(float4 *params, float4 *grads, float4 *m, float4 *v, uint32_t n,
                                                   AdamParams a) {
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
    float4 g = grads[i], mm = m[i], vv = v[i], p = params[i];
    float *gp = &g.x, *mp = &mm.x, *vp = &vv.x, *pp = &p.x;
    for (int c = 0; c < 4; c++) {
      mp[c] = a.beta1 * mp[c] + (1.0f - a.beta1) * gp[c];
      vp[c] = a.beta2 * vp[c] + (1.0f - a.beta2) * (gp[c] * gp[c]);
      const float mh = mp[c] / a.bias1, vh = vp[c] / a.bias2;
      const float q = pp[c] - a.lr * mh / (sqrtf(vh) + a.eps);
      pp[c] = fminf(fmaxf(q, 0.0f), 1.0f);
    }
    params[i] = p;
    m[i] = mm;
    v[i] = vv;
    grads[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}

}  // namespace

int launch_adam(float4 *params, float4 *grads, float4 *m, float4 *v, uint32_t n, const AdamParams &a,
                int cu_count, hipStream_t stream) {
  if (!n) return 0;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255u) / 256u, (uint64_t)cu_count * 8u);
  hipLaunchKernelGGL(adam_kernel, dim3(blocks), dim3(256), 0, stream, params, grads, m, v, n, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace svo
