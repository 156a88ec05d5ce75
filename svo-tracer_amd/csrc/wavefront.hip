// Kernel-by-kernel wavefront path: HIP implementations of the reference kernels
// init / runBaseTrace / runBlockTrace / evaluateBackground / resolveRemainders /
// drawTrace (wavefront.cl:1352-1850) over the reference's per-ray SoA buffers,
// for callers that drive the frame loop from the host through svo_enqueue
// exactly as MainWindow.RunWavefront does (MainWindow.cs:356-622).
//
// Queue appends are deterministic and order-preserving: a step kernel writes one
// code byte per work-item, then a count / scan / scatter pipeline appends the
// items to their queues in work-item order (the reference's work-group atomic
// `increment`, wavefront.cl:249-287, leaves the order to the scheduler).
#include <hip/hip_runtime.h>

#include "svo_device.hpp"
#include "svo_wavefront_kernels.hpp"

namespace svo {

using namespace dev;

namespace {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kPerThread = 16;
constexpr uint32_t kChunk = kThreads * kPerThread;

__device__ __forceinline__ void load_state(const WaveState &s, uint32_t id, float o[3], float d[3], float &fov,
                                           uint64_t &x, uint64_t &y, uint64_t &z, uint32_t &depth) {
  const float4 ov = s.origins[id], dv = s.directions[id];
  o[0] = ov.x;
  o[1] = ov.y;
  o[2] = ov.z;
  d[0] = dv.x;
  d[1] = dv.y;
  d[2] = dv.z;
  fov = s.fovs[id];
  x = s.locations[4 * (size_t)id + 0];
  y = s.locations[4 * (size_t)id + 1];
  z = s.locations[4 * (size_t)id + 2];
  depth = s.depths[id];
}

__device__ __forceinline__ void store_loc(const WaveState &s, uint32_t id, uint64_t x, uint64_t y, uint64_t z,
                                          uint32_t depth) {
  s.locations[4 * (size_t)id + 0] = x;
  s.locations[4 * (size_t)id + 1] = y;
  s.locations[4 * (size_t)id + 2] = z;
  s.depths[id] = (uint8_t)depth;
}

// getBlockAddress, wavefront.cl:296-304 (root-restart walk; the SoA state carries no stack).
__device__ __forceinline__ uint32_t walk_address(const uint4 *blocks, uint32_t n_blocks, uint32_t B, uint32_t depth,
                                                 uint64_t x, uint64_t y, uint64_t z) {
  uint32_t a = morton_top(B + 2u, x, y, z);
  for (uint32_t d = B + 2u; d < depth; d++) a = block_child(load_block(blocks, n_blocks, a)) + chunk_position(d, x, y, z);
  return a;
}

// init, wavefront.cl:1352-1398
__global__ void __launch_bounds__(kThreads) init_kernel(InitParams p) {
  const uint32_t gx = blockIdx.x * kThreads + threadIdx.x;
  const uint32_t gy = blockIdx.y;
  uint8_t code = kCodeNone;
  const uint32_t item = gx + gy * p.gsx;
  if (gx < p.gsx && gy < p.gsy) {
    const uint32_t id = gx + gy * p.W;
    const uint32_t x = gx, y = gy;
    if (id < p.s.n_rays && x < p.W && y < p.H) {
      const float cu = p.lut[x], su = p.lut[p.W + x], cv = p.lut[2u * p.W + y], sv = p.lut[2u * p.W + p.H + y];
      const float m0 = cu * cv, m1 = (0.0f - su) * cv, m2 = sv;
      float d[3], inv[3];
      for (int a = 0; a < 3; a++) {
        d[a] = (p.facing[a] * m0 + p.facing[3 + a] * m1) + p.facing[6 + a] * m2;
        inv[a] = 1.0f / d[a];
      }
      const float o[3] = {p.origin[0], p.origin[1], p.origin[2]};
      uint64_t lx = 0, ly = 0, lz = 0;
      const bool hit = trace_into_volume(o, d, inv, p.cone_fov, lx, ly, lz);
      if (hit) store_loc(p.s, id, lx, ly, lz, 1u);
      code = hit ? kQueueBase : kQueueBackground;
      p.s.origins[id] = make_float4(o[0], o[1], o[2], 0.0f);
      p.s.directions[id] = make_float4(d[0], d[1], d[2], 0.0f);
      p.s.fovs[id] = p.cone_fov;
      p.s.parents[id] = id;
      p.s.weightings[id] = 255;
      if (p.colour_r) p.colour_r[id] = 0;
      if (p.colour_g) p.colour_g[id] = 0;
      if (p.colour_b) p.colour_b[id] = 0;
      p.final_r[id] = 0;
      p.final_g[id] = 0;
      p.final_b[id] = 0;
      p.final_w[id] = 0;
      p.values[item] = id;
    }
  }
  if (gx < p.gsx && gy < p.gsy) p.codes[item] = code;
}

// runBaseTrace, wavefront.cl:1400-1463
__global__ void __launch_bounds__(kThreads) base_step_kernel(StepParams p) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= p.n) return;
  uint8_t code = kCodeNone;
  const uint32_t id = p.in_queue[i];
  p.values[i] = id;
  if (id < p.s.n_rays) {
    float o[3], d[3], fov;
    uint64_t x, y, z;
    uint32_t depth;
    load_state(p.s, id, o, d, fov, x, y, z, depth);
    const float ix = 1.0f / d[0], iy = 1.0f / d[1], iz = 1.0f / d[2];
    const uint32_t signs = (d[0] >= 0.0f ? 1u : 0u) | (d[1] >= 0.0f ? 2u : 0u) | (d[2] >= 0.0f ? 4u : 0u);
    const uint32_t B = p.B;
    if (depth >= 1u && depth <= B) {
      const uint32_t chunk = p.bases[base_offset(depth) + morton_top(depth, x, y, z)];
      if (((chunk >> (2u * chunk_position(depth, x, y, z))) & 2u) == 0u) {
        const uint64_t px = x, py = y, pz = z;
        traverse_chunk(depth, x, y, z, d[0], d[1], d[2], ix, iy, iz, signs);
        if (leaving(x, y, z, signs)) {
          depth = 0;
          code = kQueueBackground;
        } else {
          depth = compare_positions(depth, px, py, pz, x, y, z, signs);
          code = kQueueBase;
        }
      } else {
        depth = depth < B ? depth + 1u : B + 2u;
        if (depth == B + 1u) {
          depth = B;
          code = kQueueBase;
        } else {
          code = depth > B + 1u ? kQueueBlock : kQueueBase;
        }
      }
      store_loc(p.s, id, x, y, z, depth);
    }
  }
  p.codes[i] = code;
}

// runBlockTrace, wavefront.cl:1465-1566
__global__ void __launch_bounds__(kThreads) block_step_kernel(StepParams p) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= p.n) return;
  uint8_t code = kCodeNone;
  const uint32_t id = p.in_queue[i];
  p.values[i] = id;
  if (id < p.s.n_rays) {
    float o[3], d[3], fov;
    uint64_t x, y, z;
    uint32_t depth;
    load_state(p.s, id, o, d, fov, x, y, z, depth);
    const uint32_t B = p.B;
    if (depth >= B + 2u && depth <= kMaxDepth) {
      const float ix = 1.0f / d[0], iy = 1.0f / d[1], iz = 1.0f / d[2];
      const uint32_t signs = (d[0] >= 0.0f ? 1u : 0u) | (d[1] >= 0.0f ? 2u : 0u) | (d[2] >= 0.0f ? 4u : 0u);
      const uint32_t addr = walk_address(p.blocks, p.n_blocks, B, depth, x, y, z);
      const uint4 blk = load_block(p.blocks, p.n_blocks, addr);
      const uint32_t g = addr >> 3;
      if (p.usage16 && g < p.n_usages) {
        const uint16_t t = p.usage16[4u * (size_t)g];
        if (t < 65534u && t != (uint16_t)p.tick) p.usage16[4u * (size_t)g] = (uint16_t)p.tick;
      }
      const float cone = cone_at(o[0], o[1], o[2], x, y, z, fov, p.dof0, p.dof1);
      const uint32_t oct = chunk_position(depth, x, y, z);
      bool complete = false, child = false;
      if (((block_chunk(blk) >> (2u * oct)) & 2u) == 0u) {
        const uint64_t px = x, py = y, pz = z;
        traverse_chunk(depth, x, y, z, d[0], d[1], d[2], ix, iy, iz, signs);
        if (leaving(x, y, z, signs)) {
          depth = 0;
          complete = true;
        } else {
          depth = compare_positions(depth, px, py, pz, x, y, z, signs);
        }
      } else if (cone > exp2_neg(B + 2u)) {
        depth = B + 2u;
        complete = true;
      } else if (cone > exp2_neg(depth)) {
        depth--;
      } else if (block_child(blk) == kU32Max) {
        child = true;
        complete = true;
      } else if (cone < exp2_neg(depth + 1u) && depth + 1u <= kMaxDepth) {
        depth++;
      } else {
        complete = true;
      }
      if (child && p.req_staging) {
        svo_child_request rq;
        rq.address = addr;
        rq.tick = (uint16_t)p.tick;
        rq.depth = (uint8_t)depth;
        rq.pad0 = 0;
        rq.location[0] = x;
        rq.location[1] = y;
        rq.location[2] = z;
        rq.tree_size = cone <= exp2_neg(depth + 1u) ? 1 : 0;
        for (int k = 0; k < 7; k++) rq.pad1[k] = 0;
        p.req_staging[i] = rq;
      }
      if (complete) {
        code = depth == 0 ? kQueueBackground : kQueueMaterial;
      } else {
        if (depth == B + 1u) depth = B;
        code = depth > B ? kQueueBlock : kQueueBase;
      }
      if (child) code |= kCodeRequest;
      store_loc(p.s, id, x, y, z, depth);
    }
  }
  p.codes[i] = code;
}

// evaluateBackground, wavefront.cl:1741-1797 (the light mix of non-primary cones
// saturates its result to 0..255 where the reference's conversion is undefined)
__global__ void __launch_bounds__(kThreads) background_kernel(ShadeParams p) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= p.n) return;
  const uint32_t id = p.queue[i];
  if (id >= p.n_rays) return;
  const uint32_t parent = p.parents[id];
  if (parent >= p.n_pixels) return;
  const float4 dv = p.directions[id];
  const float dp = (dv.x * 0.0f + dv.y * 0.0f) + dv.z * -1.0f;
  uint32_t c[3];
  uint32_t lum;
  if (dp > 0.8f) {
    c[0] = c[1] = c[2] = 255u;
    lum = to_uchar(dp * 200.0f);
  } else {
    c[0] = 135u;
    c[1] = 206u;
    c[2] = 235u;
    lum = 1u;
  }
  const uint32_t w = p.weightings[id];
  if (w < 255u) {
    const uint32_t amb = p.ambient;
    const uint32_t boundary = lum > amb ? 255u : 0u;
    const uint32_t diff = lum > amb ? lum - amb : amb - lum;
    const float lq = (float)diff / ((float)(lum + amb) * 255.0f);
    const uint32_t base[3] = {p.colour_r ? p.colour_r[id] : 0u, p.colour_g ? p.colour_g[id] : 0u,
                              p.colour_b ? p.colour_b[id] : 0u};
    for (int k = 0; k < 3; k++) {
      const float v = (float)base[k] + lq * (float)c[k] * (float)(uint32_t)(boundary - base[k]);
      c[k] = to_uchar(v);
    }
  }
  atomicAdd(&p.final_w[parent], w);
  if (lum > 0u) {
    atomicAdd(&p.final_r[parent], c[0] * w);
    atomicAdd(&p.final_g[parent], c[1] * w);
    atomicAdd(&p.final_b[parent], c[2] * w);
  }
}

// resolveRemainders, wavefront.cl:1799-1833
__global__ void __launch_bounds__(kThreads) remainders_kernel(ShadeParams p) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= p.n) return;
  const uint32_t id = p.queue[i];
  if (id >= p.n_rays) return;
  const uint32_t parent = p.parents[id];
  if (parent >= p.n_pixels) return;
  const uint32_t w = p.weightings[id];
  const uint64_t x = p.locations[4 * (size_t)id], y = p.locations[4 * (size_t)id + 1], z = p.locations[4 * (size_t)id + 2];
  const uint32_t depth = p.depths[id];
  if (depth < p.B + 2u || depth > kMaxDepth) return;
  const uint32_t a = walk_address(p.blocks, p.n_blocks, p.B, depth, x, y, z);
  if (a == kU32Max) return;
  const uint4 blk = load_block(p.blocks, p.n_blocks, a);
  atomicAdd(&p.final_w[parent], w);
  atomicAdd(&p.final_r[parent], block_r(blk) * w);
  atomicAdd(&p.final_g[parent], block_g(blk) * w);
  atomicAdd(&p.final_b[parent], block_b(blk) * w);
}

// drawTrace, wavefront.cl:1835-1850 (into a row-major float4 buffer instead of a GL image)
__global__ void __launch_bounds__(kThreads) draw_kernel(DrawParams p) {
  const uint32_t gx = blockIdx.x * kThreads + threadIdx.x, gy = blockIdx.y;
  if (gx >= p.gsx || gy >= p.gsy) return;
  const uint32_t id = gx + gy * p.W;
  if (id >= p.n_pixels || id >= p.n_out) return;
  const uint32_t w = p.final_w[id];
  float4 c;
  if (w == 0u) {
    c = make_float4(0.75f, 0.0f, 0.0f, 1.0f);
  } else {
    const float fw = (float)w * 255.0f;
    c = make_float4(fminf((float)p.final_r[id] / fw, 1.0f), fminf((float)p.final_g[id] / fw, 1.0f),
                    fminf((float)p.final_b[id] / fw, 1.0f), 1.0f);
  }
  p.out[id] = c;
}

// ---- order-preserving append: count / scan / scatter ----

__device__ __forceinline__ void code_counts(uint8_t c, uint32_t v[kAppendTargets]) {
  const uint32_t q = c & 7u;
  for (uint32_t t = 0; t < 4; t++) v[t] += (q == t) ? 1u : 0u;
  v[4] += (c != kCodeNone && (c & kCodeRequest)) ? 1u : 0u;
}

__device__ void block_scan(uint32_t v[kAppendTargets], uint32_t total[kAppendTargets]) {
  __shared__ uint32_t s_wave[kAppendTargets][kThreads / 64];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t incl[kAppendTargets];
  for (int c = 0; c < kAppendTargets; c++) incl[c] = v[c];
  for (int c = 0; c < kAppendTargets; c++)
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t t = __shfl_up(incl[c], off);
      if (lane >= (uint32_t)off) incl[c] += t;
    }
  if (lane == 63u)
    for (int c = 0; c < kAppendTargets; c++) s_wave[c][wave] = incl[c];
  __syncthreads();
  for (int c = 0; c < kAppendTargets; c++) {
    uint32_t before = 0, all = 0;
    for (uint32_t w = 0; w < kThreads / 64; w++) {
      if (w < wave) before += s_wave[c][w];
      all += s_wave[c][w];
    }
    v[c] = before + incl[c] - v[c];
    total[c] = all;
  }
  __syncthreads();
}

__global__ void __launch_bounds__(kThreads) append_count_kernel(AppendParams p) {
  const uint32_t base = blockIdx.x * kChunk + threadIdx.x * kPerThread;
  uint32_t v[kAppendTargets] = {0, 0, 0, 0, 0};
  for (uint32_t k = 0; k < kPerThread; k++)
    if (base + k < p.n) code_counts(p.codes[base + k], v);
  uint32_t total[kAppendTargets];
  block_scan(v, total);
  if (threadIdx.x == 0)
    for (int c = 0; c < kAppendTargets; c++) p.counts[kAppendTargets * blockIdx.x + c] = total[c];
}

__global__ void __launch_bounds__(kThreads) append_scan_kernel(AppendParams p, uint32_t chunks) {
  uint32_t carry[kAppendTargets];
  for (int c = 0; c < kAppendTargets; c++) carry[c] = p.counter[c] ? *p.counter[c] : 0u;
  __syncthreads();
  for (uint32_t start = 0; start < chunks; start += kThreads) {
    const uint32_t i = start + threadIdx.x;
    uint32_t v[kAppendTargets] = {0, 0, 0, 0, 0};
    if (i < chunks)
      for (int c = 0; c < kAppendTargets; c++) v[c] = p.counts[kAppendTargets * i + c];
    uint32_t total[kAppendTargets];
    block_scan(v, total);
    if (i < chunks)
      for (int c = 0; c < kAppendTargets; c++) p.counts[kAppendTargets * i + c] = carry[c] + v[c];
    for (int c = 0; c < kAppendTargets; c++) carry[c] += total[c];
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int c = 0; c < kAppendTargets; c++)
      if (p.counter[c]) *p.counter[c] = carry[c];
}

__global__ void __launch_bounds__(kThreads) append_scatter_kernel(AppendParams p) {
  const uint32_t base = blockIdx.x * kChunk + threadIdx.x * kPerThread;
  uint32_t v[kAppendTargets] = {0, 0, 0, 0, 0};
  for (uint32_t k = 0; k < kPerThread; k++)
    if (base + k < p.n) code_counts(p.codes[base + k], v);
  uint32_t total[kAppendTargets];
  block_scan(v, total);
  uint32_t pos[kAppendTargets];
  for (int c = 0; c < kAppendTargets; c++) pos[c] = p.counts[kAppendTargets * blockIdx.x + c] + v[c];
  for (uint32_t k = 0; k < kPerThread; k++) {
    const uint32_t i = base + k;
    if (i >= p.n) break;
    const uint8_t c = p.codes[i];
    if (c == kCodeNone) continue;
    const uint32_t q = c & 7u;
    if (q < 4u) {
      if (p.queue[q] && pos[q] < p.capacity[q]) p.queue[q][pos[q]] = p.values[i];
      pos[q]++;
    }
    if (c & kCodeRequest) {
      if (p.requests && pos[4] < p.capacity[4]) p.requests[pos[4]] = p.req_staging[i];
      pos[4]++;
    }
  }
}

}  // namespace

uint32_t append_chunks(uint32_t n) { return (n + kChunk - 1u) / kChunk; }

int launch_append(const AppendParams &p, hipStream_t s) {
  const uint32_t chunks = append_chunks(p.n);
  if (chunks == 0) return 0;
  hipLaunchKernelGGL(append_count_kernel, dim3(chunks), dim3(kThreads), 0, s, p);
  hipLaunchKernelGGL(append_scan_kernel, dim3(1), dim3(kThreads), 0, s, p, chunks);
  hipLaunchKernelGGL(append_scatter_kernel, dim3(chunks), dim3(kThreads), 0, s, p);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_init(const InitParams &p, hipStream_t s) {
  hipLaunchKernelGGL(init_kernel, dim3((p.gsx + kThreads - 1) / kThreads, p.gsy), dim3(kThreads), 0, s, p);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_base_step(const StepParams &p, hipStream_t s) {
  hipLaunchKernelGGL(base_step_kernel, dim3((p.n + kThreads - 1) / kThreads), dim3(kThreads), 0, s, p);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_block_step(const StepParams &p, hipStream_t s) {
  hipLaunchKernelGGL(block_step_kernel, dim3((p.n + kThreads - 1) / kThreads), dim3(kThreads), 0, s, p);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_background(const ShadeParams &p, hipStream_t s) {
  hipLaunchKernelGGL(background_kernel, dim3((p.n + kThreads - 1) / kThreads), dim3(kThreads), 0, s, p);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_remainders(const ShadeParams &p, hipStream_t s) {
  hipLaunchKernelGGL(remainders_kernel, dim3((p.n + kThreads - 1) / kThreads), dim3(kThreads), 0, s, p);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_draw(const DrawParams &p, hipStream_t s) {
  hipLaunchKernelGGL(draw_kernel, dim3((p.gsx + kThreads - 1) / kThreads, p.gsy), dim3(kThreads), 0, s, p);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace svo
