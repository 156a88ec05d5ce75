// evaluateMaterial on the GPU (wavefront.cl:1568-1739): every cone that hit material spawns
// secondary cones around the reflection of its direction on the block's surface (a spine
// of rings), each set up one cell off the surface (setupTrace / traverseChunkNormal,
// wavefront.cl:617-636, 500-561) and saved as an accumulator trace (saveAccumulator,
// wavefront.cl:338-355).
//
// The reference hands out trace ids with a global atomic and appends to the background queue
// in arrival order.  Here the ids are canonical: material entries in queue order, each
// entry's cones in the order the reference's loops produce them.  A first pass counts the
// cones of every entry, a scan gives each entry its first id, a second pass writes them,
// and the background-queue appends are an order-preserving compaction of the cones whose
// weighting exceeds 5.
//
// Transcendentals: sin/cos come from sincos_poly below (argument reduction by
// the nearest multiple of pi/2 with a two-part constant, Cody and Waite's method, then the
// Taylor polynomials of sin and cos on [-pi/4, pi/4], all in IEEE float basic operations);
// normalize is v / sqrt(v.v) with correctly rounded division and square root.  With no FP
// contraction the host restatement gives the same bits.
#include "svo_device.hpp"
#include "svo_material_kernels.hpp"

namespace svo {
namespace {

using namespace dev;

constexpr int kWg = 256;
constexpr float kPiF = 3.14159274101257324f;      // M_PI_F
constexpr float kHalfPiF = 1.57079637050628662f;  // M_PI_2_F
// loop bounds of the definition (the reference's loops have none; a negative FovMultiplier
// can make the spine converge below pi)
constexpr uint32_t kMaxSpine = 256u, kMaxRing = 4096u;

struct V3 {
  float x, y, z;
};

__device__ __forceinline__ float dot3(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 normalize3(V3 v) {
  const float len = sqrtf(dot3(v, v));
  return {v.x / len, v.y / len, v.z / len};
}

// sin and cos of x: r = x - k pi/2 with k = rint(x 2/pi) (pi/2 = 1.5703125 + 4.838267948966e-4),
// then sin r = r + r^3(-1/6 + r^2(1/120 + r^2(-1/5040 + r^2/362880))) and
// cos r = 1 + r^2(-1/2 + r^2(1/24 + r^2(-1/720 + r^2(1/40320 - r^2/3628800)))).
__device__ __forceinline__ void sincos_poly(float x, float *s, float *c) {
  const float kf = rintf(x * 0.636619772f);
  const int k = (int)kf;
  const float r = (x - kf * 1.5703125f) - kf * 4.83826794897e-4f;
  const float r2 = r * r;
  const float ps = r + (r * r2) * (-0.166666667f + r2 * (8.33333333e-3f + r2 * (-1.98412698e-4f + r2 * 2.75573192e-6f)));
  const float pc =
      1.0f + r2 * (-0.5f + r2 * (4.16666667e-2f + r2 * (-1.38888889e-3f + r2 * (2.48015873e-5f + r2 * -2.75573192e-7f))));
  switch (k & 3) {
    case 0: *s = ps; *c = pc; break;
    case 1: *s = pc; *c = -ps; break;
    case 2: *s = -ps; *c = -pc; break;
    default: *s = -pc; *c = ps; break;
  }
}

// normalVector, wavefront.cl:235-245
__device__ __forceinline__ V3 normal_vector(int16_t pitch, int16_t yaw) {
  const float fyaw = ((float)yaw * kPiF) / 32767.0f, fpitch = ((float)pitch * kPiF) / 32767.0f;
  float sy, cy, sp, cp;
  sincos_poly(fyaw, &sy, &cy);
  sincos_poly(fpitch, &sp, &cp);
  return {cy * cp, sy * cp, sp};
}

// traverseChunkNormal, wavefront.cl:500-561: one cell of level depth+1 along the normal,
// to the farthest boundary (ties x, then y).
__device__ __forceinline__ void step_off_surface(uint32_t depth, V3 n, uint64_t &x, uint64_t &y, uint64_t &z) {
  const bool sx = n.x >= 0.0f, sy = n.y >= 0.0f, sz = n.z >= 0.0f;
  const float ix = 1.0f / n.x, iy = 1.0f / n.y, iz = 1.0f / n.z;
  uint64_t d0 = round_ulong(x, depth, sx), d1 = round_ulong(y, depth, sy), d2 = round_ulong(z, depth, sz);
  const float ax = fabsf((float)d0 * ix), ay = fabsf((float)d1 * iy), az = fabsf((float)d2 * iz);
  if (ax >= ay && ax >= az) {
    const float u = ulong_to_float(d0);
    d1 = float_to_ulong((n.y * ix) * u);
    d2 = float_to_ulong((n.z * ix) * u);
  } else if (ay >= ax && ay >= az) {
    const float u = ulong_to_float(d1);
    d0 = float_to_ulong((n.x * iy) * u);
    d2 = float_to_ulong((n.z * iy) * u);
  } else {
    const float u = ulong_to_float(d2);
    d0 = float_to_ulong((n.x * iz) * u);
    d1 = float_to_ulong((n.y * iz) * u);
  }
  x = advance_axis(x, d0, sx);
  y = advance_axis(y, d1, sy);
  z = advance_axis(z, d2, sz);
}

// Rotation of v about `axis` by the angle whose cos/sin are c/s (the reference's 12-entry
// table, written out as it is evaluated there).  `mirror` is the second ring vector, the
// transpose of the same rotation.
__device__ __forceinline__ V3 rotate(V3 axis, float c, float s, V3 v, bool mirror) {
  const float r2 = 1.0f - c;
  const float r3 = (r2 * axis.x) * axis.x, r4 = (r2 * axis.x) * axis.y, r5 = (r2 * axis.x) * axis.z;
  const float r6 = (r2 * axis.y) * axis.y, r7 = (r2 * axis.y) * axis.z, r8 = (r2 * axis.z) * axis.z;
  const float r9 = s * axis.x, r10 = s * axis.y, r11 = s * axis.z;
  if (!mirror)
    return {((r3 + c) * v.x + (r4 - r11) * v.y) + (r5 + r10) * v.z,
            ((r4 + r11) * v.x + (r6 + c) * v.y) + (r7 - r9) * v.z,
            ((r5 - r10) * v.x + (r7 + r9) * v.y) + (r8 + c) * v.z};
  return {((r3 + c) * v.x + (r4 + r11) * v.y) + (r5 - r10) * v.z,
          ((r4 - r11) * v.x + (r6 + c) * v.y) + (r7 + r9) * v.z,
          ((r5 + r10) * v.x + (r7 - r9) * v.y) + (r8 + c) * v.z};
}

struct Entry {
  bool valid;
  uint32_t trace, parent, addr;
  uint4 block;
  V3 direction;
  uint64_t x, y, z;
  uint32_t depth;
  float base_w;
};

__device__ __forceinline__ Entry load_entry(const MaterialParams &p, uint32_t i) {
  Entry e{};
  e.valid = false;
  const uint32_t id = p.queue[i];
  if (id >= p.n_root) return e;
  e.trace = id;
  e.parent = p.root_parents[id];
  const float4 d = p.root_directions[id];
  e.direction = {d.x, d.y, d.z};
  e.x = p.root_locations[4 * (size_t)id];
  e.y = p.root_locations[4 * (size_t)id + 1];
  e.z = p.root_locations[4 * (size_t)id + 2];
  e.depth = p.root_depths[id];
  e.base_w = (float)p.root_weightings[id];
  if (e.depth < p.B + 2u || e.depth > kMaxDepth) return e;
  uint32_t a = morton_top(p.B + 2u, e.x, e.y, e.z);
  for (uint32_t dd = p.B + 2u; dd < e.depth && a < p.n_blocks; dd++)
    a = block_child(load_block(p.blocks, p.n_blocks, a)) + chunk_position(dd, e.x, e.y, e.z);
  if (a >= p.n_blocks) return e;  // getBlockAddress found no block (UINT_MAX)
  e.addr = a;
  e.block = load_block(p.blocks, p.n_blocks, a);
  e.valid = true;
  return e;
}

// Runs the reference's cone loops for one entry.  emit(direction, fov, weighting) is called
// once per saved cone, in order; returns the number of cones.
template <typename Emit>
__device__ uint32_t material_cones(const MaterialParams &p, const Entry &e, V3 normal, bool *low_weight, Emit emit) {
  *low_weight = false;
  float fov = p.fov_constant, oldfov = fov, weighting = p.weighting_constant, ring_w = weighting;
  float spine = 0.0f;
  if (weighting * e.base_w < 6.0f) {
    *low_weight = true;
    return 0u;
  }
  uint32_t count = 0;
  const V3 d = e.direction;
  if (dot3(d, normal) >= 0.0f) {
    emit(d, fov, to_uchar(e.base_w));
    return 1u;
  }
  const V3 spine_axis = normalize3(cross3(d, normal));
  const float dn = 2.0f * dot3(d, normal);
  const V3 refl = {d.x - dn * normal.x, d.y - dn * normal.y, d.z - dn * normal.z};
  float dp = dot3(refl, normal);
  if (dp < fov) ring_w = (weighting * dp) / fov;
  emit(refl, fov, to_uchar(ring_w * e.base_w));
  count++;
  for (uint32_t k = 0; spine < kPiF && k < kMaxSpine; k++) {
    fov = (p.fov_multiplier * (spine + fov) + p.fov_constant) / (1.0f - p.fov_multiplier);
    spine = ((1.0f + p.fov_multiplier) * (spine + oldfov)) / (1.0f - p.fov_multiplier);
    oldfov = fov;
    weighting = spine * p.weighting_multiplier + p.weighting_constant;
    ring_w = weighting;
    float ss, cs;
    sincos_poly(0.0f - spine, &ss, &cs);
    const V3 seed = normalize3(rotate(spine_axis, cs, ss, refl, false));
    dp = dot3(seed, normal);
    if (dp < 0.0f) break;
    if (dp < fov) ring_w = (weighting * dp) / fov;
    emit(seed, fov, to_uchar(ring_w * e.base_w));
    count++;
    for (uint32_t j = 1; (float)j * fov < kHalfPiF && j <= kMaxRing; j++) {
      ring_w = weighting;
      const float ring = (float)(j * 2u) * fov;
      float sr, cr;
      sincos_poly(0.0f - ring, &sr, &cr);
      V3 rv = normalize3(rotate(refl, cr, sr, seed, false));
      dp = dot3(rv, normal);
      if (dp < 0.0f) break;
      if (dp < fov) ring_w = (weighting * dp) / fov;
      emit(rv, fov, to_uchar(ring_w * e.base_w));
      rv = normalize3(rotate(refl, cr, sr, seed, true));
      emit(rv, fov, to_uchar(ring_w * e.base_w));
      count += 2u;
    }
  }
  return count;
}

__global__ __launch_bounds__(kWg) void material_count_kernel(MaterialParams p, uint32_t *counts) {
  const uint32_t i = blockIdx.x * kWg + threadIdx.x;
  if (i >= p.n) return;
  const Entry e = load_entry(p, i);
  uint32_t c = 0;
  if (e.valid) {
    bool low;
    c = material_cones(p, e, normal_vector(block_pitch(e.block), block_yaw(e.block)), &low,
                       [](V3, float, uint32_t) {});
  }
  counts[i] = c;
}

__global__ __launch_bounds__(kWg) void material_write_kernel(MaterialParams p, const uint32_t *first) {
  const uint32_t i = blockIdx.x * kWg + threadIdx.x;
  if (i >= p.n) return;
  const Entry e = load_entry(p, i);
  if (!e.valid) return;
  // updateUsage, wavefront.cl:377-381
  const uint32_t g = e.addr >> 3;
  if (p.usage16 && g < p.n_usages) {
    const uint16_t t = p.usage16[4u * (size_t)g];
    if (t < 65534u && t != (uint16_t)p.tick) p.usage16[4u * (size_t)g] = (uint16_t)p.tick;
  }
  const V3 normal = normal_vector(block_pitch(e.block), block_yaw(e.block));
  const uint32_t base = *p.acc_counter + first[i];
  const uint32_t R = block_r(e.block), G = block_g(e.block), Bc = block_b(e.block);
  // setupTrace's location: one cell off the surface along the normal (the same for every cone)
  uint64_t lx = e.x, ly = e.y, lz = e.z;
  step_off_surface(e.depth, normal, lx, ly, lz);
  const float4 origin = make_float4(ulong_to_float(e.x), ulong_to_float(e.y), ulong_to_float(e.z), 0.0f);
  uint32_t k = 0;
  bool low;
  material_cones(p, e, normal, &low, [&](V3 dir, float fov, uint32_t w) {
    const uint32_t id = base + k;
    const uint32_t rel = first[i] + k;
    k++;
    if (rel < p.flag_capacity) p.bg_flags[rel] = (w > 5u && id < p.n_out) ? 1u : 0u;
    if (id >= p.n_out) return;  // beyond the trace arrays: counted, not stored or queued
    p.parents[id] = e.parent;
    p.weightings[id] = (uint8_t)w;
    if (w > 5u) {
      p.luminosities[id] = 0;
      p.colour_r[id] = (uint8_t)R;
      p.colour_g[id] = (uint8_t)G;
      p.colour_b[id] = (uint8_t)Bc;
      p.origins[id] = origin;
      p.directions[id] = make_float4(dir.x, dir.y, dir.z, 0.0f);
      p.fovs[id] = fov;
      p.locations[4 * (size_t)id] = lx;
      p.locations[4 * (size_t)id + 1] = ly;
      p.locations[4 * (size_t)id + 2] = lz;
      p.depths[id] = 1;
    }
  });
  if (low && p.final_w && e.trace < p.n_final) p.final_w[e.trace] = (uint32_t)e.base_w;
}

// BackgroundQueue appends of the cones with weighting > 5, in ascending trace id.
constexpr uint32_t kChunk = 1024, kPer = kChunk / kWg;

__device__ uint32_t wg_scan(uint32_t v, uint32_t *s_waves, uint32_t *total) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t incl = v;
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(incl, d, 64);
    if ((int)lane >= d) incl += o;
  }
  if (lane == 63u) s_waves[wave] = incl;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (uint32_t w = 0; w < (uint32_t)(kWg / 64); w++) {
    const uint32_t tw = s_waves[w];
    before += w < wave ? tw : 0u;
    all += tw;
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

__global__ __launch_bounds__(kWg) void chunk_sum_kernel(const uint32_t *in, uint32_t n, uint32_t *sums) {
  __shared__ uint32_t s_waves[kWg / 64];
  const uint32_t i0 = blockIdx.x * kChunk + threadIdx.x * kPer;
  uint32_t c = 0;
  for (uint32_t j = 0; j < kPer; j++) c += i0 + j < n ? in[i0 + j] : 0u;
  uint32_t all;
  wg_scan(c, s_waves, &all);
  if (threadIdx.x == 0) sums[blockIdx.x] = all;
}

// sums: chunk totals -> exclusive chunk offsets, total at sums[chunks] (one work-group)
__global__ __launch_bounds__(kWg) void chunk_offsets_kernel(uint32_t *sums, uint32_t chunks) {
  __shared__ uint32_t s_waves[kWg / 64];
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < chunks; c0 += kWg) {
    const uint32_t i = c0 + threadIdx.x;
    const uint32_t v = i < chunks ? sums[i] : 0u;
    uint32_t all;
    const uint32_t pre = wg_scan(v, s_waves, &all);
    if (i < chunks) sums[i] = carry + pre;
    carry += all;
  }
  if (threadIdx.x == 0) sums[chunks] = carry;
}

__global__ __launch_bounds__(kWg) void chunk_apply_kernel(const uint32_t *in, uint32_t n, const uint32_t *sums,
                                                          uint32_t *out) {
  __shared__ uint32_t s_waves[kWg / 64];
  const uint32_t i0 = blockIdx.x * kChunk + threadIdx.x * kPer;
  uint32_t v[kPer], c = 0;
  for (uint32_t j = 0; j < kPer; j++) {
    v[j] = i0 + j < n ? in[i0 + j] : 0u;
    c += v[j];
  }
  uint32_t all;
  uint32_t r = sums[blockIdx.x] + wg_scan(c, s_waves, &all);
  for (uint32_t j = 0; j < kPer; j++) {
    if (i0 + j < n) out[i0 + j] = r;
    r += v[j];
  }
}

__global__ __launch_bounds__(kWg) void flag_append_kernel(const uint32_t *flags, uint32_t n, const uint32_t *offsets,
                                                          const uint32_t *acc_counter, const uint32_t *bg_counter,
                                                          uint32_t *queue, uint32_t capacity) {
  __shared__ uint32_t s_waves[kWg / 64];
  const uint32_t i0 = blockIdx.x * kChunk + threadIdx.x * kPer;
  uint32_t f[kPer], c = 0;
  for (uint32_t j = 0; j < kPer; j++) {
    f[j] = i0 + j < n ? flags[i0 + j] : 0u;
    c += f[j];
  }
  uint32_t all;
  uint32_t r = *bg_counter + offsets[blockIdx.x] + wg_scan(c, s_waves, &all);
  const uint32_t base = *acc_counter;
  for (uint32_t j = 0; j < kPer; j++) {
    if (f[j] && r < capacity) queue[r] = base + i0 + j;
    r += f[j];
  }
}

// Counters last, once every pass has read them: AccumulatorID += cones, BackgroundQueueID +=
// appended cones.
__global__ void material_counters_kernel(uint32_t *acc_counter, uint32_t *bg_counter, const uint32_t *cone_total,
                                         const uint32_t *bg_total) {
  if (threadIdx.x == 0) {
    *acc_counter += *cone_total;
    if (bg_counter) *bg_counter += *bg_total;
  }
}

}  // namespace

uint32_t material_chunks(uint32_t n) { return (n + kChunk - 1) / kChunk; }

int launch_material_count(const MaterialParams &p, uint32_t *counts, hipStream_t s) {
  if (!p.n) return 0;
  hipLaunchKernelGGL(material_count_kernel, dim3((p.n + kWg - 1) / kWg), dim3(kWg), 0, s, p, counts);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_material_write(const MaterialParams &p, const uint32_t *first, hipStream_t s) {
  if (!p.n) return 0;
  hipLaunchKernelGGL(material_write_kernel, dim3((p.n + kWg - 1) / kWg), dim3(kWg), 0, s, p, first);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_material_scan(const uint32_t *in, uint32_t n, uint32_t *out, uint32_t *sums, hipStream_t s) {
  const uint32_t chunks = material_chunks(n);
  if (chunks) {
    hipLaunchKernelGGL(chunk_sum_kernel, dim3(chunks), dim3(kWg), 0, s, in, n, sums);
    if (hipGetLastError() != hipSuccess) return -1;
  }
  hipLaunchKernelGGL(chunk_offsets_kernel, dim3(1), dim3(kWg), 0, s, sums, chunks);
  if (hipGetLastError() != hipSuccess) return -1;
  if (!chunks || !out) return 0;
  hipLaunchKernelGGL(chunk_apply_kernel, dim3(chunks), dim3(kWg), 0, s, in, n, sums, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_material_append(const uint32_t *flags, uint32_t n, const uint32_t *offsets, const uint32_t *acc_counter,
                           const uint32_t *bg_counter, uint32_t *queue, uint32_t capacity, hipStream_t s) {
  const uint32_t chunks = material_chunks(n);
  if (!chunks) return 0;
  hipLaunchKernelGGL(flag_append_kernel, dim3(chunks), dim3(kWg), 0, s, flags, n, offsets, acc_counter, bg_counter,
                     queue, capacity);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_material_counters(uint32_t *acc_counter, uint32_t *bg_counter, const uint32_t *cone_total,
                             const uint32_t *bg_total, hipStream_t s) {
  hipLaunchKernelGGL(material_counters_kernel, dim3(1), dim3(64), 0, s, acc_counter, bg_counter, cone_total,
                     bg_total);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace svo
