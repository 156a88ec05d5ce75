// Fused primary-frame cone tracer for gfx950 (MI355X).
//
// One persistent kernel replaces the reference's host-driven wavefront loop
// (MainWindow.RunWavefront/RunTraces, MainWindow.cs:356-481) and its kernels
// init / runBaseTrace / runBlockTrace / evaluateBackground / resolveRemainders /
// drawTrace (wavefront.cl:1352-1850).  Each lane owns one cone and runs the
// reference step machine to completion; because a cone's steps depend only on
// its own state and the read-only tree, the per-cone result equals the
// wavefront schedule's.
//
// MI355X mapping:
//   - persistent workgroups; idle lanes are refilled from a work counter with one
//     wave-aggregated atomic (ballot + popcount + mbcnt);
//   - 8 work ranges (bands of 8x8 pixel tiles), range r served first by the
//     workgroups of XCD r (blockIdx % 8), then stolen by the others, so each
//     XCD's L2 sees a spatially coherent part of the image;
//   - the dense Bases levels live in LDS; block records are fetched with one
//     128-bit load; instead of re-walking from the dense level every step
//     (getBlockAddress, wavefront.cl:296-304) each lane keeps a short stack of
//     block addresses in LDS ([level][lane], conflict-free);
//   - per-cone outputs are written by cone id; queue order is made canonical
//     (ascending cone id) by the order-preserving compaction kernels below.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "svo_device.hpp"
#include "svo_kernels.hpp"

namespace svo {

using namespace dev;

namespace {

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// WAVES > 0 asks the compiler for that many waves per SIMD (a register budget).
// HASH: the visit-sequence hash is computed only when its output is bound.
// Camera parameters come from the table p.cams (one camera or a batch), staged in LDS:
// kernel arguments stay small, so the scalar registers are not spilled.
// MODE: kModePlain ends a cone at its first material hit (the reference's frame);
// kModeComposite makes every hit a layer composited front to back and continues past the
// cell (up to p.max_layers layers); kModeAdjoint also back-propagates the L2 image loss to
// the layers' colour and opacity.
template <int WG, int WAVES, bool HASH, int MODE = kModePlain>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(WAVES > 0 ? WAVES : 1, 8)))
trace_frame_kernel(const FrameParams p) {
  extern __shared__ uint32_t smem[];
  uint16_t *s_bases = reinterpret_cast<uint16_t *>(smem);
  const uint32_t bases_words = p.bases_in_lds ? (p.n_bases + 1u) / 2u : 0u;
  uint32_t *s_cam = smem + bases_words;
  const uint32_t cam_words = p.n_cams * kCamWords;
  uint32_t *s_ticked = s_cam + cam_words;
  uint32_t *s_stack = s_ticked + p.ticked_slots;
  if (p.bases_in_lds) {
    for (uint32_t i = threadIdx.x; i < p.n_bases; i += WG) s_bases[i] = p.bases[i];
  }
  {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(p.cams);
    for (uint32_t i = threadIdx.x; i < cam_words; i += WG) s_cam[i] = src[i];
  }
  for (uint32_t i = threadIdx.x; i < p.ticked_slots; i += WG) s_ticked[i] = kU32Max;
  __syncthreads();

  const uint32_t tid = threadIdx.x, lane = tid & 63u;
  const uint32_t B = p.B, S = p.stack_levels;
  const uint32_t xcd = blockIdx.x & 7u;
  // camera parameters: word w of camera c
  auto camf = [&](uint32_t c, uint32_t w) -> float { return __uint_as_float(s_cam[c * kCamWords + w]); };
  auto camu = [&](uint32_t c, uint32_t w) -> uint32_t { return s_cam[c * kCamWords + w]; };
  uint32_t cam = 0;  // the lane's camera
  // the frame tick travels with the camera records (word 21 of camera 0), so a captured
  // launch can be replayed for later frames
  const uint32_t tick = camu(0, 21);

  uint32_t range_step = 0;  // wave-uniform: ranges tried, in order xcd, xcd+1, ...
  bool fetching = true;     // wave-uniform
  uint32_t cur = 0, cur_end = 0, cur_t0 = 0;  // wave-uniform: the wave's claimed slice of work
  // the tile the wave took at its last refill (wave-uniform) and the steps each lane ran since
  uint32_t cur_tile = kU32Max, age = 0;
  auto record_cost = [&]() {
    uint32_t a = age;  // iterations a lane ran on the tile; the tile took the longest of them
    for (int off = 32; off > 0; off >>= 1) a = max(a, (uint32_t)__shfl_xor((int)a, off));
    if (p.tile_cost && cur_tile != kU32Max && lane == 0u) p.tile_cost[cur_tile] = a;
    cur_tile = kU32Max;
    age = 0;
  };

  bool active = false;
  uint32_t ray = 0, depth = 0, signs = 0, node_addr = 0, last_group = kU32Max, nvis = 0, nblk = 0;
  // the location in sign-canonical form (c = x on an axis the cone moves down, ~x on one it
  // moves up: every axis then moves towards 0); loc(a) gives the location itself
  uint64_t cx = 0, cy = 0, cz = 0, hash = kFnvOffset;
  float dx = 0.f, dy = 0.f, dz = 0.f, ix = 0.f, iy = 0.f, iz = 0.f;
  uint32_t nchild = 0, nchunk = 0;  // Child field and chunk of the current block (node_addr)
  // Morton code of the top min(depth, B) bits of the location: the level's Bases index is
  // powSum(depth - 1) + bm, kept up to date as the cone descends and pops instead of being
  // rebuilt from the location at every base step
  uint32_t bm = 0;
  float cone = 0.f;                // cone footprint at the current location, once cone_ok
  bool cone_ok = false;
  uint32_t tot_vis = 0, tot_blk = 0;
  // updateUsage is resolved one step late so its load never stalls the step.  Within a
  // launch every write stores the same frame tick, so a group needs its tick read (and, if
  // it is not 65534/65535 and differs, written) once: neither the deferral nor skipping a
  // group already handled changes the result.  A group counts as handled when this cone
  // visited it before (a block reached again by a pop) or when the workgroup's table of
  // ticked groups (s_ticked, direct-mapped) holds it; the lane that put it there writes it.
  uint32_t pend_g = kU32Max, pend_t = 0;
  auto resolve_usage = [&]() {
    if (pend_g != kU32Max && pend_t < 65534u && pend_t != (tick & 0xffffu))
      p.usage16[4u * (size_t)pend_g] = (uint16_t)tick;
    pend_g = kU32Max;
  };

  // The short stack holds, for each block level from B+2 down, the block the cone passed
  // through: its address, its Child field and its chunk (three [level][lane] arrays, so a
  // pop to a stacked level needs no memory access).
  uint32_t *st_child = s_stack + S * WG;
  uint16_t *st_chunk = reinterpret_cast<uint16_t *>(s_stack + 2u * S * WG);
  auto stack_put = [&](uint32_t level) {
    const uint32_t rel = level - (B + 2u);
    if (rel < S) {
      s_stack[rel * WG + tid] = node_addr;
      st_child[rel * WG + tid] = nchild;
      st_chunk[rel * WG + tid] = (uint16_t)nchunk;
    }
  };
  // The block of `level` (a level the cone is still inside): address, Child field, chunk.
  auto stack_get = [&](uint32_t level) -> uint3 {
    const uint32_t rel = level - (B + 2u);
    if (rel < S) return make_uint3(s_stack[rel * WG + tid], st_child[rel * WG + tid], st_chunk[rel * WG + tid]);
    // deeper than the stack: walk down from its last level along the current location
    uint32_t a = 0, ch = st_child[(S - 1u) * WG + tid], ck = 0;
    for (uint32_t L = B + 2u + S - 1u; L < level; L++) {
      a = ch + (chunk_position(L, cx, cy, cz) ^ signs);
      const uint2 q = load_node(p.blocks, p.n_blocks, a);
      ch = q.x;
      ck = q.y & 0xffffu;
    }
    return make_uint3(a, ch, ck);
  };
  // Descends into the block at node_addr: fetches its Child field and chunk (one 64-bit
  // load) and pushes it.
  auto enter = [&](uint32_t level) {
    const uint2 q = load_node(p.blocks, p.n_blocks, node_addr);
    nchild = q.x;
    nchunk = q.y & 0xffffu;
    stack_put(level);
  };

  // Ends the lane's cone: evaluateBackground / resolveRemainders / drawTrace for a primary
  // cone (weighting 255, parent = itself), plus the per-cone extension outputs.
  auto finish = [&](bool material, uint32_t addr, bool child, bool tree_size, bool capped) {
    float4 c;
    if (!material) {
      const float dp = (dx * 0.0f + dy * 0.0f) + dz * -1.0f;
      uint32_t r, g, b;
      if (dp > 0.8f) {
        r = g = b = 255u;
      } else {
        r = 135u;
        g = 206u;
        b = 235u;
      }
      const float fw = 255.0f * 255.0f;
      c = make_float4(fminf((float)(r * 255u) / fw, 1.0f), fminf((float)(g * 255u) / fw, 1.0f),
                      fminf((float)(b * 255u) / fw, 1.0f), 1.0f);
      addr = kU32Max;
    } else if (addr == kU32Max) {
      c = make_float4(0.75f, 0.0f, 0.0f, 1.0f);
    } else {
      const uint4 q = load_block(p.blocks, p.n_blocks, addr);
      const float fw = 255.0f * 255.0f;
      c = make_float4(fminf((float)(block_r(q) * 255u) / fw, 1.0f), fminf((float)(block_g(q) * 255u) / fw, 1.0f),
                      fminf((float)(block_b(q) * 255u) / fw, 1.0f), 1.0f);
    }
    if (p.rgba) p.rgba[ray] = c;
    if (p.hit) p.hit[ray] = addr;
    if (p.visits) p.visits[ray] = nvis;
    if (HASH) p.hashes[ray] = hash;
    if (p.status) p.status[ray] = (uint8_t)((material ? 1u : 0u) | (child ? 2u : 0u) | (capped ? 4u : 0u));
    if (p.depth_out) p.depth_out[ray] = (uint8_t)depth;
    if (p.loc_out) {
      p.loc_out[4 * (size_t)ray + 0] = cx ^ sign_mask64(signs, 0u);
      p.loc_out[4 * (size_t)ray + 1] = cy ^ sign_mask64(signs, 1u);
      p.loc_out[4 * (size_t)ray + 2] = cz ^ sign_mask64(signs, 2u);
    }
    if (child && p.req_sparse) {
      svo_child_request rq;
      rq.address = addr;
      rq.tick = (uint16_t)tick;
      rq.depth = (uint8_t)depth;
      rq.pad0 = 0;
      rq.location[0] = cx ^ sign_mask64(signs, 0u);
      rq.location[1] = cy ^ sign_mask64(signs, 1u);
      rq.location[2] = cz ^ sign_mask64(signs, 2u);
      rq.tree_size = tree_size ? 1 : 0;
      for (int k = 0; k < 7; k++) rq.pad1[k] = 0;
      p.req_sparse[ray] = rq;
      if (p.req_keys) p.req_keys[ray] = make_uint2(addr, (uint32_t)depth | (tree_size ? 256u : 0u));
    }
    resolve_usage();
    tot_vis += nvis;
    tot_blk += nblk;
    active = false;
  };

  // A cone that ends only records how (`fin`: bit0 pending, bit1 material, bit2 child
  // request, bit3 TreeSize, bit4 step cap; the block address in node_addr); its outputs are written when the wave next refills or exits, so the shading
  // and output code runs once for all lanes that ended in between instead of once per
  // step in which any lane ended.
  uint32_t fin = 0;
  auto retire = [&](bool material, uint32_t addr, bool child, bool tree_size, bool capped) {
    node_addr = addr;
    fin = 1u | (material ? 2u : 0u) | (child ? 4u : 0u) | (tree_size ? 8u : 0u) | (capped ? 16u : 0u);
    active = false;
  };
  auto flush = [&]() {
    if (fin) {
      finish((fin & 2u) != 0u, node_addr, (fin & 4u) != 0u, (fin & 8u) != 0u, (fin & 16u) != 0u);
      fin = 0;
    }
  };

  // ---- compositing state (composite modes) ----
  float cT = 1.0f, cR = 0.0f, cG = 0.0f, cB = 0.0f;  // transmittance, colour so far
  uint32_t nlay = 0, lay0 = kU32Max, lay1 = kU32Max, lay2 = kU32Max, lay3 = kU32Max;
  double loss_acc = 0.0;
  auto leaf = [&](uint32_t a) -> float4 {
    return a < p.n_blocks ? p.leaf_params[a] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  };
  auto layer_at = [&](uint32_t i) -> uint32_t { return i == 0 ? lay0 : i == 1 ? lay1 : i == 2 ? lay2 : lay3; };
  // one layer: C += T a c, T *= 1 - a
  auto composite = [&](uint32_t a) {
    const float4 q = leaf(a);
    const float w = cT * q.w;
    cR += w * q.x;
    cG += w * q.y;
    cB += w * q.z;
    cT = cT * (1.0f - q.w);
    if (nlay == 0) lay0 = a;
    else if (nlay == 1) lay1 = a;
    else if (nlay == 2) lay2 = a;
    else lay3 = a;
    nlay++;
  };
  // The cone ends: the background takes the remaining transmittance; image, layers, loss
  // and (adjoint) the gradient of every layer.
  auto end_composite = [&]() {
    const float dp = (dx * 0.0f + dy * 0.0f) + dz * -1.0f;
    const bool sun = dp > 0.8f;
    const float bg0 = sun ? 1.0f : 135.0f / 255.0f, bg1 = sun ? 1.0f : 206.0f / 255.0f,
                bg2 = sun ? 1.0f : 235.0f / 255.0f;
    const float C0 = cR + cT * bg0, C1 = cG + cT * bg1, C2 = cB + cT * bg2;
    if (p.rgba) p.rgba[ray] = make_float4(C0, C1, C2, 1.0f);
    if (p.layers_out)
      for (uint32_t i = 0; i < p.max_layers; i++) p.layers_out[(size_t)ray * p.max_layers + i] = i < nlay ? layer_at(i) : kU32Max;
    if (p.target) {
      const float4 t = p.target[ray];
      const float e0 = C0 - t.x, e1 = C1 - t.y, e2 = C2 - t.z;
      loss_acc += (double)e0 * e0 + (double)e1 * e1 + (double)e2 * e2;
      if (MODE == kModeAdjoint) {
        const float g0 = 2.0f * e0, g1 = 2.0f * e1, g2 = 2.0f * e2;
        // transmittance in front of each layer, then the composite behind it (back to front)
        float Ti[kMaxLayers];
        float T = 1.0f;
        for (uint32_t i = 0; i < kMaxLayers; i++) {
          Ti[i] = T;
          if (i < nlay) T = T * (1.0f - leaf(layer_at(i)).w);
        }
        float U0 = bg0, U1 = bg1, U2 = bg2;
        for (int i = (int)kMaxLayers - 1; i >= 0; i--) {
          if ((uint32_t)i >= nlay) continue;
          const uint32_t a = layer_at((uint32_t)i);
          const float4 q = leaf(a);
          if (a < p.n_blocks) {
            float *gr = reinterpret_cast<float *>(p.grads + a);
            const float w = Ti[i] * q.w;
            atomicAdd(gr + 0, g0 * w);
            atomicAdd(gr + 1, g1 * w);
            atomicAdd(gr + 2, g2 * w);
            atomicAdd(gr + 3, Ti[i] * ((g0 * (q.x - U0) + g1 * (q.y - U1)) + g2 * (q.z - U2)));
          }
          U0 = q.w * q.x + (1.0f - q.w) * U0;
          U1 = q.w * q.y + (1.0f - q.w) * U1;
          U2 = q.w * q.z + (1.0f - q.w) * U2;
        }
      }
    }
    resolve_usage();
    tot_vis += nvis;
    tot_blk += nblk;
    active = false;
  };

  // One step's visit (runBaseTrace / runBlockTrace up to the octant test, wavefront.cl:1400-1429,
  // 1465-1505): counts the step, ticks the block's usage group (updateUsage, wavefront.cl:377-381)
  // and tests the octant.  Returns 0 when it holds geometry (the step decides), 1 when it is
  // empty (the step advances), 2 when the cone has reached the step cap and ended.
  bool pend_adv = false;  // the current step was visited and advances
  auto visit = [&](uint32_t &oct_out) -> int {
    if (nvis >= kStepCap) {
      depth = 0;
      if (MODE == kModePlain)
        retire(false, kU32Max, false, false, true);
      else
        end_composite();
      return 2;
    }
    const uint32_t d = depth;
    const bool is_block = d > B;
    uint32_t index, chunk;
    if (!is_block) {
      // powSum(d - 1) = 8 + 64 + ... + 8^(d-1): octal 1...10, i.e. bits 3, 6, ..., 3(d-1)
      index = (0x249248u & ((1u << (3u * d)) - 1u)) + bm;
      if (p.bases_in_lds) {
        chunk = s_bases[index];
      } else {
        // a 32-bit read of the chunk pair (Bases holds an even number of chunks), so that
        // the two paths stay an LDS read and a global read
        const uint32_t w = reinterpret_cast<const uint32_t *>(p.bases)[index >> 1];
        chunk = (w >> (16u * (index & 1u))) & 0xffffu;
      }
    } else {
      index = node_addr;
      chunk = nchunk;
      const uint32_t g = index >> 3;
      if (g != last_group) {
        last_group = g;
        resolve_usage();
        if (g < p.n_usages) {
          const uint32_t slot = g & (p.ticked_slots - 1u);
          if (p.ticked_slots == 0u || s_ticked[slot] != g) {
            if (p.ticked_slots) s_ticked[slot] = g;
            pend_g = g;
            pend_t = p.usage16[4u * (size_t)g];
          }
        }
      }
    }
    const uint32_t o = chunk_position(d, cx, cy, cz) ^ signs;
    if (HASH) hash = (hash ^ (((uint64_t)d << 32) | index)) * kFnvPrime;
    nvis++;
    nblk += is_block ? 1u : 0u;
    oct_out = o;
    return ((chunk >> (2u * o)) & 2u) == 0u ? 1 : 0;
  };

  for (;;) {
    // ---- refill idle lanes with new cones (wave-uniform control flow) ----
    const uint64_t idle = __ballot(!active);
    const uint32_t n_idle = (uint32_t)__popcll(idle);
    if (fetching && (n_idle >= p.refill || n_idle == 64u)) {
      flush();
      record_cost();
      // claim a slice of `chunk` work items with one atomic when the wave's slice is used up
      while (cur >= cur_end && range_step < 8u) {
        const uint32_t r = (xcd + range_step) & 7u;
        const uint32_t limit = (p.range_begin[r + 1] - p.range_begin[r]) * 64u;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&p.counters[r * kCounterStride], p.chunk);
        base = __shfl(base, 0);
        if (base < limit) {
          cur = base;
          cur_end = min(base + p.chunk, limit);
          cur_t0 = p.range_begin[r];
        } else {
          range_step++;
        }
      }
      if (cur >= cur_end) {
        fetching = false;
      } else {
        const uint32_t k = cur + lanes_below(idle);
        const uint32_t take = min(n_idle, cur_end - cur);
        cur_tile = p.order ? p.order[cur_t0 + (cur >> 6)] : cur_t0 + (cur >> 6);
        cur += take;
        if (!active && k < cur_end) {
          uint32_t tile = cur_t0 + (k >> 6);
          if (p.order) tile = p.order[tile];
          const uint32_t l = k & 63u;
          // the camera whose tiles hold this one: the last with tile_begin <= tile
          uint32_t lo = 0, hi = p.n_cams;
          while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (camu(mid, 18) <= tile)
              lo = mid;
            else
              hi = mid;
          }
          cam = lo;
          tile -= camu(cam, 18);
          const uint32_t W = camu(cam, 15), H = camu(cam, 16), tiles_x = camu(cam, 17), ray0 = camu(cam, 19),
                         lut0 = camu(cam, 20);
          float f[9];
          for (int j = 0; j < 9; j++) f[j] = camf(cam, j);
          const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
          const uint32_t x = tx * 8u + (l & 7u), y = ty * 8u + (l >> 3);
          if (x < W && y < H) {
            // setupInitialTrace, wavefront.cl:563-595 (trig from the camera table)
            ray = ray0 + x + y * W;
            const float *lut = p.lut + lut0;
            const float cu = lut[x], su = lut[W + x];
            const float cv = lut[2u * W + y], sv = lut[2u * W + H + y];
            const float m0 = cu * cv, m1 = (0.0f - su) * cv, m2 = sv;
            dx = (f[0] * m0 + f[3] * m1) + f[6] * m2;
            dy = (f[1] * m0 + f[4] * m1) + f[7] * m2;
            dz = (f[2] * m0 + f[5] * m1) + f[8] * m2;
            ix = 1.0f / dx;
            iy = 1.0f / dy;
            iz = 1.0f / dz;
            signs = (dx >= 0.0f ? 1u : 0u) | (dy >= 0.0f ? 2u : 0u) | (dz >= 0.0f ? 4u : 0u);
            nvis = 0;
            nblk = 0;
            hash = kFnvOffset;
            last_group = kU32Max;
            cone_ok = false;
            pend_adv = false;
            if (MODE != kModePlain) {
              cT = 1.0f;
              cR = cG = cB = 0.0f;
              nlay = 0;
            }
            const float o[3] = {camf(cam, 9), camf(cam, 10), camf(cam, 11)};
            const float d[3] = {dx, dy, dz}, inv[3] = {ix, iy, iz};
            uint64_t ex, ey, ez;
            const bool entered = trace_into_volume(o, d, inv, camf(cam, 12), ex, ey, ez);
            if (!entered) ex = ey = ez = 0;
            cx = ex ^ sign_mask64(signs, 0u);
            cy = ey ^ sign_mask64(signs, 1u);
            cz = ez ^ sign_mask64(signs, 2u);
            bm = chunk_position(0u, ex, ey, ez);
            if (entered) {
              depth = 1;
              active = true;
            } else {
              depth = 0;
              if (MODE == kModePlain)
                retire(false, kU32Max, false, false, false);
              else
                end_composite();
            }
          }
        }
      }
    }
    if (__ballot(active) == 0ull) {
      if (!fetching) break;
      continue;
    }
    // Step the running lanes until enough lanes are idle to refill: a lane leaves this loop
    // when its cone ends, so the exec mask keeps finished lanes' registers untouched.
    // An iteration takes at most one advancing step and then one deciding step: a lane whose
    // step advances visits its next step at once (the block it needs is in registers or on
    // the short stack, never behind a load) and, when that step decides, takes it in the same
    // iteration; a step visited as advancing waits as pend_adv for the next iteration.  Lanes
    // then run the long advance and the decision together instead of in alternate iterations.
    while (active) {
      age++;
      bool adv = pend_adv, dec = false;
      uint32_t oct = 0;
      pend_adv = false;
      if (!adv) {
        const int v = visit(oct);
        adv = v == 1;
        dec = v == 0;
      }
      if (adv) {
        // advance to the next cell boundary, pop levels that were left
        const uint32_t dd = depth;
        const uint64_t px = cx, py = cy, pz = cz;
        traverse_canon(dd, cx, cy, cz, dx, dy, dz, ix, iy, iz, signs);
        cone_ok = false;
        const bool leave = leaving_canon(cx, cy, cz);
        uint32_t nd = compare_canon(dd, px, py, pz, cx, cy, cz);
        // a block level pops to the level it names, except that level B+1 (no blocks)
        // means the base level B; a pop to a shallower block level re-reads the stack
        const bool pop_block = !leave && dd > B && nd > B + 1u && nd < dd;
        nd = (dd > B && nd == B + 1u) ? B : nd;
        depth = leave ? 0u : nd;
        if (pop_block) {
          const uint3 q = stack_get(nd);
          node_addr = q.x;
          nchild = q.y;
          nchunk = q.z;
          last_group = q.x >> 3;  // ticked when the cone entered it
        }
        // a pop to a base level keeps the top nd bits: drop the Morton triples below them.
        // compare_canon never returns less than 1, so at nd == 1 the top bit itself may have
        // changed (the cone crossed the volume's middle plane): rebuild it from the location.
        if (nd <= B)
          bm = nd == 1u ? (chunk_position(0u, cx, cy, cz) ^ signs) : bm >> (3u * ((dd < B ? dd : B) - nd));
        if (leave) {
          if (MODE == kModePlain)
            retire(false, kU32Max, false, false, false);
          else
            end_composite();
        } else {
          // the next step: visited now; an advancing one waits for the next iteration
          const int v = visit(oct);
          pend_adv = v == 1;
          dec = v == 0;
        }
      }
      if (dec) {
        const uint32_t d = depth;
        if (d <= B) {
          // runBaseTrace's descent, wavefront.cl:1400-1463
          if (d < B) {
            depth = d + 1u;
            bm = (bm << 3) | oct;
          } else {
            // the dense level B+2: Morton code of the top B+2 bits (bm holds the top B)
            depth = B + 2u;
            node_addr = (((bm << 3) | oct) << 3) | (chunk_position(B + 1u, cx, cy, cz) ^ signs);
            enter(depth);
          }
        } else {
          // geometry in this octant: cone-footprint LOD decision, wavefront.cl:1516-1538;
          // the footprint depends on the location only, so it is computed once per location
          if (!cone_ok) {
            cone = cone_at(camf(cam, 9), camf(cam, 10), camf(cam, 11), cx ^ sign_mask64(signs, 0u),
                           cy ^ sign_mask64(signs, 1u), cz ^ sign_mask64(signs, 2u), camf(cam, 12),
                           camf(cam, 13), camf(cam, 14));
            cone_ok = true;
          }
          // the LOD rule's five outcomes: up to level B+2 and hit, up one level, request a
          // child (hit), down to the child, hit
          const bool up_top = cone > exp2_neg(B + 2u);
          const bool up_one = !up_top && cone > exp2_neg(d);
          const bool stay = !up_top && !up_one;
          const bool no_child = nchild == kU32Max;
          // requestChild, wavefront.cl:362-375; TreeSize = min(1, (uchar)ConeDepth - Depth)
          const bool child = stay && no_child;
          const bool tree_size = child && cone <= exp2_neg(d + 1u);
          const bool down = stay && !no_child && cone < exp2_neg(d + 1u) && d + 1u <= kMaxDepth;
          const bool hit = up_top || (stay && !down);
          depth = up_top ? B + 2u : (up_one ? d - 1u : (down ? d + 1u : d));
          if (up_top || up_one) {
            const uint3 q = stack_get(depth);
            node_addr = q.x;
            nchild = q.y;
            nchunk = q.z;
            last_group = q.x >> 3;  // ticked when the cone entered it
          } else if (down) {
            node_addr = nchild + oct;
            enter(depth);
          }
          if (hit) {
            if (MODE == kModePlain) {
              retire(true, node_addr, child, tree_size, false);
            } else {
              composite(node_addr);
              if (nlay >= p.max_layers)
                end_composite();
              else
                pend_adv = true;  // a composited layer: the cone continues past the cell
            }
          }
        }
      }
      const uint32_t running = (uint32_t)__popcll(__ballot(active));
      if (fetching && 64u - running >= p.refill) break;
    }
  }

  flush();
  resolve_usage();  // a tick still pending (every cone ending resolves it; kept for safety)
  record_cost();

  // frame totals: one 64-bit atomic per wave
  uint32_t vb = tot_vis - tot_blk, vk = tot_blk;
  for (int off = 32; off > 0; off >>= 1) {
    vb += __shfl_xor(vb, off);
    vk += __shfl_xor(vk, off);
  }
  if (MODE != kModePlain && p.loss) {
    for (int off = 32; off > 0; off >>= 1) loss_acc += __shfl_xor(loss_acc, off);
    if (lane == 0) atomicAdd(p.loss, loss_acc);
  }
  if (lane == 0) {
    atomicAdd(&p.totals[0], (unsigned long long)vb);
    atomicAdd(&p.totals[1], (unsigned long long)vk);
  }
}

// ---- tile schedule ----

// Cost class of a tile: floor(log2(iterations)) - 2, clamped to 0..kCostClasses-1 (tiles of
// fewer than 8 iterations, and unmeasured ones, are class 0; 512 and more, class 7).
constexpr uint32_t kCostClasses = 8;
__device__ __forceinline__ uint32_t cost_class(uint32_t c) {
  const uint32_t l = c ? 31u - (uint32_t)__clz(c) : 0u;
  return l < 3u ? 0u : min(kCostClasses - 1u, l - 2u);
}

constexpr uint32_t kOrderThreads = 1024;

// One work-group per band: a stable counting sort of the band's tiles by cost class, most
// expensive class first, so that the tiles that ran longest in the previous frame start
// first and the frame does not end waiting on an expensive tile taken late.  Thread t
// owns a contiguous run of the band's tiles.
// With cls (LDS of at least (n + 1) / 2 bytes) the band's cost classes are staged there
// first, eight loads a thread in flight together, and both passes read them from LDS; without
// it each pass reads the costs from memory, one load per tile in turn (bands of 70K tiles, as
// in a 64-camera batch, took 140 us that way).
__device__ __forceinline__ void order_band(const uint32_t *cost, uint32_t *order, uint32_t b0, uint32_t n,
                                           uint8_t *cls) {
  __shared__ uint32_t s_wave[kCostClasses][kOrderThreads / 64];
  __shared__ uint32_t s_base[kCostClasses];
  // staged classes: two tiles a byte (tile i in the low nibble of byte i/2 when i is even),
  // each byte written whole by one thread
  const uint32_t n_bytes = (n + 1u) / 2u;
  if (cls) {
    for (uint32_t j0 = 0; j0 < n_bytes; j0 += 4u * kOrderThreads) {
      uint32_t c[8];
#pragma unroll
      for (uint32_t t = 0; t < 8u; t++) {
        const uint32_t i = 2u * (j0 + (t >> 1) * kOrderThreads + threadIdx.x) + (t & 1u);
        c[t] = i < n ? cost[b0 + i] : 0u;
      }
#pragma unroll
      for (uint32_t t = 0; t < 4u; t++) {
        const uint32_t byte = j0 + t * kOrderThreads + threadIdx.x;
        if (byte < n_bytes) cls[byte] = (uint8_t)(cost_class(c[2u * t]) | (cost_class(c[2u * t + 1u]) << 4));
      }
    }
    __syncthreads();
  }
  auto class_of = [&](uint32_t i) -> uint32_t {
    return cls ? ((uint32_t)cls[i >> 1] >> (4u * (i & 1u))) & 15u : cost_class(cost[b0 + i]);
  };
  const uint32_t per = (n + kOrderThreads - 1) / kOrderThreads;
  const uint32_t i0 = min(n, threadIdx.x * per), i1 = min(n, i0 + per);
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t cnt[kCostClasses];
#pragma unroll
  for (int c = 0; c < (int)kCostClasses; c++) cnt[c] = 0;
  for (uint32_t i = i0; i < i1; i++) {
    const uint32_t k = class_of(i);
#pragma unroll
    for (int c = 0; c < (int)kCostClasses; c++) cnt[c] += (k == (uint32_t)c) ? 1u : 0u;
  }
  // exclusive scan of each class count over the threads, in thread order
  uint32_t off[kCostClasses];
#pragma unroll
  for (int c = 0; c < (int)kCostClasses; c++) {
    uint32_t v = cnt[c];
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t t = __shfl_up(v, d);
      if (lane >= (uint32_t)d) v += t;
    }
    off[c] = v - cnt[c];
    if (lane == 63u) s_wave[c][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < kCostClasses) {
    const uint32_t c = threadIdx.x;
    uint32_t run = 0;
    for (uint32_t w = 0; w < kOrderThreads / 64; w++) {
      const uint32_t t = s_wave[c][w];
      s_wave[c][w] = run;
      run += t;
    }
    s_base[c] = run;  // the class total, for now
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int c = (int)kCostClasses - 1; c >= 0; c--) {
      const uint32_t t = s_base[c];
      s_base[c] = run;
      run += t;
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < (int)kCostClasses; c++) off[c] += s_base[c] + s_wave[c][wave];
  for (uint32_t i = i0; i < i1; i++) {
    const uint32_t k = class_of(i);
    uint32_t pos = 0;
#pragma unroll
    for (int c = 0; c < (int)kCostClasses; c++) {
      if (k == (uint32_t)c) {
        pos = off[c];
        off[c]++;
      }
    }
    order[b0 + pos] = b0 + i;
  }
}

// True when every camera of the frame is within kViewMoved and kViewTurnCos of the view
// recorded in `saved` (a band's copy), which then takes the frame's views.  The band's threads
// share the cameras; the barrier inside __syncthreads_or orders every read of the copy
// before any write to it.
__device__ bool views_coherent(const CamRec *cams, uint32_t n, float *saved) {
  uint32_t far = 0;
  for (uint32_t c = threadIdx.x; c < n; c += blockDim.x) {
    const float *a = saved + (size_t)c * kViewWords;
    const CamRec &r = cams[c];
    float moved = 0.0f, dot = 0.0f, la = 0.0f, lb = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float d = a[k] - r.origin[k];
      moved += d * d;
      dot += a[3 + k] * r.facing[k];
      la += a[3 + k] * a[3 + k];
      lb += r.facing[k] * r.facing[k];
    }
    const bool near = a[6] == 1.0f && moved < kViewMoved * kViewMoved && dot > kViewTurnCos * sqrtf(la * lb);
    far |= near ? 0u : 1u;
  }
  const bool coherent = __syncthreads_or((int)far) == 0;
  for (uint32_t c = threadIdx.x; c < n; c += blockDim.x) {
    float *a = saved + (size_t)c * kViewWords;
    const CamRec &r = cams[c];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      a[k] = r.origin[k];
      a[3 + k] = r.facing[k];
    }
    a[6] = 1.0f;
    a[7] = 0.0f;
  }
  return coherent;
}

// The work a frame does before its trace launch, in one launch: work-group 0 clears the
// frame's counters and totals and writes the camera records that came in the kernel
// argument; work-group b < bands builds band b's tile order (cost order when the views
// are coherent with the band's recorded ones, raster order otherwise).  The camera poses
// are read from the argument when they came inline (work-group 0 may still be writing the
// device records), from the device records otherwise (copied before this launch).
__global__ void __launch_bounds__(kOrderThreads) frame_begin_kernel(FrameBegin f) {
  if (blockIdx.x == 0) {
    const uint32_t i = threadIdx.x;
    if (i < 8u) f.counters[i * kCounterStride] = 0u;
    if (i < 4u) f.totals[i] = 0ull;
    if (i < f.n_cams * kCamWords)
      reinterpret_cast<uint32_t *>(f.cams)[i] = reinterpret_cast<const uint32_t *>(f.cam)[i];
  }
  extern __shared__ uint8_t s_cls[];
  if (blockIdx.x < f.bands) {
    const uint32_t b0 = f.range_begin[blockIdx.x], n = f.range_begin[blockIdx.x + 1] - b0;
    const CamRec *poses = f.n_cams ? f.cam : f.cams;
    if (views_coherent(poses, f.n_views, f.views + (size_t)blockIdx.x * f.n_views * kViewWords)) {
      order_band(f.cost, f.order, b0, n, f.stage ? s_cls : nullptr);
    } else {
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) f.order[b0 + i] = b0 + i;
    }
  }
}

__global__ void zero_u32_kernel(uint32_t *p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// The device's floatToUlong / ulongToFloat (wavefront.cl:200-209) applied to an array, for
// comparing the device conversions with the host's over large samples.
__global__ void convert_kernel(uint32_t kind, const void *in, void *out, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    if (kind == 0)
      static_cast<uint64_t *>(out)[i] = float_to_ulong(static_cast<const float *>(in)[i]);
    else
      static_cast<float *>(out)[i] = ulong_to_float(static_cast<const uint64_t *>(in)[i]);
  }
}

// ---- order-preserving status compaction (ascending cone id) ----

constexpr uint32_t kCompactThreads = 256;
constexpr uint32_t kCompactPerThread = 16;
constexpr uint32_t kCompactChunk = kCompactThreads * kCompactPerThread;

__device__ __forceinline__ void load_status16(const uint8_t *status, uint32_t n, uint32_t base, uint8_t s[16]) {
  if (base + 16u <= n) {
    const uint4 v = *reinterpret_cast<const uint4 *>(status + base);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 16; i++) s[i] = (uint8_t)(w[i >> 2] >> (8 * (i & 3)));
  } else {
#pragma unroll
    for (int i = 0; i < 16; i++) s[i] = (base + i < n) ? status[base + i] : 0xffu;
  }
}

// category bits of one status byte: bit0 background, bit1 material, bit2 child request
__device__ __forceinline__ uint32_t categories(uint8_t s) {
  if (s == 0xffu) return 0u;
  return ((s & 1u) ? 2u : 1u) | ((s & 2u) ? 4u : 0u);
}

// Exclusive scan of three counters over the 256 threads of a block.
__device__ void block_scan3(uint32_t v[3], uint32_t total[3]) {
  __shared__ uint32_t s_wave[3][kCompactThreads / 64];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t incl[3] = {v[0], v[1], v[2]};
#pragma unroll
  for (int c = 0; c < 3; c++)
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t t = __shfl_up(incl[c], off);
      if (lane >= (uint32_t)off) incl[c] += t;
    }
  if (lane == 63u)
    for (int c = 0; c < 3; c++) s_wave[c][wave] = incl[c];
  __syncthreads();
  for (int c = 0; c < 3; c++) {
    uint32_t before = 0, all = 0;
    for (uint32_t w = 0; w < kCompactThreads / 64; w++) {
      if (w < wave) before += s_wave[c][w];
      all += s_wave[c][w];
    }
    v[c] = before + incl[c] - v[c];
    total[c] = all;
  }
  __syncthreads();
}

__global__ void __launch_bounds__(kCompactThreads) status_count_kernel(const uint8_t *status, uint32_t n,
                                                                        uint32_t *counts) {
  const uint32_t base = blockIdx.x * kCompactChunk + threadIdx.x * kCompactPerThread;
  uint8_t s[16];
  load_status16(status, n, base, s);
  uint32_t v[3] = {0u, 0u, 0u};
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const uint32_t c = categories(s[i]);
    v[0] += c & 1u;
    v[1] += (c >> 1) & 1u;
    v[2] += (c >> 2) & 1u;
  }
  uint32_t total[3];
  block_scan3(v, total);
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; c++) counts[4u * blockIdx.x + c] = total[c];
}

// Single-workgroup exclusive scan of the per-chunk counts; writes the queue counters.  Thread
// t sums a contiguous run of chunks, one scan over the threads gives each run its start, and
// the thread writes its run's prefixes (one barrier round however many chunks).
__global__ void __launch_bounds__(kCompactThreads) status_scan_kernel(uint32_t *counts, uint32_t chunks,
                                                                       uint32_t *bg_count, uint32_t *mat_count,
                                                                       uint32_t *req_count,
                                                                       unsigned long long *totals) {
  uint4 *cnt = reinterpret_cast<uint4 *>(counts);
  const uint32_t per = (chunks + kCompactThreads - 1u) / kCompactThreads;
  const uint32_t c0 = min(chunks, threadIdx.x * per), c1 = min(chunks, c0 + per);
  uint32_t v[3] = {0u, 0u, 0u};
  for (uint32_t i = c0; i < c1; i++) {
    const uint4 t = cnt[i];
    v[0] += t.x;
    v[1] += t.y;
    v[2] += t.z;
  }
  uint32_t carry[3];
  block_scan3(v, carry);
  for (uint32_t i = c0; i < c1; i++) {
    const uint4 t = cnt[i];
    cnt[i] = make_uint4(v[0], v[1], v[2], t.w);
    v[0] += t.x;
    v[1] += t.y;
    v[2] += t.z;
  }
  if (threadIdx.x == 0) {
    if (bg_count) *bg_count = carry[0];
    if (mat_count) *mat_count = carry[1];
    if (req_count) *req_count = carry[2];
    totals[2] = carry[2];
  }
}

// Number of set bits of m below this lane.
__device__ __forceinline__ uint32_t rank_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Wave w of a work-group writes the entries of cone ids [chunk + 1024 w, chunk + 1024 (w+1))
// in 16 rounds of 64 consecutive ids, one per lane: each round's entries of a queue go to
// consecutive positions (ranks from ballots), so the record reads and the queue writes are
// coalesced.  Positions continue the chunk offsets of the scan in ascending cone id.
// fused (at most kFusedChunks chunks): `offsets` holds the chunks' own counts, each work-group
// sums those before it, and the last one writes the queue counters and the request total.
constexpr uint32_t kFusedChunks = 1024;
__global__ void __launch_bounds__(kCompactThreads) status_scatter_kernel(
    const uint8_t *status, uint32_t n, const uint32_t *offsets, uint32_t *bg_queue, uint32_t *mat_queue,
    const svo_child_request *req_sparse, svo_child_request *requests, uint32_t *req_cones, uint32_t req_capacity,
    uint32_t fused, uint32_t *bg_count, uint32_t *mat_count, uint32_t *req_count, unsigned long long *totals) {
  uint32_t pre[3];
  if (fused) {
    uint32_t v[3] = {0u, 0u, 0u};
    for (uint32_t c = threadIdx.x; c < blockIdx.x; c += kCompactThreads)
      for (int q = 0; q < 3; q++) v[q] += offsets[4u * c + q];
    block_scan3(v, pre);
  } else {
    for (int q = 0; q < 3; q++) pre[q] = offsets[4u * blockIdx.x + q];
  }
  constexpr uint32_t kRounds = kCompactChunk / kCompactThreads;
  __shared__ uint32_t s_tot[3][kCompactThreads / 64];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t wbase = blockIdx.x * kCompactChunk + wave * (64u * kRounds);
  uint64_t cats = 0;  // 3 category bits per round
  uint32_t tot[3] = {0u, 0u, 0u};
#pragma unroll
  for (uint32_t r = 0; r < kRounds; r++) {
    const uint32_t id = wbase + r * 64u + lane;
    const uint32_t c = categories(id < n ? status[id] : (uint8_t)0xffu);
    cats |= (uint64_t)c << (3u * r);
#pragma unroll
    for (uint32_t q = 0; q < 3; q++) tot[q] += (uint32_t)__popcll(__ballot((c >> q) & 1u));
  }
  if (lane == 0)
    for (uint32_t q = 0; q < 3; q++) s_tot[q][wave] = tot[q];
  __syncthreads();
  uint32_t pos[3];
#pragma unroll
  for (uint32_t q = 0; q < 3; q++) {
    pos[q] = pre[q];
    for (uint32_t w = 0; w < wave; w++) pos[q] += s_tot[q][w];
  }
  if (fused && blockIdx.x == gridDim.x - 1u && threadIdx.x == 0) {
    uint32_t all[3] = {pre[0], pre[1], pre[2]};
    for (uint32_t q = 0; q < 3; q++)
      for (uint32_t w = 0; w < kCompactThreads / 64; w++) all[q] += s_tot[q][w];
    if (bg_count) *bg_count = all[0];
    if (mat_count) *mat_count = all[1];
    if (req_count) *req_count = all[2];
    totals[2] = all[2];
  }
#pragma unroll
  for (uint32_t r = 0; r < kRounds; r++) {
    const uint32_t id = wbase + r * 64u + lane;
    const uint32_t c = (uint32_t)(cats >> (3u * r)) & 7u;
    const uint64_t m0 = __ballot(c & 1u), m1 = __ballot(c & 2u), m2 = __ballot(c & 4u);
    if ((c & 1u) && bg_queue) bg_queue[pos[0] + rank_below(m0)] = id;
    if ((c & 2u) && mat_queue) mat_queue[pos[1] + rank_below(m1)] = id;
    if (c & 4u) {
      const uint32_t at = pos[2] + rank_below(m2);
      if (requests && at < req_capacity) requests[at] = req_sparse[id];
      if (req_cones && at < req_capacity) req_cones[at] = id;
    }
    pos[0] += (uint32_t)__popcll(m0);
    pos[1] += (uint32_t)__popcll(m1);
    pos[2] += (uint32_t)__popcll(m2);
  }
}

template <int WG, int WAVES>
int launch_trace_wg(const FrameParams &p, const LaunchShape &shape, hipStream_t stream) {
  if (p.hashes)
    hipLaunchKernelGGL((trace_frame_kernel<WG, WAVES, true>), dim3(shape.grid), dim3(WG), shape.lds_bytes, stream, p);
  else
    hipLaunchKernelGGL((trace_frame_kernel<WG, WAVES, false>), dim3(shape.grid), dim3(WG), shape.lds_bytes, stream, p);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Occupancy of the hashing variant (never fewer registers than the plain one) bounds the
// grid of both.
template <int WG, int WAVES>
int occupancy_wg(size_t lds, int *blocks_per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, trace_frame_kernel<WG, WAVES, true>, WG, lds) ==
                 hipSuccess
             ? 0
             : -1;
}

template <int WAVES>
int launch_waves(const FrameParams &p, const LaunchShape &shape, hipStream_t stream) {
  switch (shape.workgroup) {
    case 64: return launch_trace_wg<64, WAVES>(p, shape, stream);
    case 128: return launch_trace_wg<128, WAVES>(p, shape, stream);
    case 256: return launch_trace_wg<256, WAVES>(p, shape, stream);
    case 512: return launch_trace_wg<512, WAVES>(p, shape, stream);
    case 1024: return launch_trace_wg<1024, WAVES>(p, shape, stream);
    default: return -1;
  }
}

template <int WAVES>
int occupancy_waves(int wg, size_t lds, int *blocks_per_cu) {
  switch (wg) {
    case 64: return occupancy_wg<64, WAVES>(lds, blocks_per_cu);
    case 128: return occupancy_wg<128, WAVES>(lds, blocks_per_cu);
    case 256: return occupancy_wg<256, WAVES>(lds, blocks_per_cu);
    case 512: return occupancy_wg<512, WAVES>(lds, blocks_per_cu);
    case 1024: return occupancy_wg<1024, WAVES>(lds, blocks_per_cu);
    default: return -1;
  }
}

}  // namespace

int trace_occupancy(int wg, int waves, size_t lds_bytes, int *blocks_per_cu) {
  switch (waves) {
    case 4: return occupancy_waves<4>(wg, lds_bytes, blocks_per_cu);
    case 5: return occupancy_waves<5>(wg, lds_bytes, blocks_per_cu);
    case 6: return occupancy_waves<6>(wg, lds_bytes, blocks_per_cu);
    default: return -1;
  }
}

int launch_frame_begin(const FrameBegin &f, hipStream_t stream) {
  if (f.n_cams > kInlineCams || f.bands > 8u) return -1;
  if (f.bands && (!f.views || !f.n_views || (f.n_cams && f.n_cams != f.n_views))) return -1;
  // bands of up to kStageTiles tiles stage their cost classes in LDS (two tiles a byte)
  constexpr uint32_t kStageTiles = 96u * 1024u;
  uint32_t most = 0;
  for (uint32_t b = 0; b < f.bands; b++) most = std::max(most, f.range_begin[b + 1] - f.range_begin[b]);
  FrameBegin g = f;
  g.stage = most > 0u && most <= kStageTiles ? 1u : 0u;
  const size_t lds = g.stage ? (size_t)(((most + 1u) / 2u + 15u) & ~15u) : 0;
  hipLaunchKernelGGL(frame_begin_kernel, dim3(std::max(1u, f.bands)), dim3(kOrderThreads), lds, stream, g);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_zero_u32(uint32_t *p, uint32_t n, hipStream_t stream) {
  if (n == 0) return 0;
  const uint32_t blocks = std::min<uint32_t>((n + 255u) / 256u, 1024u);
  hipLaunchKernelGGL(zero_u32_kernel, dim3(blocks), dim3(256), 0, stream, p, n);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_convert(uint32_t kind, const void *in, void *out, size_t n, hipStream_t stream) {
  if (n == 0) return 0;
  const uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(convert_kernel, dim3(blocks), dim3(256), 0, stream, kind, in, out, n);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_trace_frame(const FrameParams &p, const LaunchShape &shape, hipStream_t stream) {
  switch (shape.waves) {
    case 4: return launch_waves<4>(p, shape, stream);
    case 5: return launch_waves<5>(p, shape, stream);
    case 6: return launch_waves<6>(p, shape, stream);
    default: return -1;
  }
}

// The compositing modes hold more live state per lane (layers, colour, transmittance):
// they are compiled for 4 waves per SIMD.
int composite_occupancy(size_t lds_bytes, int *blocks_per_cu) {
  int a = 0, b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, trace_frame_kernel<256, 4, false, kModeComposite>, 256,
                                                   lds_bytes) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, trace_frame_kernel<256, 4, false, kModeAdjoint>, 256,
                                                   lds_bytes) != hipSuccess)
    return -1;
  *blocks_per_cu = a < b ? a : b;
  return 0;
}

int launch_composite(const FrameParams &p, const LaunchShape &shape, int mode, hipStream_t stream) {
  if (shape.workgroup != 256 || !p.n_cams) return -1;
  if (mode == kModeComposite)
    hipLaunchKernelGGL((trace_frame_kernel<256, 4, false, kModeComposite>), dim3(shape.grid), dim3(256),
                       shape.lds_bytes, stream, p);
  else if (mode == kModeAdjoint)
    hipLaunchKernelGGL((trace_frame_kernel<256, 4, false, kModeAdjoint>), dim3(shape.grid), dim3(256),
                       shape.lds_bytes, stream, p);
  else
    return -1;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

uint32_t compaction_chunks(uint32_t n) { return (n + kCompactChunk - 1u) / kCompactChunk; }

int launch_compaction(const CompactParams &c, hipStream_t stream) {
  const uint32_t chunks = compaction_chunks(c.n);
  if (chunks == 0) return 0;
  hipLaunchKernelGGL(status_count_kernel, dim3(chunks), dim3(kCompactThreads), 0, stream, c.status, c.n, c.counts);
  if (chunks <= kFusedChunks) {
    hipLaunchKernelGGL(status_scatter_kernel, dim3(chunks), dim3(kCompactThreads), 0, stream, c.status, c.n, c.counts,
                       c.bg_queue, c.mat_queue, c.req_sparse, c.requests, c.req_cones, c.req_capacity, 1u, c.bg_count,
                       c.mat_count, c.req_count, c.totals);
    return hipGetLastError() == hipSuccess ? 0 : -1;
  }
  hipLaunchKernelGGL(status_scan_kernel, dim3(1), dim3(kCompactThreads), 0, stream, c.counts, chunks, c.bg_count,
                     c.mat_count, c.req_count, c.totals);
  hipLaunchKernelGGL(status_scatter_kernel, dim3(chunks), dim3(kCompactThreads), 0, stream, c.status, c.n, c.counts,
                     c.bg_queue, c.mat_queue, c.req_sparse, c.requests, c.req_cones, c.req_capacity, 0u, c.bg_count,
                     c.mat_count, c.req_count, c.totals);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace svo
