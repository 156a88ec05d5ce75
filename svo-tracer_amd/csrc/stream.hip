// Streaming support on the GPU: the fBm terrain (height grid + min/max pyramid), the
// seed tree (Bases and the dense block level), child-request consumption into bricks,
// and the pool's garbage collection (stale-group reclaim is the collection kernels' reclaim
// mode; compaction is here).
//
// Bricks follow TreeBuilder's rules for one level (MakeBlock/MakeChunk,
// TreeBuilder.cs:74-93, 264-284): the 8 children of a requesting block, each with the
// chunk of its octants, its colour and normal, and no children of its own.
#include <algorithm>

#include "svo_device.hpp"
#include "svo_stream_kernels.hpp"
#include "svo_shapes.hpp"
#include "svo_terrain.hpp"

namespace svo {
namespace {

using dev::kU32Max;
constexpr int kWg = 256;

__device__ __forceinline__ void footprint(const TerrainView &t, uint32_t level, uint32_t x, uint32_t y, float *lo,
                                          float *hi) {
  if (level < t.spec.lres) {
    const size_t k = t.level_offset[level] + (size_t)y * (1ull << level) + x;
    *lo = t.hmin[k];
    *hi = t.hmax[k];
  } else {
    const uint32_t sh = level - t.spec.lres;
    const size_t k = (size_t)(y >> sh) * (1ull << t.spec.lres) + (x >> sh);
    *lo = *hi = t.top[k];
  }
}

// MakeChunk, TreeBuilder.cs:264-284: per octant k, bit 2k air, bit 2k+1 geometry.
__device__ uint16_t make_chunk(const TerrainView &t, uint32_t level, uint32_t x, uint32_t y, uint32_t z) {
  uint16_t chunk = 0;
  for (uint32_t k = 0; k < 8; k++) {
    const uint32_t cx = 2 * x + (k & 1u), cy = 2 * y + ((k >> 1) & 1u), cz = 2 * z + ((k >> 2) & 1u);
    float lo, hi;
    footprint(t, level + 1, cx, cy, &lo, &hi);
    bool a, g;
    terrain::classify(level + 1, cz, lo, hi, &a, &g);
    chunk |= (uint16_t)((a ? 1u : 0u) << (2 * k));
    chunk |= (uint16_t)((g ? 1u : 0u) << (2 * k + 1));
  }
  return chunk;
}

// MakeBlock, TreeBuilder.cs:74-93 (ColourB takes colour[1] and ColourG colour[2], as there).
__device__ uint4 make_block(const TerrainView &t, uint32_t level, uint32_t x, uint32_t y, uint32_t z) {
  const uint16_t chunk = make_chunk(t, level, x, y, z);
  float lo, hi;
  footprint(t, level, x, y, &lo, &hi);
  int16_t pitch, yaw;
  uint8_t rgb[3];
  terrain::attributes(t.spec, t.top, level, x, y, z, hi, &pitch, &yaw, rgb);
  uint4 b;
  b.x = kU32Max;
  b.y = (uint32_t)chunk | ((uint32_t)(uint16_t)pitch << 16);
  // yaw @8, R @10, G @11 (= rgb[2]), B @12 (= rgb[1]), opacity @13 = 255, spare, gloss 0
  b.z = (uint32_t)(uint16_t)yaw | ((uint32_t)rgb[0] << 16) | ((uint32_t)rgb[2] << 24);
  b.w = (uint32_t)rgb[1] | (255u << 8);
  return b;
}

__device__ __forceinline__ uint32_t compact3(uint32_t v) {  // every third bit -> low bits
  v &= 0x09249249u;
  v = (v | (v >> 2)) & 0x030c30c3u;
  v = (v | (v >> 4)) & 0x0300f00fu;
  v = (v | (v >> 8)) & 0x030000ffu;
  v = (v | (v >> 16)) & 0x000003ffu;
  return v;
}

// Any builder geometry: one cell's classification and one block (MakeChunk / MakeBlock,
// TreeBuilder.cs:74-93, 264-284), by the rules the host builder shares.
__device__ void geom_classify(const GeomView &g, uint32_t level, uint32_t x, uint32_t y, uint32_t z, bool *air,
                              bool *geo) {
  switch (g.kind) {
    case SVO_GEOMETRY_FBM_HEIGHTFIELD: {
      float lo, hi;
      footprint(g.t, level, x, y, &lo, &hi);
      terrain::classify(level, z, lo, hi, air, geo);
      return;
    }
    case SVO_GEOMETRY_SPHERES:
      shapes::spheres_classify(g.shapes, g.count, shapes::cell_box(level, x, y, z), air, geo);
      return;
    case SVO_GEOMETRY_CUBES:
      shapes::cubes_classify(g.shapes, g.count, shapes::cell_box(level, x, y, z), air, geo);
      return;
    default:
      shapes::noise_classify(g.noise, level, x, y, z, air, geo);
  }
}

__device__ uint16_t geom_chunk(const GeomView &g, uint32_t level, uint32_t x, uint32_t y, uint32_t z) {
  uint16_t chunk = 0;
  for (uint32_t k = 0; k < 8; k++) {
    bool a, e;
    geom_classify(g, level + 1, 2 * x + (k & 1u), 2 * y + ((k >> 1) & 1u), 2 * z + ((k >> 2) & 1u), &a, &e);
    chunk |= (uint16_t)((a ? 1u : 0u) << (2 * k));
    chunk |= (uint16_t)((e ? 1u : 0u) << (2 * k + 1));
  }
  return chunk;
}

__device__ uint4 geom_block(const GeomView &g, uint32_t level, uint32_t x, uint32_t y, uint32_t z) {
  if (g.kind == SVO_GEOMETRY_FBM_HEIGHTFIELD) return make_block(g.t, level, x, y, z);
  const uint16_t chunk = geom_chunk(g, level, x, y, z);
  int16_t pitch = 0, yaw = 0;
  uint8_t rgb[3] = {255, 255, 255};  // boxes: white, no normal
  if (g.kind == SVO_GEOMETRY_SPHERES)
    shapes::spheres_attributes(g.shapes, g.count, shapes::cell_box(level, x, y, z), &pitch, &yaw, rgb);
  else if (g.kind != SVO_GEOMETRY_CUBES)
    shapes::noise_attributes(g.noise, level, x, y, z, &pitch, &yaw, rgb);
  uint4 b;
  b.x = kU32Max;
  b.y = (uint32_t)chunk | ((uint32_t)(uint16_t)pitch << 16);
  b.z = (uint32_t)(uint16_t)yaw | ((uint32_t)rgb[0] << 16) | ((uint32_t)rgb[2] << 24);
  b.w = (uint32_t)rgb[1] | (255u << 8);
  return b;
}

// ---------------------------------------------------------------------------
// terrain
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kWg) void terrain_top_kernel(terrain::Spec spec, float *top) {
  const uint32_t res = 1u << spec.lres;
  const size_t n = (size_t)res * res;
  for (size_t k = blockIdx.x * (size_t)kWg + threadIdx.x; k < n; k += (size_t)gridDim.x * kWg)
    top[k] = terrain::height(spec, (uint32_t)(k % res), (uint32_t)(k / res));
}

__global__ __launch_bounds__(kWg) void terrain_reduce_kernel(const float *in_min, const float *in_max, float *out_min,
                                                             float *out_max, uint32_t n) {
  const size_t total = (size_t)n * n;
  for (size_t k = blockIdx.x * (size_t)kWg + threadIdx.x; k < total; k += (size_t)gridDim.x * kWg) {
    const size_t j = k / n, i = k % n;
    const size_t a = (2 * j) * (2 * (size_t)n) + 2 * i, b = a + 1, c = a + 2 * (size_t)n, d = c + 1;
    out_min[k] = fminf(fminf(in_min[a], in_min[b]), fminf(in_min[c], in_min[d]));
    out_max[k] = fmaxf(fmaxf(in_max[a], in_max[b]), fmaxf(in_max[c], in_max[d]));
  }
}

// ---------------------------------------------------------------------------
// seed tree: Bases (levels 1..B) and the dense level B+2 (BuildBaseChunks /
// PopulateInitialBlocks, TreeBuilder.cs:152-227)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kWg) void seed_bases_kernel(TerrainView t, uint16_t *bases, uint32_t level,
                                                         uint32_t offset) {
  const uint32_t n = 1u << (3 * level);
  for (uint32_t m = blockIdx.x * kWg + threadIdx.x; m < n; m += gridDim.x * kWg)
    bases[offset + m] = make_chunk(t, level, compact3(m), compact3(m >> 1), compact3(m >> 2));
}

__global__ __launch_bounds__(kWg) void seed_dense_kernel(TerrainView t, const uint16_t *bases, uint32_t B,
                                                         uint4 *blocks) {
  const uint32_t level = B + 2, n = 1u << (3 * level);
  const uint32_t base_start = ((1u << (3u * B)) - 1u) / 7u - 1u;  // powSum(B-1)
  for (uint32_t la = blockIdx.x * kWg + threadIdx.x; la < n; la += gridDim.x * kWg) {
    const uint32_t bp = la >> 3;  // the level-(B+1) cell
    uint4 b = make_uint4(0u, 0u, 0u, 0u);
    if (((bases[base_start + (bp >> 3)] >> ((bp & 7u) * 2)) & 3u) == 3u)
      b = make_block(t, level, compact3(la), compact3(la >> 1), compact3(la >> 2));
    blocks[la] = b;
  }
}

// ---------------------------------------------------------------------------
// child requests -> bricks
// ---------------------------------------------------------------------------

// The frame's request count, capped at what the buffer stores: counters[0] = stored, [1] = all.
__global__ void request_count_kernel(const uint32_t *request_id, uint32_t stored, uint32_t *counters) {
  if (threadIdx.x != 0) return;
  const uint32_t c = *request_id;
  counters[0] = c < stored ? c : stored;
  counters[1] = c;
}

// Address, depth and TreeSize of request i (from its key when the frame kept it: 8 bytes
// instead of the 40-byte record).
struct RequestKey {
  uint32_t address, depth;
  bool tree;
};
__device__ __forceinline__ RequestKey request_key(const RequestSource &src, uint32_t i) {
  if (src.cones) {
    const uint2 k = src.keys[src.cones[i]];
    return {k.x, k.y & 0xffu, (k.y >> 8) != 0u};
  }
  const svo_child_request &r = src.records[i];
  return {r.address, r.depth, r.tree_size != 0};
}

// A request takes part in the de-duplication when its address is in Blocks and its cone needs
// a level below the block (TreeSize, wavefront.cl:362-375), and it is not right after a
// taking-part request for the same address (neighbouring cones mostly ask for the same block).
__device__ __forceinline__ bool request_inserts(const RequestSource &src, uint32_t i, uint32_t n_blocks) {
  const RequestKey k = request_key(src, i);
  if (k.address >= n_blocks || !k.tree) return false;
  if (i == 0) return true;
  const RequestKey q = request_key(src, i - 1);
  return !(q.address == k.address && q.tree);
}

// first[a] = (gen << 32) | ~i for the smallest index i of the requests for block a in this
// update (generation gen): a 64-bit maximum, so entries of earlier updates (smaller gen)
// never need clearing.
__device__ __forceinline__ unsigned long long first_key(uint32_t gen, uint32_t i) {
  return ((unsigned long long)gen << 32) | (unsigned long long)(kU32Max - i);
}

// flags[i] = 1 for the stored requests (i < *n_ptr) that take part, so that the second pass
// reads only their records; flags beyond the stored requests are not written.  Grid-stride
// over the device-side count: the grid covers at most kRequestGrid work-groups.
constexpr uint32_t kRequestGrid = 2048;
__device__ __forceinline__ uint32_t stored_count(const uint32_t *n_ptr, uint32_t n_max) {
  const uint32_t n = *n_ptr;
  return n < n_max ? n : n_max;
}
__global__ __launch_bounds__(kWg) void request_first_kernel(RequestSource req, uint32_t n_max, const uint32_t *n_ptr,
                                                            unsigned long long *first, uint32_t gen, uint32_t n_blocks,
                                                            uint32_t *flags) {
  const uint32_t n = stored_count(n_ptr, n_max);
  for (uint32_t i = blockIdx.x * kWg + threadIdx.x; i < n; i += gridDim.x * kWg) {
    const bool takes_part = request_inserts(req, i, n_blocks);
    if (takes_part) atomicMax(&first[request_key(req, i).address], first_key(gen, i));
    flags[i] = takes_part ? 1u : 0u;
  }
}

// 1 when request i is the first for its address and its block can take a child level (the
// builder refines exactly the blocks below max_depth that can have children); 0 beyond the
// stored requests.
// (after request_first_kernel) flags[i] stays 1 only for a taking-part request
__global__ __launch_bounds__(kWg) void request_flag_kernel(RequestSource req, uint32_t n_max, const uint32_t *n_ptr,
                                                           const unsigned long long *first, uint32_t gen,
                                                           const uint4 *blocks, uint32_t max_depth, uint32_t *flags) {
  const uint32_t n = stored_count(n_ptr, n_max);
  for (uint32_t i = blockIdx.x * kWg + threadIdx.x; i < n; i += gridDim.x * kWg) {
    if (flags[i] == 0u) continue;
    const RequestKey k = request_key(req, i);
    const uint32_t a = k.address, depth = k.depth;
    bool ok = first[a] == first_key(gen, i) && depth + 1u <= max_depth && depth >= 1;
    if (ok) {
      const uint4 blk = blocks[a];
      // childless, and CanHaveChildren (TreeBuilder.cs:291-302): an octant with air and geometry
      const uint32_t chunk = blk.y & 0xffffu;
      ok = blk.x == kU32Max && ((chunk & (chunk >> 1)) & 0x5555u) != 0u;
    }
    if (!ok) flags[i] = 0u;
  }
}

// Request list[k] becomes graft record k and brick blocks [8k, 8k+8); eight threads a brick.
__global__ __launch_bounds__(kWg) void brick_kernel(TerrainView t, RequestSource req, const uint32_t *list,
                                                    uint32_t nb_max, const uint32_t *unique, svo_grafting *records,
                                                    uint4 *bricks) {
  const uint32_t tid = blockIdx.x * kWg + threadIdx.x;
  const uint32_t k = tid >> 3, c = tid & 7u;
  if (k >= nb_max || k >= *unique) return;
  const uint32_t i = list[k];
  const svo_child_request r = req.records[req.cones ? req.cones[i] : i];
  const uint32_t D = r.depth, s = 64u - D;
  const uint32_t x = (uint32_t)(r.location[0] >> s), y = (uint32_t)(r.location[1] >> s),
                 z = (uint32_t)(r.location[2] >> s);
  bricks[8ull * k + c] = make_block(t, D + 1u, 2 * x + (c & 1u), 2 * y + ((c >> 1) & 1u), 2 * z + ((c >> 2) & 1u));
  if (c == 0) {
    svo_grafting g;
    g.graft_data_address = 8u * k;
    g.graft_total_size = 8u;
    g.depth = 0xffu;
    g.pad[0] = g.pad[1] = g.pad[2] = 0;
    g.graft_address = r.address;
    records[k] = g;
  }
}

// ---------------------------------------------------------------------------
// group selection and compaction
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool selected(const svo_usage *u, uint32_t g, uint32_t pred) {
  const uint32_t t = u[g].tick;
  switch (pred) {
    case kSelLive: return t != 0u;
    case kSelFree: return t == 0u;
    default: return t != 0u && t < 65534u;  // kSelMovable
  }
}

constexpr uint32_t kSelChunk = 1024, kSelPer = kSelChunk / kWg;

template <int WG>
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
  for (uint32_t w = 0; w < (uint32_t)(WG / 64); w++) {
    const uint32_t tw = s_waves[w];
    before += w < wave ? tw : 0u;
    all += tw;
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

__global__ __launch_bounds__(kWg) void select_count_kernel(const svo_usage *u, uint32_t lo, uint32_t hi, uint32_t pred,
                                                           uint32_t *counts) {
  __shared__ uint32_t s_waves[kWg / 64];
  const uint32_t g0 = lo + blockIdx.x * kSelChunk + threadIdx.x * kSelPer;
  uint32_t c = 0;
  for (uint32_t j = 0; j < kSelPer; j++) c += (g0 + j < hi && selected(u, g0 + j, pred)) ? 1u : 0u;
  uint32_t all;
  wg_scan<kWg>(c, s_waves, &all);
  if (threadIdx.x == 0) counts[blockIdx.x] = all;
}

__global__ __launch_bounds__(kWg) void select_scatter_kernel(const svo_usage *u, uint32_t lo, uint32_t hi,
                                                             uint32_t pred, const uint32_t *offsets, uint32_t *list) {
  __shared__ uint32_t s_waves[kWg / 64];
  const uint32_t g0 = lo + blockIdx.x * kSelChunk + threadIdx.x * kSelPer;
  bool is[kSelPer];
  uint32_t c = 0;
  for (uint32_t j = 0; j < kSelPer; j++) {
    is[j] = g0 + j < hi && selected(u, g0 + j, pred);
    c += is[j] ? 1u : 0u;
  }
  uint32_t all;
  uint32_t r = offsets[blockIdx.x] + wg_scan<kWg>(c, s_waves, &all);
  for (uint32_t j = 0; j < kSelPer; j++)
    if (is[j]) list[r++] = g0 + j;
}

// Exclusive ranks of the flagged entries (flags 0/1), chunk by chunk like the selection.  With
// n_dev the flags are defined only below min(n, *n_dev): later chunks count zero.
__device__ __forceinline__ uint32_t flagged_count(uint32_t n, const uint32_t *n_dev) {
  if (!n_dev) return n;
  const uint32_t d = *n_dev;
  return d < n ? d : n;
}
__global__ __launch_bounds__(kWg) void flag_count_kernel(const uint32_t *flags, uint32_t n, const uint32_t *n_dev,
                                                         uint32_t *counts) {
  __shared__ uint32_t s_waves[kWg / 64];
  n = flagged_count(n, n_dev);
  if (blockIdx.x * kSelChunk >= n) {  // uniform over the work-group
    if (threadIdx.x == 0) counts[blockIdx.x] = 0u;
    return;
  }
  const uint32_t i0 = blockIdx.x * kSelChunk + threadIdx.x * kSelPer;
  uint32_t c = 0;
  for (uint32_t j = 0; j < kSelPer; j++) c += i0 + j < n ? flags[i0 + j] : 0u;
  uint32_t all;
  wg_scan<kWg>(c, s_waves, &all);
  if (threadIdx.x == 0) counts[blockIdx.x] = all;
}

// rank[i] = flagged entries before i (when rank is given); with a list, flagged i of rank
// r < limit is list[r].  With n_dev, entries from min(n, *n_dev) on are not flagged.
__global__ __launch_bounds__(kWg) void flag_rank_kernel(const uint32_t *flags, uint32_t n, const uint32_t *n_dev,
                                                        const uint32_t *offsets, uint32_t *rank, uint32_t *list,
                                                        uint32_t limit) {
  __shared__ uint32_t s_waves[kWg / 64];
  n = flagged_count(n, n_dev);
  if (blockIdx.x * kSelChunk >= n) return;  // uniform over the work-group
  const uint32_t i0 = blockIdx.x * kSelChunk + threadIdx.x * kSelPer;
  uint32_t f[kSelPer], c = 0;
  for (uint32_t j = 0; j < kSelPer; j++) {
    f[j] = i0 + j < n ? flags[i0 + j] : 0u;
    c += f[j];
  }
  uint32_t all;
  uint32_t r = offsets[blockIdx.x] + wg_scan<kWg>(c, s_waves, &all);
  for (uint32_t j = 0; j < kSelPer; j++) {
    if (rank && i0 + j < n) rank[i0 + j] = r;
    if (list && f[j] && r < limit) list[r] = i0 + j;
    r += f[j];
  }
}

__global__ __launch_bounds__(kWg) void remap_kernel(const uint32_t *from, const uint32_t *to, uint32_t m,
                                                    uint32_t *remap) {
  const uint32_t i = blockIdx.x * kWg + threadIdx.x;
  if (i < m) remap[from[i]] = to[i];
}

__device__ __forceinline__ uint32_t remap_block(const uint32_t *remap, uint32_t b) {
  return remap[b >> 3] * 8u + (b & 7u);
}

// Copies moved groups (blocks with remapped Child fields, usage with remapped Parent).
__global__ __launch_bounds__(kWg) void move_kernel(PoolView p, const uint32_t *from, const uint32_t *to, uint32_t m,
                                                   const uint32_t *remap) {
  const uint32_t tid = blockIdx.x * kWg + threadIdx.x;
  const uint32_t i = tid >> 3, k = tid & 7u;
  if (i >= m) return;
  const uint32_t o = from[i], n = to[i];
  uint4 b = reinterpret_cast<const uint4 *>(p.blocks)[8u * o + k];
  if (b.x != kU32Max && (b.x >> 3) < p.n_usages) b.x = remap_block(remap, b.x);
  reinterpret_cast<uint4 *>(p.blocks)[8u * n + k] = b;
  if (k == 0) {
    svo_usage us = p.usages[o];
    if (us.parent < p.n_blocks && (us.parent >> 3) < p.n_usages) us.parent = remap_block(remap, us.parent);
    p.usages[n] = us;
  }
}

// Links into moved groups from groups that stay, then the old groups are freed.
__global__ __launch_bounds__(kWg) void relink_kernel(PoolView p, const uint32_t *from, const uint32_t *to, uint32_t m,
                                                     const uint32_t *remap) {
  const uint32_t tid = blockIdx.x * kWg + threadIdx.x;
  const uint32_t i = tid >> 3, k = tid & 7u;
  if (i >= m) return;
  const uint32_t o = from[i], n = to[i];
  if (k == 0) {
    const uint32_t pb = p.usages[n].parent;
    if (pb < p.n_blocks && (pb >> 3) < p.n_usages && remap[pb >> 3] == (pb >> 3)) p.blocks[pb].child = 8u * n;
  }
  const uint32_t c = p.blocks[8u * n + k].child;
  if (c == kU32Max || (c >> 3) >= p.n_usages) return;
  const uint32_t cg = c >> 3;
  const uint32_t t = p.usages[cg].tick;
  if (t < 65534u) {
    // a child that stayed and points back at the old block; a moved child's copy already
    // carries the remapped back-pointer
    if (p.usages[cg].parent == 8u * o + k) p.usages[cg].parent = 8u * n + k;
  } else if (t == 65534u && p.parents) {
    uint32_t e = p.usages[cg].parent, steps = 0;
    while (e < p.n_parents && steps++ < p.n_parents) {
      if (p.parents[e].parent_address == 8u * o + k) p.parents[e].parent_address = 8u * n + k;
      e = p.parents[e].next_element;
    }
  }
}

__global__ __launch_bounds__(kWg) void release_moved_kernel(PoolView p, const uint32_t *from, uint32_t m) {
  const uint32_t i = blockIdx.x * kWg + threadIdx.x;
  if (i >= m) return;
  p.usages[from[i]].tick = 0;
  p.usages[from[i]].parent = kU32Max;
}

__global__ __launch_bounds__(kWg) void identity_kernel(uint32_t *remap, uint32_t n) {
  for (uint32_t i = blockIdx.x * kWg + threadIdx.x; i < n; i += gridDim.x * kWg) remap[i] = i;
}

// ---------------------------------------------------------------------------
// device TreeBuilder (BuildTree, TreeBuilder.cs:102-256) for any builder geometry.
//
// TreeBuilder allocates child groups depth first: the dense level's blocks that need
// children in x-outer / z-inner order, each block's children 0..7 in order, a child's
// subtree before its next sibling, every group taken from one FIFO (the sorted empty
// level-(B+1) cells, then groups 8^(B+1)..).  The k-th allocation (its pre-order rank)
// therefore gets FIFO entry k.  Here the tree is first expanded level by level (which
// blocks need children: geometry only), then every internal block's subtree size is
// summed bottom-up, and the ranks follow top-down as prefix sums of the sizes.  Blocks
// are written in the same pass; each block's position in TreeBuilder's write sequence
// (its BlockCount rule depends on that order) is kept for the final count.
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool has_children_bits(uint32_t chunk) {  // CanHaveChildren, TreeBuilder.cs:291-302
  return ((chunk & (chunk >> 1)) & 0x5555u) != 0u;
}

__device__ __forceinline__ uint32_t spread3_dev(uint32_t v) { return dev::spread3(v); }

__global__ __launch_bounds__(kWg) void bt_bases_kernel(GeomView g, uint16_t *bases, uint32_t level, uint32_t offset) {
  const uint32_t n = 1u << (3 * level);
  for (uint32_t m = blockIdx.x * kWg + threadIdx.x; m < n; m += gridDim.x * kWg)
    bases[offset + m] = geom_chunk(g, level, compact3(m), compact3(m >> 1), compact3(m >> 2));
}

// Level-(B+1) cell bp holds both air and geometry in Bases level B.
__device__ __forceinline__ bool dense_mixed(const uint16_t *bases, uint32_t B, uint32_t bp) {
  const uint32_t base_start = ((1u << (3u * B)) - 1u) / 7u - 1u;  // powSum(B-1)
  return ((bases[base_start + (bp >> 3)] >> ((bp & 7u) * 2)) & 3u) == 3u;
}

// Dense level L0 = B+2 in x-outer / z-inner order: p = (x << 2s) | (y << s) | z.
// flag[p] = the block is written and gets children.
__global__ __launch_bounds__(kWg) void bt_dense_flags_kernel(GeomView g, const uint16_t *bases, uint32_t B,
                                                             uint32_t max_depth, uint32_t *flag) {
  const uint32_t L0 = B + 2, n = 1u << (3 * L0), m = (1u << L0) - 1u;
  for (uint32_t p = blockIdx.x * kWg + threadIdx.x; p < n; p += gridDim.x * kWg) {
    const uint32_t x = p >> (2 * L0), y = (p >> L0) & m, z = p & m;
    const uint32_t la = spread3_dev(x) | (spread3_dev(y) << 1) | (spread3_dev(z) << 2);
    bool f = false;
    if (L0 < max_depth && dense_mixed(bases, B, la >> 3)) f = has_children_bits(geom_chunk(g, L0, x, y, z));
    flag[p] = f ? 1u : 0u;
  }
}

// Level-(B+1) cells in z-outer / x-inner order (PopulateInitialBlocks' loop): mixed[q] with
// q = (z << 2s) | (y << s) | x; empty[bp] for the free-group FIFO (bp ascending).
__global__ __launch_bounds__(kWg) void bt_octant_flags_kernel(const uint16_t *bases, uint32_t B, uint32_t *mixed_zyx,
                                                              uint32_t *empty) {
  const uint32_t S = B + 1, n = 1u << (3 * S), m = (1u << S) - 1u;
  for (uint32_t q = blockIdx.x * kWg + threadIdx.x; q < n; q += gridDim.x * kWg) {
    const uint32_t z = q >> (2 * S), y = (q >> S) & m, x = q & m;
    const uint32_t bp = spread3_dev(x) | (spread3_dev(y) << 1) | (spread3_dev(z) << 2);
    const bool mx = dense_mixed(bases, B, bp);
    mixed_zyx[q] = mx ? 1u : 0u;
    empty[bp] = mx ? 0u : 1u;
  }
}

__global__ __launch_bounds__(kWg) void bt_scatter_index_kernel(const uint32_t *flag, const uint32_t *rank, uint32_t n,
                                                               uint32_t *out) {
  for (uint32_t i = blockIdx.x * kWg + threadIdx.x; i < n; i += gridDim.x * kWg)
    if (flag[i]) out[rank[i]] = i;
}

// Cells of the internal blocks of the dense level, from their x-outer positions.
__global__ __launch_bounds__(kWg) void bt_dense_cells_kernel(const uint32_t *flag, const uint32_t *rank, uint32_t L0,
                                                             uint3 *cells) {
  const uint32_t n = 1u << (3 * L0), m = (1u << L0) - 1u;
  for (uint32_t p = blockIdx.x * kWg + threadIdx.x; p < n; p += gridDim.x * kWg)
    if (flag[p]) cells[rank[p]] = make_uint3(p >> (2 * L0), (p >> L0) & m, p & m);
}

__device__ __forceinline__ uint3 child_cell(uint3 c, uint32_t i) {
  return make_uint3(2u * c.x + (i & 1u), 2u * c.y + ((i >> 1) & 1u), 2u * c.z + ((i >> 2) & 1u));
}

// Level L (> L0) nodes: node k is child k & 7 of internal block k >> 3 of level L-1.
__global__ __launch_bounds__(kWg) void bt_level_flags_kernel(GeomView g, const uint3 *parent_cells, uint32_t n,
                                                             uint32_t L, uint32_t max_depth, uint32_t *flag) {
  for (uint32_t k = blockIdx.x * kWg + threadIdx.x; k < n; k += gridDim.x * kWg) {
    bool f = false;
    if (L < max_depth) {
      const uint3 c = child_cell(parent_cells[k >> 3], k & 7u);
      f = has_children_bits(geom_chunk(g, L, c.x, c.y, c.z));
    }
    flag[k] = f ? 1u : 0u;
  }
}

__global__ __launch_bounds__(kWg) void bt_level_cells_kernel(const uint3 *parent_cells, const uint32_t *flag,
                                                             const uint32_t *rank, uint32_t n, uint3 *cells) {
  for (uint32_t k = blockIdx.x * kWg + threadIdx.x; k < n; k += gridDim.x * kWg)
    if (flag[k]) cells[rank[k]] = child_cell(parent_cells[k >> 3], k & 7u);
}

// Bottom-up: allocations (groups) and block writes in the subtree of internal block j.
__global__ __launch_bounds__(kWg) void bt_sizes_kernel(uint32_t n_int, const uint32_t *cflag, const uint32_t *crank,
                                                       const uint32_t *calloc, const uint32_t *cwrites,
                                                       uint32_t *alloc, uint32_t *writes) {
  for (uint32_t j = blockIdx.x * kWg + threadIdx.x; j < n_int; j += gridDim.x * kWg) {
    uint32_t a = 1, w = 8;
    if (cflag) {
      for (uint32_t i = 0; i < 8u; i++) {
        const uint32_t k = 8u * j + i;
        if (cflag[k]) {
          a += calloc[crank[k]];
          w += cwrites[crank[k]];
        }
      }
    }
    alloc[j] = a;
    writes[j] = w;
  }
}

// Top-down: the pre-order rank and first write position of each internal child, and the
// write position of every child block of internal block j.
__global__ __launch_bounds__(kWg) void bt_ranks_kernel(uint32_t n_int, const uint32_t *rank, const uint32_t *start,
                                                       const uint32_t *cflag, const uint32_t *crank,
                                                       const uint32_t *calloc, const uint32_t *cwrites,
                                                       uint32_t *child_rank, uint32_t *child_start,
                                                       uint32_t *child_post) {
  for (uint32_t j = blockIdx.x * kWg + threadIdx.x; j < n_int; j += gridDim.x * kWg) {
    uint32_t r = rank[j] + 1u, w = start[j];
    for (uint32_t i = 0; i < 8u; i++) {
      const uint32_t k = 8u * j + i;
      if (cflag && cflag[k]) {
        const uint32_t c = crank[k];
        child_rank[c] = r;
        child_start[c] = w;
        r += calloc[c];
        w += cwrites[c];
      }
      child_post[k] = w;
      w += 1u;
    }
  }
}

// FIFO entry r: the sorted empty level-(B+1) cells first, then groups dense_groups..
__device__ __forceinline__ uint32_t fifo_group(const GroupFifo &f, uint32_t r) {
  return r < f.n_empties ? f.empties[r] : f.dense_groups + (r - f.n_empties);
}

__device__ __forceinline__ void bt_store(const BuildOut &o, uint32_t a, uint4 b, uint32_t post) {
  o.blocks[a] = b;
  o.post[a] = post;
  atomicMax(&o.max[0], (unsigned long long)a);
}

// The dense level's blocks (written first, in PopulateInitialBlocks' order).
__global__ __launch_bounds__(kWg) void bt_write_dense_kernel(GeomView g, const uint16_t *bases, uint32_t B,
                                                             const uint32_t *flag, const uint32_t *irank,
                                                             const uint32_t *rank, const uint32_t *mixed_rank,
                                                             GroupFifo fifo, BuildOut o) {
  const uint32_t L0 = B + 2, n = 1u << (3 * L0), m = (1u << L0) - 1u, S = B + 1;
  for (uint32_t p = blockIdx.x * kWg + threadIdx.x; p < n; p += gridDim.x * kWg) {
    const uint32_t x = p >> (2 * L0), y = (p >> L0) & m, z = p & m;
    const uint32_t la = spread3_dev(x) | (spread3_dev(y) << 1) | (spread3_dev(z) << 2);
    if (!dense_mixed(bases, B, la >> 3)) continue;
    uint4 b = geom_block(g, L0, x, y, z);
    if (flag[p]) {
      const uint32_t r = rank[irank[p]];
      if (r < fifo.capacity) {
        const uint32_t g = fifo_group(fifo, r);
        b.x = g << 3;
        o.group_parent[g] = la;
      }
    }
    const uint32_t q = ((z >> 1) << (2 * S)) | ((y >> 1) << S) | (x >> 1);  // its level-(B+1) cell, z-outer
    bt_store(o, la, b, 8u * mixed_rank[q] + (la & 7u));
    atomicMax(&o.max[1], (unsigned long long)la);
  }
}

// Level L (> L0) nodes whose parent got its group.
__global__ __launch_bounds__(kWg) void bt_write_level_kernel(GeomView g, const uint3 *parent_cells,
                                                             const uint32_t *parent_rank, uint32_t n, uint32_t L,
                                                             const uint32_t *flag, const uint32_t *irank,
                                                             const uint32_t *rank, const uint32_t *post,
                                                             GroupFifo fifo, BuildOut o) {
  for (uint32_t k = blockIdx.x * kWg + threadIdx.x; k < n; k += gridDim.x * kWg) {
    const uint32_t pr = parent_rank[k >> 3];
    if (pr >= fifo.capacity) continue;  // the FIFO ran out before the parent's group
    const uint32_t a = (fifo_group(fifo, pr) << 3) + (k & 7u);
    const uint3 c = child_cell(parent_cells[k >> 3], k & 7u);
    uint4 b = geom_block(g, L, c.x, c.y, c.z);
    if (flag && flag[k]) {
      const uint32_t r = rank[irank[k]];
      if (r < fifo.capacity) {
        const uint32_t g = fifo_group(fifo, r);
        b.x = g << 3;
        o.group_parent[g] = a;
      }
    }
    bt_store(o, a, b, post[k]);
  }
}

// Initial usages of a built tree (svo_tree_initial_usages): dense groups inviolate, the
// allocated ones at `tick` with their parent, the rest free.
__global__ __launch_bounds__(kWg) void bt_usages_kernel(const uint32_t *group_parent, uint32_t n_groups,
                                                        uint32_t dense_groups, uint32_t tick, svo_usage *u) {
  for (uint32_t g = blockIdx.x * kWg + threadIdx.x; g < n_groups; g += gridDim.x * kWg) {
    const uint32_t parent = group_parent[g];
    svo_usage v;
    v.pad = 0;
    if (g < dense_groups) {
      v.tick = 65535;
      v.parent = parent;
    } else if (parent != kU32Max) {
      v.tick = (uint16_t)tick;
      v.parent = parent;
    } else {
      v.tick = 0;
      v.parent = kU32Max;
    }
    u[g] = v;
  }
}

inline uint32_t grid_for(size_t n) { return (uint32_t)std::min<size_t>((n + kWg - 1) / kWg, 65535 * 16); }
inline int rc_of() { return hipGetLastError() == hipSuccess ? 0 : -1; }

}  // namespace

int launch_terrain(const TerrainView &t, hipStream_t s) {
  const uint32_t res = 1u << t.spec.lres;
  hipLaunchKernelGGL(terrain_top_kernel, dim3(grid_for((size_t)res * res)), dim3(kWg), 0, s, t.spec, t.top);
  if (rc_of()) return -1;
  for (int L = (int)t.spec.lres - 1; L >= 0; L--) {
    const uint32_t n = 1u << L;
    const float *imn = (uint32_t)(L + 1) == t.spec.lres ? t.top : t.hmin + t.level_offset[L + 1];
    const float *imx = (uint32_t)(L + 1) == t.spec.lres ? t.top : t.hmax + t.level_offset[L + 1];
    hipLaunchKernelGGL(terrain_reduce_kernel, dim3(grid_for((size_t)n * n)), dim3(kWg), 0, s, imn, imx,
                       t.hmin + t.level_offset[L], t.hmax + t.level_offset[L], n);
    if (rc_of()) return -1;
  }
  return 0;
}

int launch_seed(const TerrainView &t, uint16_t *bases, uint32_t B, uint4 *blocks, hipStream_t s) {
  for (uint32_t d = 1; d <= B; d++) {
    const uint32_t offset = ((1u << (3u * d)) - 1u) / 7u - 1u;  // powSum(d-1)
    hipLaunchKernelGGL(seed_bases_kernel, dim3(grid_for(1ull << (3 * d))), dim3(kWg), 0, s, t, bases, d, offset);
    if (rc_of()) return -1;
  }
  hipLaunchKernelGGL(seed_dense_kernel, dim3(grid_for(1ull << (3 * (B + 2)))), dim3(kWg), 0, s, t, bases, B, blocks);
  return rc_of();
}

int launch_request_count(const uint32_t *request_id, uint32_t stored, uint32_t *counters, hipStream_t s) {
  hipLaunchKernelGGL(request_count_kernel, dim3(1), dim3(64), 0, s, request_id, stored, counters);
  return rc_of();
}

int launch_request_unique(const RequestSource &req, uint32_t n_max, const uint32_t *n, unsigned long long *first,
                          uint32_t gen, const uint4 *blocks, uint32_t n_blocks, uint32_t max_depth, uint32_t *flags,
                          hipStream_t s) {
  if (!n_max) return 0;
  const uint32_t grid = std::min((n_max + kWg - 1) / kWg, kRequestGrid);
  hipLaunchKernelGGL(request_first_kernel, dim3(grid), dim3(kWg), 0, s, req, n_max, n, first, gen, n_blocks, flags);
  if (rc_of()) return -1;
  hipLaunchKernelGGL(request_flag_kernel, dim3(grid), dim3(kWg), 0, s, req, n_max, n, first, gen, blocks, max_depth,
                     flags);
  return rc_of();
}

int launch_bricks(const TerrainView &t, const RequestSource &req, const uint32_t *list, uint32_t nb_max,
                  const uint32_t *unique, svo_grafting *records, uint4 *bricks, hipStream_t s) {
  if (!nb_max) return 0;
  hipLaunchKernelGGL(brick_kernel, dim3((uint32_t)((8ull * nb_max + kWg - 1) / kWg)), dim3(kWg), 0, s, t, req, list,
                     nb_max, unique, records, bricks);
  return rc_of();
}

int launch_rank(const uint32_t *flags, uint32_t n, uint32_t *counts, uint32_t *rank, uint32_t *total,
                uint32_t *span_sums, hipStream_t s, uint32_t *list, uint32_t limit, const uint32_t *n_dev) {
  const uint32_t chunks = (n + kSelChunk - 1) / kSelChunk;
  if (!chunks) return 0;
  hipLaunchKernelGGL(flag_count_kernel, dim3(chunks), dim3(kWg), 0, s, flags, n, n_dev, counts);
  if (rc_of()) return -1;
  if (launch_scan_u32(counts, counts, chunks, total, span_sums, s)) return -1;
  hipLaunchKernelGGL(flag_rank_kernel, dim3(chunks), dim3(kWg), 0, s, flags, n, n_dev, counts, rank, list, limit);
  return rc_of();
}

uint32_t select_chunks(uint32_t lo, uint32_t hi) { return hi > lo ? (hi - lo + kSelChunk - 1) / kSelChunk : 0u; }

int launch_select_count(const svo_usage *u, uint32_t lo, uint32_t hi, uint32_t pred, uint32_t *counts, hipStream_t s) {
  const uint32_t chunks = select_chunks(lo, hi);
  if (!chunks) return 0;
  hipLaunchKernelGGL(select_count_kernel, dim3(chunks), dim3(kWg), 0, s, u, lo, hi, pred, counts);
  return rc_of();
}

int launch_select_scatter(const svo_usage *u, uint32_t lo, uint32_t hi, uint32_t pred, const uint32_t *offsets,
                          uint32_t *list, hipStream_t s) {
  const uint32_t chunks = select_chunks(lo, hi);
  if (!chunks) return 0;
  hipLaunchKernelGGL(select_scatter_kernel, dim3(chunks), dim3(kWg), 0, s, u, lo, hi, pred, offsets, list);
  return rc_of();
}

int launch_compact_moves(const PoolView &p, const uint32_t *from, const uint32_t *to, uint32_t m, uint32_t *remap,
                         hipStream_t s) {
  hipLaunchKernelGGL(identity_kernel, dim3(grid_for(p.n_usages)), dim3(kWg), 0, s, remap, p.n_usages);
  if (rc_of() || !m) return rc_of();
  hipLaunchKernelGGL(remap_kernel, dim3((m + kWg - 1) / kWg), dim3(kWg), 0, s, from, to, m, remap);
  if (rc_of()) return -1;
  hipLaunchKernelGGL(move_kernel, dim3((8ull * m + kWg - 1) / kWg), dim3(kWg), 0, s, p, from, to, m, remap);
  if (rc_of()) return -1;
  hipLaunchKernelGGL(relink_kernel, dim3((8ull * m + kWg - 1) / kWg), dim3(kWg), 0, s, p, from, to, m, remap);
  if (rc_of()) return -1;
  hipLaunchKernelGGL(release_moved_kernel, dim3((m + kWg - 1) / kWg), dim3(kWg), 0, s, p, from, m);
  return rc_of();
}

// ---- device TreeBuilder ----

int launch_bt_bases(const GeomView &g, uint16_t *bases, uint32_t B, hipStream_t s) {
  for (uint32_t d = 1; d <= B; d++) {
    const uint32_t offset = ((1u << (3u * d)) - 1u) / 7u - 1u;  // powSum(d-1)
    hipLaunchKernelGGL(bt_bases_kernel, dim3(grid_for(1ull << (3 * d))), dim3(kWg), 0, s, g, bases, d, offset);
    if (rc_of()) return -1;
  }
  return 0;
}

int launch_bt_octant_flags(const uint16_t *bases, uint32_t B, uint32_t *mixed_zyx, uint32_t *empty, hipStream_t s) {
  hipLaunchKernelGGL(bt_octant_flags_kernel, dim3(grid_for(1ull << (3 * (B + 1)))), dim3(kWg), 0, s, bases, B,
                     mixed_zyx, empty);
  return rc_of();
}

int launch_bt_dense_flags(const GeomView &g, const uint16_t *bases, uint32_t B, uint32_t max_depth, uint32_t *flag,
                          hipStream_t s) {
  hipLaunchKernelGGL(bt_dense_flags_kernel, dim3(grid_for(1ull << (3 * (B + 2)))), dim3(kWg), 0, s, g, bases, B,
                     max_depth, flag);
  return rc_of();
}

int launch_bt_scatter_index(const uint32_t *flag, const uint32_t *rank, uint32_t n, uint32_t *out, hipStream_t s) {
  if (!n) return 0;
  hipLaunchKernelGGL(bt_scatter_index_kernel, dim3(grid_for(n)), dim3(kWg), 0, s, flag, rank, n, out);
  return rc_of();
}

int launch_bt_dense_cells(const uint32_t *flag, const uint32_t *rank, uint32_t L0, uint3 *cells, hipStream_t s) {
  hipLaunchKernelGGL(bt_dense_cells_kernel, dim3(grid_for(1ull << (3 * L0))), dim3(kWg), 0, s, flag, rank, L0, cells);
  return rc_of();
}

int launch_bt_level_flags(const GeomView &g, const uint3 *parent_cells, uint32_t n, uint32_t L, uint32_t max_depth,
                          uint32_t *flag, hipStream_t s) {
  if (!n) return 0;
  hipLaunchKernelGGL(bt_level_flags_kernel, dim3(grid_for(n)), dim3(kWg), 0, s, g, parent_cells, n, L, max_depth, flag);
  return rc_of();
}

int launch_bt_level_cells(const uint3 *parent_cells, const uint32_t *flag, const uint32_t *rank, uint32_t n,
                          uint3 *cells, hipStream_t s) {
  if (!n) return 0;
  hipLaunchKernelGGL(bt_level_cells_kernel, dim3(grid_for(n)), dim3(kWg), 0, s, parent_cells, flag, rank, n, cells);
  return rc_of();
}

int launch_bt_sizes(uint32_t n_int, const uint32_t *cflag, const uint32_t *crank, const uint32_t *calloc,
                    const uint32_t *cwrites, uint32_t *alloc, uint32_t *writes, hipStream_t s) {
  if (!n_int) return 0;
  hipLaunchKernelGGL(bt_sizes_kernel, dim3(grid_for(n_int)), dim3(kWg), 0, s, n_int, cflag, crank, calloc, cwrites,
                     alloc, writes);
  return rc_of();
}

int launch_bt_ranks(uint32_t n_int, const uint32_t *rank, const uint32_t *start, const uint32_t *cflag,
                    const uint32_t *crank, const uint32_t *calloc, const uint32_t *cwrites, uint32_t *child_rank,
                    uint32_t *child_start, uint32_t *child_post, hipStream_t s) {
  if (!n_int) return 0;
  hipLaunchKernelGGL(bt_ranks_kernel, dim3(grid_for(n_int)), dim3(kWg), 0, s, n_int, rank, start, cflag, crank, calloc,
                     cwrites, child_rank, child_start, child_post);
  return rc_of();
}

int launch_bt_write_dense(const GeomView &g, const uint16_t *bases, uint32_t B, const uint32_t *flag,
                          const uint32_t *irank, const uint32_t *rank, const uint32_t *mixed_rank,
                          const GroupFifo &fifo, const BuildOut &o, hipStream_t s) {
  hipLaunchKernelGGL(bt_write_dense_kernel, dim3(grid_for(1ull << (3 * (B + 2)))), dim3(kWg), 0, s, g, bases, B, flag,
                     irank, rank, mixed_rank, fifo, o);
  return rc_of();
}

int launch_bt_write_level(const GeomView &g, const uint3 *parent_cells, const uint32_t *parent_rank, uint32_t n,
                          uint32_t L, const uint32_t *flag, const uint32_t *irank, const uint32_t *rank,
                          const uint32_t *post, const GroupFifo &fifo, const BuildOut &o, hipStream_t s) {
  if (!n) return 0;
  hipLaunchKernelGGL(bt_write_level_kernel, dim3(grid_for(n)), dim3(kWg), 0, s, g, parent_cells, parent_rank, n, L,
                     flag, irank, rank, post, fifo, o);
  return rc_of();
}

int launch_bt_usages(const uint32_t *group_parent, uint32_t n_groups, uint32_t dense_groups, uint32_t tick,
                     svo_usage *u, hipStream_t s) {
  if (!n_groups) return 0;
  hipLaunchKernelGGL(bt_usages_kernel, dim3(grid_for(n_groups)), dim3(kWg), 0, s, group_parent, n_groups,
                     dense_groups, tick, u);
  return rc_of();
}

}  // namespace svo
