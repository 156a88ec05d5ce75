// Tree management on the GPU: prune, graft and the collection of free or stale groups
// (wavefront.cl:1097-1350, stated as deterministic batch phases).
//
// A dispatch resolves every record against the tree as it stood when the dispatch began,
// touches the ticks along the resolved paths, frees the detached subtrees level by level
// (a set operation: order-free), and then writes fields where, for every field of every
// target, the record with the highest index wins (a hash table keyed by (field, target)
// holding the largest record index).  Collection ranks candidates in ascending group
// order with an order-preserving count / scan / scatter, one pass per age threshold.

#include "svo_device.hpp"
#include "svo_tree_kernels.hpp"

namespace svo {
namespace {

using dev::kU32Max;
constexpr uint32_t kMulti = 65534u, kInviolate = 65535u, kPeriod = 65533u;
constexpr unsigned long long kEmptyKey = ~0ull;
constexpr int kFlat = 256;    // work-group size of the per-record kernels
constexpr int kScanWg = 1024; // single-work-group scans

enum : uint32_t { kClsChild = 1, kClsParent = 2, kClsTick = 3, kClsPayload = 4, kClsBase = 5, kClsGraft = 6 };

__device__ __forceinline__ unsigned long long win_key(uint32_t cls, uint32_t target) {
  return ((unsigned long long)cls << 32) | target;
}
__device__ __forceinline__ uint32_t win_hash(unsigned long long key, uint32_t mask) {
  return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 32) & mask;
}

// Records `rec` as a writer of `key`; the table keeps the largest record index.
__device__ void win_insert(unsigned long long *keys, uint32_t *vals, uint32_t mask, unsigned long long key,
                           uint32_t rec) {
  uint32_t h = win_hash(key, mask);
  for (uint32_t probe = 0; probe <= mask; probe++) {
    unsigned long long k = keys[h];
    if (k == kEmptyKey) k = atomicCAS(&keys[h], kEmptyKey, key);
    if (k == kEmptyKey || k == key) {
      atomicMax(&vals[h], rec);
      return;
    }
    h = (h + 1u) & mask;
  }
}

// The winning record of `key` (kU32Max when absent).
__device__ uint32_t win_get(const unsigned long long *keys, const uint32_t *vals, uint32_t mask,
                            unsigned long long key) {
  uint32_t h = win_hash(key, mask);
  for (uint32_t probe = 0; probe <= mask; probe++) {
    const unsigned long long k = keys[h];
    if (k == key) return vals[h];
    if (k == kEmptyKey) return kU32Max;
    h = (h + 1u) & mask;
  }
  return kU32Max;
}

__device__ __forceinline__ uint32_t tick_age(uint32_t t, uint32_t now) { return t <= now ? now - t : now + kPeriod - t; }

__device__ __forceinline__ bool stale(uint32_t t, uint32_t now, uint32_t offset) {
  if (t == 0) return true;
  if (t >= kMulti) return false;
  return tick_age(t, now) > offset;
}

__device__ __forceinline__ void touch(const PoolView &p, uint32_t g) {
  if (g < p.n_usages && p.usages[g].tick < kMulti) p.usages[g].tick = (uint16_t)p.now;
}

// findAddress's walk (wavefront.cl:1191-1211) from the dense level towards `depth`.
__device__ uint32_t walk(const PoolView &p, const uint64_t *loc, uint32_t depth, bool do_touch, uint32_t *addr,
                         uint32_t *level) {
  const uint64_t x = loc[0], y = loc[1], z = loc[2];
  uint32_t a = dev::morton_top(p.B + 2u, x, y, z);
  for (uint32_t L = p.B + 2u; L < depth; L++) {
    if (a >= p.n_blocks) return kResInvalid;
    if (do_touch) touch(p, a >> 3);
    const uint32_t c = p.blocks[a].child;
    if (c == kU32Max) {
      *addr = a;
      *level = L;
      return kResMiss;
    }
    a = c + dev::chunk_position(L, x, y, z);
  }
  if (a >= p.n_blocks) return kResInvalid;
  *addr = a;
  return kResOk;
}

__device__ ResolveOut resolve_target(const PoolView &p, uint32_t depth, uint32_t address, const uint64_t *locs,
                                     uint32_t n_locs) {
  ResolveOut r{kResInvalid, 0u, 0u, 0u};
  if (depth == 0xffu) {
    if (address >= p.n_blocks) return r;
    const uint32_t g = address >> 3;
    r.addr = address;
    r.status = (g < p.n_usages && p.usages[g].tick == 0) ? kResSkip : kResOk;
    return r;
  }
  if (address >= n_locs || depth < p.B + 2u || depth > 63u) return r;
  r.status = walk(p, locs + 3ull * address, depth, false, &r.addr, &r.level);
  return r;
}

// Frontier append (order-free: the free walk is a set operation).
// counters[slot] += 1 for every active lane, with one atomic per distinct slot per wave
// (every lane of a wave otherwise hits the same few counters).
__device__ __forceinline__ void wave_count(uint32_t *counters, uint32_t slot) {
  const uint32_t lane = threadIdx.x & 63u;
  uint64_t pending = __ballot(1);
  while (pending) {
    const uint32_t leader = (uint32_t)__ffsll((long long)pending) - 1u;
    const uint32_t s = __shfl(slot, (int)leader);
    const uint64_t same = __ballot(slot == s);
    if (lane == leader) atomicAdd(&counters[s], (uint32_t)__popcll(same));
    pending &= ~same;
  }
}

__device__ __forceinline__ void push(uint2 *front, uint32_t *count, uint32_t cap, uint32_t *overflow, uint32_t g,
                                     uint32_t pb) {
  const uint32_t i = atomicAdd(count, 1u);
  if (i < cap)
    front[i] = make_uint2(g, pb);
  else
    atomicOr(overflow, 1u);
}

// Exclusive scan of one value per thread over a work-group of WG threads; returns the
// thread's prefix and leaves the group's total in *total (all threads).
template <int WG>
__device__ uint32_t block_scan(uint32_t v, uint32_t *s_waves, uint32_t *total) {
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
    const uint32_t t = s_waves[w];
    before += w < wave ? t : 0u;
    all += t;
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

// ---------------------------------------------------------------------------
// generic
// ---------------------------------------------------------------------------

__global__ void fill_u32_kernel(uint32_t *p, uint32_t v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void fill_u64_kernel(unsigned long long *p, unsigned long long v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// Exclusive scans over spans of kScanSpan elements.  Wave w of the work-group owns 64 *
// kScanPer consecutive elements of a span and reads and writes them in kScanPer rounds of
// 64 consecutive elements (coalesced); the rounds' sums chain within the wave and the wave
// sums chain over the work-group.  in may equal out: each element is read and written by
// the same lane.
constexpr uint32_t kScanPer = 16;
constexpr uint32_t kScanSpan = kScanWg * kScanPer;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(x, d, 64);
    if ((int)lane >= d) x += t;
  }
  return x;
}

// Scans [s0, s0 + kScanSpan) clipped to n, starting from base; returns the span's sum.
__device__ uint32_t scan_span(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t s0, uint32_t base,
                              uint32_t *s_waves) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t w0 = s0 + wave * (64u * kScanPer);
  uint32_t v[kScanPer], incl[kScanPer], before[kScanPer];
#pragma unroll
  for (uint32_t r = 0; r < kScanPer; r++) {
    const uint32_t i = w0 + r * 64u + lane;
    v[r] = i < n ? in[i] : 0u;
  }
  uint32_t run = 0;
#pragma unroll
  for (uint32_t r = 0; r < kScanPer; r++) {
    incl[r] = wave_incl_scan(v[r]);
    before[r] = run;
    run += (uint32_t)__builtin_amdgcn_readlane((int)incl[r], 63);
  }
  if (lane == 0) s_waves[wave] = run;
  __syncthreads();
  uint32_t wbefore = 0, all = 0;
  for (uint32_t w = 0; w < (uint32_t)(kScanWg / 64); w++) {
    const uint32_t t = s_waves[w];
    wbefore += w < wave ? t : 0u;
    all += t;
  }
  __syncthreads();
#pragma unroll
  for (uint32_t r = 0; r < kScanPer; r++) {
    const uint32_t i = w0 + r * 64u + lane;
    if (i < n) out[i] = base + wbefore + before[r] + incl[r] - v[r];
  }
  return all;
}

// Single-work-group exclusive scan of n elements.
__global__ __launch_bounds__(kScanWg) void scan_u32_kernel(const uint32_t *in, uint32_t *out, uint32_t n,
                                                           uint32_t *total) {
  __shared__ uint32_t s_waves[kScanWg / 64];
  uint32_t carry = 0;
  for (uint32_t s0 = 0; s0 < n; s0 += kScanSpan) carry += scan_span(in, out, n, s0, carry, s_waves);
  if (threadIdx.x == 0) *total = carry;
}

// Multi-work-group scan of a large array: block b sums span b, one work-group scans the
// span sums, block b scans span b from its offset.
__global__ __launch_bounds__(kScanWg) void scan_reduce_kernel(const uint32_t *in, uint32_t n, uint32_t *partial) {
  __shared__ uint32_t s_waves[kScanWg / 64];
  const uint32_t s0 = blockIdx.x * kScanSpan;
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t r = 0; r < kScanPer; r++) {
    const uint32_t i = s0 + r * kScanWg + threadIdx.x;
    sum += i < n ? in[i] : 0u;
  }
  uint32_t all;
  (void)block_scan<kScanWg>(sum, s_waves, &all);
  if (threadIdx.x == 0) partial[blockIdx.x] = all;
}

__global__ __launch_bounds__(kScanWg) void scan_apply_kernel(const uint32_t *in, uint32_t *out, uint32_t n,
                                                             const uint32_t *offset) {
  __shared__ uint32_t s_waves[kScanWg / 64];
  (void)scan_span(in, out, n, blockIdx.x * kScanSpan, offset[blockIdx.x], s_waves);
}

// ---------------------------------------------------------------------------
// resolution, child requests, touches
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kFlat) void prune_resolve_kernel(PoolView p, const svo_pruning *rec, uint32_t n,
                                                              uint32_t n_surface, const uint64_t *locs,
                                                              uint32_t n_locs, ResolveOut *res, uint32_t *counters) {
  const uint32_t i = blockIdx.x * kFlat + threadIdx.x;
  if (i >= n) return;
  const svo_pruning r = rec[i];
  ResolveOut o{kResInvalid, 0u, 0u, 0u};
  if ((r.properties >> 4) & 1u) {
    o.status = r.address < p.n_bases ? kResBase : kResInvalid;
    o.addr = r.address;
  } else if (((r.properties >> 3) & 1u) && r.colour_address >= n_surface) {
    o.status = kResInvalid;
  } else {
    o = resolve_target(p, r.depth, r.address, locs, n_locs);
  }
  res[i] = o;
  if (o.status == kResMiss) counters[kTcMisses] = 1u;
  const uint32_t slot = (o.status == kResOk || o.status == kResBase) ? kTcApplied
                        : o.status == kResInvalid                   ? kTcInvalid
                                                                    : kTcSkipped;
  wave_count(counters, slot);
}

__global__ __launch_bounds__(kFlat) void graft_resolve_kernel(PoolView p, const svo_grafting *rec, uint32_t n,
                                                              const uint32_t *need, const uint64_t *locs,
                                                              uint32_t n_locs, ResolveOut *res, uint32_t *counters) {
  const uint32_t i = blockIdx.x * kFlat + threadIdx.x;
  if (i >= n) return;
  ResolveOut o{kResInvalid, 0u, 0u, 0u};
  if (need[i]) o = resolve_target(p, rec[i].depth, rec[i].graft_address, locs, n_locs);
  res[i] = o;
  if (o.status == kResMiss) counters[kTcMisses] = 1u;
}

struct RecordFields {
  const uint8_t *base;
  uint32_t stride, depth_off, addr_off;
  __device__ uint32_t depth(uint32_t i) const { return base[(size_t)i * stride + depth_off]; }
  __device__ uint32_t address(uint32_t i) const {
    return *reinterpret_cast<const uint32_t *>(base + (size_t)i * stride + addr_off);
  }
};

// requestChild from findAddress (wavefront.cl:1201-1205): misses append in record order
// after the current ChildRequestID; beyond capacity they are counted only.
__global__ __launch_bounds__(kScanWg) void request_scan_kernel(PoolView p, RecordFields f, uint32_t n,
                                                               const uint64_t *locs, const ResolveOut *res,
                                                               uint32_t *counters) {
  // no record missed (every graft of the streaming update addresses its block directly):
  // nothing to append, and ChildRequestID and the request count stay as they are
  if (counters[kTcMisses] == 0u) return;
  __shared__ uint32_t s_waves[kScanWg / 64];
  __shared__ uint32_t s_base;
  if (threadIdx.x == 0) s_base = p.request_count ? *p.request_count : 0u;
  __syncthreads();
  const uint32_t base = s_base;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t carry = 0;
  // wave w takes 64 * kScanPer consecutive records of each span in rounds of 64 (one a lane)
  for (uint32_t s0 = 0; s0 < n; s0 += kScanSpan) {
    const uint32_t w0 = s0 + wave * (64u * kScanPer);
    uint64_t m[kScanPer];
    uint32_t cnt = 0;
#pragma unroll
    for (uint32_t r = 0; r < kScanPer; r++) {
      const uint32_t i = w0 + r * 64u + lane;
      m[r] = __ballot(i < n && res[i].status == kResMiss);
      cnt += (uint32_t)__popcll(m[r]);
    }
    if (lane == 0) s_waves[wave] = cnt;
    __syncthreads();
    uint32_t wbefore = 0, all = 0;
    for (uint32_t w = 0; w < (uint32_t)(kScanWg / 64); w++) {
      const uint32_t t = s_waves[w];
      wbefore += w < wave ? t : 0u;
      all += t;
    }
    __syncthreads();
    uint32_t id = base + carry + wbefore;
#pragma unroll
    for (uint32_t r = 0; r < kScanPer; r++) {
      const uint32_t i = w0 + r * 64u + lane;
      const uint32_t at = id + __builtin_amdgcn_mbcnt_hi((uint32_t)(m[r] >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)m[r], 0u));
      if (((m[r] >> lane) & 1ull) && p.request_count && at < p.request_capacity) {
        const ResolveOut o = res[i];
        const uint64_t *loc = locs + 3ull * f.address(i);
        svo_child_request q;
        q.address = o.addr;
        q.tick = (uint16_t)p.now;
        q.depth = (uint8_t)o.level;
        q.pad0 = 0;
        q.location[0] = loc[0];
        q.location[1] = loc[1];
        q.location[2] = loc[2];
        q.tree_size = (uint8_t)(f.depth(i) - o.level);
        for (int k = 0; k < 7; k++) q.pad1[k] = 0;
        p.requests[at] = q;
      }
      id += (uint32_t)__popcll(m[r]);
    }
    carry += all;
  }
  if (threadIdx.x == 0) {
    if (p.request_count) *p.request_count = base + carry;
    counters[kTcRequests] += carry;
  }
}

__global__ __launch_bounds__(kFlat) void touch_kernel(PoolView p, RecordFields f, uint32_t n, const uint64_t *locs,
                                                      const ResolveOut *res) {
  const uint32_t i = blockIdx.x * kFlat + threadIdx.x;
  if (i >= n) return;
  const ResolveOut o = res[i];
  if (o.status != kResOk && o.status != kResMiss) return;
  const uint32_t depth = f.depth(i);
  if (depth == 0xffu) {
    touch(p, o.addr >> 3);
  } else {
    uint32_t a, l;
    walk(p, locs + 3ull * f.address(i), depth, true, &a, &l);
  }
}

// ---------------------------------------------------------------------------
// prune
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kFlat) void prune_winners_kernel(PoolView p, const svo_pruning *rec, uint32_t n,
                                                              const ResolveOut *res, unsigned long long *keys,
                                                              uint32_t *vals, uint32_t mask) {
  const uint32_t i = blockIdx.x * kFlat + threadIdx.x;
  if (i >= n) return;
  const ResolveOut o = res[i];
  const svo_pruning r = rec[i];
  if (o.status == kResBase) win_insert(keys, vals, mask, win_key(kClsBase, o.addr), i);
  if (o.status != kResOk) return;
  const uint32_t a = o.addr;
  if (r.properties & 1u) {
    win_insert(keys, vals, mask, win_key(kClsChild, a), i);
    if (r.child_address != kU32Max) win_insert(keys, vals, mask, win_key(kClsParent, r.child_address >> 3), i);
  }
  const uint32_t v = (r.properties >> 1) & 3u;
  if ((v == 1u || v == 3u) && (a >> 3) < p.n_usages) win_insert(keys, vals, mask, win_key(kClsTick, a >> 3), i);
  if ((r.properties >> 3) & 1u) win_insert(keys, vals, mask, win_key(kClsPayload, a), i);
}

// CullChild roots: the group each culled block holds when the dispatch begins.
__global__ __launch_bounds__(kFlat) void prune_roots_kernel(PoolView p, const svo_pruning *rec, uint32_t n,
                                                            const ResolveOut *res, uint2 *front, uint32_t *counters,
                                                            uint32_t cap) {
  const uint32_t i = blockIdx.x * kFlat + threadIdx.x;
  if (i >= n) return;
  if (res[i].status != kResOk || !(rec[i].properties & 1u)) return;
  const uint32_t a = res[i].addr;
  const uint32_t c = p.blocks[a].child;
  if (c != kU32Max) push(front, &counters[kTcFront0], cap, &counters[kTcOverflow], c >> 3, a);
}

__global__ __launch_bounds__(kFlat) void prune_child_kernel(PoolView p, const svo_pruning *rec, uint32_t n,
                                                            const ResolveOut *res, const unsigned long long *keys,
                                                            const uint32_t *vals, uint32_t mask) {
  const uint32_t i = blockIdx.x * kFlat + threadIdx.x;
  if (i >= n) return;
  if (res[i].status != kResOk || !(rec[i].properties & 1u)) return;
  const uint32_t a = res[i].addr, c = rec[i].child_address;
  if (win_get(keys, vals, mask, win_key(kClsChild, a)) == i) p.blocks[a].child = c;
  if (c == kU32Max) return;
  const uint32_t g = c >> 3;
  if (g < p.n_usages && p.usages[g].tick < kMulti && win_get(keys, vals, mask, win_key(kClsParent, g)) == i)
    p.usages[g].parent = a;
}

__global__ __launch_bounds__(kFlat) void prune_fields_kernel(PoolView p, const svo_pruning *rec, uint32_t n,
                                                             const svo_block *surface, const ResolveOut *res,
                                                             const unsigned long long *keys, const uint32_t *vals,
                                                             uint32_t mask) {
  const uint32_t i = blockIdx.x * kFlat + threadIdx.x;
  if (i >= n) return;
  const ResolveOut o = res[i];
  const svo_pruning r = rec[i];
  if (o.status == kResBase) {
    if (win_get(keys, vals, mask, win_key(kClsBase, o.addr)) == i) p.bases[o.addr] = r.chunk;
    return;
  }
  if (o.status != kResOk) return;
  const uint32_t a = o.addr, g = a >> 3;
  const uint32_t v = (r.properties >> 1) & 3u;
  if ((v == 1u || v == 3u) && g < p.n_usages && win_get(keys, vals, mask, win_key(kClsTick, g)) == i)
    p.usages[g].tick = (uint16_t)(v == 3u ? kInviolate : p.now);
  if (((r.properties >> 3) & 1u) && win_get(keys, vals, mask, win_key(kClsPayload, a)) == i) {
    // UpdateChunk: every field of the surface record except Child, with the record's chunk
    const uint4 s = reinterpret_cast<const uint4 *>(surface)[r.colour_address];
    uint32_t *w = reinterpret_cast<uint32_t *>(p.blocks + a);
    w[1] = (s.y & 0xffff0000u) | r.chunk;
    w[2] = s.z;
    w[3] = s.w;
  }
}

// ---------------------------------------------------------------------------
// the free walk (dereference / helpDereference, wavefront.cl:1097-1189)
// ---------------------------------------------------------------------------

// One level: frees the n groups of `in` and queues their children into `out`.
__device__ void free_level(const PoolView &p, const uint2 *in, uint32_t n, uint2 *out, uint32_t *out_count,
                           uint32_t cap, uint32_t *overflow, uint2 *events, uint32_t *ev_count, uint32_t ev_cap,
                           const uint32_t *stop) {
  for (uint32_t i = blockIdx.x * kFlat + threadIdx.x; i < n; i += gridDim.x * kFlat) {
    const uint2 e = in[i];
    const uint32_t g = e.x;
    if (g >= p.n_usages) continue;
    if (stop && ((stop[g >> 5] >> (g & 31u)) & 1u)) continue;
    const uint32_t t = p.usages[g].tick;
    if (t == kInviolate) continue;
    if (t == kMulti) {
      const uint32_t k = atomicAdd(ev_count, 1u);
      if (k < ev_cap) events[k] = e;
      continue;
    }
    p.usages[g].tick = 0;
    for (uint32_t k = 0; k < 8u; k++) {
      const uint32_t b = 8u * g + k;
      if (b >= p.n_blocks) break;
      const uint32_t c = p.blocks[b].child;
      if (c != kU32Max) push(out, out_count, cap, overflow, c >> 3, b);
    }
  }
}

__global__ __launch_bounds__(kFlat) void free_level_kernel(PoolView p, const uint2 *in, uint2 *out,
                                                           uint32_t *counters, uint32_t c_in, uint32_t c_out,
                                                           uint32_t c_clear, uint32_t cap, uint2 *events,
                                                           uint32_t ev_cap, const uint32_t *stop) {
  // the slot cleared here is neither read nor appended to by this level
  if (blockIdx.x == 0 && threadIdx.x == 0) counters[front_slot(c_clear)] = 0;
  const uint32_t c = counters[front_slot(c_in)];
  const uint32_t n = c < cap ? c : cap;
  if (n == 0) return;
  free_level(p, in, n, out, &counters[front_slot(c_out)], cap, &counters[kTcOverflow], events, &counters[kTcEvents],
             ev_cap, stop);
}

// Multi-parent unlink (helpDereference, wavefront.cl:1105-1148, as intended), applied in
// ascending (group, parent block) order after the free walk.
__device__ void unlink_parent(const PoolView &p, uint32_t g, uint32_t pb) {
  if (g >= p.n_usages || p.usages[g].tick != kMulti || !p.parents) return;
  uint32_t prev = kU32Max, e = p.usages[g].parent, steps = 0;
  while (e < p.n_parents && steps < p.n_parents) {
    if (p.parents[e].parent_address == pb) break;
    prev = e;
    e = p.parents[e].next_element;
    steps++;
  }
  if (e >= p.n_parents || steps >= p.n_parents) return;
  const uint32_t nx = p.parents[e].next_element;
  if (prev == kU32Max)
    p.usages[g].parent = nx;
  else
    p.parents[prev].next_element = nx;
  p.parents[e].next_element = kU32Max;
  if (p.residency) p.residency[e] = 0;
  const uint32_t h = p.usages[g].parent;
  if (h >= p.n_parents) {
    p.usages[g].tick = 0;
    p.usages[g].parent = kU32Max;
  } else if (p.parents[h].next_element == kU32Max) {
    p.usages[g].tick = (uint16_t)p.now;
    p.usages[g].parent = p.parents[h].parent_address;
    if (p.residency) p.residency[h] = 0;
  }
}

__global__ __launch_bounds__(kScanWg) void unlink_kernel(PoolView p, const uint2 *events, uint32_t cap,
                                                         uint32_t levels, uint32_t *counters) {
  __shared__ unsigned long long s[kUnlinkSortMax];
  const uint32_t n = counters[kTcEvents] < cap ? counters[kTcEvents] : cap;
  uint32_t m = 2;
  while (m < n) m <<= 1;
  for (uint32_t i = threadIdx.x; i < m; i += kScanWg)
    s[i] = i < n ? (((unsigned long long)events[i].x << 32) | events[i].y) : kEmptyKey;
  __syncthreads();
  for (uint32_t k = 2; k <= m; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < m; i += kScanWg) {
        const uint32_t ixj = i ^ j;
        if (ixj > i) {
          const bool up = (i & k) == 0;
          const unsigned long long a = s[i], b = s[ixj];
          if ((a > b) == up) {
            s[i] = b;
            s[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    uint32_t done = 0;
    for (uint32_t i = 0; i < n; i++) {
      if (i && s[i] == s[i - 1]) continue;
      unlink_parent(p, (uint32_t)(s[i] >> 32), (uint32_t)s[i]);
      done++;
    }
    counters[kTcUnlinks] += done;
    // the next walk starts with no events and empty frontiers
    if (counters[kTcEvents] > cap) counters[kTcOverflow] |= kOverflowEvents;
    if (counters[front_slot(levels % 3u)] != 0u) counters[kTcOverflow] |= kOverflowWalk;
    counters[kTcEvents] = 0;
    counters[kTcFront0] = 0;
    counters[kTcFront1] = 0;
    counters[kTcFront2] = 0;
  }
}

// ---------------------------------------------------------------------------
// collection (graft, wavefront.cl:1277-1305, as intended)
// ---------------------------------------------------------------------------

constexpr uint32_t kPerThread = kCollectChunk / kFlat;

// live_only: garbage collection's reclaim, which takes stale live groups and leaves free
// ones where they are.
__device__ __forceinline__ bool candidate(const PoolView &p, uint32_t g, uint32_t ng, uint32_t offset,
                                          const uint32_t *claimed, uint32_t live_only) {
  if (g >= ng) return false;
  if ((claimed[g >> 5] >> (g & 31u)) & 1u) return false;
  const uint32_t t = p.usages[g].tick;
  if (live_only && t == 0u) return false;
  return stale(t, p.now, offset);
}

__global__ __launch_bounds__(kFlat) void collect_count_kernel(PoolView p, uint32_t ng, uint32_t offset,
                                                              uint32_t live_only, const uint32_t *claimed,
                                                              uint32_t *chunk_counts) {
  __shared__ uint32_t s_waves[kFlat / 64];
  const uint32_t g0 = blockIdx.x * kCollectChunk + threadIdx.x * kPerThread;
  uint32_t c = 0;
  for (uint32_t j = 0; j < kPerThread; j++) c += candidate(p, g0 + j, ng, offset, claimed, live_only) ? 1u : 0u;
  uint32_t all;
  block_scan<kFlat>(c, s_waves, &all);
  if (threadIdx.x == 0) chunk_counts[blockIdx.x] = all;
}

// k = min(candidates of this pass, need - collected), taken from list position `collected`
__global__ void collect_limit_kernel(uint32_t *counters, DevCount need) {
  if (threadIdx.x != 0) return;
  const uint32_t got = counters[kTcGot], nd = need.get();
  const uint32_t want = nd > got ? nd - got : 0u, total = counters[kTcTotal];
  counters[kTcK] = total < want ? total : want;
  counters[kTcFrom] = got;
}

__global__ void collect_advance_kernel(uint32_t *counters) {
  if (threadIdx.x == 0) counters[kTcGot] += counters[kTcK];
}

__global__ __launch_bounds__(kFlat) void collect_scatter_kernel(PoolView p, uint32_t ng, uint32_t offset,
                                                                uint32_t live_only, const uint32_t *claimed,
                                                                const uint32_t *chunk_offsets,
                                                                const uint32_t *counters, uint32_t *list) {
  __shared__ uint32_t s_waves[kFlat / 64];
  const uint32_t from = counters[kTcFrom], limit = counters[kTcK];
  const uint32_t g0 = blockIdx.x * kCollectChunk + threadIdx.x * kPerThread;
  bool is[kPerThread];
  uint32_t c = 0;
  for (uint32_t j = 0; j < kPerThread; j++) {
    is[j] = candidate(p, g0 + j, ng, offset, claimed, live_only);
    c += is[j] ? 1u : 0u;
  }
  uint32_t all;
  uint32_t rank = chunk_offsets[blockIdx.x] + block_scan<kFlat>(c, s_waves, &all);
  for (uint32_t j = 0; j < kPerThread; j++) {
    if (!is[j]) continue;
    if (rank < limit) list[from + rank] = g0 + j;
    rank++;
  }
}

// Claims list[from, from+k): marks them, cuts stale live groups from their parents and
// queues what hangs below them for freeing.
__global__ __launch_bounds__(kFlat) void collect_claim_kernel(PoolView p, const uint32_t *list, uint32_t k_max,
                                                              uint32_t *claimed, uint2 *front, uint32_t *counters,
                                                              uint32_t cap) {
  const uint32_t i = blockIdx.x * kFlat + threadIdx.x;
  const uint32_t k = counters[kTcK] < k_max ? counters[kTcK] : k_max;
  if (i >= k) return;
  const uint32_t g = list[counters[kTcFrom] + i];
  atomicOr(&claimed[g >> 5], 1u << (g & 31u));
  if (p.usages[g].tick == 0) return;
  const uint32_t pb = p.usages[g].parent;
  if (pb < p.n_blocks && p.blocks[pb].child == 8u * g) p.blocks[pb].child = kU32Max;
  for (uint32_t j = 0; j < 8u; j++) {
    const uint32_t b = 8u * g + j;
    if (b >= p.n_blocks) break;
    const uint32_t c = p.blocks[b].child;
    if (c != kU32Max) push(front, &counters[kTcFront0], cap, &counters[kTcOverflow], c >> 3, b);
  }
}

// Claimed groups get `tick`: the frame tick when collected for a graft, 0 when reclaimed.
__global__ __launch_bounds__(kFlat) void collect_mark_kernel(PoolView p, const uint32_t *list, uint32_t k_max,
                                                             const uint32_t *counters, uint32_t tick) {
  const uint32_t i = blockIdx.x * kFlat + threadIdx.x;
  const uint32_t k = counters[kTcK] < k_max ? counters[kTcK] : k_max;
  if (i >= k) return;
  const uint32_t g = list[counters[kTcFrom] + i];
  p.usages[g].tick = (uint16_t)tick;
  p.usages[g].parent = kU32Max;
}

// ---------------------------------------------------------------------------
// graft (wavefront.cl:1306-1350, as intended)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kFlat) void graft_size_kernel(const svo_grafting *rec, DevCount n, uint32_t n_gb,
                                                           uint32_t *size) {
  const uint32_t i = blockIdx.x * kFlat + threadIdx.x;
  if (i >= n.v) return;
  if (i >= n.get()) {
    size[i] = 0u;  // beyond the records present: no groups, no need
    return;
  }
  const uint32_t gda = rec[i].graft_data_address, tot = rec[i].graft_total_size;
  const bool ok = tot >= 8u && (tot & 7u) == 0 && (gda & 7u) == 0 && (unsigned long long)gda + tot <= n_gb;
  size[i] = ok ? tot / 8u : 0u;
}

// Pre-order numbering of each graft's groups (the depthHeap walk of wavefront.cl:1322-1349).
__global__ __launch_bounds__(64) void graft_structure_kernel(const svo_grafting *rec, uint32_t n,
                                                             const svo_block *gb, const uint32_t *size,
                                                             const uint32_t *where, uint32_t *number,
                                                             uint32_t *parent, uint32_t *need) {
  const uint32_t r = blockIdx.x * 64u + threadIdx.x;
  if (r >= n) return;
  const uint32_t ng = size[r];
  if (!ng) {
    need[r] = 0;
    return;
  }
  const uint32_t gda = rec[r].graft_data_address, tot = ng * 8u;
  uint32_t *num = number + where[r], *par = parent + where[r];
  for (uint32_t q = 0; q < ng; q++) num[q] = kU32Max;
  uint32_t stack_q[64], stack_k[64];
  uint32_t top = 1, next = 1;
  bool ok = true;
  num[0] = 0;
  par[0] = kU32Max;
  stack_q[0] = 0;
  stack_k[0] = 0;
  while (top && ok) {
    const uint32_t q = stack_q[top - 1];
    if (stack_k[top - 1] == 8u) {
      top--;
      continue;
    }
    const uint32_t k = stack_k[top - 1]++;
    const uint32_t c = gb[gda + 8u * q + k].child;
    if (c == kU32Max) continue;
    if (c < gda || c - gda >= tot || ((c - gda) & 7u)) {
      ok = false;
      break;
    }
    const uint32_t q2 = (c - gda) / 8u;
    if (num[q2] != kU32Max || top == 64u) {
      ok = false;
      break;
    }
    num[q2] = next++;
    par[q2] = 8u * q + k;
    stack_q[top] = q2;
    stack_k[top] = 0;
    top++;
  }
  need[r] = (ok && next == ng) ? ng : 0u;
}

__global__ __launch_bounds__(kFlat) void graft_winners_kernel(const ResolveOut *res, const uint32_t *start,
                                                              const uint32_t *need, uint32_t n, DevCount got_count,
                                                              unsigned long long *keys, uint32_t *vals,
                                                              uint32_t mask) {
  const uint32_t i = blockIdx.x * kFlat + threadIdx.x;
  if (i >= n) return;
  const uint32_t got = got_count.get();
  if (res[i].status == kResOk && start[i] + need[i] <= got) win_insert(keys, vals, mask, win_key(kClsGraft, res[i].addr), i);
}

// Decides which grafts are linked and queues what their targets held before.
__global__ __launch_bounds__(kFlat) void graft_roots_kernel(PoolView p, ResolveOut *res, const uint32_t *start,
                                                            const uint32_t *need, DevCount n, DevCount got_count,
                                                            const unsigned long long *keys, const uint32_t *vals,
                                                            uint32_t mask, uint2 *front, uint32_t *counters,
                                                            uint32_t cap) {
  const uint32_t i = blockIdx.x * kFlat + threadIdx.x;
  if (i >= n.v) return;
  if (i >= n.get()) {  // beyond the records present: not placed, not counted
    res[i].placed = 0u;
    return;
  }
  const uint32_t got = got_count.get();
  const ResolveOut o = res[i];
  const bool placed = o.status == kResOk && start[i] + need[i] <= got &&
                      win_get(keys, vals, mask, win_key(kClsGraft, o.addr)) == i;
  res[i].placed = placed ? 1u : 0u;
  const uint32_t slot = placed ? kTcPlaced : o.status == kResInvalid ? kTcInvalid : kTcSkipped;
  wave_count(counters, slot);
  if (!placed) return;
  wave_count(counters, kTcApplied);
  const uint32_t c = p.blocks[o.addr].child;
  if (c != kU32Max) push(front, &counters[kTcFront0], cap, &counters[kTcOverflow], c >> 3, o.addr);
}

// One thread per needed block of all grafts: copies the blocks into the collected groups
// (Child fields remapped) and sets the usage of every group.  Thread t places block t & 7
// of needed group g = t / 8, which belongs to the last graft r with start[r] <= g (start is
// the exclusive scan of need, so g < start[r] + need[r] for that r).
__global__ __launch_bounds__(kFlat) void graft_place_kernel(PoolView p, const svo_grafting *rec, const svo_block *gb,
                                                            const ResolveOut *res, const uint32_t *start,
                                                            const uint32_t *where, const uint32_t *number,
                                                            const uint32_t *parent, const uint32_t *list, uint32_t n,
                                                            DevCount need_count) {
  const uint32_t t = blockIdx.x * kFlat + threadIdx.x;
  const uint32_t g = t >> 3;
  if (g >= need_count.v || g >= need_count.get()) return;
  uint32_t lo = 0, hi = n;  // start[lo] <= g < start[hi] (start[n] taken as total_need)
  while (hi - lo > 1u) {
    const uint32_t mid = lo + (hi - lo) / 2u;
    if (start[mid] <= g) lo = mid; else hi = mid;
  }
  const uint32_t r = lo;
  if (!res[r].placed) return;
  const uint32_t gda = rec[r].graft_data_address, S = start[r], W = where[r];
  {
    const uint32_t b = 8u * (g - S) + (t & 7u);
    const uint32_t q = b >> 3, k = b & 7u;
    const uint32_t P = list[S + number[W + q]];
    uint4 v = reinterpret_cast<const uint4 *>(gb)[gda + b];
    if (v.x != kU32Max) v.x = 8u * list[S + number[W + (v.x - gda) / 8u]];
    if (8u * P + k < p.n_blocks) reinterpret_cast<uint4 *>(p.blocks)[8u * P + k] = v;
    if (k == 0) {
      const uint32_t pq = parent[W + q];
      p.usages[P].tick = (uint16_t)p.now;
      p.usages[P].parent = pq == kU32Max ? res[r].addr : 8u * list[S + number[W + pq / 8u]] + (pq & 7u);
    }
  }
}

__global__ __launch_bounds__(kFlat) void graft_link_kernel(PoolView p, const ResolveOut *res, const uint32_t *start,
                                                           const uint32_t *list, uint32_t n) {
  const uint32_t i = blockIdx.x * kFlat + threadIdx.x;
  if (i >= n || !res[i].placed) return;
  p.blocks[res[i].addr].child = 8u * list[start[i]];
}

// Collected groups no linked graft uses go back to the free pool.
__global__ __launch_bounds__(kFlat) void graft_release_kernel(PoolView p, const ResolveOut *res, const uint32_t *start,
                                                              uint32_t n, const uint32_t *list, DevCount got) {
  const uint32_t i = blockIdx.x * kFlat + threadIdx.x;
  if (i >= got.v || i >= got.get()) return;
  // owner: the last record whose run starts at or before i
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (start[mid] <= i)
      lo = mid;
    else
      hi = mid;
  }
  if (n && res[lo].placed) return;
  const uint32_t g = list[i];
  p.usages[g].tick = 0;
  p.usages[g].parent = kU32Max;
}

inline uint32_t blocks_for(uint32_t n) { return (n + kFlat - 1u) / kFlat; }
// 0 when the last launch was accepted
inline int launch_rc() { return hipGetLastError() == hipSuccess ? 0 : -1; }

}  // namespace

int tree_fill_u32(uint32_t *p, uint32_t value, size_t n, hipStream_t s) {
  if (!n) return 0;
  const uint32_t g = (uint32_t)std::min<size_t>((n + kFlat - 1) / kFlat, 4096);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(g), dim3(kFlat), 0, s, p, value, n);
  return launch_rc();
}

int tree_fill_u64(unsigned long long *p, unsigned long long value, size_t n, hipStream_t s) {
  if (!n) return 0;
  const uint32_t g = (uint32_t)std::min<size_t>((n + kFlat - 1) / kFlat, 4096);
  hipLaunchKernelGGL(fill_u64_kernel, dim3(g), dim3(kFlat), 0, s, p, value, n);
  return launch_rc();
}

size_t scan_span_sums_bytes(uint32_t n) {
  return n <= 2u * kScanSpan ? 0 : (size_t)(((uint64_t)n + kScanSpan - 1) / kScanSpan) * 4;
}

// Arrays above two spans take three launches: span sums (into the caller's span_sums, which
// belongs to one context, so scans on different contexts never share it), a one-work-group
// scan of the sums, span scans from their offsets.
int launch_scan_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *total, uint32_t *span_sums,
                    hipStream_t s) {
  if (n <= 2u * kScanSpan) {
    hipLaunchKernelGGL(scan_u32_kernel, dim3(1), dim3(kScanWg), 0, s, in, out, n, total);
    return launch_rc();
  }
  uint32_t *partial = span_sums;
  if (!partial) return -1;
  const uint32_t spans = (uint32_t)(((uint64_t)n + kScanSpan - 1) / kScanSpan);
  hipLaunchKernelGGL(scan_reduce_kernel, dim3(spans), dim3(kScanWg), 0, s, in, n, partial);
  if (launch_rc()) return -1;
  hipLaunchKernelGGL(scan_u32_kernel, dim3(1), dim3(kScanWg), 0, s, partial, partial, spans, total);
  if (launch_rc()) return -1;
  hipLaunchKernelGGL(scan_apply_kernel, dim3(spans), dim3(kScanWg), 0, s, in, out, n, partial);
  return launch_rc();
}

int launch_prune_resolve(const PoolView &p, const svo_pruning *rec, uint32_t n, uint32_t n_surface,
                         const uint64_t *locs, uint32_t n_locs, const TreeScratch &t, hipStream_t s) {
  if (!n) return 0;
  hipLaunchKernelGGL(prune_resolve_kernel, dim3(blocks_for(n)), dim3(kFlat), 0, s, p, rec, n, n_surface, locs, n_locs,
                     t.res, t.counters);
  return launch_rc();
}

int launch_graft_resolve(const PoolView &p, const svo_grafting *rec, uint32_t n, const uint64_t *locs,
                         uint32_t n_locs, const TreeScratch &t, hipStream_t s) {
  if (!n) return 0;
  hipLaunchKernelGGL(graft_resolve_kernel, dim3(blocks_for(n)), dim3(kFlat), 0, s, p, rec, n, t.need, locs, n_locs,
                     t.res, t.counters);
  return launch_rc();
}

int launch_requests_and_touches(const PoolView &p, const uint8_t *depths, uint32_t depth_stride,
                                const uint32_t *addresses, uint32_t address_stride, uint32_t n,
                                const uint64_t *locs, const TreeScratch &t, hipStream_t s) {
  if (!n) return 0;
  // both fields live in the same record array: express them as offsets from the depth byte
  RecordFields f;
  f.base = depths;
  f.stride = depth_stride;
  f.depth_off = 0;
  f.addr_off = (uint32_t)(reinterpret_cast<const uint8_t *>(addresses) - depths);
  (void)address_stride;
  hipLaunchKernelGGL(request_scan_kernel, dim3(1), dim3(kScanWg), 0, s, p, f, n, locs, t.res, t.counters);
  if (launch_rc()) return -1;
  hipLaunchKernelGGL(touch_kernel, dim3(blocks_for(n)), dim3(kFlat), 0, s, p, f, n, locs, t.res);
  return launch_rc();
}

int launch_prune_winners(const PoolView &p, const svo_pruning *rec, uint32_t n, const TreeScratch &t,
                         hipStream_t s) {
  if (!n) return 0;
  hipLaunchKernelGGL(prune_winners_kernel, dim3(blocks_for(n)), dim3(kFlat), 0, s, p, rec, n, t.res, t.keys, t.vals,
                     t.table_mask);
  return launch_rc();
}

int launch_prune_roots(const PoolView &p, const svo_pruning *rec, uint32_t n, const TreeScratch &t, hipStream_t s) {
  if (!n) return 0;
  hipLaunchKernelGGL(prune_roots_kernel, dim3(blocks_for(n)), dim3(kFlat), 0, s, p, rec, n, t.res, t.front[0],
                     t.counters, t.front_capacity);
  return launch_rc();
}

int launch_free_level(const PoolView &p, const TreeScratch &t, uint32_t level, bool use_stop, hipStream_t s) {
  const uint32_t grid = 1024;  // grid-stride over the device-side count
  hipLaunchKernelGGL(free_level_kernel, dim3(grid), dim3(kFlat), 0, s, p, t.front[level & 1u], t.front[(level + 1u) & 1u],
                     t.counters, level % 3u, (level + 1u) % 3u, (level + 2u) % 3u, t.front_capacity, t.events,
                     t.event_capacity, use_stop ? t.claimed : nullptr);
  return launch_rc();
}

int launch_unlink(const PoolView &p, const TreeScratch &t, uint32_t levels, hipStream_t s) {
  hipLaunchKernelGGL(unlink_kernel, dim3(1), dim3(kScanWg), 0, s, p, t.events, t.event_capacity, levels, t.counters);
  return launch_rc();
}

int launch_prune_apply(const PoolView &p, const svo_pruning *rec, uint32_t n, const svo_block *surface,
                       const TreeScratch &t, hipStream_t s) {
  if (!n) return 0;
  hipLaunchKernelGGL(prune_child_kernel, dim3(blocks_for(n)), dim3(kFlat), 0, s, p, rec, n, t.res, t.keys, t.vals,
                     t.table_mask);
  if (launch_rc()) return -1;
  hipLaunchKernelGGL(prune_fields_kernel, dim3(blocks_for(n)), dim3(kFlat), 0, s, p, rec, n, surface, t.res, t.keys,
                     t.vals, t.table_mask);
  return launch_rc();
}

int launch_collect_count(const PoolView &p, uint32_t n_groups, uint32_t offset, bool live_only,
                         const TreeScratch &t, hipStream_t s) {
  const uint32_t chunks = (n_groups + kCollectChunk - 1u) / kCollectChunk;
  if (!chunks) return 0;
  hipLaunchKernelGGL(collect_count_kernel, dim3(chunks), dim3(kFlat), 0, s, p, n_groups, offset,
                     live_only ? 1u : 0u, t.claimed, t.chunk_counts);
  if (launch_rc()) return -1;
  return launch_scan_u32(t.chunk_counts, t.chunk_counts, chunks, &t.counters[kTcTotal], t.scan_sums, s);
}

int launch_collect_limit(const TreeScratch &t, DevCount need, hipStream_t s) {
  hipLaunchKernelGGL(collect_limit_kernel, dim3(1), dim3(64), 0, s, t.counters, need);
  return launch_rc();
}

int launch_collect_advance(const TreeScratch &t, hipStream_t s) {
  hipLaunchKernelGGL(collect_advance_kernel, dim3(1), dim3(64), 0, s, t.counters);
  return launch_rc();
}

int launch_collect_scatter(const PoolView &p, uint32_t n_groups, uint32_t offset, bool live_only,
                           const TreeScratch &t, hipStream_t s) {
  const uint32_t chunks = (n_groups + kCollectChunk - 1u) / kCollectChunk;
  if (!chunks) return 0;
  hipLaunchKernelGGL(collect_scatter_kernel, dim3(chunks), dim3(kFlat), 0, s, p, n_groups, offset,
                     live_only ? 1u : 0u, t.claimed, t.chunk_counts, t.counters, t.list);
  return launch_rc();
}

int launch_collect_claim(const PoolView &p, uint32_t k_max, const TreeScratch &t, hipStream_t s) {
  if (!k_max) return 0;
  hipLaunchKernelGGL(collect_claim_kernel, dim3(blocks_for(k_max)), dim3(kFlat), 0, s, p, t.list, k_max, t.claimed,
                     t.front[0], t.counters, t.front_capacity);
  return launch_rc();
}

int launch_collect_mark(const PoolView &p, uint32_t k_max, uint32_t tick, const TreeScratch &t, hipStream_t s) {
  if (!k_max) return 0;
  hipLaunchKernelGGL(collect_mark_kernel, dim3(blocks_for(k_max)), dim3(kFlat), 0, s, p, t.list, k_max, t.counters,
                     tick);
  return launch_rc();
}

int launch_graft_structure(const svo_grafting *rec, uint32_t n, const svo_block *gb, uint32_t n_gb,
                           const TreeScratch &t, hipStream_t s) {
  (void)n_gb;
  if (!n) return 0;
  hipLaunchKernelGGL(graft_structure_kernel, dim3((n + 63u) / 64u), dim3(64), 0, s, rec, n, gb, t.size, t.where,
                     t.number, t.parent, t.need);
  return launch_rc();
}

int launch_graft_sizes(const svo_grafting *rec, DevCount n, uint32_t n_gb, const TreeScratch &t, hipStream_t s) {
  if (!n.v) return 0;
  hipLaunchKernelGGL(graft_size_kernel, dim3(blocks_for(n.v)), dim3(kFlat), 0, s, rec, n, n_gb, t.size);
  return launch_rc();
}

int launch_graft_winners(const PoolView &p, uint32_t n, DevCount got, const TreeScratch &t, hipStream_t s) {
  (void)p;
  if (!n) return 0;
  hipLaunchKernelGGL(graft_winners_kernel, dim3(blocks_for(n)), dim3(kFlat), 0, s, t.res, t.start, t.need, n, got,
                     t.keys, t.vals, t.table_mask);
  return launch_rc();
}

int launch_graft_roots(const PoolView &p, DevCount n, DevCount got, const TreeScratch &t, hipStream_t s) {
  if (!n.v) return 0;
  hipLaunchKernelGGL(graft_roots_kernel, dim3(blocks_for(n.v)), dim3(kFlat), 0, s, p, t.res, t.start, t.need, n, got,
                     t.keys, t.vals, t.table_mask, t.front[0], t.counters, t.front_capacity);
  return launch_rc();
}

int launch_graft_place(const PoolView &p, const svo_grafting *rec, uint32_t n, DevCount total_need,
                       const svo_block *gb, const TreeScratch &t, hipStream_t s) {
  if (!n) return 0;
  if (total_need.v) {
    const uint32_t blocks = (uint32_t)((8ull * total_need.v + kFlat - 1) / kFlat);
    hipLaunchKernelGGL(graft_place_kernel, dim3(blocks), dim3(kFlat), 0, s, p, rec, gb, t.res, t.start, t.where,
                       t.number, t.parent, t.list, n, total_need);
  }
  if (launch_rc()) return -1;
  hipLaunchKernelGGL(graft_link_kernel, dim3(blocks_for(n)), dim3(kFlat), 0, s, p, t.res, t.start, t.list, n);
  return launch_rc();
}

int launch_graft_release(const PoolView &p, uint32_t n, DevCount got, const TreeScratch &t, hipStream_t s) {
  if (!got.v) return 0;
  hipLaunchKernelGGL(graft_release_kernel, dim3(blocks_for(got.v)), dim3(kFlat), 0, s, p, t.res, t.start, n, t.list,
                     got);
  return launch_rc();
}

}  // namespace svo
