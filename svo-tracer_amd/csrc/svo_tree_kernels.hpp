// Launch interface of the tree-management kernels (prune, graft, collect).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "svo.h"

namespace svo {

// The mutable pool one management dispatch works on (device pointers).
struct PoolView {
  uint16_t *bases;
  uint32_t n_bases;
  svo_block *blocks;
  uint32_t n_blocks;
  svo_usage *usages;
  uint32_t n_usages;
  svo_parent *parents;  // null when no parent list is bound
  uint32_t n_parents;
  uint8_t *residency;   // one byte per parent element, may be null
  svo_child_request *requests;
  uint32_t request_capacity;
  uint32_t *request_count;  // null: requests are only counted
  uint32_t B;
  uint32_t now;
};

// Per-record resolution of a target (findAddress, wavefront.cl:1191-1211).
struct ResolveOut {
  uint32_t status;
  uint32_t addr;
  uint32_t level;  // kResMiss: level of the childless block
  uint32_t placed; // graft: linked into the tree
};
enum : uint32_t { kResOk = 0, kResMiss = 1, kResSkip = 2, kResInvalid = 3, kResBase = 4 };

// Slots of the u32 counter block of a management dispatch.
enum : uint32_t {
  kTcFront0 = 0,      // frontier sizes (ping-pong)
  kTcFront1 = 1,
  kTcEvents = 2,      // multi-parent unlink events
  kTcOverflow = 3,    // frontier overflow flag
  kTcTotal = 4,       // scan totals read by the host
  kTcFront2 = 5,      // third frontier count (level L reads count L % 3, appends to (L+1) % 3)
  kTcApplied = 8,     // statistics (svo_tree_stats order)
  kTcSkipped = 9,
  kTcInvalid = 10,
  kTcRequests = 11,
  kTcCollected = 12,
  kTcPlaced = 13,
  kTcUnlinks = 14,
  kTcGot = 16,        // collection: groups collected so far
  kTcNeed = 17,       // graft: groups the grafts need (scan total of need)
  kTcK = 18,          // collection: groups this pass claims
  kTcFrom = 19,       // collection: where this pass's groups start in the list
  kTcMisses = 20,     // resolve: 1 when some record's walk met a childless block
  kTcSlots = 64
};
// kTcOverflow bits
enum : uint32_t { kOverflowFront = 1, kOverflowEvents = 2, kOverflowWalk = 4 };
// free-walk levels launched before the host first reads the frontier, and launched between
// later reads while the frontier is not empty
constexpr uint32_t kSpecLevels = 3;
constexpr uint32_t kMoreLevels = 4;
// The counter slot of frontier count i (0..2) of the free walk's rotation.
__host__ __device__ inline uint32_t front_slot(uint32_t i) {
  return i == 0u ? (uint32_t)kTcFront0 : (i == 1u ? (uint32_t)kTcFront1 : (uint32_t)kTcFront2);
}

// A count a launch needs: the value itself, or (p set) a device location holding it, with v
// then its host-known upper bound (grids are sized by v; kernels read min(*p, v)).
struct DevCount {
  uint32_t v;
  const uint32_t *p;
  __device__ __forceinline__ uint32_t get() const { return p ? (*p < v ? *p : v) : v; }
};
inline DevCount host_count(uint32_t v) { return DevCount{v, nullptr}; }

// Scratch of a management dispatch (device pointers; sizes in elements).
struct TreeScratch {
  uint32_t *counters;            // kTcSlots
  ResolveOut *res;               // per record
  unsigned long long *keys;      // winner table
  uint32_t *vals;
  uint32_t table_mask;
  uint2 *front[2];
  uint32_t front_capacity;
  uint2 *events;
  uint32_t event_capacity;       // power of two, at most kUnlinkSortMax
  uint32_t *claimed;             // bitmap over groups
  uint32_t claimed_words;
  uint32_t *chunk_counts;        // collection scan, one per chunk of groups
  uint32_t *scan_sums;           // span sums of the collection scan (scan_span_sums_bytes)
  uint32_t *list;                // collected groups
  uint32_t list_capacity;
  uint32_t *size, *where, *need, *start;  // graft: per record
  uint32_t *number, *parent;     // graft: per graft group (DFS numbering)
};

constexpr uint32_t kUnlinkSortMax = 4096;
constexpr uint32_t kCollectChunk = 1024;  // groups per collection work-group
constexpr uint32_t kMaxFreeLevels = 64;

int tree_fill_u32(uint32_t *p, uint32_t value, size_t n, hipStream_t s);
int tree_fill_u64(unsigned long long *p, unsigned long long value, size_t n, hipStream_t s);

int launch_prune_resolve(const PoolView &p, const svo_pruning *rec, uint32_t n, uint32_t n_surface,
                         const uint64_t *locs, uint32_t n_locs, const TreeScratch &t, hipStream_t s);
int launch_graft_resolve(const PoolView &p, const svo_grafting *rec, uint32_t n, const uint64_t *locs,
                         uint32_t n_locs, const TreeScratch &t, hipStream_t s);
// child requests of the missed records, in record order, then the ticks along every walk
int launch_requests_and_touches(const PoolView &p, const uint8_t *depths, uint32_t depth_stride,
                                const uint32_t *addresses, uint32_t address_stride, uint32_t n,
                                const uint64_t *locs, const TreeScratch &t, hipStream_t s);
int launch_prune_winners(const PoolView &p, const svo_pruning *rec, uint32_t n, const TreeScratch &t,
                         hipStream_t s);
int launch_prune_roots(const PoolView &p, const svo_pruning *rec, uint32_t n, const TreeScratch &t,
                       hipStream_t s);
// Level `level` of the free walk from the groups queued in front[0] (count kTcFront0): reads
// front[level % 2] with count front_slot(level % 3), appends to front[(level + 1) % 2] with
// count front_slot((level + 1) % 3) and clears front_slot((level + 2) % 3) for the level
// after; a level whose count is 0 does nothing.
int launch_free_level(const PoolView &p, const TreeScratch &t, uint32_t level, bool use_stop, hipStream_t s);
// After `levels` levels: multi-parent unlinks queued by the walk, in order; flags in
// counters[kTcOverflow] more events than the table holds and a walk whose frontier after the
// last level is not empty; clears the event and frontier counts for the next walk.
int launch_unlink(const PoolView &p, const TreeScratch &t, uint32_t levels, hipStream_t s);
int launch_prune_apply(const PoolView &p, const svo_pruning *rec, uint32_t n, const svo_block *surface,
                       const TreeScratch &t, hipStream_t s);

// collection: one pass with `offset`; candidates ranked in ascending group order.  With
// live_only (reclaim) only stale live groups are candidates; mark gives claimed groups `tick`.
int launch_collect_count(const PoolView &p, uint32_t n_groups, uint32_t offset, bool live_only,
                         const TreeScratch &t, hipStream_t s);
// this pass takes k = min(candidates, need - collected) groups from list position collected
// (counters kTcK, kTcFrom); advance adds k to the collected count
int launch_collect_limit(const TreeScratch &t, DevCount need, hipStream_t s);
int launch_collect_scatter(const PoolView &p, uint32_t n_groups, uint32_t offset, bool live_only,
                           const TreeScratch &t, hipStream_t s);
// k_max: a bound on this pass's k (the list capacity)
int launch_collect_claim(const PoolView &p, uint32_t k_max, const TreeScratch &t, hipStream_t s);
int launch_collect_mark(const PoolView &p, uint32_t k_max, uint32_t tick, const TreeScratch &t, hipStream_t s);
int launch_collect_advance(const TreeScratch &t, hipStream_t s);

// graft
// n.v records are processed; those at or beyond n.get() count as absent (size and need 0)
int launch_graft_sizes(const svo_grafting *rec, DevCount n, uint32_t n_gb, const TreeScratch &t, hipStream_t s);
int launch_graft_structure(const svo_grafting *rec, uint32_t n, const svo_block *gb, uint32_t n_gb,
                           const TreeScratch &t, hipStream_t s);
// Exclusive scan of n u32; the sum goes to *total.  Above 2 x 16384 elements the scan needs
// span_sums, device scratch of scan_span_sums_bytes(n) bytes owned by the caller.
size_t scan_span_sums_bytes(uint32_t n);
int launch_scan_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *total, uint32_t *span_sums,
                    hipStream_t s);
// got: the groups the collection gathered
int launch_graft_winners(const PoolView &p, uint32_t n, DevCount got, const TreeScratch &t, hipStream_t s);
int launch_graft_roots(const PoolView &p, DevCount n, DevCount got, const TreeScratch &t, hipStream_t s);
// total_need: the sum of need over the n records (the scan total of need into start).
int launch_graft_place(const PoolView &p, const svo_grafting *rec, uint32_t n, DevCount total_need,
                       const svo_block *gb, const TreeScratch &t, hipStream_t s);
int launch_graft_release(const PoolView &p, uint32_t n, DevCount got, const TreeScratch &t, hipStream_t s);

}  // namespace svo
