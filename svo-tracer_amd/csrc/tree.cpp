// Host side of tree management: svo_prune, svo_graft, svo_collect and the Prune / Graft
// entries of svo_enqueue (KernelName.Prune / KernelName.Graft, ComputeManager.cs:313-318).
//
// Every phase runs on the device and sizes itself from device-side counts; the host sequences
// the kernels and reads the counters back once at the end of a dispatch, once per collection
// pass, and once per free walk that has not ended within its first kSpecLevels levels (each
// level reads its frontier size on the device, so the first levels are launched unread).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>

#include "svo_internal.hpp"
#include "svo_runtime.hpp"
#include "svo_tree_kernels.hpp"

namespace svo {

namespace {

enum TreeBuf {
  kTbCounters,
  kTbRes,
  kTbKeys,
  kTbVals,
  kTbFront0,
  kTbFront1,
  kTbEvents,
  kTbClaimed,
  kTbChunks,
  kTbList,
  kTbSize,
  kTbWhere,
  kTbNeed,
  kTbStart,
  kTbNumber,
  kTbParent,
  kTbRecords,
  kTbSide,
  kTbLocations,
  kTbCount
};

struct Pool {
  struct B {
    void *ptr = nullptr;
    size_t bytes = 0;
  };
  B bases, blocks, usages, parents, residency, requests, request_id, holding, holding_count;
};

uint32_t count_of(const Pool::B &b, size_t elem) { return (uint32_t)std::min<size_t>(b.bytes / elem, 0xffffffffu); }

uint32_t pow2_at_least(uint64_t v) {
  uint64_t p = 64;
  while (p < v) p <<= 1;
  return (uint32_t)std::min<uint64_t>(p, 1ull << 31);
}

}  // namespace

// The statistics of a dispatch from its counter block (kTcSlots words read back), after the
// overflow flags the device may have raised.
int stats_from(const uint32_t *c, svo_tree_stats *out) {
  if (c[kTcOverflow] & kOverflowFront) return fail(SVO_ERR_CAPACITY, "tree: free walk frontier overflow");
  if (c[kTcOverflow] & kOverflowEvents)
    return fail(SVO_ERR_CAPACITY, "tree: more than 4096 multi-parent unlinks in one dispatch");
  if (c[kTcOverflow] & kOverflowWalk)
    return fail(SVO_ERR_CAPACITY, "tree: a free walk expected to be empty was not (streamed graft targets with children)");
  if (!out) return SVO_OK;
  std::memset(out, 0, sizeof *out);
  out->applied = c[kTcApplied];
  out->skipped = c[kTcSkipped];
  out->invalid = c[kTcInvalid];
  out->requests = c[kTcRequests];
  out->collected = c[kTcGot];
  out->placed = c[kTcPlaced];
  out->unlinks = c[kTcUnlinks];
  return SVO_OK;
}

namespace {

struct Op {
  Context *ctx;
  PoolView pv{};
  TreeScratch ts{};
  uint32_t *host = nullptr;  // pinned counter mirror

  std::shared_ptr<DeviceBuffer> &buf(int k) { return ctx->tree_scratch[k]; }

  int grow(int k, size_t bytes, void **out) {
    int rc = ensure_scratch(buf(k), bytes);
    if (rc) return rc;
    *out = buf(k)->ptr;
    return SVO_OK;
  }

  int init(const Pool &pool, const svo_update_input_data &in) {
    if (!pool.blocks.ptr) return fail(SVO_ERR_MISSING_ARGUMENT, "tree management: Blocks is not bound");
    if (!pool.usages.ptr) return fail(SVO_ERR_MISSING_ARGUMENT, "tree management: Usages is not bound");
    if (in.base_depth < 1 || in.base_depth > 8)
      return fail(SVO_ERR_INVALID_ARGUMENT, "tree management: BaseDepth must be 1..8");
    pv.bases = (uint16_t *)pool.bases.ptr;
    pv.n_bases = count_of(pool.bases, 2);
    pv.blocks = (svo_block *)pool.blocks.ptr;
    pv.n_blocks = count_of(pool.blocks, 16);
    pv.usages = (svo_usage *)pool.usages.ptr;
    pv.n_usages = count_of(pool.usages, 8);
    pv.parents = (svo_parent *)pool.parents.ptr;
    pv.n_parents = pool.parents.ptr ? count_of(pool.parents, 8) : 0;
    pv.residency = pool.residency.bytes >= pv.n_parents ? (uint8_t *)pool.residency.ptr : nullptr;
    const bool have_requests = pool.request_id.ptr && pool.request_id.bytes >= 4;
    pv.requests = have_requests ? (svo_child_request *)pool.requests.ptr : nullptr;
    pv.request_capacity = have_requests ? std::min(in.max_child_request_id, count_of(pool.requests, 40)) : 0;
    pv.request_count = have_requests ? (uint32_t *)pool.request_id.ptr : nullptr;
    pv.B = in.base_depth;
    pv.now = in.tick;
    if (!ctx->tree_host) {
      hipError_t e = hipHostMalloc((void **)&ctx->tree_host, kTcSlots * sizeof(uint32_t), hipHostMallocDefault);
      if (e != hipSuccess) return hip_fail(e, "hipHostMalloc");
    }
    host = ctx->tree_host;
    void *p;
    int rc = grow(kTbCounters, kTcSlots * 4, &p);
    if (rc) return rc;
    ts.counters = (uint32_t *)p;
    // the free walk's frontiers hold at most every group once per level, plus duplicates
    // from records that name the same target
    ts.front_capacity = (uint32_t)std::min<uint64_t>((uint64_t)pv.n_usages + 65536u, 0xffffffffu / 8u);
    if ((rc = grow(kTbFront0, (size_t)ts.front_capacity * 8, &p))) return rc;
    ts.front[0] = (uint2 *)p;
    if ((rc = grow(kTbFront1, (size_t)ts.front_capacity * 8, &p))) return rc;
    ts.front[1] = (uint2 *)p;
    ts.event_capacity = kUnlinkSortMax;
    if ((rc = grow(kTbEvents, (size_t)ts.event_capacity * 8, &p))) return rc;
    ts.events = (uint2 *)p;
    ts.claimed_words = (pv.n_usages + 31u) / 32u;
    if ((rc = grow(kTbClaimed, (size_t)ts.claimed_words * 4 + 4, &p))) return rc;
    ts.claimed = (uint32_t *)p;
    if ((rc = grow(kTbChunks, ((size_t)pv.n_usages / kCollectChunk + 2) * 4, &p))) return rc;
    ts.chunk_counts = (uint32_t *)p;
    if ((rc = context_scan_sums(ctx, scan_span_sums_bytes(pv.n_usages / kCollectChunk + 1), &ts.scan_sums)))
      return rc;
    return tree_fill_u32(ts.counters, 0, kTcSlots, ctx->stream) ? fail(SVO_ERR_HIP, "tree: fill") : SVO_OK;
  }

  int records(uint32_t n, uint32_t table_entries) {
    void *p;
    int rc = grow(kTbRes, (size_t)std::max(n, 1u) * sizeof(ResolveOut), &p);
    if (rc) return rc;
    ts.res = (ResolveOut *)p;
    const uint32_t cap = pow2_at_least(2ull * std::max(table_entries, 1u));
    if ((rc = grow(kTbKeys, (size_t)cap * 8, &p))) return rc;
    ts.keys = (unsigned long long *)p;
    if ((rc = grow(kTbVals, (size_t)cap * 4, &p))) return rc;
    ts.vals = (uint32_t *)p;
    ts.table_mask = cap - 1u;
    if (tree_fill_u64(ts.keys, ~0ull, cap, ctx->stream) || tree_fill_u32(ts.vals, 0, cap, ctx->stream))
      return fail(SVO_ERR_HIP, "tree: fill");
    return SVO_OK;
  }

  // Reads counter slots [first, first+count) into host[first...].
  int read(uint32_t first, uint32_t count) {
    hipError_t e = hipMemcpyAsync(host + first, ts.counters + first, count * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? SVO_OK : hip_fail(e, "tree: counter read");
  }

  int launch(int r, const char *what) { return r ? fail(SVO_ERR_HIP, std::string("tree: launch ") + what) : SVO_OK; }

  // Frees the subtrees queued in front[0] (count in its counter), level by level, then
  // applies the multi-parent unlink events in order.  The first kSpecLevels levels are
  // launched without reading anything back; with `finish_walk` the host then reads the
  // frontier and adds levels until it is empty.  Without it (a walk known to end within those
  // levels) an unfinished walk is flagged on the device, as are overflows of the frontier or
  // of the event table, and finish() reports them.
  int free_walk(bool use_stop, bool finish_walk = true) {
    int rc = walk_levels(use_stop, 0, kSpecLevels);
    if (rc) return rc;
    uint32_t level = kSpecLevels;
    if (finish_walk) {
      if ((rc = read(front_slot(level % 3u), 1))) return rc;
      if ((rc = walk_on(use_stop, &level))) return rc;
    }
    return end_walk(level);
  }
  // After a read of the frontier of `*level`: while it is not empty, the next kMoreLevels
  // levels go out unread and the frontier after them is read (a walk that outlasts the
  // unread levels mostly runs several more: one read for them instead of one a level).
  int walk_on(bool use_stop, uint32_t *level) {
    while (*level < kMaxFreeLevels && host[front_slot(*level % 3u)] != 0) {
      const uint32_t to = std::min(*level + kMoreLevels, kMaxFreeLevels);
      int rc = walk_levels(use_stop, *level, to);
      if (!rc) rc = read(front_slot(to % 3u), 1);
      if (rc) return rc;
      *level = to;
    }
    return SVO_OK;
  }
  int walk_levels(bool use_stop, uint32_t from, uint32_t to) {
    for (uint32_t level = from; level < to; level++) {
      int rc = launch(launch_free_level(pv, ts, level, use_stop, ctx->stream), "free level");
      if (rc) return rc;
    }
    return SVO_OK;
  }
  int end_walk(uint32_t levels) { return launch(launch_unlink(pv, ts, levels, ctx->stream), "unlink"); }

  // Collection (graft, wavefront.cl:1277-1305): up to `need` groups (at most need_max) into
  // ts.list; the count collected accumulates in counters[kTcGot].  One pass per age
  // threshold (offset, offset/2, ..., 0); after each pass that may be followed by another the
  // host reads whether the need is met.
  int collect(const svo_update_input_data &in, DevCount need, uint32_t need_max) {
    void *p;
    int rc = grow(kTbList, (size_t)std::max(need_max, 1u) * 4, &p);
    if (rc) return rc;
    ts.list = (uint32_t *)p;
    ts.list_capacity = need_max;
    if (!need_max) return SVO_OK;
    if ((rc = launch(tree_fill_u32(ts.claimed, 0, ts.claimed_words + 1, ctx->stream), "fill"))) return rc;
    uint32_t ng = pv.n_usages;
    if (in.memory_size) ng = (uint32_t)std::min<uint64_t>(((uint64_t)in.memory_size + 7u) / 8u, ng);
    for (uint32_t offset = in.offset;; offset >>= 1) {
      if ((rc = launch(launch_collect_count(pv, ng, offset, false, ts, ctx->stream), "collect count"))) return rc;
      if ((rc = launch(launch_collect_limit(ts, need, ctx->stream), "collect limit"))) return rc;
      if ((rc = launch(launch_collect_scatter(pv, ng, offset, false, ts, ctx->stream), "collect scatter"))) return rc;
      if ((rc = launch(launch_collect_claim(pv, need_max, ts, ctx->stream), "collect claim"))) return rc;
      // the pass's groups count as collected now (mark and the next pass read kTcFrom / kTcK)
      if ((rc = launch(launch_collect_advance(ts, ctx->stream), "collect advance"))) return rc;
      if ((rc = walk_levels(true, 0, kSpecLevels))) return rc;
      // one read: the walk's frontier, the groups collected so far and (device need) the need
      if ((rc = read(0, kTcSlots))) return rc;
      uint32_t level = kSpecLevels;
      if ((rc = walk_on(true, &level))) return rc;
      if ((rc = end_walk(level))) return rc;
      if ((rc = launch(launch_collect_mark(pv, need_max, pv.now, ts, ctx->stream), "collect mark"))) return rc;
      const uint32_t nd = need.p ? std::min(host[kTcNeed], need.v) : need.v;
      if (offset == 0 || host[kTcGot] >= nd) break;
    }
    return SVO_OK;
  }

  // Garbage collection's reclaim: every stale live group (age > in.offset) is cut from its
  // parent, what hangs below it is freed and the group itself gets tick 0.
  int reclaim(const svo_update_input_data &in, uint32_t *got) {
    *got = 0;
    int rc = launch(tree_fill_u32(ts.claimed, 0, ts.claimed_words + 1, ctx->stream), "fill");
    if (rc) return rc;
    uint32_t ng = pv.n_usages;
    if (in.memory_size) ng = (uint32_t)std::min<uint64_t>(((uint64_t)in.memory_size + 7u) / 8u, ng);
    if ((rc = launch(launch_collect_count(pv, ng, in.offset, true, ts, ctx->stream), "reclaim count"))) return rc;
    if ((rc = read(kTcTotal, 1))) return rc;
    const uint32_t k = host[kTcTotal];
    if (!k) return finish(nullptr);
    void *p;
    if ((rc = grow(kTbList, (size_t)k * 4, &p))) return rc;
    ts.list = (uint32_t *)p;
    ts.list_capacity = k;
    if ((rc = launch(launch_collect_limit(ts, host_count(k), ctx->stream), "reclaim limit"))) return rc;
    if ((rc = launch(launch_collect_scatter(pv, ng, in.offset, true, ts, ctx->stream), "reclaim scatter"))) return rc;
    if ((rc = launch(launch_collect_claim(pv, k, ts, ctx->stream), "reclaim claim"))) return rc;
    if ((rc = free_walk(true))) return rc;
    if ((rc = launch(launch_collect_mark(pv, k, 0u, ts, ctx->stream), "reclaim mark"))) return rc;
    if ((rc = finish(nullptr))) return rc;
    *got = k;
    return SVO_OK;
  }

  // Reads the counters once, reports a flagged overflow and fills `out` (may be null).
  int finish(svo_tree_stats *out) {
    int rc = read(0, kTcSlots);
    if (rc) return rc;
    return stats_from(host, out);
  }
};

int prune_impl(Context *ctx, const Pool &pool, const svo_update_input_data &in, const svo_pruning *rec, uint32_t n,
               const svo_block *surface, uint32_t n_surface, const uint64_t *locs, uint32_t n_locs,
               svo_tree_stats *stats) {
  Op op{ctx};
  int rc = op.init(pool, in);
  if (rc) return rc;
  if (n && !rec) return fail(SVO_ERR_INVALID_ARGUMENT, "prune: records missing");
  if ((rc = op.records(n, 4u * n + 1u))) return rc;
  hipStream_t s = ctx->stream;
  const PoolView &pv = op.pv;
  const TreeScratch &ts = op.ts;
  const uint8_t *rb = reinterpret_cast<const uint8_t *>(rec);
  if ((rc = op.launch(launch_prune_resolve(pv, rec, n, n_surface, locs, n_locs, ts, s), "prune resolve"))) return rc;
  if ((rc = op.launch(launch_requests_and_touches(pv, rb + offsetof(svo_pruning, depth), sizeof(svo_pruning),
                                                  reinterpret_cast<const uint32_t *>(rb + offsetof(svo_pruning, address)),
                                                  sizeof(svo_pruning), n, locs, ts, s),
                      "prune requests")))
    return rc;
  if ((rc = op.launch(launch_prune_winners(pv, rec, n, ts, s), "prune winners"))) return rc;
  if ((rc = op.launch(launch_prune_roots(pv, rec, n, ts, s), "prune roots"))) return rc;
  if ((rc = op.free_walk(false))) return rc;
  if ((rc = op.launch(launch_prune_apply(pv, rec, n, surface, ts, s), "prune apply"))) return rc;
  return op.finish(stats);
}

// dev_records: when set, the number of records lives there on the device (at most n) and the
// dispatch reads nothing back: its counters stay in the op's counter block for the caller
// (graft_counters) to read with its own.
int graft_impl(Context *ctx, const Pool &pool, const svo_update_input_data &in, const svo_grafting *rec, uint32_t n,
               const svo_block *gb, uint32_t n_gb, const uint64_t *locs, uint32_t n_locs, svo_tree_stats *stats,
               const uint32_t *dev_records = nullptr) {
  Op op{ctx};
  int rc = op.init(pool, in);
  if (rc) return rc;
  if (n && !rec) return fail(SVO_ERR_INVALID_ARGUMENT, "graft: records missing");
  if ((rc = op.records(n, n + 1u))) return rc;
  hipStream_t s = ctx->stream;
  void *p;
  const size_t per_rec = (size_t)std::max(n, 1u) * 4;
  if ((rc = op.grow(kTbSize, per_rec, &p))) return rc;
  op.ts.size = (uint32_t *)p;
  if ((rc = op.grow(kTbWhere, per_rec, &p))) return rc;
  op.ts.where = (uint32_t *)p;
  if ((rc = op.grow(kTbNeed, per_rec, &p))) return rc;
  op.ts.need = (uint32_t *)p;
  if ((rc = op.grow(kTbStart, per_rec, &p))) return rc;
  op.ts.start = (uint32_t *)p;
  const DevCount records{n, dev_records};
  // structure: sizes, where each graft's numbering lives, pre-order numbering, needs
  if ((rc = op.launch(launch_graft_sizes(rec, records, n_gb, op.ts, s), "graft sizes"))) return rc;
  uint32_t *sums = nullptr;
  if ((rc = context_scan_sums(ctx, scan_span_sums_bytes(n), &sums))) return rc;
  if ((rc = op.launch(launch_scan_u32(op.ts.size, op.ts.where, n, op.ts.counters + kTcTotal, sums, s), "scan")))
    return rc;
  // groups of all grafts: read back for host records; for device records every graft's
  // groups lie in its own part of the graft blocks, so n_gb / 8 bounds them
  uint32_t groups = n_gb / 8u;
  if (!dev_records) {
    if ((rc = op.read(kTcTotal, 1))) return rc;
    groups = op.host[kTcTotal];
  }
  if ((rc = op.grow(kTbNumber, ((size_t)groups + 1) * 4, &p))) return rc;
  op.ts.number = (uint32_t *)p;
  if ((rc = op.grow(kTbParent, ((size_t)groups + 1) * 4, &p))) return rc;
  op.ts.parent = (uint32_t *)p;
  if ((rc = op.launch(launch_graft_structure(rec, n, gb, n_gb, op.ts, s), "graft structure"))) return rc;
  if ((rc = op.launch(launch_scan_u32(op.ts.need, op.ts.start, n, op.ts.counters + kTcNeed, sums, s), "scan")))
    return rc;
  // every need is at most its graft's size, so `groups` bounds the total need
  const DevCount need{groups, op.ts.counters + kTcNeed};
  if ((rc = op.collect(in, need, groups))) return rc;
  const DevCount got{groups, op.ts.counters + kTcGot};
  const PoolView &pv = op.pv;
  const TreeScratch &ts = op.ts;
  const uint8_t *rb = reinterpret_cast<const uint8_t *>(rec);
  if ((rc = op.launch(launch_graft_resolve(pv, rec, n, locs, n_locs, ts, s), "graft resolve"))) return rc;
  if ((rc = op.launch(launch_requests_and_touches(pv, rb + offsetof(svo_grafting, depth), sizeof(svo_grafting),
                                                  reinterpret_cast<const uint32_t *>(rb + offsetof(svo_grafting, graft_address)),
                                                  sizeof(svo_grafting), n, locs, ts, s),
                      "graft requests")))
    return rc;
  if ((rc = op.launch(launch_graft_winners(pv, n, got, ts, s), "graft winners"))) return rc;
  if ((rc = op.launch(launch_graft_roots(pv, records, got, ts, s), "graft roots"))) return rc;
  // device records come from the streaming update, whose targets are childless blocks: their
  // walk is empty, and one that is not is flagged rather than read back and finished
  if ((rc = op.free_walk(true, !dev_records))) return rc;
  if ((rc = op.launch(launch_graft_place(pv, rec, n, need, gb, ts, s), "graft place"))) return rc;
  if ((rc = op.launch(launch_graft_release(pv, n, got, ts, s), "graft release"))) return rc;
  if (dev_records) return SVO_OK;
  return op.finish(stats);
}

Pool pool_from_context(Context *ctx) {
  Pool p;
  auto get = [&](int name, Pool::B *b) {
    auto d = ctx->get(name);
    if (d) {
      b->ptr = d->ptr;
      b->bytes = d->bytes;
    }
  };
  get(SVO_BUF_BASES, &p.bases);
  get(SVO_BUF_BLOCKS, &p.blocks);
  get(SVO_BUF_USAGES, &p.usages);
  get(SVO_BUF_PARENTS, &p.parents);
  get(SVO_BUF_PARENT_RESIDENCY, &p.residency);
  get(SVO_BUF_CHILD_REQUESTS, &p.requests);
  get(SVO_BUF_CHILD_REQUEST_ID, &p.request_id);
  return p;
}

// Copies a host array into a context scratch buffer; returns its device address.
int stage(Context *ctx, int slot, const void *src, size_t bytes, const void **out) {
  *out = nullptr;
  if (!bytes) return SVO_OK;
  if (!src) return fail(SVO_ERR_INVALID_ARGUMENT, "tree management: null host array");
  int rc = ensure_scratch(ctx->tree_scratch[slot], bytes);
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(ctx->tree_scratch[slot]->ptr, src, bytes, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(e, "tree management: upload");
  *out = ctx->tree_scratch[slot]->ptr;
  return SVO_OK;
}

// SetArg lookups for the Prune / Graft entries of svo_enqueue.
struct Bound {
  Context *ctx;
  int kernel;
  Pool::B buf(const char *param) const {
    Pool::B b;
    auto k = ctx->args.find(kernel);
    if (k == ctx->args.end()) return b;
    auto a = k->second.find(param);
    if (a == k->second.end() || !a->second.buffer) return b;
    b.ptr = a->second.buffer->ptr;
    b.bytes = a->second.buffer->bytes;
    return b;
  }
  bool input(svo_update_input_data *out) const {
    auto k = ctx->args.find(kernel);
    if (k == ctx->args.end()) return false;
    auto a = k->second.find("inputData");
    if (a == k->second.end() || a->second.bytes.size() != sizeof *out) return false;
    std::memcpy(out, a->second.bytes.data(), sizeof *out);
    return true;
  }
};

}  // namespace

int graft_device(Context *ctx, const svo_update_input_data &in, const svo_grafting *records, uint32_t n,
                 const svo_block *graft_blocks, uint32_t n_graft_blocks, svo_tree_stats *stats) {
  return graft_impl(ctx, pool_from_context(ctx), in, records, n, graft_blocks, n_graft_blocks, nullptr, 0, stats);
}

int graft_device_counted(Context *ctx, const svo_update_input_data &in, const svo_grafting *records, uint32_t n_max,
                         const uint32_t *n_dev, const svo_block *graft_blocks, uint32_t n_graft_blocks,
                         const uint32_t **counters) {
  int rc = graft_impl(ctx, pool_from_context(ctx), in, records, n_max, graft_blocks, n_graft_blocks, nullptr, 0,
                      nullptr, n_dev);
  if (rc) return rc;
  *counters = (const uint32_t *)ctx->tree_scratch[kTbCounters]->ptr;
  return SVO_OK;
}

int reserve_tree(Context *ctx, uint32_t base_depth, uint32_t records) {
  svo_update_input_data in;
  std::memset(&in, 0, sizeof in);
  in.base_depth = (uint8_t)base_depth;
  Op op{ctx};
  int rc = op.init(pool_from_context(ctx), in);
  if (rc || !records) return rc;
  // what a graft of `records` one-group records grows: resolution and winner table, the
  // per-record and per-group arrays, the collection list
  if ((rc = op.records(records, records + 1u))) return rc;
  void *p;
  for (int k : {kTbSize, kTbWhere, kTbNeed, kTbStart, kTbList})
    if ((rc = op.grow(k, (size_t)records * 4, &p))) return rc;
  for (int k : {kTbNumber, kTbParent})
    if ((rc = op.grow(k, ((size_t)records + 1) * 4, &p))) return rc;
  return SVO_OK;
}

int reclaim_device(Context *ctx, const svo_update_input_data &in, uint32_t *reclaimed) {
  Op op{ctx};
  int rc = op.init(pool_from_context(ctx), in);
  if (rc) return rc;
  return op.reclaim(in, reclaimed);
}

int enqueue_tree(Context *ctx, int kernel, uint32_t n) {
  Bound b{ctx, kernel};
  svo_update_input_data in;
  const char *label = kernel == SVO_KERNEL_PRUNE ? "prune" : "graft";
  if (!b.input(&in))
    return fail(SVO_ERR_MISSING_ARGUMENT, std::string("EnqueueNDRangeKernel: ") + label +
                                              " argument inputData (UpdateInputData) is not set");
  Pool pool;
  pool.blocks = b.buf("blocks");
  pool.usages = b.buf("usage");
  pool.parents = b.buf("parents");
  pool.residency = b.buf("parentResidency");
  pool.requests = b.buf("childRequests");
  pool.request_id = b.buf("childRequestId");
  if (kernel == SVO_KERNEL_PRUNE) {
    pool.bases = b.buf("bases");
    const Pool::B rec = b.buf("pruning"), side = b.buf("pruningBlockData"), loc = b.buf("pruningAddresses");
    if (!rec.ptr || rec.bytes / sizeof(svo_pruning) < n)
      return fail(SVO_ERR_MISSING_ARGUMENT, "EnqueueNDRangeKernel: prune argument pruning is not set or too small");
    return prune_impl(ctx, pool, in, (const svo_pruning *)rec.ptr, n, (const svo_block *)side.ptr,
                      count_of(side, 16), (const uint64_t *)loc.ptr, count_of(loc, 24), nullptr);
  }
  const Pool::B rec = b.buf("grafting"), gb = b.buf("graftingBlocks"), loc = b.buf("graftingAddresses");
  if (!rec.ptr || rec.bytes / sizeof(svo_grafting) < n)
    return fail(SVO_ERR_MISSING_ARGUMENT, "EnqueueNDRangeKernel: graft argument grafting is not set or too small");
  return graft_impl(ctx, pool, in, (const svo_grafting *)rec.ptr, n, (const svo_block *)gb.ptr, count_of(gb, 16),
                    (const uint64_t *)loc.ptr, count_of(loc, 24), nullptr);
}

}  // namespace svo

using namespace svo;

extern "C" int svo_prune(svo_context *ctx, const svo_update_input_data *input, const svo_pruning *records, uint32_t n,
                         const svo_block *surface, uint32_t n_surface, const uint64_t *locations, uint32_t n_locations,
                         svo_tree_stats *stats) {
  if (!ctx || !input) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_prune: null argument");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(SVO_ERR_HIP, "hipSetDevice");
  const void *d_rec, *d_side, *d_loc;
  int rc = stage(ctx, kTbRecords, records, (size_t)n * sizeof(svo_pruning), &d_rec);
  if (!rc) rc = stage(ctx, kTbSide, surface, (size_t)n_surface * sizeof(svo_block), &d_side);
  if (!rc) rc = stage(ctx, kTbLocations, locations, (size_t)n_locations * 24, &d_loc);
  if (rc) return rc;
  return prune_impl(ctx, pool_from_context(ctx), *input, (const svo_pruning *)d_rec, n, (const svo_block *)d_side,
                    n_surface, (const uint64_t *)d_loc, n_locations, stats);
}

extern "C" int svo_graft(svo_context *ctx, const svo_update_input_data *input, const svo_grafting *records, uint32_t n,
                         const svo_block *graft_blocks, uint32_t n_graft_blocks, const uint64_t *locations,
                         uint32_t n_locations, svo_tree_stats *stats) {
  if (!ctx || !input) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_graft: null argument");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(SVO_ERR_HIP, "hipSetDevice");
  const void *d_rec, *d_gb, *d_loc;
  int rc = stage(ctx, kTbRecords, records, (size_t)n * sizeof(svo_grafting), &d_rec);
  if (!rc) rc = stage(ctx, kTbSide, graft_blocks, (size_t)n_graft_blocks * sizeof(svo_block), &d_gb);
  if (!rc) rc = stage(ctx, kTbLocations, locations, (size_t)n_locations * 24, &d_loc);
  if (rc) return rc;
  return graft_impl(ctx, pool_from_context(ctx), *input, (const svo_grafting *)d_rec, n, (const svo_block *)d_gb,
                    n_graft_blocks, (const uint64_t *)d_loc, n_locations, stats);
}

extern "C" int svo_reclaim(svo_context *ctx, const svo_update_input_data *input, uint32_t *reclaimed) {
  if (!ctx || !input) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_reclaim: null argument");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(SVO_ERR_HIP, "hipSetDevice");
  uint32_t got = 0;
  const int rc = reclaim_device(ctx, *input, &got);
  if (!rc && reclaimed) *reclaimed = got;
  return rc;
}

extern "C" int svo_collect(svo_context *ctx, const svo_update_input_data *input, uint32_t need, uint32_t *groups,
                           uint32_t *count) {
  if (!ctx || !input) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_collect: null argument");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(SVO_ERR_HIP, "hipSetDevice");
  Op op{ctx};
  int rc = op.init(pool_from_context(ctx), *input);
  if (rc) return rc;
  if ((rc = op.collect(*input, host_count(need), need))) return rc;
  if ((rc = op.finish(nullptr))) return rc;
  const uint32_t got = op.host[kTcGot];
  if (groups && got) {
    hipError_t e = hipMemcpyAsync(groups, op.ts.list, (size_t)got * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "svo_collect: read");
  }
  if (count) *count = got;
  return SVO_OK;
}
