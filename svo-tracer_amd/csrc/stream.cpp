// Streaming on the device: the procedural terrain, the seed tree, the per-frame update that
// turns the frame's child requests into grafted bricks (the role of WorldManager.UpdateGrafting
// and MainWindow.RunGraft, WorldManager.cs:84-103 and MainWindow.cs:307-337, fed by
// requestChild, wavefront.cl:362-375), and the pool's garbage collection (reclaim of stale
// groups, then compaction).
//
// Everything runs on the context's stream.  The per-frame update sizes its launches from
// device-side counts (the request count, the distinct blocks, the grafts' needs) and reads its
// counters back once at the end, plus once after each collection pass that may be followed by
// another; the garbage collection, which runs rarely, reads its list sizes back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>

#include "svo_internal.hpp"
#include "svo_runtime.hpp"
#include "svo_stream_kernels.hpp"

namespace svo {
terrain::Spec terrain_spec(const svo_geometry &g);
shapes::Noise noise_spec(const svo_geometry &g);
}

using namespace svo;

struct svo_terrain {
  int device = 0;
  std::shared_ptr<DeviceBuffer> top, hmin, hmax;
  TerrainView view{};
};

namespace {

enum StreamBuf { kSbFirst, kSbState, kSbFlags, kSbList, kSbCounts, kSbFrom, kSbTo, kSbRemap, kSbCount };
// kShTotal: scan totals of select(); the update's own counters (kSbState): stored requests,
// all requests, distinct requesting blocks, free groups
enum : uint32_t { kShTotal = 0, kShStored = 0, kShRequests = 1, kShUnique = 2, kShFree = 3, kShSlots = 16 };
// the pinned mirror holds the update's counters, then a tree-management counter block
constexpr uint32_t kHostWords = kShSlots + kTcSlots;

constexpr uint64_t kUsageInviolate = 0xFFFFFFFF0000FFFFull;  // tick 65535, parent UINT32_MAX
constexpr uint64_t kUsageFree = 0xFFFFFFFF00000000ull;       // tick 0, parent UINT32_MAX

struct Stream {
  Context *ctx;
  uint32_t *counters = nullptr;  // device, kShSlots (in the counts scratch tail)

  int grow(int k, size_t bytes, void **out) {
    int rc = ensure_scratch(ctx->stream_scratch[k], bytes);
    if (rc) return rc;
    *out = ctx->stream_scratch[k]->ptr;
    return SVO_OK;
  }
  int host_mirror() {
    if (ctx->stream_host) return SVO_OK;
    hipError_t e = hipHostMalloc((void **)&ctx->stream_host, kHostWords * sizeof(uint32_t), hipHostMallocDefault);
    return e == hipSuccess ? SVO_OK : hip_fail(e, "hipHostMalloc");
  }
  int read_u32(const uint32_t *dev, uint32_t *out) {
    hipError_t e = hipMemcpyAsync(ctx->stream_host, dev, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "stream: counter read");
    *out = ctx->stream_host[0];
    return SVO_OK;
  }
  int launch(int r, const char *what) { return r ? fail(SVO_ERR_HIP, std::string("stream: launch ") + what) : SVO_OK; }

  // The number of groups of [lo, hi) satisfying `pred` into *total (device); nothing read back.
  int count_into(const svo_usage *u, uint32_t lo, uint32_t hi, uint32_t pred, uint32_t *total) {
    const uint32_t chunks = select_chunks(lo, hi);
    void *p;
    int rc = grow(kSbCounts, ((size_t)chunks + kShSlots) * 4, &p);
    if (rc) return rc;
    if (!chunks) return launch(tree_fill_u32(total, 0, 1, ctx->stream), "fill");
    uint32_t *counts = (uint32_t *)p;
    if ((rc = launch(launch_select_count(u, lo, hi, pred, counts, ctx->stream), "select count"))) return rc;
    uint32_t *sums = nullptr;
    if ((rc = context_scan_sums(ctx, scan_span_sums_bytes(chunks), &sums))) return rc;
    return launch(launch_scan_u32(counts, counts, chunks, total, sums, ctx->stream), "scan");
  }

  // Ascending list of the groups of [lo, hi) satisfying `pred` into scratch slot `slot`;
  // returns its length.  counts / scan use the kSbCounts slot.
  int select(const svo_usage *u, uint32_t lo, uint32_t hi, uint32_t pred, int slot, uint32_t **list, uint32_t *n) {
    *n = 0;
    const uint32_t chunks = select_chunks(lo, hi);
    void *p;
    int rc = grow(kSbCounts, ((size_t)chunks + kShSlots) * 4, &p);
    if (rc) return rc;
    uint32_t *counts = (uint32_t *)p;
    counters = counts + chunks;
    if (!chunks) {
      *list = nullptr;
      return SVO_OK;
    }
    if ((rc = launch(launch_select_count(u, lo, hi, pred, counts, ctx->stream), "select count"))) return rc;
    uint32_t *sums = nullptr;
    if ((rc = context_scan_sums(ctx, scan_span_sums_bytes(chunks), &sums))) return rc;
    if ((rc = launch(launch_scan_u32(counts, counts, chunks, counters + kShTotal, sums, ctx->stream), "scan")))
      return rc;
    uint32_t total = 0;
    if ((rc = read_u32(counters + kShTotal, &total))) return rc;
    if (slot < 0 || !total) {
      *n = total;
      *list = nullptr;
      return SVO_OK;
    }
    if ((rc = grow(slot, (size_t)total * 4, &p))) return rc;
    *list = (uint32_t *)p;
    if ((rc = launch(launch_select_scatter(u, lo, hi, pred, counts, *list, ctx->stream), "select scatter")))
      return rc;
    *n = total;
    return SVO_OK;
  }
};

struct PoolBufs {
  std::shared_ptr<DeviceBuffer> bases, blocks, usages, parents, residency, requests, request_id;
};

PoolBufs pool_bufs(Context *ctx) {
  PoolBufs b;
  b.bases = ctx->get(SVO_BUF_BASES);
  b.blocks = ctx->get(SVO_BUF_BLOCKS);
  b.usages = ctx->get(SVO_BUF_USAGES);
  b.parents = ctx->get(SVO_BUF_PARENTS);
  b.residency = ctx->get(SVO_BUF_PARENT_RESIDENCY);
  b.requests = ctx->get(SVO_BUF_CHILD_REQUESTS);
  b.request_id = ctx->get(SVO_BUF_CHILD_REQUEST_ID);
  return b;
}

PoolView view_of(const PoolBufs &b, uint32_t B, uint32_t now) {
  PoolView pv{};
  pv.bases = b.bases ? (uint16_t *)b.bases->ptr : nullptr;
  pv.n_bases = b.bases ? (uint32_t)std::min<size_t>(b.bases->bytes / 2, 0xffffffffu) : 0;
  pv.blocks = (svo_block *)b.blocks->ptr;
  pv.n_blocks = (uint32_t)std::min<size_t>(b.blocks->bytes / 16, 0xffffffffu);
  pv.usages = (svo_usage *)b.usages->ptr;
  pv.n_usages = (uint32_t)std::min<size_t>(b.usages->bytes / 8, 0xffffffffu);
  pv.parents = b.parents ? (svo_parent *)b.parents->ptr : nullptr;
  pv.n_parents = b.parents ? (uint32_t)(b.parents->bytes / 8) : 0;
  pv.residency = b.residency && b.residency->bytes >= pv.n_parents ? (uint8_t *)b.residency->ptr : nullptr;
  pv.B = B;
  pv.now = now;
  return pv;
}

// A named buffer of at least `bytes` (kept when large enough, replaced otherwise).
int ensure_named(Context *ctx, int name, size_t bytes, std::shared_ptr<DeviceBuffer> *out) {
  auto cur = ctx->get(name);
  if (cur && cur->bytes >= bytes) {
    *out = cur;
    return SVO_OK;
  }
  if (cur) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  }
  std::shared_ptr<DeviceBuffer> nb;
  int rc = new_buffer(std::max<size_t>(bytes, 256), &nb);
  if (rc) return rc;
  ctx->buffers[name] = nb;
  *out = nb;
  return SVO_OK;
}

int compact_impl(Context *ctx, const PoolView &pv, uint32_t *moved) {
  *moved = 0;
  Stream st{ctx};
  int rc = st.host_mirror();
  if (rc) return rc;
  // groups whose 8 blocks all lie in Blocks (the remap table still covers every usage)
  const uint32_t n = std::min(pv.n_usages, pv.n_blocks / 8u);
  uint32_t live = 0, nf = 0, nt = 0, *from = nullptr, *to = nullptr;
  if ((rc = st.select(pv.usages, 0, n, kSelLive, -1, &from, &live))) return rc;
  if ((rc = st.select(pv.usages, live, n, kSelMovable, kSbFrom, &from, &nf))) return rc;
  if (!nf) return SVO_OK;
  if ((rc = st.select(pv.usages, 0, live, kSelFree, kSbTo, &to, &nt))) return rc;
  const uint32_t m = std::min(nf, nt);
  void *p;
  if ((rc = st.grow(kSbRemap, (size_t)pv.n_usages * 4, &p))) return rc;
  if ((rc = st.launch(launch_compact_moves(pv, from, to, m, (uint32_t *)p, ctx->stream), "compact"))) return rc;
  *moved = m;
  return SVO_OK;
}

int free_groups(Context *ctx, const PoolView &pv, uint32_t *n_free) {
  Stream st{ctx};
  int rc = st.host_mirror();
  if (rc) return rc;
  uint32_t *unused = nullptr;
  return st.select(pv.usages, 0, pv.n_usages, kSelFree, -1, &unused, n_free);
}

int check_pool(const PoolBufs &b, const char *fn) {
  if (!b.blocks || !b.blocks->ptr) return fail(SVO_ERR_MISSING_ARGUMENT, std::string(fn) + ": Blocks is not bound");
  if (!b.usages || !b.usages->ptr) return fail(SVO_ERR_MISSING_ARGUMENT, std::string(fn) + ": Usages is not bound");
  return SVO_OK;
}

// Requests the update reads at most: the ChildRequests records (capped by MaxChildRequestId).
uint32_t stored_requests(const PoolBufs &b, uint32_t max_child_request_id) {
  if (!b.requests || !b.request_id || b.request_id->bytes < 4) return 0;
  uint32_t stored = (uint32_t)std::min<size_t>(b.requests->bytes / sizeof(svo_child_request), 0xffffffffu);
  return max_child_request_id ? std::min(stored, max_child_request_id) : stored;
}

// The first-request table of the pool (u64 per block); a table created here starts cleared,
// at generation 0.
int first_table(Context *ctx, Stream &st, uint32_t n_blocks, unsigned long long **first) {
  const std::shared_ptr<DeviceBuffer> before = ctx->stream_scratch[kSbFirst];
  void *p;
  int rc = st.grow(kSbFirst, (size_t)n_blocks * 8, &p);
  if (rc) return rc;
  *first = (unsigned long long *)p;
  if (ctx->stream_scratch[kSbFirst] == before) return SVO_OK;
  ctx->stream_gen = 0;
  return st.launch(tree_fill_u64(*first, 0ull, ctx->stream_scratch[kSbFirst]->bytes / 8, ctx->stream), "fill");
}

// Everything an update of this pool and request buffer allocates, so that the frame loop does
// not allocate: the first-request table, the per-request flags, the brick list,
// Grafting / GraftingBlocks for max_bricks and the tree-management scratch.
int reserve_update(Context *ctx, const PoolBufs &b, uint32_t base_depth, uint32_t max_bricks) {
  Stream st{ctx};
  const PoolView pv = view_of(b, base_depth, 0);
  unsigned long long *first;
  int rc = first_table(ctx, st, pv.n_blocks, &first);
  void *p;
  if (!rc) rc = st.grow(kSbState, kShSlots * 4, &p);
  const uint32_t stored = stored_requests(b, 0);
  if (!rc && stored) {
    const uint32_t nb_max = std::min(stored, max_bricks);
    std::shared_ptr<DeviceBuffer> rec, gb;
    if (!rc) rc = st.grow(kSbFlags, (size_t)stored * 4, &p);
    if (!rc) rc = st.grow(kSbList, (size_t)nb_max * 4, &p);
    if (!rc) rc = st.grow(kSbCounts, ((size_t)(stored + 1023u) / 1024u + kShSlots) * 4, &p);
    if (!rc) rc = ensure_named(ctx, SVO_BUF_GRAFTING, (size_t)nb_max * sizeof(svo_grafting), &rec);
    if (!rc) rc = ensure_named(ctx, SVO_BUF_GRAFTING_BLOCKS, (size_t)nb_max * 8 * sizeof(svo_block), &gb);
  }
  if (!rc) rc = reserve_tree(ctx, base_depth, std::min(stored, max_bricks));
  return rc;
}

}  // namespace

extern "C" int svo_terrain_create(svo_context *ctx, const svo_geometry *geometry, svo_terrain **out) {
  if (!ctx || !geometry || !out) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_terrain_create: null argument");
  *out = nullptr;
  if (geometry->kind != SVO_GEOMETRY_FBM_HEIGHTFIELD)
    return fail(SVO_ERR_INVALID_ARGUMENT, "svo_terrain_create: only SVO_GEOMETRY_FBM_HEIGHTFIELD streams");
  const terrain::Spec spec = terrain_spec(*geometry);
  if (spec.lres > kMaxTerrainLres)
    return fail(SVO_ERR_INVALID_ARGUMENT, "svo_terrain_create: resolution above 32768 samples per side");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(SVO_ERR_HIP, "hipSetDevice");
  auto *t = new svo_terrain();
  t->device = ctx->device;
  const size_t res = (size_t)1 << spec.lres;
  size_t pyramid = 0;
  for (uint32_t L = 0; L < spec.lres; L++) {
    t->view.level_offset[L] = pyramid;
    pyramid += (size_t)1 << (2 * L);
  }
  int rc = new_buffer(res * res * 4, &t->top);
  if (!rc) rc = new_buffer(std::max<size_t>(pyramid, 1) * 4, &t->hmin);
  if (!rc) rc = new_buffer(std::max<size_t>(pyramid, 1) * 4, &t->hmax);
  if (rc) {
    delete t;
    return rc;
  }
  t->view.spec = spec;
  t->view.top = (float *)t->top->ptr;
  t->view.hmin = (float *)t->hmin->ptr;
  t->view.hmax = (float *)t->hmax->ptr;
  if (launch_terrain(t->view, ctx->stream)) {
    delete t;
    return fail(SVO_ERR_HIP, "svo_terrain_create: launch");
  }
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    delete t;
    return hip_fail(e, "svo_terrain_create");
  }
  *out = t;
  return SVO_OK;
}

extern "C" int svo_terrain_destroy(svo_terrain *t) {
  if (!t) return SVO_OK;
  (void)hipSetDevice(t->device);
  (void)hipDeviceSynchronize();
  delete t;
  return SVO_OK;
}

extern "C" int svo_terrain_heights(const svo_terrain *t, float *dst, size_t count) {
  if (!t || !dst) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_terrain_heights: null argument");
  const size_t n = t->top->bytes / 4;
  if (count < n) return fail(SVO_ERR_BUFFER_TOO_SMALL, "svo_terrain_heights: destination too small");
  if (hipSetDevice(t->device) != hipSuccess) return fail(SVO_ERR_HIP, "hipSetDevice");
  hipError_t e = hipMemcpy(dst, t->top->ptr, n * 4, hipMemcpyDeviceToHost);
  return e == hipSuccess ? SVO_OK : hip_fail(e, "svo_terrain_heights");
}

extern "C" int svo_stream_seed(svo_context *ctx, const svo_terrain *t, uint32_t base_depth) {
  if (!ctx || !t) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_stream_seed: null argument");
  if (base_depth < 1 || base_depth > 8) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_stream_seed: BaseDepth must be 1..8");
  if (t->device != ctx->device) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_stream_seed: terrain of another device");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(SVO_ERR_HIP, "hipSetDevice");
  const PoolBufs b = pool_bufs(ctx);
  int rc = check_pool(b, "svo_stream_seed");
  if (rc) return rc;
  if (!b.bases || b.bases->bytes < (size_t)pow_sum(base_depth) * 2)
    return fail(SVO_ERR_BUFFER_TOO_SMALL, "svo_stream_seed: Bases smaller than PowSum(BaseDepth) chunks");
  const size_t dense_blocks = (size_t)1 << (3 * (base_depth + 2));
  const size_t dense_groups = dense_blocks / 8;
  if (b.blocks->bytes < dense_blocks * 16)
    return fail(SVO_ERR_BUFFER_TOO_SMALL, "svo_stream_seed: Blocks smaller than the dense level 8^(BaseDepth+2)");
  if (b.usages->bytes < dense_groups * 8)
    return fail(SVO_ERR_BUFFER_TOO_SMALL, "svo_stream_seed: Usages smaller than the dense groups 8^(BaseDepth+1)");
  hipStream_t s = ctx->stream;
  if (launch_seed(t->view, (uint16_t *)b.bases->ptr, base_depth, (uint4 *)b.blocks->ptr, s))
    return fail(SVO_ERR_HIP, "svo_stream_seed: launch");
  const size_t rest = b.blocks->bytes - dense_blocks * 16;
  if (rest && hipMemsetAsync((char *)b.blocks->ptr + dense_blocks * 16, 0, rest, s) != hipSuccess)
    return fail(SVO_ERR_HIP, "svo_stream_seed: clear");
  const size_t n_usages = b.usages->bytes / 8;
  auto *u = (unsigned long long *)b.usages->ptr;
  if (tree_fill_u64(u, kUsageInviolate, dense_groups, s) ||
      tree_fill_u64(u + dense_groups, kUsageFree, n_usages - dense_groups, s))
    return fail(SVO_ERR_HIP, "svo_stream_seed: usages");
  if ((rc = reserve_update(ctx, b, base_depth, 65536u))) return rc;
  hipError_t e = hipStreamSynchronize(s);
  return e == hipSuccess ? SVO_OK : hip_fail(e, "svo_stream_seed");
}

extern "C" int svo_compact(svo_context *ctx, const svo_update_input_data *input, uint32_t *moved) {
  if (!ctx || !input) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_compact: null argument");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(SVO_ERR_HIP, "hipSetDevice");
  const PoolBufs b = pool_bufs(ctx);
  int rc = check_pool(b, "svo_compact");
  if (rc) return rc;
  uint32_t m = 0;
  rc = compact_impl(ctx, view_of(b, input->base_depth, input->tick), &m);
  if (!rc) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "svo_compact");
  }
  if (!rc && moved) *moved = m;
  return rc;
}

extern "C" int svo_stream_update(svo_context *ctx, const svo_terrain *t, const svo_stream_options *o,
                                 svo_stream_stats *stats) {
  if (!ctx || !t || !o) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_stream_update: null argument");
  if (o->base_depth < 1 || o->base_depth > 8)
    return fail(SVO_ERR_INVALID_ARGUMENT, "svo_stream_update: BaseDepth must be 1..8");
  if (o->max_depth <= o->base_depth + 2u || o->max_depth > 60)
    return fail(SVO_ERR_INVALID_ARGUMENT, "svo_stream_update: max_depth must be in BaseDepth+3..60");
  if (t->device != ctx->device) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_stream_update: terrain of another device");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(SVO_ERR_HIP, "hipSetDevice");
  const PoolBufs b = pool_bufs(ctx);
  int rc = check_pool(b, "svo_stream_update");
  if (rc) return rc;
  svo_stream_stats out;
  std::memset(&out, 0, sizeof out);
  Stream st{ctx};
  if ((rc = st.host_mirror())) return rc;
  hipStream_t s = ctx->stream;
  const PoolView pv = view_of(b, o->base_depth, o->tick);
  const uint32_t max_bricks = o->max_bricks ? o->max_bricks : 65536u;

  void *p;
  if ((rc = st.grow(kSbState, kShSlots * 4, &p))) return rc;
  uint32_t *state = (uint32_t *)p;
  if ((rc = st.launch(tree_fill_u32(state, 0, kShSlots, s), "fill"))) return rc;
  const uint32_t *tree_counters = nullptr;
  svo_update_input_data in;
  std::memset(&in, 0, sizeof in);
  in.base_depth = (uint8_t)o->base_depth;
  in.tick = o->tick;
  in.offset = o->prune_age;

  // 1. the frame's child requests (the count keeps counting beyond what was stored): at most
  // `stored`, the exact number on the device.  Requests the last direct frame kept are read
  // where it kept them, under the same limits as ChildRequests.
  uint32_t stored = 0;
  RequestSource src{nullptr, nullptr, nullptr};
  if (ctx->kept.valid) {
    if (b.request_id && b.request_id->bytes >= 4) {
      stored = ctx->kept.capacity;
      if (o->max_child_request_id) stored = std::min(stored, o->max_child_request_id);
      if (b.requests)
        stored = (uint32_t)std::min<size_t>(stored, b.requests->bytes / sizeof(svo_child_request));
    }
    src = RequestSource{ctx->kept.records, ctx->kept.cones, ctx->kept.keys};
  } else {
    stored = stored_requests(b, o->max_child_request_id);
    if (stored) src.records = (const svo_child_request *)b.requests->ptr;
  }
  if (stored &&
      (rc = st.launch(launch_request_count((const uint32_t *)b.request_id->ptr, stored, state, s), "request count")))
    return rc;
  // 2. distinct requesting blocks that can take a child level, ranked in request order
  if (stored) {
    unsigned long long *first;
    if ((rc = first_table(ctx, st, pv.n_blocks, &first))) return rc;
    // every update takes the next generation; entries of earlier ones never match
    if (++ctx->stream_gen == 0) {
      if ((rc = st.launch(tree_fill_u64(first, 0ull, ctx->stream_scratch[kSbFirst]->bytes / 8, s), "fill"))) return rc;
      ctx->stream_gen = 1;
    }
    if ((rc = st.grow(kSbFlags, (size_t)stored * 4, &p))) return rc;
    uint32_t *flags = (uint32_t *)p;
    if ((rc = st.grow(kSbList, (size_t)std::max(1u, std::min(stored, max_bricks)) * 4, &p))) return rc;
    uint32_t *list = (uint32_t *)p;
    if ((rc = st.launch(launch_request_unique(src, stored, state + kShStored,
                                              first, ctx->stream_gen, (const uint4 *)pv.blocks, pv.n_blocks,
                                              o->max_depth, flags, s),
                        "request unique")))
      return rc;
    const uint32_t chunks = (stored + 1023u) / 1024u;
    if ((rc = st.grow(kSbCounts, ((size_t)chunks + kShSlots) * 4, &p))) return rc;
    uint32_t *sums = nullptr;
    if ((rc = context_scan_sums(ctx, scan_span_sums_bytes(chunks), &sums))) return rc;
    if ((rc = st.launch(launch_rank(flags, stored, (uint32_t *)p, nullptr, state + kShUnique, sums, s, list,
                                    max_bricks, state + kShStored),
                        "rank")))
      return rc;
    // 3. records and bricks of the first min(unique, max_bricks) into Grafting /
    // GraftingBlocks (sized for max_bricks), then the graft
    const uint32_t nb_max = std::min(stored, max_bricks);
    in.graft_size = 8u * nb_max;
    std::shared_ptr<DeviceBuffer> rec, gb;
    if ((rc = ensure_named(ctx, SVO_BUF_GRAFTING, (size_t)nb_max * sizeof(svo_grafting), &rec))) return rc;
    if ((rc = ensure_named(ctx, SVO_BUF_GRAFTING_BLOCKS, (size_t)nb_max * 8 * sizeof(svo_block), &gb))) return rc;
    if ((rc = st.launch(launch_bricks(t->view, src, list, nb_max,
                                      state + kShUnique, (svo_grafting *)rec->ptr, (uint4 *)gb->ptr, s),
                        "bricks")))
      return rc;
    if ((rc = graft_device_counted(ctx, in, (const svo_grafting *)rec->ptr, nb_max, state + kShUnique,
                                   (const svo_block *)gb->ptr, 8u * nb_max, &tree_counters)))
      return rc;
  }
  // the free groups, then every counter of the update in one read
  if ((rc = st.count_into(pv.usages, 0, pv.n_usages, kSelFree, state + kShFree))) return rc;
  uint32_t *host = ctx->stream_host;
  hipError_t e = hipMemcpyAsync(host, state, kShSlots * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && tree_counters)
    e = hipMemcpyAsync(host + kShSlots, tree_counters, kTcSlots * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "svo_stream_update: counter read");
  out.requests = stored ? host[kShRequests] : 0u;
  out.unique = host[kShUnique];
  if (tree_counters) {
    svo_tree_stats ts;
    if ((rc = stats_from(host + kShSlots, &ts))) return rc;
    out.bricks = ts.placed;
    out.collected = ts.collected;
  }
  uint32_t n_free = host[kShFree];
  // 4. garbage collection when free groups fall below gc_free_percent of the pool
  const bool want_gc = !(o->flags & SVO_STREAM_NO_GC) &&
                       ((o->flags & SVO_STREAM_FORCE_GC) ||
                        (uint64_t)n_free * 100u < (uint64_t)o->gc_free_percent * pv.n_usages);
  if (want_gc) {
    out.gc_runs = 1;
    if ((rc = reclaim_device(ctx, in, &out.reclaimed))) return rc;
    if ((rc = compact_impl(ctx, pv, &out.moved))) return rc;
    if ((rc = free_groups(ctx, pv, &n_free))) return rc;
  }
  out.free_groups = n_free;
  out.pool_groups = pv.n_usages;
  if (want_gc && (e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "svo_stream_update");
  if (stats) *stats = out;
  return SVO_OK;
}

// ---------------------------------------------------------------------------
// svo_terrain_build_tree: TreeBuilder (BuildTree, TreeBuilder.cs:102-256) on the device
// ---------------------------------------------------------------------------

namespace {

struct BuildLevel {
  uint32_t n = 0;      // nodes (dense level: x-outer positions, 8^(B+2))
  uint32_t n_int = 0;  // nodes that get children
  std::shared_ptr<DeviceBuffer> flag, irank, cells, alloc, writes, rank, start, post;
};

struct Builder {
  Context *ctx;
  hipStream_t s;
  std::vector<std::shared_ptr<DeviceBuffer>> keep;

  int buf(size_t bytes, std::shared_ptr<DeviceBuffer> *out) { return new_buffer(std::max<size_t>(bytes, 4), out); }
  template <class T>
  T *ptr(const std::shared_ptr<DeviceBuffer> &b) {
    return b ? (T *)b->ptr : nullptr;
  }
  int launched(int r, const char *what) {
    return r ? fail(SVO_ERR_HIP, std::string("svo_terrain_build_tree: launch ") + what) : SVO_OK;
  }
  int read(const void *dev, void *host, size_t bytes) {
    hipError_t e = hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? SVO_OK : hip_fail(e, "svo_terrain_build_tree: read back");
  }
  // Exclusive ranks of a 0/1 flag array and the number of ones.
  int rank(const uint32_t *flags, uint32_t n, std::shared_ptr<DeviceBuffer> *out, uint32_t *total) {
    std::shared_ptr<DeviceBuffer> counts;
    int rc = buf(((size_t)n / 1024 + 2) * 4, &counts);
    if (!rc) rc = buf((size_t)n * 4, out);
    if (rc) return rc;
    uint32_t *c = ptr<uint32_t>(counts);
    const uint32_t chunks = (n + 1023) / 1024;
    uint32_t *sums = nullptr;
    if ((rc = context_scan_sums(ctx, scan_span_sums_bytes(chunks), &sums))) return rc;
    if ((rc = launched(launch_rank(flags, n, c, ptr<uint32_t>(*out), c + chunks, sums, s), "rank"))) return rc;
    *total = 0;
    return n ? read(c + chunks, total, 4) : SVO_OK;
  }
  int fill(const std::shared_ptr<DeviceBuffer> &b, int byte) {
    hipError_t e = hipMemsetAsync(b->ptr, byte, b->bytes, s);
    return e == hipSuccess ? SVO_OK : hip_fail(e, "svo_terrain_build_tree: clear");
  }
};

}  // namespace

namespace {

int build_tree_device(Context *ctx, const GeomView &g, uint32_t base_depth, uint32_t max_depth, uint32_t max_size,
                      int with_usages, uint16_t usage_tick, svo_tree **out, uint32_t *block_count) {
  Builder bd{ctx, ctx->stream, {}};
  const uint32_t B = base_depth, L0 = B + 2;
  const uint32_t dense_groups = 1u << (3 * (B + 1)), n0 = 1u << (3 * L0);
  // FIFO length, as TreeBuilder.BuildTree sizes it (maxSize at least the dense level)
  const uint64_t dense_min = 1ull << (3 * B + 6);
  const uint64_t msize = std::max<uint64_t>(max_size, dense_min);
  const uint64_t range_end = std::max<uint64_t>(msize >> 3, dense_groups);
  int rc;

  // Bases, and the level-(B+1) cells: mixed ones in PopulateInitialBlocks' write order,
  // empty ones (ascending) as the head of the free-group FIFO
  std::shared_ptr<DeviceBuffer> bases, mixed, empty, mixed_rank, empty_rank, empties;
  if ((rc = bd.buf((size_t)pow_sum(B) * 2, &bases)) || (rc = bd.buf((size_t)dense_groups * 4, &mixed)) ||
      (rc = bd.buf((size_t)dense_groups * 4, &empty)))
    return rc;
  if ((rc = bd.launched(launch_bt_bases(g, bd.ptr<uint16_t>(bases), B, bd.s), "bases"))) return rc;
  if ((rc = bd.launched(launch_bt_octant_flags(bd.ptr<uint16_t>(bases), B, bd.ptr<uint32_t>(mixed),
                                               bd.ptr<uint32_t>(empty), bd.s),
                        "octants")))
    return rc;
  uint32_t n_mixed = 0, n_empties = 0;
  if ((rc = bd.rank(bd.ptr<uint32_t>(mixed), dense_groups, &mixed_rank, &n_mixed))) return rc;
  if ((rc = bd.rank(bd.ptr<uint32_t>(empty), dense_groups, &empty_rank, &n_empties))) return rc;
  if ((rc = bd.buf((size_t)n_empties * 4, &empties))) return rc;
  if ((rc = bd.launched(launch_bt_scatter_index(bd.ptr<uint32_t>(empty), bd.ptr<uint32_t>(empty_rank), dense_groups,
                                                bd.ptr<uint32_t>(empties), bd.s),
                        "empties")))
    return rc;
  const uint64_t capacity64 = (uint64_t)n_empties + (range_end - dense_groups);
  const uint32_t capacity = (uint32_t)std::min<uint64_t>(capacity64, 0xffffffffu);

  // expansion, level by level: which blocks get children (geometry only)
  std::vector<BuildLevel> lv(1);
  lv[0].n = n0;
  if ((rc = bd.buf((size_t)n0 * 4, &lv[0].flag))) return rc;
  if ((rc = bd.launched(launch_bt_dense_flags(g, bd.ptr<uint16_t>(bases), B, max_depth,
                                              bd.ptr<uint32_t>(lv[0].flag), bd.s),
                        "dense flags")))
    return rc;
  if ((rc = bd.rank(bd.ptr<uint32_t>(lv[0].flag), n0, &lv[0].irank, &lv[0].n_int))) return rc;
  if ((rc = bd.buf((size_t)lv[0].n_int * 12, &lv[0].cells))) return rc;
  if ((rc = bd.launched(launch_bt_dense_cells(bd.ptr<uint32_t>(lv[0].flag), bd.ptr<uint32_t>(lv[0].irank), L0,
                                              bd.ptr<uint3>(lv[0].cells), bd.s),
                        "dense cells")))
    return rc;
  while (lv.back().n_int > 0) {
    const BuildLevel &p = lv.back();
    const uint32_t L = L0 + (uint32_t)lv.size();
    const uint64_t n = 8ull * p.n_int;
    if (n > 0xffffffffull) return fail(SVO_ERR_CAPACITY, "svo_terrain_build_tree: a level above 2^32 blocks");
    BuildLevel c;
    c.n = (uint32_t)n;
    if ((rc = bd.buf((size_t)n * 4, &c.flag))) return rc;
    if ((rc = bd.launched(launch_bt_level_flags(g, bd.ptr<uint3>(p.cells), c.n, L, max_depth,
                                                bd.ptr<uint32_t>(c.flag), bd.s),
                          "level flags")))
      return rc;
    if ((rc = bd.rank(bd.ptr<uint32_t>(c.flag), c.n, &c.irank, &c.n_int))) return rc;
    if ((rc = bd.buf((size_t)c.n_int * 12, &c.cells))) return rc;
    if ((rc = bd.launched(launch_bt_level_cells(bd.ptr<uint3>(p.cells), bd.ptr<uint32_t>(c.flag),
                                                bd.ptr<uint32_t>(c.irank), c.n, bd.ptr<uint3>(c.cells), bd.s),
                          "level cells")))
      return rc;
    lv.push_back(std::move(c));
  }

  // bottom-up: groups and block writes under every internal block
  for (size_t i = lv.size(); i-- > 0;) {
    BuildLevel &l = lv[i];
    if ((rc = bd.buf((size_t)l.n_int * 4, &l.alloc)) || (rc = bd.buf((size_t)l.n_int * 4, &l.writes))) return rc;
    const BuildLevel *c = i + 1 < lv.size() ? &lv[i + 1] : nullptr;
    if ((rc = bd.launched(launch_bt_sizes(l.n_int, c ? bd.ptr<uint32_t>(c->flag) : nullptr,
                                          c ? bd.ptr<uint32_t>(c->irank) : nullptr,
                                          c ? bd.ptr<uint32_t>(c->alloc) : nullptr,
                                          c ? bd.ptr<uint32_t>(c->writes) : nullptr, bd.ptr<uint32_t>(l.alloc),
                                          bd.ptr<uint32_t>(l.writes), bd.s),
                          "sizes")))
      return rc;
  }
  // top-down: pre-order ranks (FIFO entries) and write positions; the dense level's
  // internal blocks are the roots, in x-outer / z-inner order
  std::shared_ptr<DeviceBuffer> totals;
  if ((rc = bd.buf(16, &totals))) return rc;
  uint32_t *tot = bd.ptr<uint32_t>(totals);
  if ((rc = bd.buf((size_t)lv[0].n_int * 4, &lv[0].rank)) || (rc = bd.buf((size_t)lv[0].n_int * 4, &lv[0].start)))
    return rc;
  uint32_t root_total[2] = {0, 0};
  if (lv[0].n_int) {
    uint32_t *sums = nullptr;
    if ((rc = context_scan_sums(ctx, scan_span_sums_bytes(lv[0].n_int), &sums))) return rc;
    if ((rc = bd.launched(launch_scan_u32(bd.ptr<uint32_t>(lv[0].alloc), bd.ptr<uint32_t>(lv[0].rank), lv[0].n_int,
                                          tot, sums, bd.s),
                          "root ranks")) ||
        (rc = bd.launched(launch_scan_u32(bd.ptr<uint32_t>(lv[0].writes), bd.ptr<uint32_t>(lv[0].start),
                                          lv[0].n_int, tot + 1, sums, bd.s),
                          "root writes")))
      return rc;
    if ((rc = bd.read(tot, root_total, 8))) return rc;
    uint32_t last_rank = 0;
    if ((rc = bd.read(bd.ptr<uint32_t>(lv[0].rank) + lv[0].n_int - 1, &last_rank, 4))) return rc;
    // TreeBuilder fails when the FIFO is empty at a dense-level block (TreeBuilder.cs:131-136)
    if (last_rank >= capacity) return fail(SVO_ERR_CAPACITY, "svo_terrain_build_tree: free address queue exhausted");
  }
  for (size_t i = 0; i + 1 < lv.size(); i++) {
    BuildLevel &l = lv[i], &c = lv[i + 1];
    if ((rc = bd.buf((size_t)c.n_int * 4, &c.rank)) || (rc = bd.buf((size_t)c.n_int * 4, &c.start)) ||
        (rc = bd.buf((size_t)c.n * 4, &c.post)))
      return rc;
    if ((rc = bd.launched(launch_bt_ranks(l.n_int, bd.ptr<uint32_t>(l.rank), bd.ptr<uint32_t>(l.start),
                                          bd.ptr<uint32_t>(c.flag), bd.ptr<uint32_t>(c.irank),
                                          bd.ptr<uint32_t>(c.alloc), bd.ptr<uint32_t>(c.writes),
                                          bd.ptr<uint32_t>(c.rank), bd.ptr<uint32_t>(c.start),
                                          bd.ptr<uint32_t>(c.post), bd.s),
                          "ranks")))
      return rc;
  }

  // the blocks: groups in use are the dense level and the FIFO entries that were taken
  const uint64_t taken = std::min<uint64_t>(root_total[0], capacity);
  const uint64_t groups64 = dense_groups + (taken > n_empties ? taken - n_empties : 0);
  if (groups64 * 8 > 0xffffffffull) return fail(SVO_ERR_CAPACITY, "svo_terrain_build_tree: more than 2^32 blocks");
  const uint32_t n_groups = (uint32_t)groups64;
  std::shared_ptr<DeviceBuffer> blocks, post, gparent, maxes;
  if ((rc = bd.buf((size_t)n_groups * 128, &blocks)) || (rc = bd.buf((size_t)n_groups * 32, &post)) ||
      (rc = bd.buf((size_t)n_groups * 4, &gparent)) || (rc = bd.buf(16, &maxes)))
    return rc;
  if ((rc = bd.fill(blocks, 0)) || (rc = bd.fill(post, 0xff)) || (rc = bd.fill(gparent, 0xff)) ||
      (rc = bd.fill(maxes, 0)))
    return rc;
  GroupFifo fifo{bd.ptr<uint32_t>(empties), n_empties, dense_groups, capacity};
  BuildOut o{bd.ptr<uint4>(blocks), bd.ptr<uint32_t>(post), bd.ptr<uint32_t>(gparent),
             bd.ptr<unsigned long long>(maxes)};
  if ((rc = bd.launched(launch_bt_write_dense(g, bd.ptr<uint16_t>(bases), B, bd.ptr<uint32_t>(lv[0].flag),
                                              bd.ptr<uint32_t>(lv[0].irank), bd.ptr<uint32_t>(lv[0].rank),
                                              bd.ptr<uint32_t>(mixed_rank), fifo, o, bd.s),
                        "write dense")))
    return rc;
  for (size_t i = 1; i < lv.size(); i++) {
    const BuildLevel &p = lv[i - 1], &c = lv[i];
    const bool deeper = i + 1 < lv.size() || c.n_int;
    if ((rc = bd.launched(launch_bt_write_level(g, bd.ptr<uint3>(p.cells), bd.ptr<uint32_t>(p.rank), c.n,
                                                L0 + (uint32_t)i, deeper ? bd.ptr<uint32_t>(c.flag) : nullptr,
                                                bd.ptr<uint32_t>(c.irank), bd.ptr<uint32_t>(c.rank),
                                                bd.ptr<uint32_t>(c.post), fifo, o, bd.s),
                          "write level")))
      return rc;
  }

  // BlockCount (TreeBuilder.cs:145, 222, 254).  The dense level counts every block it writes
  // (max address + 1); BuildBlocks then sets count = a + 1 only when count < a, so an
  // address written when count equals it leaves count unchanged.  The count therefore ends
  // at M + 1 or M (M the largest address); which one follows from the run of "records"
  // (writes above everything written before) that ends at M: the first record of the run
  // sets count = a + 1, the next (a + 1) finds count equal to it and leaves it, the next
  // sets it again, and so on.
  unsigned long long mx[2] = {0, 0};
  if ((rc = bd.read(bd.ptr<unsigned long long>(maxes), mx, 16))) return rc;
  const uint32_t M = (uint32_t)mx[0], M1 = (uint32_t)mx[1];
  uint32_t count = n_mixed ? M1 + 1 : 0u;  // nothing is written without a mixed level-(B+1) cell
  if (n_mixed && M > M1) {
    // records among the DFS writes: a is one when every written address above it was
    // written later (post(a) < post(b) for all written b > a); all addresses above M1 are
    // DFS writes, so positions compare within one phase
    uint64_t window = 4096;
    for (;;) {
      const uint32_t lo = ((uint64_t)(M - M1) > window) ? (uint32_t)(M - window + 1) : M1 + 1;
      std::vector<uint32_t> pw(M - lo + 1);
      if ((rc = bd.read(bd.ptr<uint32_t>(post) + lo, pw.data(), pw.size() * 4))) return rc;
      uint32_t suf = pw.back(), a = M;  // a: lowest record of the run so far
      while (a > lo) {
        const uint32_t q = pw[a - 1 - lo];
        if (q == 0xffffffffu || q >= suf) break;  // a - 1 unwritten or written after a higher address
        suf = q;
        a--;
      }
      if (a > lo || lo == M1 + 1) {
        const uint32_t run = M - a + 1;
        // the run's first record: count becomes a + 1 unless it directly follows the dense
        // level's count (a == M1 + 1, count already a), in which case it stays a
        bool plus_one = !(a == M1 + 1);
        if (run % 2 == 0) plus_one = !plus_one;
        count = plus_one ? M + 1 : M;
        break;
      }
      window *= 4;  // the run reaches the window's bottom: look further down
    }
  }
  if (block_count) *block_count = count;

  // outputs: the context's Bases / Blocks (/ Usages) and, if asked, a host tree
  const uint32_t out_groups = (count + 7) / 8;
  std::shared_ptr<DeviceBuffer> cb, ck;
  if ((rc = new_buffer((size_t)pow_sum(B) * 2, &cb)) || (rc = new_buffer((size_t)count * 16, &ck))) return rc;
  cb->flags = SVO_BUFFER_READ_ONLY;
  ck->flags = SVO_BUFFER_READ_ONLY;
  hipError_t e = hipMemcpyAsync(cb->ptr, bases->ptr, cb->bytes, hipMemcpyDeviceToDevice, bd.s);
  if (e == hipSuccess && count) e = hipMemcpyAsync(ck->ptr, blocks->ptr, ck->bytes, hipMemcpyDeviceToDevice, bd.s);
  if (e != hipSuccess) return hip_fail(e, "svo_terrain_build_tree: copy");
  std::shared_ptr<DeviceBuffer> cu;
  if (with_usages) {
    if ((rc = new_buffer((size_t)out_groups * sizeof(svo_usage), &cu))) return rc;
    cu->flags = SVO_BUFFER_READ_WRITE;
    if ((rc = bd.launched(launch_bt_usages(bd.ptr<uint32_t>(gparent), out_groups, dense_groups, usage_tick,
                                           (svo_usage *)cu->ptr, bd.s),
                          "usages")))
      return rc;
  }
  if (out) {
    std::unique_ptr<svo_tree> tree(new svo_tree());
    tree->t.base_depth = B;
    tree->t.max_depth = max_depth;
    tree->t.block_count = count;
    tree->t.bases.resize(pow_sum(B));
    tree->t.blocks.resize(count);
    tree->t.group_parent.resize(out_groups);
    if ((rc = bd.read(bases->ptr, tree->t.bases.data(), tree->t.bases.size() * 2))) return rc;
    if (count && (rc = bd.read(blocks->ptr, tree->t.blocks.data(), (size_t)count * 16))) return rc;
    if (out_groups && (rc = bd.read(gparent->ptr, tree->t.group_parent.data(), (size_t)out_groups * 4))) return rc;
    *out = tree.release();
  }
  e = hipStreamSynchronize(bd.s);
  if (e != hipSuccess) return hip_fail(e, "svo_terrain_build_tree");
  ctx->buffers[SVO_BUF_BASES] = cb;
  ctx->buffers[SVO_BUF_BLOCKS] = ck;
  if (cu) ctx->buffers[SVO_BUF_USAGES] = cu;
  ctx->tree_max_depth = max_depth;
  return SVO_OK;
}

int check_build_args(svo_context *ctx, uint32_t base_depth, uint32_t max_depth, svo_tree **out, const char *what) {
  if (out) *out = nullptr;
  if (base_depth < 1 || base_depth > 7) return fail(SVO_ERR_INVALID_ARGUMENT, std::string(what) + ": base_depth must be 1..7");
  if (max_depth > 20) return fail(SVO_ERR_INVALID_ARGUMENT, std::string(what) + ": max_depth must be <= 20");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(SVO_ERR_HIP, "hipSetDevice");
  return SVO_OK;
}

}  // namespace

extern "C" int svo_terrain_build_tree(svo_context *ctx, const svo_terrain *t, uint32_t base_depth, uint32_t max_depth,
                                      uint32_t max_size, int with_usages, uint16_t usage_tick, svo_tree **out,
                                      uint32_t *block_count) {
  if (!ctx || !t) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_terrain_build_tree: null argument");
  int rc = check_build_args(ctx, base_depth, max_depth, out, "svo_terrain_build_tree");
  if (rc) return rc;
  if (t->device != ctx->device)
    return fail(SVO_ERR_INVALID_ARGUMENT, "svo_terrain_build_tree: terrain of another device");
  GeomView g{};
  g.kind = SVO_GEOMETRY_FBM_HEIGHTFIELD;
  g.t = t->view;
  return build_tree_device(ctx, g, base_depth, max_depth, max_size, with_usages, usage_tick, out, block_count);
}

extern "C" int svo_tree_build_device(svo_context *ctx, const svo_geometry *geometry, uint32_t base_depth,
                                     uint32_t max_depth, uint32_t max_size, int with_usages, uint16_t usage_tick,
                                     svo_tree **out, uint32_t *block_count) {
  if (!ctx || !geometry) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_tree_build_device: null argument");
  int rc = check_build_args(ctx, base_depth, max_depth, out, "svo_tree_build_device");
  if (rc) return rc;
  GeomView g{};
  g.kind = geometry->kind;
  switch (geometry->kind) {
    case SVO_GEOMETRY_FBM_HEIGHTFIELD: {
      svo_terrain *t = nullptr;
      if ((rc = svo_terrain_create(ctx, geometry, &t))) return rc;
      g.t = t->view;
      rc = build_tree_device(ctx, g, base_depth, max_depth, max_size, with_usages, usage_tick, out, block_count);
      svo_terrain_destroy(t);
      return rc;
    }
    case SVO_GEOMETRY_SPHERES:
    case SVO_GEOMETRY_CUBES: {
      const size_t per = geometry->kind == SVO_GEOMETRY_SPHERES ? 4 : 6;
      if (!geometry->shapes && geometry->count) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_tree_build_device: no shapes");
      std::shared_ptr<DeviceBuffer> shp;
      if ((rc = new_buffer(std::max<size_t>(per * geometry->count * 4, 4), &shp))) return rc;
      if (geometry->count) {
        hipError_t e = hipMemcpy(shp->ptr, geometry->shapes, per * geometry->count * 4, hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "svo_tree_build_device: shapes upload");
      }
      g.shapes = (const float *)shp->ptr;
      g.count = geometry->count;
      return build_tree_device(ctx, g, base_depth, max_depth, max_size, with_usages, usage_tick, out, block_count);
    }
    case SVO_GEOMETRY_SPHERE_NOISE:
      g.noise = noise_spec(*geometry);
      return build_tree_device(ctx, g, base_depth, max_depth, max_size, with_usages, usage_tick, out, block_count);
    default:
      return fail(SVO_ERR_INVALID_ARGUMENT, "svo_tree_build_device: unknown geometry kind");
  }
}
