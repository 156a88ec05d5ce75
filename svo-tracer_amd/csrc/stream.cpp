// Streaming on the device: the procedural terrain, the seed tree, the per-frame update that
// turns the frame's child requests into grafted bricks (the role of WorldManager.UpdateGrafting
// and MainWindow.RunGraft, WorldManager.cs:84-103 and MainWindow.cs:307-337, fed by
// requestChild, wavefront.cl:362-375), and the pool's garbage collection (reclaim of stale
// groups, then compaction).
//
// Everything runs on the context's stream; the host reads a few totals back (request count,
// distinct blocks, list sizes) to size the launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>

#include "svo_internal.hpp"
#include "svo_runtime.hpp"
#include "svo_stream_kernels.hpp"

namespace svo {
terrain::Spec terrain_spec(const svo_geometry &g);
}

using namespace svo;

struct svo_terrain {
  int device = 0;
  std::shared_ptr<DeviceBuffer> top, hmin, hmax;
  TerrainView view{};
};

namespace {

enum StreamBuf { kSbKeys, kSbVals, kSbFlags, kSbRank, kSbCounts, kSbFrom, kSbTo, kSbRemap, kSbCount };
enum : uint32_t { kShTotal = 0, kShSlots = 16 };

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
    hipError_t e = hipHostMalloc((void **)&ctx->stream_host, kShSlots * sizeof(uint32_t), hipHostMallocDefault);
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
    if ((rc = launch(launch_scan_u32(counts, counts, chunks, counters + kShTotal, ctx->stream), "scan"))) return rc;
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

  // 1. the frame's child requests (the count keeps counting beyond what was stored)
  uint32_t n = 0;
  if (b.requests && b.request_id && b.request_id->bytes >= 4) {
    uint32_t count = 0;
    if ((rc = st.read_u32((const uint32_t *)b.request_id->ptr, &count))) return rc;
    const uint32_t stored = (uint32_t)std::min<size_t>(b.requests->bytes / sizeof(svo_child_request), 0xffffffffu);
    n = std::min(count, std::min(stored, o->max_child_request_id ? o->max_child_request_id : stored));
    out.requests = count;
  }
  // 2. distinct requesting blocks that can take a child level, ranked in request order
  uint32_t unique = 0, *flags = nullptr, *rank = nullptr;
  if (n) {
    uint32_t cap = 64;
    while (cap < 2ull * n && cap < (1u << 31)) cap <<= 1;
    void *p;
    if ((rc = st.grow(kSbKeys, (size_t)cap * 4, &p))) return rc;
    uint32_t *keys = (uint32_t *)p;
    if ((rc = st.grow(kSbVals, (size_t)cap * 4, &p))) return rc;
    uint32_t *vals = (uint32_t *)p;
    if ((rc = st.grow(kSbFlags, (size_t)n * 4, &p))) return rc;
    flags = (uint32_t *)p;
    if ((rc = st.grow(kSbRank, (size_t)n * 4, &p))) return rc;
    rank = (uint32_t *)p;
    if (tree_fill_u32(keys, 0xffffffffu, cap, s) || tree_fill_u32(vals, 0xffffffffu, cap, s))
      return fail(SVO_ERR_HIP, "stream: fill");
    if ((rc = st.launch(launch_request_unique((const svo_child_request *)b.requests->ptr, n, keys, vals, cap - 1u,
                                              (const uint4 *)pv.blocks, pv.n_blocks, o->max_depth, flags, s),
                        "request unique")))
      return rc;
    const uint32_t chunks = (n + 1023u) / 1024u;
    if ((rc = st.grow(kSbCounts, ((size_t)chunks + kShSlots) * 4, &p))) return rc;
    st.counters = (uint32_t *)p + chunks;
    if ((rc = st.launch(launch_rank(flags, n, (uint32_t *)p, rank, st.counters + kShTotal, s), "rank"))) return rc;
    if ((rc = st.read_u32(st.counters + kShTotal, &unique))) return rc;
  }
  out.unique = unique;
  // 3. records and bricks into Grafting / GraftingBlocks, then the graft
  const uint32_t nb = std::min(unique, max_bricks);
  svo_update_input_data in;
  std::memset(&in, 0, sizeof in);
  in.base_depth = (uint8_t)o->base_depth;
  in.tick = o->tick;
  in.offset = o->prune_age;
  in.graft_size = 8u * nb;
  if (nb) {
    std::shared_ptr<DeviceBuffer> rec, gb;
    if ((rc = ensure_named(ctx, SVO_BUF_GRAFTING, (size_t)nb * sizeof(svo_grafting), &rec))) return rc;
    if ((rc = ensure_named(ctx, SVO_BUF_GRAFTING_BLOCKS, (size_t)nb * 8 * sizeof(svo_block), &gb))) return rc;
    if ((rc = st.launch(launch_bricks(t->view, (const svo_child_request *)b.requests->ptr, n, flags, rank, nb,
                                      (svo_grafting *)rec->ptr, (uint4 *)gb->ptr, s),
                        "bricks")))
      return rc;
    svo_tree_stats ts;
    if ((rc = graft_device(ctx, in, (const svo_grafting *)rec->ptr, nb, (const svo_block *)gb->ptr, 8u * nb, &ts)))
      return rc;
    out.bricks = ts.placed;
    out.collected = ts.collected;
  }
  // 4. garbage collection when free groups fall below gc_free_percent of the pool
  uint32_t n_free = 0;
  if ((rc = free_groups(ctx, pv, &n_free))) return rc;
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
  hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "svo_stream_update");
  if (stats) *stats = out;
  return SVO_OK;
}
