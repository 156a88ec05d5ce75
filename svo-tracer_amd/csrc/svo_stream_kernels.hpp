// Launch interface of the streaming kernels: GPU terrain, seed tree, child-request
// consumption into bricks, group selection and pool compaction.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "svo.h"
#include "svo_shapes.hpp"
#include "svo_terrain.hpp"
#include "svo_tree_kernels.hpp"

namespace svo {

constexpr uint32_t kMaxTerrainLres = 15;

// Device view of a terrain: the finest height grid and its min/max pyramid.
struct TerrainView {
  terrain::Spec spec;
  float *top;                                // 2^lres x 2^lres heights, row-major (j * res + i)
  float *hmin, *hmax;                        // pyramid levels 0..lres-1, concatenated
  size_t level_offset[kMaxTerrainLres + 1];  // element offset of level L in hmin / hmax
};

// Group predicates of the selection kernels.
enum : uint32_t { kSelLive = 0, kSelFree = 1, kSelMovable = 2 };

int launch_terrain(const TerrainView &t, hipStream_t s);
int launch_seed(const TerrainView &t, uint16_t *bases, uint32_t B, uint4 *blocks, hipStream_t s);
// flags[i] = 1 when request i asks for a level (TreeSize >= 1), is the first such request for
// its block address, and the block can take a child level (address in range, childless,
// 1 <= depth < max_depth, an octant with both air and geometry).  keys / vals are a
// table of mask+1 entries, filled with UINT32_MAX beforehand.
// counters[0] = min(*request_id, stored), counters[1] = *request_id
int launch_request_count(const uint32_t *request_id, uint32_t stored, uint32_t *counters, hipStream_t s);
// The frame's child requests in request order: request i is records[i] (ChildRequests), or,
// for requests a frame kept, records[cones[i]] with its address, depth and TreeSize also in
// keys[cones[i]] (x = address, y = depth | TreeSize << 8).
struct RequestSource {
  const svo_child_request *records;
  const uint32_t *cones;  // null: records in request order
  const uint2 *keys;
};
// flags[i] (i < min(*n, n_max), not written beyond) = 1 for the first request of each block
// that can take a child level among the *n stored requests; first: u64 per block of the
// pool, entries of generations below gen are ignored (zero-filled when allocated)
int launch_request_unique(const RequestSource &req, uint32_t n_max, const uint32_t *n, unsigned long long *first,
                          uint32_t gen, const uint4 *blocks, uint32_t n_blocks, uint32_t max_depth, uint32_t *flags,
                          hipStream_t s);
// bricks of the first min(*unique, nb_max) listed requests
int launch_bricks(const TerrainView &t, const RequestSource &req, const uint32_t *list, uint32_t nb_max,
                  const uint32_t *unique, svo_grafting *records, uint4 *bricks, hipStream_t s);
// rank[i] = number of flagged entries before i (rank may be null); counts holds one entry per
// 1024 flags.  With a list, the flagged entries of rank below limit are listed in rank order.
// span_sums: scratch of the scan over the counts (scan_span_sums_bytes of the count of
// chunks).  n_dev: the flags are defined only below min(n, *n_dev) (a device-side count).
int launch_rank(const uint32_t *flags, uint32_t n, uint32_t *counts, uint32_t *rank, uint32_t *total,
                uint32_t *span_sums, hipStream_t s, uint32_t *list = nullptr, uint32_t limit = 0,
                const uint32_t *n_dev = nullptr);
// Groups of [lo, hi) that satisfy a predicate, in ascending order: per-chunk counts, then the
// scatter of the list given the scanned counts.
uint32_t select_chunks(uint32_t lo, uint32_t hi);
int launch_select_count(const svo_usage *u, uint32_t lo, uint32_t hi, uint32_t pred, uint32_t *counts, hipStream_t s);
int launch_select_scatter(const svo_usage *u, uint32_t lo, uint32_t hi, uint32_t pred, const uint32_t *offsets,
                          uint32_t *list, hipStream_t s);
// Moves groups from[i] -> to[i] (i < m), fixing Child fields, Usage.Parent back-pointers and
// parent-list entries; `remap` has one entry per group.
int launch_compact_moves(const PoolView &p, const uint32_t *from, const uint32_t *to, uint32_t m, uint32_t *remap,
                         hipStream_t s);

// ---- device TreeBuilder (BuildTree, TreeBuilder.cs:102-256) ----
// The builder's geometry: a terrain, spheres (centre, radius), boxes (min, max) or the
// sphere-plus-noise SDF, classified by the rules the host builder shares.
struct GeomView {
  uint32_t kind;  // SVO_GEOMETRY_*
  TerrainView t;
  const float *shapes;
  uint32_t count;
  shapes::Noise noise;
};
// Free-group FIFO: the n_empties sorted empty level-(B+1) cells, then dense_groups, ... up
// to `capacity` entries in all.
struct GroupFifo {
  const uint32_t *empties;
  uint32_t n_empties, dense_groups, capacity;
};
// Outputs of the write kernels.
struct BuildOut {
  uint4 *blocks;
  uint32_t *post;           // per block address: position in TreeBuilder's write sequence (phase-local)
  uint32_t *group_parent;   // per group: the block whose children it holds
  unsigned long long *max;  // [0] largest address written, [1] largest dense-level address written
};
int launch_bt_bases(const GeomView &g, uint16_t *bases, uint32_t B, hipStream_t s);
// Level-(B+1) cells: mixed_zyx in z-outer / x-inner order, empty by Morton index.
int launch_bt_octant_flags(const uint16_t *bases, uint32_t B, uint32_t *mixed_zyx, uint32_t *empty, hipStream_t s);
// Dense level in x-outer / z-inner order: the written blocks that get children.
int launch_bt_dense_flags(const GeomView &g, const uint16_t *bases, uint32_t B, uint32_t max_depth, uint32_t *flag,
                          hipStream_t s);
int launch_bt_scatter_index(const uint32_t *flag, const uint32_t *rank, uint32_t n, uint32_t *out, hipStream_t s);
int launch_bt_dense_cells(const uint32_t *flag, const uint32_t *rank, uint32_t L0, uint3 *cells, hipStream_t s);
// Level L's nodes (child k & 7 of internal block k >> 3 of level L-1): those that get children.
int launch_bt_level_flags(const GeomView &g, const uint3 *parent_cells, uint32_t n, uint32_t L, uint32_t max_depth,
                          uint32_t *flag, hipStream_t s);
int launch_bt_level_cells(const uint3 *parent_cells, const uint32_t *flag, const uint32_t *rank, uint32_t n,
                          uint3 *cells, hipStream_t s);
// Subtree allocations and writes of each internal block (cflag null: no internal children).
int launch_bt_sizes(uint32_t n_int, const uint32_t *cflag, const uint32_t *crank, const uint32_t *calloc,
                    const uint32_t *cwrites, uint32_t *alloc, uint32_t *writes, hipStream_t s);
// Ranks and write positions of the next level from this level's.
int launch_bt_ranks(uint32_t n_int, const uint32_t *rank, const uint32_t *start, const uint32_t *cflag,
                    const uint32_t *crank, const uint32_t *calloc, const uint32_t *cwrites, uint32_t *child_rank,
                    uint32_t *child_start, uint32_t *child_post, hipStream_t s);
int launch_bt_write_dense(const GeomView &g, const uint16_t *bases, uint32_t B, const uint32_t *flag,
                          const uint32_t *irank, const uint32_t *rank, const uint32_t *mixed_rank,
                          const GroupFifo &fifo, const BuildOut &o, hipStream_t s);
int launch_bt_write_level(const GeomView &g, const uint3 *parent_cells, const uint32_t *parent_rank, uint32_t n,
                          uint32_t L, const uint32_t *flag, const uint32_t *irank, const uint32_t *rank,
                          const uint32_t *post, const GroupFifo &fifo, const BuildOut &o, hipStream_t s);
int launch_bt_usages(const uint32_t *group_parent, uint32_t n_groups, uint32_t dense_groups, uint32_t tick,
                     svo_usage *u, hipStream_t s);

}  // namespace svo
