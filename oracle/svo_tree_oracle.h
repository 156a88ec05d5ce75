/*
 * svo_tree_oracle.h -- CPU oracle for on-GPU tree management (prune, graft, collect).
 *
 * TEST INFRASTRUCTURE ONLY, like svo_oracle.h: used by tests/ as the checker and never
 * linked or loaded by the product.
 *
 * The reference kernels (prune, graft, dereference/helpDereference, findAddress in
 * SvoTracer.Kernel/wavefront.cl:1097-1350) cannot run as written: the stale-group scan
 * never starts, the DFS never terminates, the dereference walk is serialised by a spin
 * lock whose lanes can livelock, and the host never dispatches them.  This file states
 * the intended operation as a batch with phases, so that a parallel implementation has
 * one deterministic result:
 *
 *   - every record of one dispatch resolves its target against the tree as it stood
 *     when the dispatch began (read-only), then the ticks along every resolved path are
 *     touched, then subtrees are freed, then fields are written;
 *   - when several records write the same field of the same target, the record with the
 *     highest index wins (what a sequential loop over the records would leave);
 *   - freeing is a set operation: the detached child group and every descendant group
 *     that is neither inviolate (65535) nor multi-parent (65534) gets tick 0;
 *     multi-parent groups are unlinked from their parent lists instead, in ascending
 *     (group, parent block) order;
 *   - collection takes free or stale groups in ascending group order, pass by pass with
 *     the age threshold halved after each pass; grafts take consecutive runs of the
 *     collected list in record order, each graft's groups in pre-order DFS order.
 */
#ifndef SVO_TREE_ORACLE_H
#define SVO_TREE_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#include "svo_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Parent, wavefront.cl:19-22 */
typedef struct {
  uint32_t parent_address;
  uint32_t next_element;
} orc_parent;

/* Pruning, wavefront.cl:24-36 (natural layout, 20 bytes) */
typedef struct {
  uint8_t properties;
  uint8_t depth;
  uint16_t pad0;
  uint32_t address;
  uint16_t chunk;
  uint16_t pad1;
  uint32_t colour_address;
  uint32_t child_address;
} orc_pruning;

/* Grafting, wavefront.cl:38-43 (natural layout, 16 bytes) */
typedef struct {
  uint32_t graft_data_address;
  uint32_t graft_total_size;
  uint8_t depth;
  uint8_t pad[3];
  uint32_t graft_address;
} orc_grafting;

/* UpdateInputData, wavefront.cl:77-84 (natural layout, 20 bytes) */
typedef struct {
  uint8_t base_depth;
  uint8_t pad0;
  uint16_t tick;
  uint32_t max_child_request_id;
  uint32_t memory_size;
  uint32_t offset;
  uint32_t graft_size;
} orc_update_input;

/* The mutable pool a management dispatch works on. */
typedef struct {
  uint16_t *bases;
  uint32_t n_bases;
  orc_block *blocks;
  uint32_t n_blocks;
  orc_usage *usages;
  uint32_t n_usages;
  orc_parent *parents; /* may be NULL when n_parents == 0 */
  uint32_t n_parents;
  uint8_t *residency;  /* one byte per parent element */
  orc_child_request *requests;
  uint32_t request_capacity;
  uint32_t *request_count;
} orc_pool;

typedef struct {
  uint32_t applied;   /* records whose target resolved and whose effects were applied */
  uint32_t skipped;   /* target freed (tick 0) or missing (a child request was made) */
  uint32_t invalid;   /* an index out of range or a malformed graft */
  uint32_t requests;  /* child requests made (also those beyond capacity) */
  uint32_t collected; /* groups collected */
  uint32_t placed;    /* grafts linked into the tree */
  uint32_t unlinks;   /* multi-parent unlink events */
} orc_tree_stats;

/* prune, wavefront.cl:1215-1266. */
int orc_prune(orc_pool *pool, const orc_update_input *in, const orc_pruning *records, uint32_t n,
              const orc_block *surface, uint32_t n_surface, const uint64_t *locations, uint32_t n_locations,
              orc_tree_stats *stats);

/* The collection step of graft, wavefront.cl:1277-1305: claims `need` groups.  The
 * claimed group numbers are written to out_groups (capacity `need`); returns how many. */
uint32_t orc_collect(orc_pool *pool, const orc_update_input *in, uint32_t need, uint32_t *out_groups);

/* graft, wavefront.cl:1268-1350 (collection included). */
int orc_graft(orc_pool *pool, const orc_update_input *in, const orc_grafting *records, uint32_t n,
              const orc_block *graft_blocks, uint32_t n_graft_blocks, const uint64_t *locations,
              uint32_t n_locations, orc_tree_stats *stats);

/* Garbage collection: frees every stale group (tick age > offset; cut from its parent, its
 * subtree freed); returns the number reclaimed. */
uint32_t orc_reclaim(orc_pool *pool, const orc_update_input *in);
/* Compaction: with L live groups, live single-parent groups at or above L move in ascending
 * order into the free groups below L in ascending order; links follow.  Returns the number
 * moved. */
uint32_t orc_compact(orc_pool *pool, const orc_update_input *in);

/* Age of a tick at frame `now` on the 1..65533 tick cycle (StateManager.cs:113-122). */
uint32_t orc_tick_age(uint16_t tick, uint16_t now);

#ifdef __cplusplus
}
#endif
#endif
