/*
 * svo_tree_oracle.c -- CPU oracle for prune / graft / collect (TEST INFRASTRUCTURE ONLY).
 *
 * A sequential statement of the batch semantics described in svo_tree_oracle.h.  Each
 * phase is a plain loop over the records in index order, so "the highest record index
 * wins" falls out of the loop order; the GPU implementation reaches the same state with
 * winner tables and level-synchronous passes.
 */
#include "svo_tree_oracle.h"

#include <stdlib.h>
#include <string.h>

#define U32MAX 0xffffffffu
#define TICK_MULTI 65534u
#define TICK_INVIOLATE 65535u
#define TICK_PERIOD 65533u
#define MAX_FREE_LEVELS 64u
#define MAX_GRAFT_DEPTH 64u

enum { RES_OK = 0, RES_MISS = 1, RES_SKIP = 2, RES_INVALID = 3, RES_BASE = 4 };

typedef struct {
  uint32_t g, pb;
} link_t;

typedef struct {
  link_t *v;
  size_t n, cap;
} links_t;

static void links_push(links_t *l, uint32_t g, uint32_t pb) {
  if (l->n == l->cap) {
    l->cap = l->cap ? 2 * l->cap : 64;
    l->v = (link_t *)realloc(l->v, l->cap * sizeof(link_t));
  }
  l->v[l->n].g = g;
  l->v[l->n].pb = pb;
  l->n++;
}

/* Tick age on the 1..65533 cycle (IncrementTick, StateManager.cs:113-122): the reference's
 * wrap-aware test at wavefront.cl:1283-1285 intends age = frames since `tick`. */
uint32_t orc_tick_age(uint16_t tick, uint16_t now) {
  if (tick <= now) return (uint32_t)now - tick;
  return (uint32_t)now + TICK_PERIOD - tick;
}

static int stale(uint16_t t, uint16_t now, uint32_t offset) {
  if (t == 0) return 1;
  if (t >= TICK_MULTI) return 0;
  return orc_tick_age(t, now) > offset;
}

/* updateUsage-style touch on a walked group (findAddress, wavefront.cl:1197-1199). */
static void touch(orc_pool *p, uint32_t g, uint16_t now) {
  if (g < p->n_usages && p->usages[g].tick < TICK_MULTI) p->usages[g].tick = now;
}

/* findAddress, wavefront.cl:1191-1211, read-only: walk from the dense level to `depth`.
 * Returns RES_OK (addr = target), RES_MISS (addr = childless block, *level = its level)
 * or RES_INVALID. */
static int walk(const orc_pool *p, uint32_t B, const uint64_t loc[3], uint32_t depth, uint32_t *addr,
                uint32_t *level) {
  uint32_t a = orc_base_location(B + 2u, loc);
  for (uint32_t L = B + 2u; L < depth; L++) {
    if (a >= p->n_blocks) return RES_INVALID;
    const uint32_t c = p->blocks[a].child;
    if (c == U32MAX) {
      *addr = a;
      *level = L;
      return RES_MISS;
    }
    a = c + orc_chunk_position(L, loc);
  }
  if (a >= p->n_blocks) return RES_INVALID;
  *addr = a;
  return RES_OK;
}

/* Ticks along a walk that resolved or missed (the same path: only ticks changed since). */
static void touch_walk(orc_pool *p, uint32_t B, const uint64_t loc[3], uint32_t depth, uint16_t now) {
  uint32_t a = orc_base_location(B + 2u, loc);
  for (uint32_t L = B + 2u; L < depth; L++) {
    if (a >= p->n_blocks) return;
    touch(p, a >> 3, now);
    const uint32_t c = p->blocks[a].child;
    if (c == U32MAX) return;
    a = c + orc_chunk_position(L, loc);
  }
}

typedef struct {
  int status;
  uint32_t addr;
  uint32_t level; /* RES_MISS: level of the childless block */
} res_t;

/* Target resolution shared by prune and graft records. */
static res_t resolve(const orc_pool *p, const orc_update_input *in, uint8_t depth, uint32_t address,
                     const uint64_t *locations, uint32_t n_locations) {
  res_t r = {RES_INVALID, 0, 0};
  const uint32_t B = in->base_depth;
  if (depth == 0xffu) {
    if (address >= p->n_blocks) return r;
    const uint32_t g = address >> 3;
    r.addr = address;
    r.status = (g < p->n_usages && p->usages[g].tick == 0) ? RES_SKIP : RES_OK;
    return r;
  }
  if (address >= n_locations || depth < B + 2u || depth > 63u) return r;
  r.status = walk(p, B, locations + 3 * (size_t)address, depth, &r.addr, &r.level);
  return r;
}

static void apply_touches(orc_pool *p, const orc_update_input *in, uint8_t depth, uint32_t address,
                          const uint64_t *locations, const res_t *r) {
  if (r->status != RES_OK && r->status != RES_MISS) return;
  if (depth == 0xffu)
    touch(p, r->addr >> 3, in->tick);
  else
    touch_walk(p, in->base_depth, locations + 3 * (size_t)address, depth, in->tick);
}

/* requestChild from findAddress (wavefront.cl:1201-1205): appended in record order; beyond
 * capacity the request is dropped and the counter still counts it. */
static void request(orc_pool *p, const orc_update_input *in, const res_t *r, uint8_t depth, const uint64_t *loc) {
  const uint32_t id = *p->request_count;
  *p->request_count = id + 1u;
  if (id >= p->request_capacity || !p->requests) return;
  orc_child_request q;
  memset(&q, 0, sizeof q);
  q.address = r->addr;
  q.tick = in->tick;
  q.depth = (uint8_t)r->level;
  q.location[0] = loc[0];
  q.location[1] = loc[1];
  q.location[2] = loc[2];
  q.tree_size = (uint8_t)(depth - r->level);
  p->requests[id] = q;
}

static int cmp_link(const void *a, const void *b) {
  const link_t *x = (const link_t *)a, *y = (const link_t *)b;
  if (x->g != y->g) return x->g < y->g ? -1 : 1;
  if (x->pb != y->pb) return x->pb < y->pb ? -1 : 1;
  return 0;
}

/* Removes parent block `pb` from the parent list of multi-parent group g
 * (helpDereference, wavefront.cl:1105-1148, as intended: the element is unlinked; when one
 * parent remains the group becomes single-parent with tick `now`; when none remains the
 * group is free). */
static void unlink_parent(orc_pool *p, uint32_t g, uint32_t pb, uint16_t now) {
  if (g >= p->n_usages || p->usages[g].tick != TICK_MULTI || !p->parents) return;
  uint32_t prev = U32MAX, e = p->usages[g].parent, steps = 0;
  while (e < p->n_parents && steps < p->n_parents) {
    if (p->parents[e].parent_address == pb) break;
    prev = e;
    e = p->parents[e].next_element;
    steps++;
  }
  if (e >= p->n_parents || steps >= p->n_parents) return;
  const uint32_t nx = p->parents[e].next_element;
  if (prev == U32MAX)
    p->usages[g].parent = nx;
  else
    p->parents[prev].next_element = nx;
  p->parents[e].next_element = U32MAX;
  if (p->residency) p->residency[e] = 0;
  const uint32_t h = p->usages[g].parent;
  if (h >= p->n_parents) {
    p->usages[g].tick = 0;
    p->usages[g].parent = U32MAX;
  } else if (p->parents[h].next_element == U32MAX) {
    p->usages[g].tick = now;
    p->usages[g].parent = p->parents[h].parent_address;
    if (p->residency) p->residency[h] = 0;
  }
}

/* dereference / helpDereference (wavefront.cl:1097-1189) as a set operation over the
 * subtrees below `roots` ((group, parent block) pairs).  Groups flagged in `stop` are left
 * alone and not descended into. */
static uint32_t free_subtrees(orc_pool *p, const links_t *roots, const uint8_t *stop, uint16_t now) {
  links_t cur = {0, 0, 0}, next = {0, 0, 0}, events = {0, 0, 0};
  for (size_t i = 0; i < roots->n; i++) links_push(&cur, roots->v[i].g, roots->v[i].pb);
  for (uint32_t level = 0; level < MAX_FREE_LEVELS && cur.n; level++) {
    next.n = 0;
    for (size_t i = 0; i < cur.n; i++) {
      const uint32_t g = cur.v[i].g;
      if (g >= p->n_usages || (stop && stop[g])) continue;
      const uint16_t t = p->usages[g].tick;
      if (t == TICK_INVIOLATE) continue;
      if (t == TICK_MULTI) {
        links_push(&events, g, cur.v[i].pb);
        continue;
      }
      p->usages[g].tick = 0;
      for (uint32_t k = 0; k < 8; k++) {
        const uint32_t b = 8u * g + k;
        if (b >= p->n_blocks) break;
        const uint32_t c = p->blocks[b].child;
        if (c != U32MAX) links_push(&next, c >> 3, b);
      }
    }
    links_t t = cur;
    cur = next;
    next = t;
  }
  uint32_t n_events = 0;
  if (events.n) {
    qsort(events.v, events.n, sizeof(link_t), cmp_link);
    for (size_t i = 0; i < events.n; i++) {
      if (i && events.v[i].g == events.v[i - 1].g && events.v[i].pb == events.v[i - 1].pb) continue;
      unlink_parent(p, events.v[i].g, events.v[i].pb, now);
      n_events++;
    }
  }
  free(cur.v);
  free(next.v);
  free(events.v);
  return n_events;
}

/* ------------------------------------------------------------------------ */
/* prune                                                                    */
/* ------------------------------------------------------------------------ */

int orc_prune(orc_pool *p, const orc_update_input *in, const orc_pruning *rec, uint32_t n,
              const orc_block *surface, uint32_t n_surface, const uint64_t *locations, uint32_t n_locations,
              orc_tree_stats *stats) {
  orc_tree_stats st;
  memset(&st, 0, sizeof st);
  res_t *res = (res_t *)calloc(n ? n : 1, sizeof(res_t));
  if (!res) return -1;
  const uint16_t now = in->tick;

  /* resolve (read-only) */
  for (uint32_t i = 0; i < n; i++) {
    const orc_pruning *r = &rec[i];
    if ((r->properties >> 4) & 1u) {
      res[i].status = r->address < p->n_bases ? RES_BASE : RES_INVALID;
      res[i].addr = r->address;
    } else if (((r->properties >> 3) & 1u) && r->colour_address >= n_surface) {
      res[i].status = RES_INVALID;
    } else {
      res[i] = resolve(p, in, r->depth, r->address, locations, n_locations);
    }
  }
  /* child requests (record order), then ticks along every walked path */
  for (uint32_t i = 0; i < n; i++) {
    if (res[i].status == RES_MISS) {
      request(p, in, &res[i], rec[i].depth, locations + 3 * (size_t)rec[i].address);
      st.requests++;
    }
    switch (res[i].status) {
      case RES_OK:
      case RES_BASE: st.applied++; break;
      case RES_INVALID: st.invalid++; break;
      default: st.skipped++;
    }
  }
  for (uint32_t i = 0; i < n; i++) apply_touches(p, in, rec[i].depth, rec[i].address, locations, &res[i]);

  /* CullChild: free the old subtrees, then link the new children */
  links_t roots = {0, 0, 0};
  for (uint32_t i = 0; i < n; i++) {
    if (res[i].status != RES_OK || !(rec[i].properties & 1u)) continue;
    const uint32_t c = p->blocks[res[i].addr].child;
    if (c != U32MAX) links_push(&roots, c >> 3, res[i].addr);
  }
  st.unlinks = free_subtrees(p, &roots, NULL, now);
  free(roots.v);
  for (uint32_t i = 0; i < n; i++) {
    if (res[i].status != RES_OK || !(rec[i].properties & 1u)) continue;
    p->blocks[res[i].addr].child = rec[i].child_address;
  }
  for (uint32_t i = 0; i < n; i++) {
    if (res[i].status != RES_OK || !(rec[i].properties & 1u)) continue;
    const uint32_t c = rec[i].child_address;
    if (c == U32MAX) continue;
    const uint32_t g = c >> 3;
    if (g < p->n_usages && p->usages[g].tick < TICK_MULTI) p->usages[g].parent = res[i].addr;
  }

  /* AlterViolability / MakeInviolate, UpdateChunk, BaseBlock */
  for (uint32_t i = 0; i < n; i++) {
    const orc_pruning *r = &rec[i];
    if (res[i].status == RES_BASE) {
      p->bases[res[i].addr] = r->chunk;
      continue;
    }
    if (res[i].status != RES_OK) continue;
    const uint32_t a = res[i].addr, g = a >> 3;
    const uint32_t v = (r->properties >> 1) & 3u;
    if (g < p->n_usages && (v == 3u || v == 1u)) p->usages[g].tick = v == 3u ? TICK_INVIOLATE : now;
    if ((r->properties >> 3) & 1u) {
      const uint32_t child = p->blocks[a].child;
      p->blocks[a] = surface[r->colour_address];
      p->blocks[a].child = child;
      p->blocks[a].chunk = r->chunk;
    }
  }
  free(res);
  if (stats) *stats = st;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* collect                                                                  */
/* ------------------------------------------------------------------------ */

static uint32_t pool_groups(const orc_pool *p, const orc_update_input *in) {
  if (in->memory_size == 0) return p->n_usages;
  const uint32_t g = (uint32_t)(((uint64_t)in->memory_size + 7u) / 8u);
  return g < p->n_usages ? g : p->n_usages;
}

static uint32_t collect_into(orc_pool *p, const orc_update_input *in, uint32_t need, uint32_t *out,
                             uint8_t *claimed) {
  const uint16_t now = in->tick;
  const uint32_t ng = pool_groups(p, in);
  uint32_t count = 0, offset = in->offset;
  uint32_t *pass = (uint32_t *)malloc(sizeof(uint32_t) * (need ? need : 1));
  while (count < need) {
    uint32_t k = 0;
    for (uint32_t g = 0; g < ng && count + k < need; g++)
      if (!claimed[g] && stale(p->usages[g].tick, now, offset)) pass[k++] = g;
    links_t roots = {0, 0, 0};
    for (uint32_t i = 0; i < k; i++) {
      const uint32_t g = pass[i];
      claimed[g] = 1;
      out[count + i] = g;
      if (p->usages[g].tick == 0) continue;
      /* a stale live group: cut it from its parent and free what hangs below it */
      const uint32_t pb = p->usages[g].parent;
      if (pb < p->n_blocks && p->blocks[pb].child == 8u * g) p->blocks[pb].child = U32MAX;
      for (uint32_t j = 0; j < 8; j++) {
        const uint32_t b = 8u * g + j;
        if (b >= p->n_blocks) break;
        const uint32_t c = p->blocks[b].child;
        if (c != U32MAX) links_push(&roots, c >> 3, b);
      }
    }
    free_subtrees(p, &roots, claimed, now);
    free(roots.v);
    for (uint32_t i = 0; i < k; i++) {
      p->usages[pass[i]].tick = now;
      p->usages[pass[i]].parent = U32MAX;
    }
    count += k;
    if (offset == 0) break;
    offset >>= 1;
  }
  free(pass);
  return count;
}

uint32_t orc_collect(orc_pool *p, const orc_update_input *in, uint32_t need, uint32_t *out_groups) {
  uint8_t *claimed = (uint8_t *)calloc(p->n_usages ? p->n_usages : 1, 1);
  const uint32_t n = collect_into(p, in, need, out_groups, claimed);
  free(claimed);
  return n;
}

/* ------------------------------------------------------------------------ */
/* graft                                                                    */
/* ------------------------------------------------------------------------ */

/* Pre-order numbering of a graft's groups (the depthHeap walk of wavefront.cl:1322-1349).
 * number[q] = DFS index of graft-relative group q; parent[q] = graft-relative parent block
 * (U32MAX for the top group).  Returns 1 when every group is reached exactly once. */
static int graft_structure(const orc_block *gb, uint32_t gda, uint32_t total, uint32_t *number, uint32_t *parent) {
  const uint32_t ng = total / 8u;
  for (uint32_t q = 0; q < ng; q++) number[q] = U32MAX;
  uint32_t stack_q[MAX_GRAFT_DEPTH], stack_k[MAX_GRAFT_DEPTH], top = 0, next = 1;
  number[0] = 0;
  parent[0] = U32MAX;
  stack_q[0] = 0;
  stack_k[0] = 0;
  top = 1;
  while (top) {
    const uint32_t q = stack_q[top - 1];
    if (stack_k[top - 1] == 8u) {
      top--;
      continue;
    }
    const uint32_t k = stack_k[top - 1]++;
    const uint32_t c = gb[gda + 8u * q + k].child;
    if (c == U32MAX) continue;
    if (c < gda || c - gda >= total || ((c - gda) & 7u)) return 0;
    const uint32_t q2 = (c - gda) / 8u;
    if (number[q2] != U32MAX || top == MAX_GRAFT_DEPTH) return 0;
    number[q2] = next++;
    parent[q2] = 8u * q + k;
    stack_q[top] = q2;
    stack_k[top] = 0;
    top++;
  }
  return next == ng;
}

int orc_graft(orc_pool *p, const orc_update_input *in, const orc_grafting *rec, uint32_t n,
              const orc_block *gb, uint32_t n_gb, const uint64_t *locations, uint32_t n_locations,
              orc_tree_stats *stats) {
  orc_tree_stats st;
  memset(&st, 0, sizeof st);
  const uint16_t now = in->tick;
  /* per-record storage of the DFS numbering: records may share graft data */
  uint32_t *size = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
  uint32_t *where = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
  uint64_t total_size = 0;
  for (uint32_t r = 0; r < n; r++) {
    const uint32_t gda = rec[r].graft_data_address, tot = rec[r].graft_total_size;
    size[r] = (tot >= 8u && (tot & 7u) == 0 && (gda & 7u) == 0 && (uint64_t)gda + tot <= n_gb) ? tot / 8u : 0;
    where[r] = (uint32_t)total_size;
    total_size += size[r];
  }
  uint32_t *number = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(total_size + 1u));
  uint32_t *parent = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(total_size + 1u));
  uint32_t *need = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
  uint32_t *start = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
  res_t *res = (res_t *)calloc(n ? n : 1, sizeof(res_t));
  uint8_t *placed = (uint8_t *)calloc(n ? n : 1, 1);

  /* structure of every graft, and the groups it needs */
  uint32_t total_need = 0;
  for (uint32_t r = 0; r < n; r++) {
    const uint32_t gda = rec[r].graft_data_address, tot = rec[r].graft_total_size;
    const int ok = size[r] && graft_structure(gb, gda, tot, number + where[r], parent + where[r]);
    need[r] = ok ? size[r] : 0;
    start[r] = total_need;
    total_need += need[r];
  }

  /* collect */
  uint8_t *claimed = (uint8_t *)calloc(p->n_usages ? p->n_usages : 1, 1);
  uint32_t *L = (uint32_t *)malloc(sizeof(uint32_t) * (total_need ? total_need : 1));
  const uint32_t got = collect_into(p, in, total_need, L, claimed);
  st.collected = got;

  /* resolve targets against the collected tree, requests, touches */
  for (uint32_t r = 0; r < n; r++) {
    if (!need[r]) {
      res[r].status = RES_INVALID;
      continue;
    }
    res[r] = resolve(p, in, rec[r].depth, rec[r].graft_address, locations, n_locations);
  }
  for (uint32_t r = 0; r < n; r++) {
    if (res[r].status == RES_MISS) {
      request(p, in, &res[r], rec[r].depth, locations + 3 * (size_t)rec[r].graft_address);
      st.requests++;
    }
  }
  for (uint32_t r = 0; r < n; r++) apply_touches(p, in, rec[r].depth, rec[r].graft_address, locations, &res[r]);

  /* which grafts are linked: target resolved, groups available, last record for its target */
  for (uint32_t r = 0; r < n; r++)
    placed[r] = res[r].status == RES_OK && start[r] + need[r] <= got;
  for (uint32_t r = 0; r < n; r++) {
    if (!placed[r]) continue;
    for (uint32_t s = r + 1; s < n; s++)
      if (placed[s] && res[s].addr == res[r].addr) {
        placed[r] = 0;
        break;
      }
  }
  for (uint32_t r = 0; r < n; r++) {
    if (placed[r])
      st.placed++;
    else if (res[r].status == RES_INVALID)
      st.invalid++;
    else
      st.skipped++;
  }
  st.applied = st.placed;

  /* free what the targets held before */
  links_t roots = {0, 0, 0};
  for (uint32_t r = 0; r < n; r++) {
    if (!placed[r]) continue;
    const uint32_t c = p->blocks[res[r].addr].child;
    if (c != U32MAX) links_push(&roots, c >> 3, res[r].addr);
  }
  st.unlinks = free_subtrees(p, &roots, claimed, now);
  free(roots.v);

  /* copy the graft blocks into their groups */
  for (uint32_t r = 0; r < n; r++) {
    if (!placed[r]) continue;
    const uint32_t gda = rec[r].graft_data_address, ng = need[r];
    const uint32_t *num = number + where[r], *par = parent + where[r];
    for (uint32_t q = 0; q < ng; q++) {
      const uint32_t P = L[start[r] + num[q]];
      for (uint32_t k = 0; k < 8; k++) {
        orc_block b = gb[gda + 8u * q + k];
        if (b.child != U32MAX) b.child = 8u * L[start[r] + num[(b.child - gda) / 8u]];
        if (8u * P + k < p->n_blocks) p->blocks[8u * P + k] = b;
      }
      p->usages[P].tick = now;
      if (par[q] == U32MAX)
        p->usages[P].parent = res[r].addr;
      else
        p->usages[P].parent = 8u * L[start[r] + num[par[q] / 8u]] + (par[q] & 7u);
    }
  }
  /* links after every copy (a target may sit in a group another graft filled) */
  for (uint32_t r = 0; r < n; r++)
    if (placed[r]) p->blocks[res[r].addr].child = 8u * L[start[r]];

  /* groups collected but not used go back to the free pool */
  uint8_t *used = (uint8_t *)calloc(got ? got : 1, 1);
  for (uint32_t r = 0; r < n; r++)
    if (placed[r])
      for (uint32_t j = 0; j < need[r]; j++) used[start[r] + j] = 1;
  for (uint32_t i = 0; i < got; i++)
    if (!used[i]) {
      p->usages[L[i]].tick = 0;
      p->usages[L[i]].parent = U32MAX;
    }

  free(used);
  free(L);
  free(claimed);
  free(number);
  free(parent);
  free(size);
  free(where);
  free(need);
  free(start);
  free(res);
  free(placed);
  if (stats) *stats = st;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* garbage collection: reclaim stale groups, compact the pool                 */
/* ------------------------------------------------------------------------ */

/* Frees every stale group (age > offset, as collection defines it): cut from its parent,
 * its subtree freed, its own tick 0.  Returns the number of groups reclaimed. */
uint32_t orc_reclaim(orc_pool *p, const orc_update_input *in) {
  uint8_t *claimed = (uint8_t *)calloc(p->n_usages ? p->n_usages : 1, 1);
  uint32_t *list = (uint32_t *)malloc(sizeof(uint32_t) * (p->n_usages ? p->n_usages : 1));
  const uint16_t now = in->tick;
  const uint32_t ng = pool_groups(p, in);
  uint32_t k = 0;
  for (uint32_t g = 0; g < ng; g++) {
    const uint16_t t = p->usages[g].tick;
    if (t != 0 && t < TICK_MULTI && orc_tick_age(t, now) > in->offset) list[k++] = g;
  }
  links_t roots = {0, 0, 0};
  for (uint32_t i = 0; i < k; i++) {
    const uint32_t g = list[i];
    claimed[g] = 1;
    const uint32_t pb = p->usages[g].parent;
    if (pb < p->n_blocks && p->blocks[pb].child == 8u * g) p->blocks[pb].child = U32MAX;
    for (uint32_t j = 0; j < 8; j++) {
      const uint32_t b = 8u * g + j;
      if (b >= p->n_blocks) break;
      const uint32_t c = p->blocks[b].child;
      if (c != U32MAX) links_push(&roots, c >> 3, b);
    }
  }
  free_subtrees(p, &roots, claimed, now);
  free(roots.v);
  for (uint32_t i = 0; i < k; i++) {
    p->usages[list[i]].tick = 0;
    p->usages[list[i]].parent = U32MAX;
  }
  free(list);
  free(claimed);
  return k;
}

/* Compaction (over the groups whose blocks lie in the pool): with L live groups (tick != 0),
 * the live single-parent groups at or above L
 * move, in ascending order, into the free groups below L, in ascending order; blocks move
 * with their groups, Child fields and Usage.Parent back-pointers follow.  Returns the number
 * of groups moved. */
uint32_t orc_compact(orc_pool *p, const orc_update_input *in) {
  (void)in;
  /* groups whose 8 blocks all lie in the pool */
  const uint32_t n = p->n_usages < p->n_blocks / 8u ? p->n_usages : p->n_blocks / 8u;
  uint32_t live = 0;
  for (uint32_t g = 0; g < n; g++) live += p->usages[g].tick != 0;
  uint32_t *remap = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *from = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *to = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  for (uint32_t g = 0; g < n; g++) remap[g] = g;
  uint32_t nf = 0, nt = 0;
  for (uint32_t g = live; g < n; g++) {
    const uint16_t t = p->usages[g].tick;
    if (t != 0 && t < TICK_MULTI) from[nf++] = g;
  }
  for (uint32_t g = 0; g < live && g < n; g++)
    if (p->usages[g].tick == 0) to[nt++] = g;
  const uint32_t m = nf < nt ? nf : nt;
  for (uint32_t i = 0; i < m; i++) remap[from[i]] = to[i];
#define REMAP_BLOCK(b) (remap[(b) >> 3] * 8u + ((b)&7u))
  for (uint32_t i = 0; i < m; i++) {
    const uint32_t o = from[i], nw = to[i];
    for (uint32_t k = 0; k < 8; k++) {
      orc_block b = p->blocks[8u * o + k];
      if (b.child != U32MAX && (b.child >> 3) < n) b.child = REMAP_BLOCK(b.child);
      p->blocks[8u * nw + k] = b;
    }
    orc_usage u = p->usages[o];
    if (u.parent < p->n_blocks) u.parent = REMAP_BLOCK(u.parent);
    p->usages[nw] = u;
  }
  /* links into moved groups from groups that stay: the parent block's Child, the
   * children's Usage.Parent (or their parent-list entries when they are multi-parent) */
  uint8_t *target = (uint8_t *)calloc(n ? n : 1, 1);
  for (uint32_t i = 0; i < m; i++) target[to[i]] = 1;
  for (uint32_t i = 0; i < m; i++) {
    const uint32_t o = from[i], nw = to[i];
    const uint32_t pb = p->usages[nw].parent;
    if (pb < p->n_blocks && remap[pb >> 3] == (pb >> 3)) p->blocks[pb].child = 8u * nw;
    for (uint32_t k = 0; k < 8; k++) {
      const uint32_t c = p->blocks[8u * nw + k].child;
      if (c == U32MAX || (c >> 3) >= n) continue;
      const uint32_t cg = c >> 3;
      if (target[cg]) continue;
      if (p->usages[cg].tick < TICK_MULTI) {
        if (p->usages[cg].parent == 8u * o + k) p->usages[cg].parent = 8u * nw + k;
      } else if (p->usages[cg].tick == TICK_MULTI && p->parents) {
        uint32_t e = p->usages[cg].parent, steps = 0;
        while (e < p->n_parents && steps++ < p->n_parents) {
          if (p->parents[e].parent_address == 8u * o + k) p->parents[e].parent_address = 8u * nw + k;
          e = p->parents[e].next_element;
        }
      }
    }
  }
  free(target);
  for (uint32_t i = 0; i < m; i++) {
    p->usages[from[i]].tick = 0;
    p->usages[from[i]].parent = U32MAX;
  }
#undef REMAP_BLOCK
  free(remap);
  free(from);
  free(to);
  return m;
}
