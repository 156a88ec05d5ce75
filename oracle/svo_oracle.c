/*
 * svo_oracle.c -- CPU oracle (test infrastructure only; see svo_oracle.h).
 *
 * Defined floating-point semantics (identical to the GPU product by design):
 *   - no FP contraction (built with -ffp-contract=off);
 *   - '/' and sqrtf correctly rounded; native_divide(a,b) -> a/b;
 *     fast_length(v) -> sqrtf((x*x + y*y) + z*z);
 *   - u64 -> f32 rounds to nearest even, f32 -> u64 truncates (saturating,
 *     NaN -> 0, where the reference is undefined);
 *   - camera sin/cos from a host LUT: (float)sin((double)u) per column/row;
 *   - ConeDepth = -half_log2(cone) is never formed: every comparison against an
 *     integer depth k is evaluated exactly as a comparison of `cone` with 2^-k.
 */
#include "svo_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define U64MAX UINT64_MAX
#define U32MAX UINT32_MAX

enum { Q_NONE = 0, Q_BASE = 1, Q_BLOCK = 2, Q_BACKGROUND = 3, Q_MATERIAL = 4 };

/* Defined termination bound: a cone that has taken this many steps ends as background
 * (status bit 2).  The reference has no bound; no scene used here comes near it. */
#define STEP_CAP (1u << 20)
/* Defined depth bound: a cone never descends below this level (64-bit positions; the
 * reference's shifts are undefined from depth 63 on). */
#define MAX_DEPTH 60u

/* ------------------------------------------------------------------------ */
/* fixed-point helpers                                                       */
/* ------------------------------------------------------------------------ */

/* floatToUlong, wavefront.cl:200-207.  Pinned: tests/test_oracle_helpers.py
 * (exact power-of-two scaling against Python integer arithmetic). */
uint64_t orc_float_to_ulong(float x) {
  if (x >= 0.999999911f) return U64MAX;
  float a = fabsf(x) * 18446744073709551616.0f; /* (float)ULONG_MAX == 2^64 */
  if (a != a) return 0;
  if (a >= 18446744073709551616.0f) return U64MAX;
  return (uint64_t)a;
}

/* ulongToFloat, wavefront.cl:209: (float)x / 2^64, exact scaling.  Pinned as above. */
float orc_ulong_to_float(uint64_t x) { return (float)x * 5.42101086242752217e-20f; }

/* Array forms of the two conversions (for the exhaustive tests). */
void orc_float_to_ulong_n(const float *x, uint64_t *out, size_t n) {
  for (size_t i = 0; i < n; i++) out[i] = orc_float_to_ulong(x[i]);
}
void orc_ulong_to_float_n(const uint64_t *x, float *out, size_t n) {
  for (size_t i = 0; i < n; i++) out[i] = orc_ulong_to_float(x[i]);
}

/* roundUlong, wavefront.cl:211-221: distance to the next (round_up) or just past the
 * previous level-(depth+1) cell boundary.  Pinned: boundary-landing property test. */
uint64_t orc_round_ulong(uint64_t value, uint32_t depth, int round_up) {
  uint64_t keep = U64MAX - (U64MAX >> (depth + 1));
  if (round_up) return ((value & keep) + ((U64MAX - (U64MAX >> 1)) >> depth)) - value;
  uint64_t out = value - ((value & keep) - 1);
  if (out >= value) return value;
  return out;
}

/* powSum, wavefront.cl:224 / TreeBuilder.cs:336-342 (loop form, valid past d=11).
 * Pinned: closed form (8^(d+1)-8)/7. */
uint32_t orc_pow_sum(uint32_t depth) {
  uint32_t s = 0, p = 1;
  for (uint32_t i = 1; i <= depth; i++) {
    p *= 8u;
    s += p;
  }
  return s;
}

/* chunkPosition, wavefront.cl:428-430.  Pinned: exhaustive Morton test vs TreeBuilder.Interleave. */
uint32_t orc_chunk_position(uint32_t depth, const uint64_t loc[3]) {
  uint32_t s = 63u - depth;
  return (uint32_t)(((loc[0] >> s) & 1u) | (((loc[1] >> s) & 1u) << 1) | (((loc[2] >> s) & 1u) << 2));
}

/* baseLocation, wavefront.cl:405-411.  Pinned: as chunkPosition. */
uint32_t orc_base_location(uint32_t depth, const uint64_t loc[3]) {
  uint32_t out = 0;
  for (uint32_t i = 0; i < depth; i++) out = (out << 3) + orc_chunk_position(i, loc);
  return out;
}

static int crossed(uint32_t prev_oct, uint32_t new_oct, const uint8_t sign[3]) {
  for (int a = 0; a < 3; a++) {
    uint32_t p = (prev_oct >> a) & 1u, n = (new_oct >> a) & 1u;
    if (p == sign[a] && n != sign[a]) return 1;
  }
  return 0;
}

/* comparePositions, wavefront.cl:413-426.  Pinned: pop-depth invariant (the returned
 * level's cell is unchanged, the next finer one changed) in tests. */
uint32_t orc_compare_positions(uint32_t depth, const uint64_t prev[3], const uint64_t loc[3],
                               const uint8_t sign[3]) {
  uint32_t n = orc_chunk_position(depth, loc), p = orc_chunk_position(depth, prev);
  while (crossed(p, n, sign)) {
    if (depth == 1) break;
    depth--;
    n = orc_chunk_position(depth, loc);
    p = orc_chunk_position(depth, prev);
  }
  return depth;
}

/* leaving, wavefront.cl:433-437. */
static int leaving(const uint64_t loc[3], const uint8_t sign[3]) {
  for (int a = 0; a < 3; a++) {
    if (!sign[a] && loc[a] == 0) return 1;
    if (sign[a] && loc[a] == U64MAX) return 1;
  }
  return 0;
}

/* traverseChunk, wavefront.cl:440-497 (minus the trailing setConeDepth, which the
 * oracle evaluates where it is consumed).  Pinned: step invariants + brute force. */
void orc_traverse_chunk(uint32_t depth, uint64_t loc[3], const float dir[3], const float inv[3],
                        const uint8_t sign[3]) {
  uint64_t d[3];
  float a[3];
  for (int k = 0; k < 3; k++) {
    d[k] = orc_round_ulong(loc[k], depth, sign[k]);
    a[k] = fabsf((float)d[k] * inv[k]);
  }
  int m;
  if (a[0] <= a[1] && a[0] <= a[2])
    m = 0;
  else if (a[1] <= a[0] && a[1] <= a[2])
    m = 1;
  else
    m = 2;
  float ud = orc_ulong_to_float(d[m]);
  for (int k = 0; k < 3; k++) {
    if (k == m) continue;
    float t = dir[k] * inv[m];
    d[k] = orc_float_to_ulong(t * ud);
  }
  for (int k = 0; k < 3; k++) {
    if (sign[k])
      loc[k] += d[k];
    else
      loc[k] -= d[k];
  }
  for (int k = 0; k < 3; k++) {
    if (sign[k] && loc[k] == 0)
      loc[k] = U64MAX;
    else if (!sign[k] && loc[k] == U64MAX)
      loc[k] = 0;
  }
}

/* coneSize, wavefront.cl:387-391 (OpenCL max(x,y) = x < y ? y : x). */
static float cone_size(float m, float fov, const float dof[2]) {
  float eye = fov * m;
  float f = fabsf(dof[0] * (dof[1] - m));
  return eye < f ? f : eye;
}

/* setConeDepth, wavefront.cl:394-399, returning the cone instead of -log2(cone). */
static float cone_at(const float origin[3], const uint64_t loc[3], float fov, const float dof[2]) {
  float x = origin[0] - orc_ulong_to_float(loc[0]);
  float y = origin[1] - orc_ulong_to_float(loc[1]);
  float z = origin[2] - orc_ulong_to_float(loc[2]);
  float m = sqrtf((x * x + y * y) + z * z);
  return cone_size(fabsf(m), fov, dof);
}

/* 2^-k as a float (k <= 126). */
static float thr(int k) { return ldexpf(1.0f, -k); }

/* ------------------------------------------------------------------------ */
/* camera + volume entry                                                     */
/* ------------------------------------------------------------------------ */

static float cam_u(const orc_trace_input *in, uint32_t x) {
  float w = (float)in->screen_size[0];
  return in->fov[0] * ((w / 2.0f - (float)x) / w);
}
static float cam_v(const orc_trace_input *in, uint32_t y) {
  float h = (float)in->screen_size[1];
  return in->fov[1] * ((h / 2.0f - (float)y) / h);
}

/* Separable camera LUT replacing native_sin/native_cos of setupInitialTrace
 * (wavefront.cl:566-572).  Parity of the reference's device trig is unpinned. */
void orc_camera_lut(const orc_trace_input *in, float *cos_u, float *sin_u, float *cos_v, float *sin_v) {
  for (uint32_t x = 0; x < in->screen_size[0]; x++) {
    double u = (double)cam_u(in, x);
    cos_u[x] = (float)cos(u);
    sin_u[x] = (float)sin(u);
  }
  for (uint32_t y = 0; y < in->screen_size[1]; y++) {
    double v = (double)cam_v(in, y);
    cos_v[y] = (float)cos(v);
    sin_v[y] = (float)sin(v);
  }
}

/* traceIntoVolume, wavefront.cl:639-1088, stated per axis: among the axes on whose
 * outside the origin lies, the entry plane is the first (x,y,z order) whose slab
 * distance is >= every other candidate's, else the last candidate.  Pinned:
 * brute-force slab test in tests. */
static int trace_into_volume(const float origin[3], const float dir[3], const float inv[3], float fov,
                             uint64_t loc[3]) {
  int lo[3], hi[3], out = 0;
  for (int a = 0; a < 3; a++) {
    lo[a] = origin[a] < 0.0f;
    hi[a] = origin[a] > 1.0f;
    out |= lo[a] | hi[a];
  }
  if (!out) {
    for (int a = 0; a < 3; a++) loc[a] = orc_float_to_ulong(origin[a]);
    return 1;
  }
  for (int a = 0; a < 3; a++)
    if ((lo[a] | hi[a]) && dir[a] == 0.0f) return 0;
  for (int a = 0; a < 3; a++) {
    int pos = dir[a] >= 0.0f;
    if (lo[a] && !pos) return 0;
    if (hi[a] && pos) return 0;
  }
  int cand[3], nc = 0;
  float m_of[3] = {0, 0, 0};
  for (int a = 0; a < 3; a++) {
    if (lo[a] | hi[a]) {
      float bound = lo[a] ? 0.0f : 1.0f;
      m_of[a] = fabsf((bound - origin[a]) * inv[a]);
      cand[nc++] = a;
    }
  }
  int pick = cand[nc - 1];
  for (int i = 0; i < nc - 1; i++) {
    int a = cand[i], ok = 1;
    for (int j = 0; j < nc; j++)
      if (cand[j] != a && !(m_of[a] >= m_of[cand[j]])) ok = 0;
    if (ok) {
      pick = a;
      break;
    }
  }
  float m = m_of[pick];
  float lf[3];
  for (int a = 0; a < 3; a++) {
    if (a == pick)
      lf[a] = lo[a] ? 0.0f : 1.0f;
    else
      lf[a] = origin[a] + dir[a] * m;
  }
  for (int a = 0; a < 3; a++) loc[a] = orc_float_to_ulong(lf[a]);
  const float zero_dof[2] = {0.0f, 0.0f};
  float c = cone_size(m, fov, zero_dof);
  for (int a = 0; a < 3; a++)
    if (lf[a] < -c || lf[a] > 1.0f + c) return 0;
  return 1;
}

typedef struct {
  float origin[3], dir[3], inv[3], fov, dof[2];
  uint8_t sign[3];
  uint64_t loc[3];
  uint32_t depth;
} ray_t;

/* setupInitialTrace + traceIntoVolume, wavefront.cl:563-595, 1372-1381, with the
 * trig of pixel (x,y) supplied as (cu,su) for its column and (cv,sv) for its row. */
static int init_ray_trig(const orc_trace_input *in, float aperture, float cu, float su, float cv, float sv,
                         ray_t *r) {
  float m0 = cu * cv;
  float m1 = (0.0f - su) * cv;
  float m2 = sv;
  const float *F = in->facing;
  r->dir[0] = (F[0] * m0 + F[3] * m1) + F[6] * m2;
  r->dir[1] = (F[1] * m0 + F[4] * m1) + F[7] * m2;
  r->dir[2] = (F[2] * m0 + F[5] * m1) + F[8] * m2;
  for (int a = 0; a < 3; a++) {
    r->inv[a] = 1.0f / r->dir[a];
    r->sign[a] = r->dir[a] >= 0.0f;
    r->origin[a] = in->origin[a];
  }
  /* cone aperture: `aperture` pixels (1 in the reference; (1*FoV)/W == FoV/W exactly) */
  r->fov = (aperture * in->fov[0]) / (float)in->screen_size[0];
  r->dof[0] = in->dof[0];
  r->dof[1] = in->dof[1];
  r->loc[0] = r->loc[1] = r->loc[2] = 0;
  r->depth = 1;
  if (trace_into_volume(r->origin, r->dir, r->inv, r->fov, r->loc)) return 1;
  /* a cone that misses the volume keeps Depth 0 and Location 0 (init leaves them unwritten) */
  r->depth = 0;
  r->loc[0] = r->loc[1] = r->loc[2] = 0;
  return 0;
}

static int init_ray_lut(const orc_trace_input *in, float aperture, const float *cu, const float *su,
                        const float *cv, const float *sv, uint32_t x, uint32_t y, ray_t *r) {
  return init_ray_trig(in, aperture, cu[x], su[x], cv[y], sv[y], r);
}

int orc_init_ray(const orc_trace_input *in, uint32_t x, uint32_t y, float dir[3], float inv[3],
                 uint64_t loc[3], float *cone_fov) {
  double u = (double)cam_u(in, x), v = (double)cam_v(in, y);
  ray_t r;
  int hit = init_ray_trig(in, 1.0f, (float)cos(u), (float)sin(u), (float)cos(v), (float)sin(v), &r);
  for (int a = 0; a < 3; a++) {
    dir[a] = r.dir[a];
    inv[a] = r.inv[a];
    loc[a] = hit ? r.loc[a] : 0;
  }
  *cone_fov = r.fov;
  return hit;
}

/* ------------------------------------------------------------------------ */
/* the step machine                                                          */
/* ------------------------------------------------------------------------ */

typedef struct {
  const uint16_t *bases;
  uint32_t n_bases;
  const orc_block *blocks;
  uint32_t n_blocks;
  const orc_usage *usages; /* initial ticks (read-only while tracing) */
  uint32_t n_usages;
  uint8_t *touched; /* per usage group: set when updateUsage would write */
  uint32_t B;
  uint16_t tick;
} tree_t;

typedef struct {
  uint32_t visits, base_steps, block_steps;
  uint64_t hash;
  uint64_t *log;
  uint32_t log_stride;
} acc_t;

static const orc_block ZERO_BLOCK;

/* Out-of-range block reads return an all-zero record (defined; the reference reads
 * past the buffer). */
static orc_block read_block(const tree_t *t, uint32_t a) { return a < t->n_blocks ? t->blocks[a] : ZERO_BLOCK; }

/* getBlockAddress, wavefront.cl:296-304: root-restart walk from the dense level. */
static uint32_t block_address(const tree_t *t, uint32_t depth, const uint64_t loc[3]) {
  uint32_t a = orc_base_location(t->B + 2, loc);
  for (uint32_t d = t->B + 2; d < depth; d++) a = read_block(t, a).child + orc_chunk_position(d, loc);
  return a;
}

/* updateUsage, wavefront.cl:377-381.  Within one frame every write stores the frame
 * tick, so the final state only depends on which groups were visited. */
static void update_usage(const tree_t *t, uint32_t group) {
  if (group >= t->n_usages) return;
  uint16_t tk = t->usages[group].tick;
  if (tk < 65534u && tk != t->tick) __atomic_store_n(&t->touched[group], (uint8_t)1, __ATOMIC_RELAXED);
}

static void visit(acc_t *acc, uint32_t depth, uint32_t index) {
  uint64_t w = ((uint64_t)depth << 32) | index;
  if (acc->log && acc->visits < acc->log_stride) acc->log[acc->visits] = w;
  acc->hash = (acc->hash ^ w) * 0x100000001b3ull;
  acc->visits++;
}

/* runBaseTrace, wavefront.cl:1400-1463 (one work-item). */
static int base_step(const tree_t *t, ray_t *r, acc_t *acc) {
  uint32_t d = r->depth;
  uint32_t idx = orc_pow_sum(d - 1) + orc_base_location(d, r->loc);
  uint16_t chunk = idx < t->n_bases ? t->bases[idx] : 0;
  visit(acc, d, idx);
  acc->base_steps++;
  if (((chunk >> (orc_chunk_position(d, r->loc) * 2)) & 2u) != 2u) {
    uint64_t prev[3] = {r->loc[0], r->loc[1], r->loc[2]};
    orc_traverse_chunk(d, r->loc, r->dir, r->inv, r->sign);
    if (leaving(r->loc, r->sign)) {
      r->depth = 0;
      return Q_BACKGROUND;
    }
    r->depth = orc_compare_positions(d, prev, r->loc, r->sign);
    return Q_BASE;
  }
  r->depth = d < t->B ? d + 1 : t->B + 2;
  if (r->depth == t->B + 1) {
    r->depth = t->B;
    return Q_BASE;
  }
  return r->depth > t->B + 1 ? Q_BLOCK : Q_BASE;
}

/* runBlockTrace, wavefront.cl:1465-1566 (one work-item).  On a child request the
 * request fields are returned through *req. */
static int block_step(const tree_t *t, ray_t *r, acc_t *acc, int *child, orc_child_request *req) {
  uint32_t D = r->depth, B = t->B;
  uint32_t addr = block_address(t, D, r->loc);
  orc_block blk = read_block(t, addr);
  update_usage(t, addr >> 3);
  float cone = cone_at(r->origin, r->loc, r->fov, r->dof);
  visit(acc, D, addr);
  acc->block_steps++;
  int complete = 0;
  *child = 0;
  if (((blk.chunk >> (orc_chunk_position(D, r->loc) * 2)) & 2u) != 2u) {
    uint64_t prev[3] = {r->loc[0], r->loc[1], r->loc[2]};
    orc_traverse_chunk(D, r->loc, r->dir, r->inv, r->sign);
    if (leaving(r->loc, r->sign)) {
      D = 0;
      complete = 1;
    } else {
      D = orc_compare_positions(D, prev, r->loc, r->sign);
    }
  } else {
    if (cone > thr((int)B + 2)) { /* ConeDepth < BaseDepth+2 */
      D = B + 2;
      complete = 1;
    } else if (cone > thr((int)D)) { /* ConeDepth < Depth */
      D--;
    } else if (blk.child == U32MAX) {
      *child = 1;
      complete = 1;
    } else if (cone < thr((int)D + 1) && D + 1 <= MAX_DEPTH) { /* ConeDepth > Depth+1 */
      D++;
    } else {
      complete = 1;
    }
  }
  if (*child) {
    memset(req, 0, sizeof *req);
    req->address = addr;
    req->tick = t->tick;
    req->depth = (uint8_t)D;
    req->location[0] = r->loc[0];
    req->location[1] = r->loc[1];
    req->location[2] = r->loc[2];
    req->tree_size = cone <= thr((int)D + 1) ? 1 : 0; /* min(1, (uchar)ConeDepth - Depth) */
  }
  r->depth = D;
  if (complete) return D == 0 ? Q_BACKGROUND : Q_MATERIAL;
  if (D == B + 1) r->depth = D = B;
  return D > B ? Q_BLOCK : Q_BASE;
}

/* evaluateBackground (wavefront.cl:1741-1797) + resolveRemainders (1799-1833) +
 * drawTrace (1835-1850) for a primary ray (weighting 255, parent = itself). */
static void shade(const tree_t *t, const ray_t *r, int q, uint32_t *hit_out, float rgba[4]) {
  uint32_t W = 0, C[3] = {0, 0, 0};
  uint32_t hit = U32MAX;
  if (q == Q_BACKGROUND) {
    float dp = (r->dir[0] * 0.0f + r->dir[1] * 0.0f) + r->dir[2] * -1.0f;
    uint32_t c[3];
    uint8_t lum;
    if (dp > 0.8f) {
      c[0] = c[1] = c[2] = 255;
      lum = (uint8_t)(dp * 200.0f);
    } else {
      c[0] = 135;
      c[1] = 206;
      c[2] = 235;
      lum = 1;
    }
    W += 255;
    if (lum > 0)
      for (int k = 0; k < 3; k++) C[k] += c[k] * 255u;
  } else {
    uint32_t a = block_address(t, r->depth, r->loc);
    if (a != U32MAX) {
      orc_block b = read_block(t, a);
      hit = a;
      W += 255;
      C[0] += (uint32_t)b.r * 255u;
      C[1] += (uint32_t)b.g * 255u;
      C[2] += (uint32_t)b.b * 255u;
    }
  }
  if (hit_out) *hit_out = hit;
  if (W == 0) {
    rgba[0] = 0.75f;
    rgba[1] = 0.0f;
    rgba[2] = 0.0f;
    rgba[3] = 1.0f;
  } else {
    float fw = (float)W * 255.0f;
    for (int k = 0; k < 3; k++) rgba[k] = fminf((float)C[k] / fw, 1.0f);
    rgba[3] = 1.0f;
  }
}

static void setup_tree(tree_t *t, const orc_trace_input *in, const uint16_t *bases, uint32_t n_bases,
                       const orc_block *blocks, uint32_t n_blocks, const orc_usage *usages, uint32_t n_usages,
                       uint8_t *touched) {
  t->bases = bases;
  t->n_bases = n_bases;
  t->blocks = blocks;
  t->n_blocks = n_blocks;
  t->usages = usages;
  t->n_usages = usages ? n_usages : 0;
  t->touched = touched;
  t->B = in->base_depth;
  t->tick = in->tick;
}

static void acc_init(acc_t *acc, const orc_frame_out *out, uint32_t id) {
  acc->visits = acc->base_steps = acc->block_steps = 0;
  acc->hash = 0xcbf29ce484222325ull;
  acc->log = (out && out->visit_log) ? out->visit_log + (size_t)id * out->log_stride : NULL;
  acc->log_stride = out ? out->log_stride : 0;
}

static void write_ray(orc_frame_out *out, uint32_t id, const ray_t *r, int q, int child, int capped,
                      const acc_t *acc, uint32_t hit, const float rgba[4]) {
  if (!out) return;
  if (out->hit) out->hit[id] = hit;
  if (out->visits) out->visits[id] = acc->visits;
  if (out->hash) out->hash[id] = acc->hash;
  if (out->rgba) memcpy(out->rgba + 4 * (size_t)id, rgba, 4 * sizeof(float));
  if (out->status) out->status[id] = (uint8_t)((q == Q_MATERIAL ? 1 : 0) | (child ? 2 : 0) | (capped ? 4 : 0));
  if (out->depth) out->depth[id] = (uint8_t)r->depth;
  if (out->location) memcpy(out->location + 3 * (size_t)id, r->loc, 3 * sizeof(uint64_t));
}

static int check_input(const orc_trace_input *in, uint32_t n_bases) {
  if (!in || in->screen_size[0] == 0 || in->screen_size[1] == 0) return -1;
  if (in->base_depth < 1 || in->base_depth > 8) return -2;
  if (n_bases < orc_pow_sum(in->base_depth)) return -3;
  return 0;
}

static void apply_usage(orc_usage *usages, uint32_t n_usages, const uint8_t *touched, uint16_t tick,
                        uint64_t *count) {
  uint64_t c = 0;
  if (usages && touched)
    for (uint32_t g = 0; g < n_usages; g++)
      if (touched[g]) {
        usages[g].tick = tick;
        c++;
      }
  if (count) *count = c;
}

/* One camera of a frame: traces every cone, writes outputs at `offset + id`, flags child
 * requests in req_flag/req_tmp at the same index and marks touched usage groups. */
static int frame_core(const orc_trace_input *in, float aperture, const tree_t *t, size_t offset, uint8_t *req_flag,
                      orc_child_request *req_tmp, orc_frame_out *out, int threads, uint64_t *base_sum,
                      uint64_t *block_sum) {
  const uint32_t W = in->screen_size[0], H = in->screen_size[1];
  const size_t N = (size_t)W * H;
  float *lut = (float *)malloc(sizeof(float) * 2 * ((size_t)W + H));
  if (!lut) return -10;
  float *cu = lut, *su = lut + W, *cv = lut + 2 * W, *sv = lut + 2 * W + H;
  orc_camera_lut(in, cu, su, cv, sv);
  uint64_t base_total = 0, block_total = 0;
#ifdef _OPENMP
  if (threads < 1) threads = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads) reduction(+ : base_total, block_total)
#endif
  for (long long i = 0; i < (long long)N; i++) {
    uint32_t x = (uint32_t)i % W, y = (uint32_t)i / W;
    size_t id = offset + (size_t)i;
    ray_t r;
    acc_t acc;
    acc_init(&acc, out, (uint32_t)id);
    int q = init_ray_lut(in, aperture, cu, su, cv, sv, x, y, &r) ? Q_BASE : Q_BACKGROUND;
    int child = 0, capped = 0;
    while (q == Q_BASE || q == Q_BLOCK) {
      if (acc.visits >= STEP_CAP) {
        r.depth = 0;
        q = Q_BACKGROUND;
        capped = 1;
        break;
      }
      if (q == Q_BASE) {
        q = base_step(t, &r, &acc);
      } else {
        q = block_step(t, &r, &acc, &child, &req_tmp[id]);
        if (child) req_flag[id] = 1;
      }
    }
    uint32_t hit;
    float rgba[4];
    shade(t, &r, q, &hit, rgba);
    write_ray(out, (uint32_t)id, &r, q, child, capped, &acc, hit, rgba);
    base_total += acc.base_steps;
    block_total += acc.block_steps;
  }
  *base_sum += base_total;
  *block_sum += block_total;
  free(lut);
  return 0;
}

/* Several cameras in one frame (the batched form): outputs concatenated camera by camera,
 * child requests in ascending concatenated cone index, one usage tick for all. */
int orc_trace_batch(const orc_trace_input *ins, const float *apertures, uint32_t n_cams, const uint16_t *bases,
                    uint32_t n_bases, const orc_block *blocks, uint32_t n_blocks, orc_usage *usages,
                    uint32_t n_usages, orc_child_request *requests, uint32_t request_capacity,
                    uint32_t *request_count, orc_frame_out *out, int threads) {
  if (!ins || n_cams == 0) return -1;
  size_t N = 0;
  for (uint32_t c = 0; c < n_cams; c++) {
    int rc = check_input(&ins[c], n_bases);
    if (rc) return rc;
    if (ins[c].tick != ins[0].tick || ins[c].base_depth != ins[0].base_depth) return -4;
    N += (size_t)ins[c].screen_size[0] * ins[c].screen_size[1];
  }
  uint8_t *touched = usages ? (uint8_t *)calloc(n_usages ? n_usages : 1, 1) : NULL;
  uint8_t *req_flag = (uint8_t *)calloc(N ? N : 1, 1);
  orc_child_request *req_tmp = (orc_child_request *)malloc(sizeof(orc_child_request) * (N ? N : 1));
  if ((usages && !touched) || !req_flag || !req_tmp) {
    free(touched);
    free(req_flag);
    free(req_tmp);
    return -10;
  }
  tree_t t;
  setup_tree(&t, &ins[0], bases, n_bases, blocks, n_blocks, usages, n_usages, touched);
  uint64_t base_total = 0, block_total = 0;
  size_t offset = 0;
  int rc = 0;
  for (uint32_t c = 0; c < n_cams && !rc; c++) {
    rc = frame_core(&ins[c], apertures ? apertures[c] : 1.0f, &t, offset, req_flag, req_tmp, out, threads,
                    &base_total, &block_total);
    offset += (size_t)ins[c].screen_size[0] * ins[c].screen_size[1];
  }
  if (!rc) {
    uint32_t n_req = 0;
    for (size_t i = 0; i < N; i++) {
      if (!req_flag[i]) continue;
      if (requests && n_req < request_capacity) requests[n_req] = req_tmp[i];
      n_req++;
    }
    if (request_count) *request_count = n_req;
    uint64_t groups = 0;
    apply_usage(usages, n_usages, touched, ins[0].tick, &groups);
    if (out) {
      out->base_steps = base_total;
      out->block_steps = block_total;
      out->groups_touched = groups;
    }
  }
  free(touched);
  free(req_flag);
  free(req_tmp);
  return rc;
}

int orc_trace_frame(const orc_trace_input *in, const uint16_t *bases, uint32_t n_bases,
                    const orc_block *blocks, uint32_t n_blocks, orc_usage *usages, uint32_t n_usages,
                    orc_child_request *requests, uint32_t request_capacity, uint32_t *request_count,
                    orc_frame_out *out, int threads) {
  return orc_trace_batch(in, NULL, 1, bases, n_bases, blocks, n_blocks, usages, n_usages, requests,
                         request_capacity, request_count, out, threads);
}

int orc_trace_rays(const orc_trace_input *in, const uint16_t *bases, uint32_t n_bases,
                   const orc_block *blocks, uint32_t n_blocks, const uint32_t *ray_ids, uint32_t n_rays,
                   uint32_t *hit, uint32_t *visits, int threads) {
  int rc = check_input(in, n_bases);
  if (rc) return rc;
  const uint32_t W = in->screen_size[0], H = in->screen_size[1];
  float *lut = (float *)malloc(sizeof(float) * 2 * ((size_t)W + H));
  if (!lut) return -10;
  float *cu = lut, *su = lut + W, *cv = lut + 2 * W, *sv = lut + 2 * W + H;
  orc_camera_lut(in, cu, su, cv, sv);
  tree_t t;
  setup_tree(&t, in, bases, n_bases, blocks, n_blocks, NULL, 0, NULL);
#ifdef _OPENMP
  if (threads < 1) threads = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
#endif
  for (long long i = 0; i < (long long)n_rays; i++) {
    uint32_t id = ray_ids[i], x = id % W, y = id / W;
    ray_t r;
    acc_t acc;
    acc_init(&acc, NULL, id);
    orc_child_request req;
    int child = 0;
    int q = init_ray_lut(in, 1.0f, cu, su, cv, sv, x, y, &r) ? Q_BASE : Q_BACKGROUND;
    while (q == Q_BASE || q == Q_BLOCK) {
      if (acc.visits >= STEP_CAP) {
        r.depth = 0;
        q = Q_BACKGROUND;
        break;
      }
      q = q == Q_BASE ? base_step(&t, &r, &acc) : block_step(&t, &r, &acc, &child, &req);
    }
    uint32_t h;
    float rgba[4];
    shade(&t, &r, q, &h, rgba);
    if (hit) hit[i] = h;
    if (visits) visits[i] = acc.visits;
  }
  free(lut);
  return 0;
}

/* Exact emulation of MainWindow.RunWavefront / RunTraces (MainWindow.cs:356-481):
 * init over the (W,H) NDRange, then {ResetIDBuffer(Base) + swap + runBaseTrace;
 * ResetIDBuffer(Block) + swap + runBlockTrace} until both counters read zero, then
 * evaluateBackground, resolveRemainders, drawTrace.  Every kernel appends in
 * work-item order (the deterministic order the product's wavefront kernels keep). */
int orc_wavefront_frame(const orc_trace_input *in, const uint16_t *bases, uint32_t n_bases,
                        const orc_block *blocks, uint32_t n_blocks, orc_usage *usages, uint32_t n_usages,
                        orc_child_request *requests, uint32_t request_capacity, uint32_t *request_count,
                        uint32_t *background_queue, uint32_t *background_count,
                        uint32_t *material_queue, uint32_t *material_count,
                        uint32_t *iterations, orc_frame_out *out) {
  int rc = check_input(in, n_bases);
  if (rc) return rc;
  const uint32_t W = in->screen_size[0], H = in->screen_size[1];
  const size_t N = (size_t)W * H;
  float *lut = (float *)malloc(sizeof(float) * 2 * ((size_t)W + H));
  ray_t *rays = (ray_t *)malloc(sizeof(ray_t) * (N ? N : 1));
  acc_t *accs = (acc_t *)malloc(sizeof(acc_t) * (N ? N : 1));
  uint8_t *childf = (uint8_t *)calloc(N ? N : 1, 1);
  uint8_t *cappedf = (uint8_t *)calloc(N ? N : 1, 1);
  uint32_t *qa = (uint32_t *)malloc(sizeof(uint32_t) * (N ? N : 1)); /* base queue buffer A */
  uint32_t *qb = (uint32_t *)malloc(sizeof(uint32_t) * (N ? N : 1)); /* base queue buffer B */
  uint32_t *ka = (uint32_t *)malloc(sizeof(uint32_t) * (N ? N : 1)); /* block queue buffer A */
  uint32_t *kb = (uint32_t *)malloc(sizeof(uint32_t) * (N ? N : 1)); /* block queue buffer B */
  uint8_t *touched = usages ? (uint8_t *)calloc(n_usages ? n_usages : 1, 1) : NULL;
  if (!lut || !rays || !accs || !childf || !cappedf || !qa || !qb || !ka || !kb || (usages && !touched)) {
    free(lut), free(rays), free(accs), free(childf), free(cappedf), free(qa), free(qb), free(ka), free(kb), free(touched);
    return -10;
  }
  float *cu = lut, *su = lut + W, *cv = lut + 2 * W, *sv = lut + 2 * W + H;
  orc_camera_lut(in, cu, su, cv, sv);
  tree_t t;
  setup_tree(&t, in, bases, n_bases, blocks, n_blocks, usages, n_usages, touched);

  uint32_t *base_q = qa, *base_in = qb, *block_q = ka, *block_in = kb;
  uint32_t n_base = 0, n_block = 0, n_bg = 0, n_mat = 0, n_req = 0, iters = 0;
  /* init */
  for (uint32_t id = 0; id < N; id++) {
    acc_init(&accs[id], out, id);
    if (init_ray_lut(in, 1.0f, cu, su, cv, sv, id % W, id / W, &rays[id]))
      base_q[n_base++] = id;
    else
      background_queue[n_bg++] = id;
  }
  for (;;) {
    uint32_t base_size = n_base;
    n_base = 0;
    int ran = 0;
    if (base_size > 0) {
      uint32_t *tmp = base_in;
      base_in = base_q;
      base_q = tmp;
      for (uint32_t i = 0; i < base_size; i++) {
        uint32_t id = base_in[i];
        if (accs[id].visits >= STEP_CAP) {
          rays[id].depth = 0;
          cappedf[id] = 1;
          background_queue[n_bg++] = id;
          continue;
        }
        int q = base_step(&t, &rays[id], &accs[id]);
        if (q == Q_BACKGROUND)
          background_queue[n_bg++] = id;
        else if (q == Q_BASE)
          base_q[n_base++] = id;
        else
          block_q[n_block++] = id;
      }
      ran = 1;
    }
    uint32_t block_size = n_block;
    n_block = 0;
    if (block_size > 0) {
      uint32_t *tmp = block_in;
      block_in = block_q;
      block_q = tmp;
      for (uint32_t i = 0; i < block_size; i++) {
        uint32_t id = block_in[i];
        if (accs[id].visits >= STEP_CAP) {
          rays[id].depth = 0;
          cappedf[id] = 1;
          background_queue[n_bg++] = id;
          continue;
        }
        int child = 0;
        orc_child_request req;
        int q = block_step(&t, &rays[id], &accs[id], &child, &req);
        if (child) {
          childf[id] = 1;
          if (requests && n_req < request_capacity) requests[n_req] = req;
          n_req++;
        }
        if (q == Q_BACKGROUND)
          background_queue[n_bg++] = id;
        else if (q == Q_MATERIAL)
          material_queue[n_mat++] = id;
        else if (q == Q_BLOCK)
          block_q[n_block++] = id;
        else
          base_q[n_base++] = id;
      }
      ran = 1;
    }
    if (!ran) break;
    iters++;
  }
  /* shading: every ray is in exactly one of the two final queues */
  uint64_t base_total = 0, block_total = 0;
  for (uint32_t i = 0; i < n_bg; i++) {
    uint32_t id = background_queue[i], hit;
    float rgba[4];
    shade(&t, &rays[id], Q_BACKGROUND, &hit, rgba);
    write_ray(out, id, &rays[id], Q_BACKGROUND, 0, cappedf[id], &accs[id], hit, rgba);
  }
  for (uint32_t i = 0; i < n_mat; i++) {
    uint32_t id = material_queue[i], hit;
    float rgba[4];
    shade(&t, &rays[id], Q_MATERIAL, &hit, rgba);
    write_ray(out, id, &rays[id], Q_MATERIAL, childf[id], 0, &accs[id], hit, rgba);
  }
  for (size_t id = 0; id < N; id++) {
    base_total += accs[id].base_steps;
    block_total += accs[id].block_steps;
  }
  uint64_t groups = 0;
  apply_usage(usages, n_usages, touched, in->tick, &groups);
  if (out) {
    out->base_steps = base_total;
    out->block_steps = block_total;
    out->groups_touched = groups;
  }
  if (request_count) *request_count = n_req;
  if (background_count) *background_count = n_bg;
  if (material_count) *material_count = n_mat;
  if (iterations) *iterations = iters;
  free(lut), free(rays), free(accs), free(childf), free(cappedf), free(qa), free(qb), free(ka), free(kb), free(touched);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* differentiable compositing (no reference counterpart)                     */
/* ------------------------------------------------------------------------ */

/* Continues a composited cone past the cell it stopped in: the empty-octant transition
 * of runBlockTrace (wavefront.cl:1505-1514) at the cone's depth. */
static int pass_cell(const tree_t *t, ray_t *r) {
  uint32_t D = r->depth, B = t->B;
  uint64_t prev[3] = {r->loc[0], r->loc[1], r->loc[2]};
  orc_traverse_chunk(D, r->loc, r->dir, r->inv, r->sign);
  if (leaving(r->loc, r->sign)) {
    r->depth = 0;
    return Q_BACKGROUND;
  }
  D = orc_compare_positions(D, prev, r->loc, r->sign);
  if (D == B + 1) D = B;
  r->depth = D;
  return D > B ? Q_BLOCK : Q_BASE;
}

/* Sky or sun colour of a direction (evaluateBackground, wavefront.cl:1762-1772), in [0,1]. */
static void background_colour(const float dir[3], float bg[3]) {
  float dp = (dir[0] * 0.0f + dir[1] * 0.0f) + dir[2] * -1.0f;
  if (dp > 0.8f) {
    bg[0] = bg[1] = bg[2] = 1.0f;
  } else {
    bg[0] = 135.0f / 255.0f;
    bg[1] = 206.0f / 255.0f;
    bg[2] = 235.0f / 255.0f;
  }
}

int orc_composite_layers(const orc_trace_input *ins, const float *apertures, uint32_t n_views, const uint16_t *bases,
                         uint32_t n_bases, const orc_block *blocks, uint32_t n_blocks, uint32_t max_layers,
                         uint32_t *layers, float *bg, int threads) {
  if (!ins || n_views == 0 || max_layers == 0) return -1;
  size_t offset = 0;
  for (uint32_t v = 0; v < n_views; v++) {
    const orc_trace_input *in = &ins[v];
    int rc = check_input(in, n_bases);
    if (rc) return rc;
    const uint32_t W = in->screen_size[0], H = in->screen_size[1];
    const size_t N = (size_t)W * H;
    const float aperture = apertures ? apertures[v] : 1.0f;
    float *lut = (float *)malloc(sizeof(float) * 2 * ((size_t)W + H));
    if (!lut) return -10;
    float *cu = lut, *su = lut + W, *cv = lut + 2 * W, *sv = lut + 2 * W + H;
    orc_camera_lut(in, cu, su, cv, sv);
    tree_t t;
    setup_tree(&t, in, bases, n_bases, blocks, n_blocks, NULL, 0, NULL);
#ifdef _OPENMP
    if (threads < 1) threads = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
#endif
    for (long long i = 0; i < (long long)N; i++) {
      const size_t id = offset + (size_t)i;
      uint32_t *L = layers + id * max_layers;
      for (uint32_t k = 0; k < max_layers; k++) L[k] = U32MAX;
      ray_t r;
      acc_t acc;
      acc_init(&acc, NULL, 0);
      int q = init_ray_lut(in, aperture, cu, su, cv, sv, (uint32_t)i % W, (uint32_t)i / W, &r) ? Q_BASE : Q_BACKGROUND;
      uint32_t n = 0;
      int child = 0;
      orc_child_request req;
      while (q != Q_BACKGROUND && n < max_layers) {
        if (acc.visits >= STEP_CAP) break;
        if (q == Q_BASE)
          q = base_step(&t, &r, &acc);
        else if (q == Q_BLOCK)
          q = block_step(&t, &r, &acc, &child, &req);
        if (q == Q_MATERIAL) {
          L[n++] = block_address(&t, r.depth, r.loc);
          q = pass_cell(&t, &r);
        }
      }
      background_colour(r.dir, bg + 3 * id);
    }
    free(lut);
    offset += N;
  }
  return 0;
}

/* Front-to-back compositing of the layer lists: C = sum_i T_i a_i c_i + T_end bg with
 * T_1 = 1, T_{i+1} = T_i (1 - a_i); params are (r, g, b, a) per block.  Returns the L2
 * loss sum over rays and channels of (C - target)^2 over the listed rays (all when rays is
 * NULL), optionally the image and the gradient (accumulated into grads, 4 per block):
 * dC/dc_i = T_i a_i, dC/da_i = T_i (c_i - U_i) with U_i the composite behind layer i. */
double orc_composite_loss(const uint32_t *layers, const float *bg, const uint32_t *rays, uint32_t n_rays,
                          uint32_t max_layers, const double *params, uint32_t n_blocks, const double *target,
                          double *image, double *grads) {
  double loss = 0.0;
  double T[64], U[3];
  if (max_layers > 63) return -1.0;
  for (uint32_t j = 0; j < n_rays; j++) {
    const size_t id = rays ? rays[j] : j;
    const uint32_t *L = layers + id * max_layers;
    const float *b = bg + 3 * id;
    double C[3] = {0.0, 0.0, 0.0};
    uint32_t n = 0;
    T[0] = 1.0;
    for (; n < max_layers && L[n] != U32MAX; n++) {
      const uint32_t a = L[n];
      const double *p = a < n_blocks ? params + 4 * (size_t)a : NULL;
      const double al = p ? p[3] : 0.0;
      for (int c = 0; c < 3; c++) C[c] += T[n] * al * (p ? p[c] : 0.0);
      T[n + 1] = T[n] * (1.0 - al);
    }
    for (int c = 0; c < 3; c++) C[c] += T[n] * (double)b[c];
    double g[3];
    for (int c = 0; c < 3; c++) {
      const double e = C[c] - (target ? target[3 * id + c] : 0.0);
      loss += e * e;
      g[c] = 2.0 * e;
      if (image) image[3 * id + c] = C[c];
    }
    if (!grads) continue;
    for (int c = 0; c < 3; c++) U[c] = (double)b[c];
    for (int i = (int)n - 1; i >= 0; i--) {
      const uint32_t a = L[i];
      if (a >= n_blocks) continue;
      const double *p = params + 4 * (size_t)a;
      double *gr = grads + 4 * (size_t)a;
      double da = 0.0;
      for (int c = 0; c < 3; c++) {
        gr[c] += g[c] * T[i] * p[3];
        da += g[c] * T[i] * (p[c] - U[c]);
      }
      gr[3] += da;
      for (int c = 0; c < 3; c++) U[c] = p[3] * p[c] + (1.0 - p[3]) * U[c];
    }
  }
  return loss;
}
