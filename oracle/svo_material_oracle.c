/* CPU oracle of evaluateMaterial (wavefront.cl:1568-1739) with setupTrace (617-636),
 * traverseChunkNormal (500-561), normalVector (235-245) and saveAccumulator (338-355), and of
 * the shading kernels evaluateBackground (1741-1797), resolveRemainders (1799-1833) and
 * drawTrace (1835-1850).
 * TEST INFRASTRUCTURE ONLY.  The definitions are listed in svo_material_oracle.h. */
#include "svo_material_oracle.h"

#include <math.h>
#include <stddef.h>

#define U32MAX 0xFFFFFFFFu
#define U64MAX 0xFFFFFFFFFFFFFFFFull
#define PI_F 3.14159274101257324f   /* M_PI_F */
#define HALF_PI_F 1.57079637050628662f /* M_PI_2_F */
#define MAX_SPINE 256u
#define MAX_RING 4096u
#define MAX_DEPTH 60u

typedef struct {
  float x, y, z;
} v3;

static float dot3(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

static v3 cross3(v3 a, v3 b) {
  v3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
  return r;
}

static v3 normalize3(v3 v) {
  float len = sqrtf(dot3(v, v));
  v3 r = {v.x / len, v.y / len, v.z / len};
  return r;
}

void orc_sincos(float x, float *s, float *c) {
  float kf = rintf(x * 0.636619772f);
  int k = (int)kf;
  float r = (x - kf * 1.5703125f) - kf * 4.83826794897e-4f;
  float r2 = r * r;
  float ps = r + (r * r2) * (-0.166666667f + r2 * (8.33333333e-3f + r2 * (-1.98412698e-4f + r2 * 2.75573192e-6f)));
  float pc = 1.0f + r2 * (-0.5f + r2 * (4.16666667e-2f + r2 * (-1.38888889e-3f + r2 * (2.48015873e-5f + r2 * -2.75573192e-7f))));
  switch (k & 3) {
    case 0: *s = ps; *c = pc; break;
    case 1: *s = pc; *c = -ps; break;
    case 2: *s = -ps; *c = -pc; break;
    default: *s = -pc; *c = ps; break;
  }
}

static uint32_t to_uchar(float f) {
  if (!(f > 0.0f)) return 0u;
  if (f >= 255.0f) return 255u;
  return (uint32_t)f;
}

/* one axis of the move: add or subtract by the sign; wrap-around pins to the face */
static uint64_t move_axis(uint64_t v, uint64_t n, int pos) {
  v = pos ? v + n : v - n;
  if (pos && v == 0) return U64MAX;
  if (!pos && v == U64MAX) return 0;
  return v;
}

/* traverseChunkNormal: to the farthest level-(depth+1) boundary along the normal */
static void traverse_chunk_normal(uint32_t depth, v3 n, uint64_t loc[3]) {
  int sx = n.x >= 0.0f, sy = n.y >= 0.0f, sz = n.z >= 0.0f;
  float ix = 1.0f / n.x, iy = 1.0f / n.y, iz = 1.0f / n.z;
  uint64_t dx = orc_round_ulong(loc[0], depth, sx), dy = orc_round_ulong(loc[1], depth, sy),
           dz = orc_round_ulong(loc[2], depth, sz);
  float ax = fabsf((float)dx * ix), ay = fabsf((float)dy * iy), az = fabsf((float)dz * iz);
  if (ax >= ay && ax >= az) {
    float u = orc_ulong_to_float(dx);
    dy = orc_float_to_ulong((n.y * ix) * u);
    dz = orc_float_to_ulong((n.z * ix) * u);
  } else if (ay >= ax && ay >= az) {
    float u = orc_ulong_to_float(dy);
    dx = orc_float_to_ulong((n.x * iy) * u);
    dz = orc_float_to_ulong((n.z * iy) * u);
  } else {
    float u = orc_ulong_to_float(dz);
    dx = orc_float_to_ulong((n.x * iz) * u);
    dy = orc_float_to_ulong((n.y * iz) * u);
  }
  loc[0] = move_axis(loc[0], dx, sx);
  loc[1] = move_axis(loc[1], dy, sy);
  loc[2] = move_axis(loc[2], dz, sz);
}

/* The 12-entry rotation table of the reference, applied to v (mirror: the second ring vector) */
static v3 rotate(v3 a, float c, float s, v3 v, int mirror) {
  float t[12];
  t[0] = c;
  t[1] = s;
  t[2] = 1.0f - c;
  t[3] = t[2] * a.x * a.x;
  t[4] = t[2] * a.x * a.y;
  t[5] = t[2] * a.x * a.z;
  t[6] = t[2] * a.y * a.y;
  t[7] = t[2] * a.y * a.z;
  t[8] = t[2] * a.z * a.z;
  t[9] = t[1] * a.x;
  t[10] = t[1] * a.y;
  t[11] = t[1] * a.z;
  v3 r;
  if (!mirror) {
    r.x = (t[3] + t[0]) * v.x + (t[4] - t[11]) * v.y + (t[5] + t[10]) * v.z;
    r.y = (t[4] + t[11]) * v.x + (t[6] + t[0]) * v.y + (t[7] - t[9]) * v.z;
    r.z = (t[5] - t[10]) * v.x + (t[7] + t[9]) * v.y + (t[8] + t[0]) * v.z;
  } else {
    r.x = (t[3] + t[0]) * v.x + (t[4] + t[11]) * v.y + (t[5] - t[10]) * v.z;
    r.y = (t[4] - t[11]) * v.x + (t[6] + t[0]) * v.y + (t[7] + t[9]) * v.z;
    r.z = (t[5] + t[10]) * v.x + (t[7] - t[9]) * v.y + (t[8] + t[0]) * v.z;
  }
  return r;
}

typedef struct {
  orc_material_io *io;
  uint32_t parent;
  uint8_t r, g, b;
  float origin[3];
  uint64_t loc[3];
} acc_t;

/* saveAccumulator for a cone set up by setupTrace (Depth 1, location one cell off the surface) */
static void save(acc_t *a, v3 dir, float fov, uint32_t w) {
  orc_material_io *io = a->io;
  uint32_t id = (*io->acc_counter)++;
  if (id >= io->n_out) return;
  io->parents[id] = a->parent;
  io->weightings[id] = (uint8_t)w;
  if (w <= 5u) return;
  if (*io->bg_counter < io->bg_capacity) io->bg_queue[*io->bg_counter] = id;
  (*io->bg_counter)++;
  io->luminosities[id] = 0;
  io->colour_r[id] = a->r;
  io->colour_g[id] = a->g;
  io->colour_b[id] = a->b;
  for (int k = 0; k < 3; k++) io->origins[4 * (size_t)id + k] = a->origin[k];
  io->origins[4 * (size_t)id + 3] = 0.0f;
  io->directions[4 * (size_t)id + 0] = dir.x;
  io->directions[4 * (size_t)id + 1] = dir.y;
  io->directions[4 * (size_t)id + 2] = dir.z;
  io->directions[4 * (size_t)id + 3] = 0.0f;
  io->fovs[id] = fov;
  for (int k = 0; k < 3; k++) io->locations[4 * (size_t)id + k] = a->loc[k];
  io->depths[id] = 1;
}

int orc_evaluate_material(const orc_trace_input *in, orc_material_io *io) {
  if (in->fov_constant <= 0.0f || in->fov_multiplier >= 1.0f || in->fov_multiplier <= -1.0f) return 0;
  const uint32_t B = in->base_depth;
  for (uint32_t i = 0; i < io->n; i++) {
    uint32_t id = io->queue[i];
    if (id >= io->n_root) continue;
    uint64_t loc[3] = {io->root_locations[4 * (size_t)id], io->root_locations[4 * (size_t)id + 1],
                       io->root_locations[4 * (size_t)id + 2]};
    uint32_t depth = io->root_depths[id];
    if (depth < B + 2 || depth > MAX_DEPTH) continue;
    /* getBlockAddress, wavefront.cl:296-304 */
    uint32_t a = orc_base_location(B + 2, loc);
    for (uint32_t d = B + 2; d < depth && a < io->n_blocks; d++) a = io->blocks[a].child + orc_chunk_position(d, loc);
    if (a >= io->n_blocks) continue;
    const orc_block blk = io->blocks[a];
    /* updateUsage, wavefront.cl:377-381 */
    uint32_t g = a >> 3;
    if (io->usages && g < io->n_usages && io->usages[g].tick < 65534u && io->usages[g].tick != in->tick)
      io->usages[g].tick = in->tick;
    v3 d = {io->root_directions[4 * (size_t)id], io->root_directions[4 * (size_t)id + 1],
            io->root_directions[4 * (size_t)id + 2]};
    float base_w = (float)io->root_weightings[id];
    /* normalVector, wavefront.cl:235-245 */
    float fyaw = ((float)blk.normal_yaw * PI_F) / 32767.0f, fpitch = ((float)blk.normal_pitch * PI_F) / 32767.0f;
    float sy, cy, sp, cp;
    orc_sincos(fyaw, &sy, &cy);
    orc_sincos(fpitch, &sp, &cp);
    v3 normal = {cy * cp, sy * cp, sp};

    acc_t acc;
    acc.io = io;
    acc.parent = io->root_parents[id];
    acc.r = blk.r;
    acc.g = blk.g;
    acc.b = blk.b;
    for (int k = 0; k < 3; k++) acc.origin[k] = orc_ulong_to_float(loc[k]);
    acc.loc[0] = loc[0];
    acc.loc[1] = loc[1];
    acc.loc[2] = loc[2];
    traverse_chunk_normal(depth, normal, acc.loc);

    float fov = in->fov_constant, oldfov = fov, weighting = in->weighting_constant, ring_w = weighting;
    float spine = 0.0f;
    if (weighting * base_w < 6.0f) {
      if (io->final_w && id < io->n_final) io->final_w[id] = (uint32_t)base_w;
      continue;
    }
    if (dot3(d, normal) >= 0.0f) {
      save(&acc, d, fov, to_uchar(base_w));
      continue;
    }
    v3 axis = normalize3(cross3(d, normal));
    float dn = 2.0f * dot3(d, normal);
    v3 refl = {d.x - dn * normal.x, d.y - dn * normal.y, d.z - dn * normal.z};
    float dp = dot3(refl, normal);
    if (dp < fov) ring_w = (weighting * dp) / fov;
    save(&acc, refl, fov, to_uchar(ring_w * base_w));
    for (uint32_t k = 0; spine < PI_F && k < MAX_SPINE; k++) {
      fov = (in->fov_multiplier * (spine + fov) + in->fov_constant) / (1.0f - in->fov_multiplier);
      spine = ((1.0f + in->fov_multiplier) * (spine + oldfov)) / (1.0f - in->fov_multiplier);
      oldfov = fov;
      weighting = spine * in->weighting_multiplier + in->weighting_constant;
      ring_w = weighting;
      float ss, cs;
      orc_sincos(0.0f - spine, &ss, &cs);
      v3 seed = normalize3(rotate(axis, cs, ss, refl, 0));
      dp = dot3(seed, normal);
      if (dp < 0.0f) break;
      if (dp < fov) ring_w = (weighting * dp) / fov;
      save(&acc, seed, fov, to_uchar(ring_w * base_w));
      for (uint32_t j = 1; (float)j * fov < HALF_PI_F && j <= MAX_RING; j++) {
        ring_w = weighting;
        float ring = (float)(j * 2u) * fov;
        float sr, cr;
        orc_sincos(0.0f - ring, &sr, &cr);
        v3 rv = normalize3(rotate(refl, cr, sr, seed, 0));
        dp = dot3(rv, normal);
        if (dp < 0.0f) break;
        if (dp < fov) ring_w = (weighting * dp) / fov;
        save(&acc, rv, fov, to_uchar(ring_w * base_w));
        rv = normalize3(rotate(refl, cr, sr, seed, 1));
        save(&acc, rv, fov, to_uchar(ring_w * base_w));
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* shading kernels of the kernel-by-kernel frame                             */
/* ------------------------------------------------------------------------ */

/* evaluateBackground, wavefront.cl:1741-1797 (the light mix at 1775-1788). */
int orc_evaluate_background(uint8_t ambient, orc_shade_io *io) {
  for (uint32_t i = 0; i < io->n; i++) {
    const uint32_t id = io->queue[i];
    if (id >= io->n_rays) continue;
    const uint32_t parent = io->parents[id];
    if (parent >= io->n_pixels) continue;
    const float *dir = io->directions + 4 * (size_t)id;
    const float dp = (dir[0] * 0.0f + dir[1] * 0.0f) + dir[2] * -1.0f; /* dot(dir, (0, 0, -1)) */
    uint32_t c[3], lum;
    if (dp > 0.8f) {
      c[0] = c[1] = c[2] = 255u;
      lum = to_uchar(dp * 200.0f);
    } else {
      c[0] = 135u;
      c[1] = 206u;
      c[2] = 235u;
      lum = 1u;
    }
    const uint32_t w = io->weightings[id];
    if (w < 255u) {
      const uint32_t amb = ambient;
      const uint32_t boundary = lum > amb ? 255u : 0u;
      const uint32_t diff = lum > amb ? lum - amb : amb - lum; /* abs_diff */
      const float q = (float)diff / ((float)(lum + amb) * 255.0f);
      const uint8_t *base[3] = {io->colour_r, io->colour_g, io->colour_b};
      for (int k = 0; k < 3; k++) {
        const uint32_t b = base[k] ? base[k][id] : 0u;
        const float v = (float)b + (q * (float)c[k]) * (float)(uint32_t)(boundary - b);
        c[k] = to_uchar(v);
      }
    }
    io->final_w[parent] += w;
    if (lum > 0u) {
      io->final_r[parent] += c[0] * w;
      io->final_g[parent] += c[1] * w;
      io->final_b[parent] += c[2] * w;
    }
  }
  return 0;
}

/* resolveRemainders, wavefront.cl:1799-1833. */
int orc_resolve_remainders(uint32_t base_depth, const orc_block *blocks, uint32_t n_blocks, orc_shade_io *io) {
  const uint32_t B = base_depth;
  for (uint32_t i = 0; i < io->n; i++) {
    const uint32_t id = io->queue[i];
    if (id >= io->n_rays) continue;
    const uint32_t parent = io->parents[id];
    if (parent >= io->n_pixels) continue;
    const uint32_t w = io->weightings[id];
    const uint64_t loc[3] = {io->locations[4 * (size_t)id], io->locations[4 * (size_t)id + 1],
                             io->locations[4 * (size_t)id + 2]};
    const uint32_t depth = io->depths[id];
    if (depth < B + 2 || depth > MAX_DEPTH) continue;
    /* getBlockAddress, wavefront.cl:296-304 */
    uint32_t a = orc_base_location(B + 2, loc);
    for (uint32_t d = B + 2; d < depth; d++) a = (a < n_blocks ? blocks[a].child : 0u) + orc_chunk_position(d, loc);
    if (a == U32MAX) continue;
    const orc_block zero = {0};
    const orc_block blk = a < n_blocks ? blocks[a] : zero;
    io->final_w[parent] += w;
    io->final_r[parent] += (uint32_t)blk.r * w;
    io->final_g[parent] += (uint32_t)blk.g * w;
    io->final_b[parent] += (uint32_t)blk.b * w;
  }
  return 0;
}

/* drawTrace, wavefront.cl:1835-1850. */
void orc_draw_trace(uint32_t n_pixels, const uint32_t *final_r, const uint32_t *final_g, const uint32_t *final_b,
                    const uint32_t *final_w, float *rgba) {
  for (uint32_t i = 0; i < n_pixels; i++) {
    float *o = rgba + 4 * (size_t)i;
    const uint32_t w = final_w[i];
    if (w == 0u) {
      o[0] = 0.75f;
      o[1] = 0.0f;
      o[2] = 0.0f;
    } else {
      const float fw = (float)w * 255.0f;
      o[0] = fminf((float)final_r[i] / fw, 1.0f);
      o[1] = fminf((float)final_g[i] / fw, 1.0f);
      o[2] = fminf((float)final_b[i] / fw, 1.0f);
    }
    o[3] = 1.0f;
  }
}
