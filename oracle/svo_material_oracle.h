/*
 * svo_material_oracle.h -- CPU oracle of the second bounce: evaluateMaterial
 * (wavefront.cl:1568-1739) and the kernel-by-kernel shading it feeds: evaluateBackground
 * with the light mix of non-primary cones (1741-1797), resolveRemainders (1799-1833) and
 * drawTrace (1835-1850).
 *
 * TEST INFRASTRUCTURE ONLY, like svo_oracle.h: used by tests/ as the checker and never
 * linked or loaded by the product.
 *
 * Sequential restatement: material queue entries in order; each entry's secondary cones in
 * the order of the reference's loops (the reflection, then per spine step the ring seed and
 * the ring pairs); accumulator ids from *acc_counter on; background-queue appends (cones with
 * weighting > 5 whose id fits the trace arrays) in id order.  The functions the reference
 * leaves to the OpenCL driver are defined here:
 *   - sin/cos: k = rint(x * 2/pi), r = (x - k*1.5703125) - k*4.83826794897e-4, then the Taylor
 *     polynomials sin r = r + r^3(-1/6 + r^2(1/120 + r^2(-1/5040 + r^2/362880))) and
 *     cos r = 1 + r^2(-1/2 + r^2(1/24 + r^2(-1/720 + r^2(1/40320 - r^2/3628800)))), quadrant
 *     by k mod 4 (Cody-Waite reduction; float basic operations, no contraction);
 *   - normalize(v) = v / sqrt(dot(v, v)), dot(a, b) = (a.x b.x + a.y b.y) + a.z b.z,
 *     cross as the textbook formula;
 *   - (uchar)f truncates and clamps to [0, 255], NaN -> 0;
 *   - loop bounds: at most 256 spine steps and 4096 ring steps per spine step;
 *   - getBlockAddress that leaves Blocks (address >= n_blocks) skips the entry.
 */
#ifndef SVO_MATERIAL_ORACLE_H
#define SVO_MATERIAL_ORACLE_H
#include <stdint.h>

#include "svo_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  /* inputs */
  const uint32_t *queue;
  uint32_t n;
  const float *root_directions; /* 4 floats per trace */
  const uint64_t *root_locations; /* 4 u64 per trace */
  const uint8_t *root_depths, *root_weightings;
  const uint32_t *root_parents;
  uint32_t n_root;
  const orc_block *blocks;
  uint32_t n_blocks;
  orc_usage *usages; /* may be NULL */
  uint32_t n_usages;
  /* accumulator traces (n_out entries each) */
  uint32_t *acc_counter;
  float *origins, *directions; /* 4 floats per trace */
  float *fovs;
  uint64_t *locations; /* 4 u64 per trace */
  uint8_t *depths, *weightings, *colour_r, *colour_g, *colour_b, *luminosities;
  uint32_t *parents;
  uint32_t n_out;
  uint32_t *final_w;
  uint32_t n_final;
  uint32_t *bg_queue, *bg_counter;
  uint32_t bg_capacity;
} orc_material_io;

/* evaluateMaterial over queue[0, n) with TraceInput `in` (BaseDepth, Tick, Fov*, Weighting*). */
int orc_evaluate_material(const orc_trace_input *in, orc_material_io *io);

/* The shading kernels over the per-trace SoA buffers of MainWindow (the product's
 * kernel-by-kernel path): queue[0, n) of trace ids into the per-pixel accumulators of their
 * parents.  Ids >= n_rays and parents >= n_pixels are skipped (the reference indexes past
 * its buffers).  Accumulator sums wrap as u32 additions. */
typedef struct {
  const uint32_t *queue;
  uint32_t n;
  uint32_t n_rays;
  const float *directions;   /* 4 floats per trace */
  const uint32_t *parents;
  const uint8_t *weightings;
  const uint8_t *colour_r, *colour_g, *colour_b; /* may be NULL: read as 0 */
  const uint64_t *locations; /* 4 u64 per trace (resolveRemainders) */
  const uint8_t *depths;
  uint32_t *final_r, *final_g, *final_b, *final_w;
  uint32_t n_pixels;
} orc_shade_io;

/* evaluateBackground, wavefront.cl:1741-1797.  Sky (135, 206, 235) with luminosity 1, or the
 * sun (255, 255, 255) with luminosity (uchar)(200 d) where d = dot(direction, (0, 0, -1)) >
 * 0.8.  A cone of weighting below 255 mixes its base colour towards the light:
 *   boundary = luminosity > ambient ? 255 : 0                                  (u32)
 *   q = |luminosity - ambient| / ((luminosity + ambient) * 255)                (float; the
 *       reference's native_divide is a correctly rounded division here)
 *   colour = base + (q * light) * (float)(u32)(boundary - base)                (the u32
 *       subtraction wraps when base > boundary, as in the reference)
 * and the float result is converted to 0..255 by truncation with saturation (NaN -> 0; the
 * reference's conversion of an out-of-range float to uchar is undefined).  W[parent] +=
 * weighting; C[parent] += colour * weighting when the luminosity is non-zero. */
int orc_evaluate_background(uint8_t ambient, orc_shade_io *io);

/* resolveRemainders, wavefront.cl:1799-1833: the block at (Location, Depth) by the
 * root-restart walk (getBlockAddress; entries with Depth outside [B+2, 60] or an address of
 * UINT_MAX are skipped, reads past Blocks give a zero block), then W[parent] += weighting,
 * C[parent] += colour * weighting. */
int orc_resolve_remainders(uint32_t base_depth, const orc_block *blocks, uint32_t n_blocks, orc_shade_io *io);

/* drawTrace, wavefront.cl:1835-1850: W = 0 -> (0.75, 0, 0, 1), else min(C / (W * 255), 1)
 * per channel and alpha 1, into rgba (4 floats per pixel). */
void orc_draw_trace(uint32_t n_pixels, const uint32_t *final_r, const uint32_t *final_g, const uint32_t *final_b,
                    const uint32_t *final_w, float *rgba);

/* The defined sin/cos above (exposed for tests). */
void orc_sincos(float x, float *s, float *c);

#ifdef __cplusplus
}
#endif
#endif
