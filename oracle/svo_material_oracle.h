/*
 * svo_material_oracle.h -- CPU oracle of evaluateMaterial (wavefront.cl:1568-1739).
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

/* The defined sin/cos above (exposed for tests). */
void orc_sincos(float x, float *s, float *c);

#ifdef __cplusplus
}
#endif
#endif
