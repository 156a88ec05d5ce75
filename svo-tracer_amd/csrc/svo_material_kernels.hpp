// Launch interface of evaluateMaterial (material.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace svo {

struct MaterialParams {
  // MaterialQueue entries [0, n) name cones by id in the Root* arrays (n_root of them)
  const uint32_t *queue;
  uint32_t n;
  const float4 *root_directions;
  const uint64_t *root_locations;  // 32-B stride
  const uint8_t *root_depths;
  const uint8_t *root_weightings;
  const uint32_t *root_parents;
  uint32_t n_root;
  const uint4 *blocks;
  uint32_t n_blocks;
  uint16_t *usage16;  // Usages as u16 words (tick at word 4g), may be null
  uint32_t n_usages;
  uint32_t B, tick;
  float fov_multiplier, fov_constant, weighting_multiplier, weighting_constant;
  // accumulator traces: ids from *acc_counter on, arrays of n_out entries
  uint32_t *acc_counter;
  float4 *origins, *directions;
  float *fovs;
  uint64_t *locations;  // 32-B stride
  uint8_t *depths, *weightings, *colour_r, *colour_g, *colour_b, *luminosities;
  uint32_t *parents;
  uint32_t n_out;
  uint32_t *final_w;  // FinalWeightings, indexed by the material cone's id
  uint32_t n_final;
  uint32_t *bg_flags;  // one per accumulator trace of this launch
  uint32_t flag_capacity;
};

uint32_t material_chunks(uint32_t n);
int launch_material_count(const MaterialParams &p, uint32_t *counts, hipStream_t s);
// Exclusive scan of n u32 values in chunks of 1024: sums gets one entry per chunk (+1 for the
// total, written at sums[chunks]).
int launch_material_scan(const uint32_t *in, uint32_t n, uint32_t *out, uint32_t *sums, hipStream_t s);
int launch_material_write(const MaterialParams &p, const uint32_t *first, hipStream_t s);
int launch_material_append(const uint32_t *flags, uint32_t n, const uint32_t *offsets, const uint32_t *acc_counter,
                           const uint32_t *bg_counter, uint32_t *queue, uint32_t capacity, hipStream_t s);
int launch_material_counters(uint32_t *acc_counter, uint32_t *bg_counter, const uint32_t *cone_total,
                             const uint32_t *bg_total, hipStream_t s);

}  // namespace svo
