// Launch interface of the kernel-by-kernel wavefront path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "svo.h"

namespace svo {

// queue codes written by the step kernels (low 3 bits), plus the child-request flag
constexpr uint8_t kQueueBase = 0;
constexpr uint8_t kQueueBlock = 1;
constexpr uint8_t kQueueBackground = 2;
constexpr uint8_t kQueueMaterial = 3;
constexpr uint8_t kCodeRequest = 8;
constexpr uint8_t kCodeNone = 0xff;
constexpr int kAppendTargets = 5;  // four queues + child requests

// The reference's per-cone SoA buffers (float3 with a 16-byte stride, ulong3 with 32).
struct WaveState {
  float4 *origins;
  float4 *directions;
  float *fovs;
  uint64_t *locations;
  uint8_t *depths;
  uint8_t *weightings;
  uint32_t *parents;
  uint32_t n_rays;  // cones every bound state buffer can hold
};

struct InitParams {
  WaveState s;
  const float *lut;
  float facing[9];
  float origin[3];
  float cone_fov;
  uint32_t W, H, gsx, gsy;
  uint8_t *colour_r, *colour_g, *colour_b;
  uint32_t *final_r, *final_g, *final_b, *final_w;
  uint8_t *codes;
  uint32_t *values;
};

struct StepParams {
  WaveState s;
  const uint32_t *in_queue;
  uint32_t n;
  const uint16_t *bases;
  uint32_t B;
  const uint4 *blocks;
  uint32_t n_blocks;
  uint16_t *usage16;
  uint32_t n_usages;
  uint32_t tick;
  float dof0, dof1;
  uint8_t *codes;
  uint32_t *values;
  svo_child_request *req_staging;
};

struct ShadeParams {
  const uint32_t *queue;
  uint32_t n;
  uint32_t n_rays;
  const float4 *directions;
  const uint32_t *parents;
  const uint8_t *weightings;
  const uint8_t *colour_r, *colour_g, *colour_b;
  uint32_t *final_r, *final_g, *final_b, *final_w;
  uint32_t n_pixels;
  uint32_t ambient;
  const uint4 *blocks;
  uint32_t n_blocks;
  const uint64_t *locations;
  const uint8_t *depths;
  uint32_t B;
};

struct DrawParams {
  const uint32_t *final_r, *final_g, *final_b, *final_w;
  uint32_t n_pixels;
  float4 *out;
  uint32_t n_out;
  uint32_t W, gsx, gsy;
};

struct AppendParams {
  const uint8_t *codes;
  const uint32_t *values;
  uint32_t n;
  uint32_t *queue[4];
  uint32_t capacity[kAppendTargets];
  uint32_t *counter[kAppendTargets];
  const svo_child_request *req_staging;
  svo_child_request *requests;
  uint32_t *counts;  // kAppendTargets per chunk
};

uint32_t append_chunks(uint32_t n);
int launch_append(const AppendParams &p, hipStream_t s);
int launch_init(const InitParams &p, hipStream_t s);
int launch_base_step(const StepParams &p, hipStream_t s);
int launch_block_step(const StepParams &p, hipStream_t s);
int launch_background(const ShadeParams &p, hipStream_t s);
int launch_remainders(const ShadeParams &p, hipStream_t s);
int launch_draw(const DrawParams &p, hipStream_t s);

}  // namespace svo
