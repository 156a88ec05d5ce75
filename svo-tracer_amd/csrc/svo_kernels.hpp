// Launch interface of the HIP kernels (plain structs, usable from host code).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "svo.h"

namespace svo {

// work counters sit on separate 128-byte lines (u32 stride)
constexpr uint32_t kCounterStride = 32;
// LDS bytes per short-stack entry of the trace kernel: block address and Child (u32 each),
// chunk (u16)
constexpr uint32_t kStackEntryBytes = 10;
// Most slots of the trace kernel's per-workgroup LDS table of usage groups already ticked in
// the launch (direct-mapped on the group index, u32 each); a launch uses a power of two up to
// this many, or none, so that the table never lowers the work-groups a CU holds
constexpr uint32_t kMinTickedSlots = 256;  // what the default stack depth leaves room for
constexpr uint32_t kTickedSlots = 1024;

// One camera of a batched frame (96 bytes; read into LDS by the trace kernel).
struct CamRec {
  float facing[9];
  float origin[3];
  float cone_fov;
  float dof[2];
  uint32_t W, H, tiles_x, tile_begin, ray_begin, lut_offset;
  uint32_t pad[3];
};
static_assert(sizeof(CamRec) == 96, "CamRec is 24 words");
constexpr uint32_t kCamWords = 24;
constexpr uint32_t kMaxBatchCameras = 256;

// Everything one primary frame needs; passed by value as the kernel argument.
struct FrameParams {
  // tree
  const uint16_t *bases;
  uint32_t n_bases;
  uint32_t bases_in_lds;
  const uint4 *blocks;
  uint32_t n_blocks;
  uint16_t *usage16;  // svo_usage array viewed as u16: tick of group g at [4g]
  uint32_t n_usages;
  // camera
  const float *lut;  // per camera: cos_u[W], sin_u[W], cos_v[H], sin_v[H] at its lut_offset
  uint32_t B;
  // work distribution
  uint32_t range_begin[9];
  uint32_t *counters;
  unsigned long long *totals;  // [0] base steps [1] block steps [2] child requests
  // outputs (any may be null)
  float4 *rgba;
  uint32_t *hit;
  uint32_t *visits;
  uint64_t *hashes;
  uint8_t *status;
  uint8_t *depth_out;
  uint64_t *loc_out;  // ulong3 with a 32-byte stride
  svo_child_request *req_sparse;
  uint2 *req_keys;  // (kept requests) per requesting cone: address, depth | TreeSize << 8
  // cameras (one or a batch): the tiles of all cameras form one work space, outputs are
  // indexed by the concatenated cone index
  const CamRec *cams;
  uint32_t n_cams;
  // differentiable compositing (composite modes only)
  const float4 *leaf_params;  // (r, g, b, opacity) per block
  const float4 *target;       // per cone, may be null
  float4 *grads;              // per block (adjoint)
  uint32_t *layers_out;       // max_layers per cone, may be null
  double *loss;               // sum of squared errors, may be null
  uint32_t max_layers;        // 1..4
  // tuning
  uint32_t stack_levels;
  uint32_t ticked_slots;  // slots of the LDS table of ticked usage groups (0: no table)
  uint32_t refill;
  uint32_t chunk;  // work items a wave claims per atomic (multiple of 64)
  // tile schedule (may be null): position k of a band's work range holds tile order[k]; a
  // wave stores the loop iterations it spent on a tile in tile_cost[tile]
  const uint32_t *order;
  uint32_t *tile_cost;
};

struct LaunchShape {
  int workgroup;
  int waves;  // register budget: waves per SIMD the kernel is compiled for (6)
  int grid;
  size_t lds_bytes;
  uint32_t ticked_slots;  // the ticked-group table within lds_bytes
};

struct CompactParams {
  const uint8_t *status;
  uint32_t n;
  uint32_t *counts;  // 4 per chunk
  uint32_t *bg_queue, *mat_queue, *bg_count, *mat_count, *req_count;
  const svo_child_request *req_sparse;
  svo_child_request *requests;  // the first req_capacity requests are copied here, or
  uint32_t *req_cones;          // (kept requests) the cone ids of the first req_capacity written here
  uint32_t req_capacity;
  unsigned long long *totals;
};

int trace_occupancy(int wg, int waves, size_t lds_bytes, int *blocks_per_cu);
// What a frame does before its trace launch, in one launch: clear the counters and totals,
// write up to kInlineCams camera records carried in the argument (more are uploaded by a
// copy beforehand), and, with bands = 8, build each band's tile order.  A band's tiles are
// ordered by the cost its waves measured in the previous frame (most expensive first, in
// power-of-two classes, stable within a class) when every camera moved less than
// kViewMoved of the volume and turned less than kViewTurnCos (as a cosine) since the views
// the band recorded last; otherwise the band's order is raster order.  Either way the
// band then records the frame's views.
constexpr uint32_t kInlineCams = 4;
constexpr uint32_t kViewWords = 8;  // origin[3], forward direction[3], recorded flag (1.0f), unused
constexpr float kViewMoved = 0.25f;
constexpr float kViewTurnCos = 0.8660f;  // cos 30 degrees
struct FrameBegin {
  uint32_t *counters;
  unsigned long long *totals;
  CamRec *cams;     // the frame's camera records (destination of the inline records)
  uint32_t n_cams;  // inline records to write, 0..kInlineCams
  uint32_t bands;   // 0 (no tile order) or 8
  uint32_t stage;   // set by the launcher: stage the cost classes in LDS
  uint32_t n_views; // cameras of the frame (all inline when n_cams > 0)
  const uint32_t *cost;
  uint32_t *order;
  float *views;     // per band, n_views records of kViewWords floats
  uint32_t range_begin[9];
  CamRec cam[kInlineCams];
};
int launch_frame_begin(const FrameBegin &f, hipStream_t stream);
// Sets n words to zero (a kernel, so it may be captured in a graph).
int launch_zero_u32(uint32_t *p, uint32_t n, hipStream_t stream);
// kind 0: floatToUlong of n floats into n u64; kind 1: ulongToFloat of n u64 into n floats.
int launch_convert(uint32_t kind, const void *in, void *out, size_t n, hipStream_t stream);
int launch_trace_frame(const FrameParams &p, const LaunchShape &shape, hipStream_t stream);
// Compositing modes of the trace kernel (camera table, 256-lane work-groups).
enum { kModePlain = 0, kModeComposite = 1, kModeAdjoint = 2 };
constexpr uint32_t kMaxLayers = 4;
int launch_composite(const FrameParams &p, const LaunchShape &shape, int mode, hipStream_t stream);
int composite_occupancy(size_t lds_bytes, int *blocks_per_cu);
uint32_t compaction_chunks(uint32_t n);
int launch_compaction(const CompactParams &c, hipStream_t stream);

}  // namespace svo
