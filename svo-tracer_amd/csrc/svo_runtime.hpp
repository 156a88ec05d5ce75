// Context state of libsvo (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "svo.h"
#include "svo_kernels.hpp"

namespace svo {

struct DeviceBuffer {
  void *ptr = nullptr;
  size_t bytes = 0;
  uint32_t flags = 0;
  ~DeviceBuffer();
};

// A bound kernel argument: either a buffer allocation (captured at bind time, as
// clSetKernelArg does) or by-value bytes.
struct Arg {
  std::shared_ptr<DeviceBuffer> buffer;
  int buffer_name = -1;
  std::vector<uint8_t> bytes;
};

// Layout of the small per-context scratch block.
constexpr size_t kCountersOffset = 0;    // 8 work counters, 128 bytes apart
constexpr size_t kTotalsOffset = 1024;   // 4 x u64 frame totals
constexpr size_t kSmallScratchBytes = 2048;

int hip_fail(hipError_t e, const char *what);
int new_buffer(size_t bytes, std::shared_ptr<DeviceBuffer> *out);
int ensure_scratch(std::shared_ptr<DeviceBuffer> &b, size_t bytes);
struct Context;
// The context's span-sum scratch, grown to `bytes` (null when bytes is 0).
int context_scan_sums(Context *ctx, size_t bytes, uint32_t **out);
const std::vector<std::string> &kernel_params(int kernel);
bool valid_buffer_name(int name);

// Buffers one fused frame reads and writes (pointer + size; null = not bound).
struct FrameBuffers {
  struct B {
    void *ptr = nullptr;
    size_t bytes = 0;
  };
  B bases, blocks, usages, requests, request_id;
  B rgba, hit, visits, hashes, status;
  B locations, depths, bg_queue, mat_queue, bg_count, mat_count;
};

struct Context;
int update_camera_table(Context *ctx, const svo_trace_input &in);
// The context's named buffers a frame reads and writes.
FrameBuffers frame_buffers(Context *ctx);
// KernelName.Prune / KernelName.Graft behind svo_enqueue (arguments bound with SetArg).
int enqueue_tree(Context *ctx, int kernel, uint32_t n);
// Graft of device-resident records and graft blocks against the context's pool.
int graft_device(Context *ctx, const svo_update_input_data &in, const svo_grafting *records, uint32_t n,
                 const svo_block *graft_blocks, uint32_t n_graft_blocks, svo_tree_stats *stats);
// The same with the record count on the device (at most n_max records): nothing is read
// back; *counters gets the device address of the dispatch's counter block (kTcSlots u32),
// whose statistics stats_from() reports once the caller has read it.
int graft_device_counted(Context *ctx, const svo_update_input_data &in, const svo_grafting *records, uint32_t n_max,
                         const uint32_t *n_dev, const svo_block *graft_blocks, uint32_t n_graft_blocks,
                         const uint32_t **counters);
int stats_from(const uint32_t *counters, svo_tree_stats *out);
// Allocates the tree-management scratch of the context's pool, and what a graft of up to
// `records` one-group records needs, ahead of the first dispatch.
int reserve_tree(Context *ctx, uint32_t base_depth, uint32_t records);
// Garbage collection's reclaim of stale live groups (age > in.offset) in the context's pool.
int reclaim_device(Context *ctx, const svo_update_input_data &in, uint32_t *reclaimed);
int trace_frame_impl(Context *ctx, const svo_trace_input &in, const svo_frame_options &o, const FrameBuffers &fb);
// A frame prepared for submission (svo_trace_frame / svo_trace_batch / captured graphs).
struct BatchPlan {
  FrameParams p;
  LaunchShape shape;
  CompactParams compact;
  bool compaction = false;
  size_t cones = 0;
  uint32_t n_cams = 0;  // camera records uploaded before the kernel (0: single-camera kernel)
  float *views = nullptr;  // with a tile order: the views each band recorded (FrameBegin::views)
};
// Device scratch of frames: one set per context for direct calls, one per captured graph
// (a graph keeps its own so later frames never move memory a replay writes to).
struct FrameScratch {
  std::shared_ptr<DeviceBuffer> status, requests, small, counts, cams;
  // kept requests (SVO_FRAME_KEEP_REQUESTS): 8-byte keys per cone, requesting cone ids in order
  std::shared_ptr<DeviceBuffer> req_keys, req_cones;
  // With an arena (device memory the caller owns), the frame's scratch is carved from it in
  // 256-byte aligned pieces instead of being allocated, and the tile schedule is not kept.
  uint8_t *arena = nullptr;
  size_t arena_bytes = 0, arena_used = 0;
  // tile schedule: the per-tile cost of the last frame and the order built from it, valid
  // for the tiling (band ranges) they were measured on, and per band the views it recorded
  // last (kept on the device, so that graph replays with new poses decide there)
  std::shared_ptr<DeviceBuffer> tile_cost, tile_order, tile_views;
  uint32_t sched_ranges[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t sched_cams = 0;  // cameras the recorded views are laid out for
};
// One launch before a planned frame's trace: clears its counters and totals, writes the first
// n_inline of host_cams (n_inline <= kInlineCams) to the device records, and orders the
// tiles when the plan has a tile order.
int launch_begin(const BatchPlan &plan, const CamRec *host_cams, uint32_t n_inline, hipStream_t stream);
// `bytes` of device scratch for a frame: grown in b, or carved from sc's arena when it has one
// (SVO_ERR_CAPACITY when the arena cannot hold it).
int frame_scratch(FrameScratch &sc, std::shared_ptr<DeviceBuffer> &b, size_t bytes, void **out);
int plan_batch(Context *ctx, FrameScratch &sc, const svo_trace_input *ins, const float *apertures, uint32_t n_cams,
               const svo_frame_options &o, const FrameBuffers &fb, CamRec *host_cams, CamRec *dev_cams,
               const float *lut_dev, const uint32_t *lut_offsets, BatchPlan *plan);
int submit_batch(Context *ctx, const BatchPlan &plan, const CamRec *host_cams, bool eager);
int validate_batch(const svo_trace_input *ins, const float *apertures, uint32_t n_cams, size_t *cones);
void fill_camera_poses(const svo_trace_input *ins, const float *apertures, uint32_t n_cams, CamRec *cams);
// Several cameras in one persistent launch; outputs concatenated camera by camera.
int trace_batch_impl(Context *ctx, const svo_trace_input *ins, const float *apertures, uint32_t n_cams,
                     const svo_frame_options &o, const FrameBuffers &fb);

struct Context {
  int device = 0;
  int cu_count = 0;
  std::string arch;
  hipStream_t stream = nullptr;
  std::map<int, std::shared_ptr<DeviceBuffer>> buffers;
  std::map<int, std::map<std::string, Arg>> args;
  uint32_t tree_max_depth = 0;

  // scratch owned by the context
  FrameScratch frame_scratch;
  std::shared_ptr<DeviceBuffer> scratch_lut;
  // scratch of the kernel-by-kernel path
  std::shared_ptr<DeviceBuffer> wave_codes, wave_values, wave_staging, wave_counts;
  // scratch of tree management (prune / graft / collect) and its pinned counter mirror
  std::shared_ptr<DeviceBuffer> tree_scratch[24];
  uint32_t *tree_host = nullptr;
  // scratch of streaming (request consumption, compaction) and its pinned counter mirror
  std::shared_ptr<DeviceBuffer> stream_scratch[12];
  // scratch of evaluateMaterial: per-entry counts and first ids, chunk sums, queue flags
  std::shared_ptr<DeviceBuffer> material_scratch[5];
  uint32_t *stream_host = nullptr;
  // generation of the streaming update's first-request table (0: never used)
  uint32_t stream_gen = 0;
  // The child requests of the context's last frame, when it was a direct frame that kept them
  // (SVO_FRAME_KEEP_REQUESTS): the first `capacity` requesting cone ids in ascending order, a
  // key per requesting cone and the full records by cone id, all in frame_scratch.  Every
  // other frame (direct without the flag, a graph replay, the allocation-free frame, a trace
  // kernel through svo_enqueue) clears `valid`: its requests went to ChildRequests.
  struct KeptRequests {
    bool valid = false;
    const uint32_t *cones = nullptr;
    const uint2 *keys = nullptr;
    const svo_child_request *records = nullptr;
    uint32_t capacity = 0;
  } kept;
  // span sums of the multi-work-group scans (tree management, streaming, device builds)
  std::shared_ptr<DeviceBuffer> scan_sums;
  uint32_t *id_host = nullptr;  // pinned word for the queue-counter read-backs (svo_reset_id)
  // batched cameras: concatenated camera tables (cache key), camera records
  std::shared_ptr<DeviceBuffer> scratch_batch_lut;
  std::vector<float> batch_lut_key;
  bool batch_lut_valid = false;
  CamRec *cam_host = nullptr;  // pinned staging of the camera records
  hipEvent_t cam_copied = nullptr;
  // camera table cache key
  uint32_t lut_w = 0, lut_h = 0;
  float lut_fov0 = 0.f, lut_fov1 = 0.f;
  bool lut_valid = false;

  uint64_t last_rays = 0;
  const void *last_totals = nullptr;  // frame totals of the last submitted frame (device)

  // timing
  hipEvent_t timer_begin = nullptr, timer_end = nullptr;
  bool kernel_timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> kernel_events;
  size_t kernel_events_used = 0;

  std::shared_ptr<DeviceBuffer> get(int name) const;
};

}  // namespace svo

struct svo_context : svo::Context {};
