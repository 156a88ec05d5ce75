/*
 * svo.h -- C ABI of the MI355X (gfx950) wavefront sparse-voxel-octree cone tracer.
 *
 * This is the drop-in boundary for the operator surface of SvoTracer.Kernel
 * (ComputeManager + KernelName + BufferName, ComputeManager.cs:10-367) and the
 * buffer/argument layouts of SvoTracer.Domain (the Models directory).  A C# caller binds
 * these symbols with [DllImport("svo")]; the shim source is in
 * svo-tracer_amd/csharp/SvoNative.cs.
 *
 * Conventions
 *   - every call returns an int status (SVO_OK == 0, negative on error) and sets
 *     a thread-local message readable with svo_last_error();
 *   - sizes are in bytes everywhere (the reference mixes bytes and elements);
 *   - the library copies host data during the call and never keeps host
 *     pointers;
 *   - one context per host thread; all work of a context is serialised on one
 *     HIP stream (the reference's single in-order queue).
 *   - struct layouts are the natural OpenCL layouts of wavefront.cl:1-84
 *     (the reference's host serialisers pack some of them differently).
 */
#ifndef SVO_H
#define SVO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SVO_ABI_VERSION 1

/* ---- status codes (HandleResultCode, ComputeManager.cs:283-287, throws on any of these) ---- */
enum {
  SVO_OK = 0,
  SVO_ERR_INVALID_ARGUMENT = -1,
  SVO_ERR_NO_DEVICE = -2,
  SVO_ERR_HIP = -3,
  SVO_ERR_UNKNOWN_BUFFER = -4, /* "Buffer {name} not initializied", ComputeManager.cs:130 */
  SVO_ERR_UNKNOWN_KERNEL = -5, /* "Kernel name {name} not found", ComputeManager.cs:304,347 */
  SVO_ERR_UNKNOWN_PARAMETER = -6,
  SVO_ERR_BUFFER_TOO_SMALL = -7,
  SVO_ERR_OUT_OF_MEMORY = -8,
  SVO_ERR_NOT_SUPPORTED = -9,
  SVO_ERR_CAPACITY = -10,
  SVO_ERR_IO = -11,
  SVO_ERR_MISSING_ARGUMENT = -12
};

/* ---- KernelName ordinals, KernelName.cs:11-21 ---- */
enum {
  SVO_KERNEL_PRUNE = 0,
  SVO_KERNEL_GRAFT = 1,
  SVO_KERNEL_TRACE = 2,
  SVO_KERNEL_TEST = 3,
  SVO_KERNEL_INIT = 4,
  SVO_KERNEL_RUN_BASE_TRACE = 5,
  SVO_KERNEL_RUN_BLOCK_TRACE = 6,
  SVO_KERNEL_EVALUATE_BACKGROUND = 7,
  SVO_KERNEL_EVALUATE_MATERIAL = 8,
  SVO_KERNEL_RESOLVE_REMAINDERS = 9,
  SVO_KERNEL_DRAW_TRACE = 10,
  SVO_KERNEL_COUNT = 11
};

/* ---- BufferName ordinals, BufferName.cs:11-63 ---- */
enum {
  SVO_BUF_BASES = 0,
  SVO_BUF_BLOCKS,
  SVO_BUF_USAGES,
  SVO_BUF_CHILD_REQUEST_ID,
  SVO_BUF_CHILD_REQUESTS,
  SVO_BUF_ORIGINS,
  SVO_BUF_FOVS,
  SVO_BUF_DIRECTIONS,
  SVO_BUF_LOCATIONS,
  SVO_BUF_DEPTHS,
  SVO_BUF_WEIGHTINGS,
  SVO_BUF_PARENT_TRACES,
  SVO_BUF_ROOT_DIRECTIONS,
  SVO_BUF_ROOT_LOCATIONS,
  SVO_BUF_ROOT_DEPTHS,
  SVO_BUF_ROOT_WEIGHTINGS,
  SVO_BUF_ROOT_PARENT_TRACES,
  SVO_BUF_BASE_TRACES,
  SVO_BUF_BLOCK_TRACES,
  SVO_BUF_BASE_TRACE_QUEUE,
  SVO_BUF_BASE_TRACE_QUEUE_ID,
  SVO_BUF_BLOCK_TRACE_QUEUE,
  SVO_BUF_BLOCK_TRACE_QUEUE_ID,
  SVO_BUF_BACKGROUND_QUEUE,
  SVO_BUF_BACKGROUND_QUEUE_ID,
  SVO_BUF_MATERIAL_QUEUE,
  SVO_BUF_MATERIAL_QUEUE_ID,
  SVO_BUF_COLOUR_RS,
  SVO_BUF_COLOUR_GS,
  SVO_BUF_COLOUR_BS,
  SVO_BUF_LUMINOSITIES,
  SVO_BUF_RAY_LENGTHS,
  SVO_BUF_FINAL_COLOUR_RS,
  SVO_BUF_FINAL_COLOUR_GS,
  SVO_BUF_FINAL_COLOUR_BS,
  SVO_BUF_FINAL_WEIGHTINGS,
  SVO_BUF_ACCUMULATOR_ID,
  SVO_BUF_PARENT_SIZE,
  SVO_BUF_PARENT_RESIDENCY,
  SVO_BUF_PARENTS,
  SVO_BUF_DEREFERENCE_QUEUE,
  SVO_BUF_DEREFERENCE_REMAINING,
  SVO_BUF_SEMAPHOR,
  SVO_BUF_PRUNING,
  SVO_BUF_PRUNING_BLOCK_DATA,
  SVO_BUF_PRUNING_ADDRESSES,
  SVO_BUF_GRAFTING,
  SVO_BUF_GRAFTING_BLOCKS,
  SVO_BUF_GRAFTING_ADDRESSES,
  SVO_BUF_HOLDING_ADDRESSES,
  SVO_BUF_ADDRESS_POSITION,
  SVO_BUF_REFERENCE_COUNT, /* = 51, number of reference buffer names */

  /* Extensions with no BufferName counterpart: per-ray outputs of the frame-level
   * entry points (the reference renders into a GL image instead, wavefront.cl:1835). */
  SVO_BUF_OUTPUT_IMAGE = 64, /* float4 RGBA per ray (drawTrace's image) */
  SVO_BUF_HIT_INDICES = 65,  /* u32 per ray: block address hit, UINT32_MAX = background */
  SVO_BUF_VISIT_COUNTS = 66, /* u32 per ray: base + block steps taken */
  SVO_BUF_VISIT_HASHES = 67, /* u64 per ray: FNV-1a-64 of the visit words (depth<<32 | index) */
  SVO_BUF_RAY_STATUS = 68,   /* u8 per ray: bit0 material, bit1 child requested, bit2 step cap */
  /* differentiable compositing (svo_composite, svo_adam_step) */
  SVO_BUF_LEAF_PARAMS = 69,    /* float4 per block: colour r, g, b and opacity in [0, 1] */
  SVO_BUF_TARGET_IMAGE = 70,   /* float4 per cone: target colour of the L2 loss */
  SVO_BUF_LEAF_GRADIENTS = 71, /* float4 per block: accumulated dLoss/dLeafParams */
  SVO_BUF_LAYERS = 72,         /* max_layers u32 per cone: composited block addresses */
  SVO_BUF_ADAM_M = 73,         /* float4 per block: Adam first moment */
  SVO_BUF_ADAM_V = 74,         /* float4 per block: Adam second moment */
  SVO_BUF_NAME_LIMIT = 80
};

/* buffer flags for svo_buffer_init (InitBuffer/InitDeviceBuffer/InitReadBuffer/InitWriteBuffer,
 * ComputeManager.cs:135-215; all map to device memory, the flags are advisory) */
enum { SVO_BUFFER_READ_WRITE = 0, SVO_BUFFER_READ_ONLY = 1, SVO_BUFFER_WRITE_ONLY = 2 };

/* ---- domain structs (natural layouts) ---- */

/* Block, wavefront.cl:1-12 / Block.cs:7-21: 16 bytes */
typedef struct {
  uint32_t child;      /* first of 8 consecutive children, UINT32_MAX = none */
  uint16_t chunk;      /* per octant k: bit 2k air, bit 2k+1 geometry */
  int16_t normal_pitch;
  int16_t normal_yaw;
  uint8_t colour_r, colour_g, colour_b, opacity;
  uint8_t spare; /* Specularity in Block.cs */
  uint8_t gloss;
} svo_block;

/* Usage, wavefront.cl:14-17: 8 bytes, one per group of 8 blocks */
typedef struct {
  uint16_t tick; /* 0 free, 1..65533 last used, 65534 multi-parent, 65535 inviolate */
  uint16_t pad;
  uint32_t parent;
} svo_usage;

/* Parent, wavefront.cl:19-22 */
typedef struct {
  uint32_t parent_address;
  uint32_t next_element;
} svo_parent;

/* Pruning, wavefront.cl:24-36: 20 bytes */
typedef struct {
  uint8_t properties; /* bit0 CullChild, bit1 AlterViolability, bit2 MakeInviolate, bit3 UpdateChunk, bit4 BaseBlock */
  uint8_t depth;      /* 0xFF: address is a block address; else address indexes pruning addresses (Location) */
  uint16_t pad0;
  uint32_t address;
  uint16_t chunk;
  uint16_t pad1;
  uint32_t colour_address;
  uint32_t child_address;
} svo_pruning;

/* Grafting, wavefront.cl:38-43: 16 bytes */
typedef struct {
  uint32_t graft_data_address;
  uint32_t graft_total_size;
  uint8_t depth;
  uint8_t pad[3];
  uint32_t graft_address;
} svo_grafting;

/* ChildRequest, wavefront.cl:45-51: 40 bytes */
typedef struct {
  uint32_t address;
  uint16_t tick;
  uint8_t depth;
  uint8_t pad0;
  uint64_t location[3];
  uint8_t tree_size;
  uint8_t pad1[7];
} svo_child_request;

/* TraceInput, wavefront.cl:53-75 / TraceInput.cs:42-61: 104 bytes */
typedef struct {
  float origin[3];
  float facing[9]; /* row-major; rows: forward, right, up (camera basis) */
  float fov[2];
  float dof[2];
  uint32_t screen_size[2];
  uint8_t max_opacity;
  uint8_t base_depth;
  uint16_t tick;
  uint32_t max_child_request_id;
  float fov_multiplier;
  float fov_constant;
  float weighting_multiplier;
  float weighting_constant;
  uint8_t ambient_light_level;
  uint8_t pad[7];
} svo_trace_input;

/* UpdateInputData, wavefront.cl:77-84: 20 bytes */
typedef struct {
  uint8_t base_depth;
  uint8_t pad0;
  uint16_t tick;
  uint32_t max_child_request_id;
  uint32_t memory_size;
  uint32_t offset;
  uint32_t graft_size;
} svo_update_input_data;

/* ---- context: ComputeManagerFactory.Build (ComputeManagerFactory.cs:12-89) / Dispose (ComputeManager.cs:357-367) ---- */
typedef struct svo_context svo_context;

int svo_device_count(int *count);
int svo_create(int device_ordinal, svo_context **out);
int svo_destroy(svo_context *ctx);
const char *svo_last_error(void);
int svo_abi_version(void);

/* ---- buffers ---- */
/* InitBuffer<T> / InitDeviceBuffer / InitReadBuffer / InitWriteBuffer, ComputeManager.cs:135-215.
 * Re-initialising a name releases the old allocation, as the reference does.
 * host_data may be NULL (device-only buffer, zero-filled). */
int svo_buffer_init(svo_context *ctx, int buffer_name, size_t bytes, const void *host_data, uint32_t flags);
/* Release one named buffer (part of Dispose, ComputeManager.cs:359-360). */
int svo_buffer_release(svo_context *ctx, int buffer_name);
int svo_buffer_size(svo_context *ctx, int buffer_name, size_t *bytes);
/* WriteBuffer<T>, ComputeManager.cs:217-223 (the copy is complete when the call returns). */
int svo_buffer_write(svo_context *ctx, int buffer_name, size_t offset, const void *src, size_t bytes);
/* ReadBuffer<T>, ComputeManager.cs:225-230 (blocking). */
int svo_buffer_read(svo_context *ctx, int buffer_name, size_t offset, void *dst, size_t bytes);
/* FillBuffer<T>, ComputeManager.cs:258-273 (pattern repeated over [offset, offset+bytes)). */
int svo_buffer_fill(svo_context *ctx, int buffer_name, const void *pattern, size_t pattern_bytes, size_t offset,
                    size_t bytes);
/* Device-to-device copy between two named buffers (the clEnqueueCopyBuffer role; the reference's
 * ComputeManager has no counterpart).  Asynchronous on the context's stream. */
int svo_buffer_copy(svo_context *ctx, int dst_name, size_t dst_offset, int src_name, size_t src_offset, size_t bytes);
/* ResetIDBuffer, ComputeManager.cs:232-241: read a u32 counter, then set it to 0 (blocking). */
int svo_reset_id(svo_context *ctx, int buffer_name, uint32_t *count_out);
/* ZeroIDBuffer, ComputeManager.cs:243-249. */
int svo_zero_id(svo_context *ctx, int buffer_name);
/* Swap, ComputeManager.cs:251-256: exchanges the handles; kernel arguments bound to either
 * name must be re-bound afterwards, as in the reference. */
int svo_swap(svo_context *ctx, int buffer_a, int buffer_b);
/* Device address of a named buffer (for zero-copy interop; valid until re-init/release). */
int svo_buffer_device_pointer(svo_context *ctx, int buffer_name, void **device_ptr);

/* ---- kernels: SetArg / Enqueue / Wait / Flush, ComputeManager.cs:72-120, 277-281 ---- */
/* SetArg(KernelName, string, BufferName), ComputeManager.cs:94-100; parameter names are those of
 * the argument tables at ComputeManager.cs:313-344. */
int svo_set_arg_buffer(svo_context *ctx, int kernel_name, const char *param, int buffer_name);
/* SetArg<T>(KernelName, string, T / T[]), ComputeManager.cs:80-92 (by-value structs such as
 * TraceInput / UpdateInputData, size checked against the natural layout). */
int svo_set_arg_bytes(svo_context *ctx, int kernel_name, const char *param, const void *value, size_t bytes);
/* SetLocalArg(KernelName, string, BufferName), ComputeManager.cs:101-107: a local-memory
 * argument of the named buffer's size.  The kernel, the parameter name and the buffer are
 * checked as SetArg checks them; no kernel of this surface declares a local-memory parameter
 * (the reference's kernels keep their local arrays inside, wavefront.cl), so a valid call
 * fails with SVO_ERR_INVALID_ARGUMENT, as clSetKernelArg rejects a local size for a global
 * pointer argument. */
int svo_set_local_arg(svo_context *ctx, int kernel_name, const char *param, int buffer_name);
/* Enqueue(KernelName, nuint[] workSize), ComputeManager.cs:72-78: global size only.  Queue
 * appends are in work-item order.  EvaluateMaterial (wavefront.cl:1568-1739) hands out
 * accumulator ids from AccumulatorID in canonical order (material entries in queue order,
 * each entry's cones in the order of the reference's loops) and appends the cones with
 * weighting > 5 to BackgroundQueue in id order; sin/cos/normalize are defined functions
 * (Cody-Waite reduction + Taylor polynomials; v / sqrt(v.v)), cones beyond the trace arrays
 * are counted but not stored, and the spine/ring loops stop after 256/4096 steps. */
int svo_enqueue(svo_context *ctx, int kernel_name, uint32_t dims, const size_t *global_size);
/* Flush, ComputeManager.cs:277-281. */
int svo_flush(svo_context *ctx);
/* Wait(CLEvent[]), ComputeManager.cs:109-120: waits for everything enqueued so far. */
int svo_finish(svo_context *ctx);

/* ---- frame-level entry points (keep MainWindow.RunWavefront's loop on the device) ---- */

/* Options of svo_trace_frame; zero-initialised means defaults. */
typedef struct {
  uint32_t flags;          /* SVO_FRAME_* */
  uint32_t max_tree_depth; /* deepest block level that can be reached (stack sizing), 0 = 24 */
  uint32_t workgroup_size; /* 0 = tuned default; 64, 128, 256, 512 or 1024 */
  uint32_t refill_lanes;   /* lanes idle before a wave fetches new cones, 0 = 64 (whole wave) */
  uint32_t stack_levels;   /* LDS short-stack levels, 0 = from max_tree_depth */
  uint32_t work_chunk;     /* cones a wave claims per atomic (rounded up to 64), 0 = 64 (one 8x8 tile) */
  uint32_t waves_per_simd; /* register budget of the trace kernel: 4, 5 or 6 waves per SIMD; 0 = 5 */
  uint32_t reserved[1];
} svo_frame_options;

enum {
  SVO_FRAME_STATE = 1u << 0,  /* also write final Locations (ulong3, 32-B stride) and Depths */
  SVO_FRAME_QUEUES = 1u << 1, /* also write BackgroundQueue/MaterialQueue (+ their ID counters) */
  SVO_FRAME_NO_XCD_SPLIT = 1u << 2,
  SVO_FRAME_CAMERA_TABLE = 1u << 3, /* accepted for compatibility: camera parameters always come from a
                                       device table (one record per camera) */
  SVO_FRAME_RASTER_ORDER = 1u << 4,  /* take tiles in raster order.  By default the tiles of every band
                                        are taken most expensive first, by the cost measured in the
                                        previous frame of the same tiling (the same cameras repeated, as
                                        in a graph replay or an optimisation loop, gain most) when every
                                        camera moved less than 0.25 of the volume and turned less than
                                        30 degrees since, in raster order otherwise; the outputs are the
                                        same either way */
  SVO_FRAME_KEEP_REQUESTS = 1u << 5  /* keep the child requests for the next svo_stream_update of the
                                        context instead of writing them to ChildRequests: the frame
                                        writes their count to ChildRequestID as usual and keeps, in the
                                        context, the requesting cones in ascending order (as many as
                                        ChildRequests would have stored: MaxChildRequestId, and the
                                        buffer's size when it is bound) with an 8-byte key each;
                                        ChildRequests is not written.  For svo_trace_frame and
                                        svo_trace_batch (SVO_ERR_NOT_SUPPORTED elsewhere); needs
                                        ChildRequestID.  Any later frame that does not keep its own
                                        drops them: a direct frame without the flag, a frame graph
                                        launch, svo_trace_frame_static, or Init, Trace, RunBaseTrace
                                        or RunBlockTrace through svo_enqueue. */
};

/* One primary frame (MainWindow.RunWavefront, MainWindow.cs:356-426: init, the base/block
 * trace loop, evaluateBackground, resolveRemainders, drawTrace) in a persistent kernel.
 * Reads Bases, Blocks; updates Usages ticks (if bound); appends ChildRequests and sets
 * ChildRequestID (if bound); writes the per-ray extension buffers that are bound.
 * Queue contents are in ascending ray-id order. */
int svo_trace_frame(svo_context *ctx, const svo_trace_input *input, const svo_frame_options *options);

/* svo_trace_frame without any allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same inputs, outputs and results as svo_trace_frame,
 * except that tiles are taken in raster order (no tile schedule is kept between frames).  All
 * of the frame's device scratch -- work counters and step totals, the camera record and
 * table, one status byte per cone, 40 bytes per cone of staged child requests when
 * ChildRequests or ChildRequestID is bound, the compaction counts -- comes from one static
 * device array of 96 MiB per GPU (enough for a 1920x1080 frame with child requests and
 * queues), held from the first call on that GPU until the process ends; a frame that needs
 * more fails with SVO_ERR_CAPACITY before anything is launched.  The camera table is
 * uploaded from host stack memory (one synchronisation).  Every call on a GPU shares that
 * GPU's array, and calls that overlap on one GPU (several threads or contexts) race. */
int svo_trace_frame_static(svo_context *ctx, const svo_trace_input *input, const svo_frame_options *options);

/* Several cameras of uneven resolution in one persistent launch (no reference
 * counterpart; the batched form of svo_trace_frame).  The per-cone buffers hold the
 * cameras' cones concatenated in camera order (camera c's cone (x, y) at
 * sum_{j<c} W_j*H_j + x + y*W_c); child requests and queues are in ascending concatenated
 * index.  apertures[c] (NULL: all 1) is the cone aperture in pixels: cone FoV =
 * (apertures[c] * FoV.x) / W_c.  All cameras share BaseDepth and Tick; at most 256. */
int svo_trace_batch(svo_context *ctx, const svo_trace_input *inputs, const float *apertures, uint32_t n_cameras,
                    const svo_frame_options *options);

/* A frame's submission captured into a HIP graph (no reference counterpart): the camera
 * upload, counter reset, trace kernel and compaction of svo_trace_batch (per_camera = 0) or
 * of one svo_trace_batch per camera (per_camera = 1, each camera's cones at index 0), with
 * scratch owned by the graph.  Replays take new poses and a new tick; resolutions, FoVs,
 * BaseDepth, MaxChildRequestId and the bound buffers must stay as captured. */
typedef struct svo_frame_graph svo_frame_graph;
int svo_frame_graph_create(svo_context *ctx, const svo_trace_input *inputs, const float *apertures,
                           uint32_t n_cameras, const svo_frame_options *options, uint32_t per_camera,
                           svo_frame_graph **out);
int svo_frame_graph_launch(svo_context *ctx, svo_frame_graph *graph, const svo_trace_input *inputs);
int svo_frame_graph_destroy(svo_context *ctx, svo_frame_graph *graph);

/* Picks svo_frame_options for a (batched) frame by timing `frames` real frames per
 * candidate (coordinate descent over workgroup_size 64..1024, work_chunk 64..1024 cones per
 * claim, refill_lanes 16..64 and stack_levels 4..24), starting from *base (may be NULL).  All candidates give
 * the same results, and repeating a frame leaves the same state, so the caller's frame can
 * be tuned in place.  No reference counterpart. */
int svo_autotune(svo_context *ctx, const svo_trace_input *inputs, const float *apertures, uint32_t n_cameras,
                 uint32_t frames, const svo_frame_options *base, svo_frame_options *best, float *best_ms);

/* ---- differentiable compositing (no reference counterpart) ----
 * Every material hit of a cone (the block the reference would shade) becomes a layer with
 * the block's LeafParams (r, g, b, a); the cone continues past the hit cell, as it would
 * past an empty octant, until it has max_layers layers or leaves the volume.  Colour =
 * sum_i T_i a_i c_i + T_end * sky, T_1 = 1, T_{i+1} = T_i (1 - a_i), in FP32; with every
 * opacity 1 it is the reference's first-hit colour.  OutputImage receives the colour; with
 * a TargetImage the L2 loss sum (C - target)^2 is returned; SVO_COMPOSITE_ADJOINT adds
 * dLoss/dLeafParams to LeafGradients (the traversal is recomputed, no per-sample state is
 * stored); SVO_COMPOSITE_LAYERS writes each cone's layer addresses to Layers. */
typedef struct {
  uint32_t max_layers; /* 1..4, 0 = 4 */
  uint32_t flags;      /* SVO_COMPOSITE_* */
  uint32_t reserved[2];
} svo_composite_options;
enum { SVO_COMPOSITE_ADJOINT = 1u << 0, SVO_COMPOSITE_LAYERS = 1u << 1 };
int svo_composite(svo_context *ctx, const svo_trace_input *views, const float *apertures, uint32_t n_views,
                  const svo_composite_options *options, double *loss);

/* One Adam step over LeafParams from LeafGradients (moments in AdamM / AdamV, zero them
 * before step 1); parameters are clamped to [0, 1] and LeafGradients cleared. */
typedef struct {
  float learning_rate, beta1, beta2, epsilon;
  uint32_t step; /* 1, 2, ... */
} svo_adam_options;
int svo_adam_step(svo_context *ctx, const svo_adam_options *options);

/* Frame totals of the last svo_trace_frame: base steps, block steps, rays, cones. */
typedef struct {
  uint64_t base_steps;
  uint64_t block_steps;
  uint64_t rays;
  uint64_t child_requests;
} svo_frame_stats;
int svo_frame_stats_get(svo_context *ctx, svo_frame_stats *stats);

/* ---- tree management: prune / graft / collect on the device ----
 *
 * The reference kernels (prune, graft, dereference, findAddress: wavefront.cl:1097-1350)
 * run as one deterministic batch per call:
 *   1. every record resolves its target against the tree as it stood when the call began
 *      (Depth 0xFF: the address is a block; otherwise it indexes `locations`, three u64
 *      per entry, and the target is found by walking from the dense level to Depth; a
 *      childless block on the way appends a ChildRequest, as findAddress does, and the
 *      record is skipped; a direct target whose group is free (tick 0) is skipped);
 *   2. the ticks of every group on a resolved path are set to input->tick (inviolate and
 *      multi-parent groups keep theirs);
 *   3. detached subtrees are freed: the child group and every descendant that is neither
 *      inviolate (65535) nor multi-parent (65534) gets tick 0; a multi-parent group is
 *      removed from that parent's list instead (ascending (group, parent block) order),
 *      and becomes single-parent when one parent remains;
 *   4. fields are written; when several records write the same field of the same target,
 *      the record with the highest index wins.
 * Operates on the named buffers Bases, Blocks, Usages (required), Parents and
 * ParentResidency (optional), ChildRequests + ChildRequestID (optional; capacity
 * min(MaxChildRequestId, buffer size); requests beyond it are counted, not stored).
 * Host arrays are copied during the call. */
typedef struct {
  uint32_t applied;   /* records applied (grafts: linked) */
  uint32_t skipped;   /* target free, missing (child requested) or out-run by a later record */
  uint32_t invalid;   /* an index out of range, BaseDepth outside 1..8, or a malformed graft */
  uint32_t requests;  /* child requests made, also those beyond capacity */
  uint32_t collected; /* groups collected (graft, collect) */
  uint32_t placed;    /* grafts linked into the tree */
  uint32_t unlinks;   /* multi-parent unlink events */
  uint32_t reserved;
} svo_tree_stats;

/* prune, wavefront.cl:1215-1266.  Per record (Pruning.Properties): bit4 BaseBlock writes
 * Bases[Address] = Chunk and nothing else; bit0 CullChild frees the target's subtree and
 * sets Child = ChildAddress (and that group's Usage.Parent to the target); bits1-2 == 3 make
 * the target's group inviolate, == 1 set its tick; bit3 UpdateChunk copies every field of
 * surface[ColourAddress] except Child and sets Chunk. */
int svo_prune(svo_context *ctx, const svo_update_input_data *input, const svo_pruning *records, uint32_t n,
              const svo_block *surface, uint32_t n_surface, const uint64_t *locations, uint32_t n_locations,
              svo_tree_stats *stats);

/* graft, wavefront.cl:1268-1350.  Each record links a subtree packed in graft_blocks
 * [GraftDataAddress, +GraftTotalSize) (whole groups; Child fields index graft_blocks, as
 * WorldManager.UpdateGrafting packs them, WorldManager.cs:84-103) below its target: groups
 * are collected first (svo_collect, sum of all grafts' groups), grafts take consecutive
 * runs of the collected list in record order, each graft's groups in pre-order DFS order;
 * the target's previous subtree is freed.  Collected groups that no linked graft uses go
 * back to the pool. */
int svo_graft(svo_context *ctx, const svo_update_input_data *input, const svo_grafting *records, uint32_t n,
              const svo_block *graft_blocks, uint32_t n_graft_blocks, const uint64_t *locations,
              uint32_t n_locations, svo_tree_stats *stats);

/* The collection step of graft (wavefront.cl:1277-1305): claims up to `need` groups that are
 * free (tick 0) or stale (1..65533 and older than input->offset frames on the 1..65533 tick
 * cycle), in ascending group order, pass by pass with the offset halved after each pass
 * until it reaches 0.  A stale live group is cut from its parent block and its descendants
 * are freed.  Claimed groups get tick input->tick and parent UINT32_MAX.  input->memory_size
 * (blocks; 0 = all of Usages) bounds the groups scanned.  `groups` (host, may be NULL)
 * receives the claimed group numbers in claim order. */
int svo_collect(svo_context *ctx, const svo_update_input_data *input, uint32_t need, uint32_t *groups,
                uint32_t *count);

/* Garbage collection, step 1 (no reference counterpart; the reference's only eviction is the
 * stale scan inside graft, wavefront.cl:1277-1305): every stale live group (tick age >
 * input->offset, tick not 65534/65535) is cut from its parent block, what hangs below it is
 * freed as prune frees a subtree, and the group gets tick 0.  `reclaimed` (may be NULL)
 * receives the number of groups reclaimed. */
int svo_reclaim(svo_context *ctx, const svo_update_input_data *input, uint32_t *reclaimed);
/* Garbage collection, step 2 (no reference counterpart): with L groups live (tick != 0), the
 * live single-parent groups at or above L move, in ascending order, into the free groups below
 * L, in ascending order; blocks move with their group, and Child fields, Usage.Parent
 * back-pointers and parent-list entries follow.  `moved` (may be NULL) receives the count. */
int svo_compact(svo_context *ctx, const svo_update_input_data *input, uint32_t *moved);

/* ---- timing on the context's stream (hipEvent based) ---- */
int svo_timer_begin(svo_context *ctx);
int svo_timer_end(svo_context *ctx, float *elapsed_ms);
/* When enabled, svo_trace_frame brackets its main kernel with events; the accumulated
 * kernel time and launch count can be read (and reset) here. */
int svo_kernel_timing_enable(svo_context *ctx, int enable);
int svo_kernel_timing_get(svo_context *ctx, double *total_ms, uint64_t *launches, int reset);

/* ---- host-side scene construction: TreeBuilder.BuildTree (TreeBuilder.cs:102-150) ---- */

enum {
  SVO_GEOMETRY_SPHERES = 0,        /* SphereDefinition list (SphereDefinition.cs), TreeBuilder rules */
  SVO_GEOMETRY_CUBES = 1,          /* CubeDefinition list, axis-aligned boxes */
  SVO_GEOMETRY_SPHERE_NOISE = 2,   /* unit sphere SDF + value-noise displacement, world [-s,s]^3 */
  SVO_GEOMETRY_FBM_HEIGHTFIELD = 3 /* fBm heightfield over a square grid, z up */
};

typedef struct {
  uint32_t kind;
  uint32_t count;       /* SPHERES/CUBES: number of shapes */
  const float *shapes;  /* SPHERES: (cx,cy,cz,r)*count; CUBES: (minx,miny,minz,maxx,maxy,maxz)*count */
  uint64_t seed;
  uint32_t octaves;
  uint32_t resolution;  /* FBM_HEIGHTFIELD: samples per side */
  float lacunarity;
  float gain;
  float amplitude;      /* SPHERE_NOISE: displacement amplitude */
  float frequency;      /* base frequency of the first octave */
  float half_extent;    /* SPHERE_NOISE: octree cube covers world [-s, s]^3 */
  float z_min, z_max;   /* FBM_HEIGHTFIELD: height range inside the unit cube */
} svo_geometry;

typedef struct svo_tree svo_tree;

typedef struct {
  uint32_t base_depth;
  uint32_t max_depth;
  uint32_t block_count; /* trimmed exactly as TreeBuilder.cs:145 (including the BuildBlocks count rule) */
  uint32_t base_count;  /* PowSum(base_depth) */
  uint32_t group_count; /* ceil(block_count / 8) */
  uint32_t allocated_groups;
} svo_tree_info;

int svo_tree_build(const svo_geometry *geometry, uint32_t base_depth, uint32_t max_depth, uint32_t max_size,
                   svo_tree **out);

/* The builder's geometry on the host, for MakeBlock of arbitrary cells (TreeBuilder.MakeBlock /
 * MakeChunk, TreeBuilder.cs:74-93, 264-284: the block the builder writes for that cell, with
 * Child = UINT32_MAX).  cells holds n (level, x, y, z) quadruples, level <= 30; out receives n
 * blocks.  No reference counterpart: the host side of on-demand refinement (the bricks the
 * device generator makes in svo_stream_update equal these blocks), e.g. for CPU tools and
 * checks.  A heightfield builds its full grid and min/max pyramid once, in svo_geometry_create. */
typedef struct svo_geometry_host svo_geometry_host;
int svo_geometry_create(const svo_geometry *geometry, svo_geometry_host **out);
int svo_geometry_destroy(svo_geometry_host *geometry);
int svo_geometry_blocks(const svo_geometry_host *geometry, const uint32_t *cells, uint32_t n, svo_block *out);
int svo_tree_info_get(const svo_tree *tree, svo_tree_info *info);
int svo_tree_copy_bases(const svo_tree *tree, uint16_t *dst, size_t count);
int svo_tree_copy_blocks(const svo_tree *tree, svo_block *dst, size_t count);
/* Initial usage table (MainWindow.cs:129-141 with its defect D11 resolved): groups of the dense
 * level (g < 8^(B+1)) are inviolate (65535, parent UINT32_MAX); every other group gets
 * `tick` and the address of the block that points to it (UINT32_MAX when unreferenced). */
int svo_tree_initial_usages(const svo_tree *tree, uint16_t tick, svo_usage *dst, size_t count);
/* .oct persistence (OctreeSerializer.cs:13-45). */
int svo_tree_save_oct(const svo_tree *tree, const char *path);
int svo_tree_load_oct(const char *path, svo_tree **out);
int svo_tree_free(svo_tree *tree);
/* Uploads Bases, Blocks and (optionally) initial Usages of a tree into the named buffers. */
int svo_context_load_tree(svo_context *ctx, const svo_tree *tree, int with_usages, uint16_t usage_tick);

/* No reference counterpart; a self-test hook: the device's floatToUlong (kind 0: n floats in,
 * n u64 out) or ulongToFloat (kind 1: n u64 in, n floats out), wavefront.cl:200-209, applied
 * to host arrays, so the device conversions can be compared with any host restatement. */
int svo_convert_probe(svo_context *ctx, int kind, const void *in, void *out, size_t n);

/* Separable camera table used for ray generation (replaces native_sin/native_cos of
 * setupInitialTrace, wavefront.cl:563-595): cos_u/sin_u per column, cos_v/sin_v per row. */
int svo_camera_lut(const svo_trace_input *input, float *cos_u, float *sin_u, float *cos_v, float *sin_v);

/* ---- streaming: procedural terrain generated on the device on demand ---- */

typedef struct svo_terrain svo_terrain;

/* Builds the height grid and its min/max pyramid of an SVO_GEOMETRY_FBM_HEIGHTFIELD
 * description on the context's device (the same function the host builder samples). */
int svo_terrain_create(svo_context *ctx, const svo_geometry *geometry, svo_terrain **out);
int svo_terrain_destroy(svo_terrain *terrain);
/* Copies the finest height grid (resolution^2 floats, row-major) to the host. */
int svo_terrain_heights(const svo_terrain *terrain, float *dst, size_t count);
/* TreeBuilder.BuildTree(baseDepth, maxDepth, maxSize) (TreeBuilder.cs:102-256) of the terrain,
 * built on the context's device: the same Bases, the same blocks at the same addresses (the
 * depth-first allocation order from the free-group queue), the same BlockCount and group
 * parents as svo_tree_build with the terrain's geometry.  The tree becomes the context's
 * Bases and Blocks (and Usages, as svo_context_load_tree sets them, when with_usages);
 * *out (optional) receives a host copy, *block_count (optional) the block count.  As there,
 * a free-group queue that runs out at a dense-level block is SVO_ERR_CAPACITY, and one that
 * runs out deeper leaves the remaining blocks childless.  base_depth 1..7, max_depth <= 20. */
int svo_terrain_build_tree(svo_context *ctx, const svo_terrain *terrain, uint32_t base_depth, uint32_t max_depth,
                           uint32_t max_size, int with_usages, uint16_t usage_tick, svo_tree **out,
                           uint32_t *block_count);
/* The same for any builder geometry (svo_tree_build's): spheres, boxes, the sphere + noise SDF
 * or a heightfield (whose terrain is made and released inside the call). */
int svo_tree_build_device(svo_context *ctx, const svo_geometry *geometry, uint32_t base_depth, uint32_t max_depth,
                          uint32_t max_size, int with_usages, uint16_t usage_tick, svo_tree **out,
                          uint32_t *block_count);
/* Seed tree in the bound Bases / Blocks / Usages: Bases of levels 1..base_depth and the dense
 * block level base_depth+2 exactly as svo_tree_build with max_depth = base_depth+2 makes them
 * (BuildBaseChunks / PopulateInitialBlocks, TreeBuilder.cs:152-227); the rest of Blocks is
 * zeroed; the dense groups (g < 8^(base_depth+1)) are inviolate and every other group free. */
int svo_stream_seed(svo_context *ctx, const svo_terrain *terrain, uint32_t base_depth);

enum {
  SVO_STREAM_NO_GC = 1,   /* never run garbage collection in this update */
  SVO_STREAM_FORCE_GC = 2 /* run it whatever the free share */
};

typedef struct {
  uint32_t base_depth;
  uint32_t max_depth;            /* deepest block level the terrain is refined to */
  uint16_t tick;                 /* frame tick (1..65533) */
  uint16_t pad;
  uint32_t max_bricks;           /* grafts per update; 0 = 65536 */
  uint32_t max_child_request_id; /* requests stored by the frame (TraceInput.MaxChildRequestId); 0 = buffer */
  uint32_t prune_age;            /* a group unused for more frames than this is stale */
  uint32_t gc_free_percent;      /* collect garbage when free groups < this % of the pool */
  uint32_t flags;                /* SVO_STREAM_* */
} svo_stream_options;

typedef struct {
  uint32_t requests;    /* child requests the frame made (also those beyond capacity) */
  uint32_t unique;      /* distinct requesting blocks that can take a child level */
  uint32_t bricks;      /* bricks grafted */
  uint32_t collected;   /* groups collected for the grafts */
  uint32_t gc_runs;     /* 1 when garbage collection ran */
  uint32_t reclaimed;   /* groups reclaimed */
  uint32_t moved;       /* groups moved by compaction */
  uint32_t free_groups; /* free groups after the update */
  uint32_t pool_groups;
  uint32_t reserved[3];
} svo_stream_stats;

/* One streaming step after a traced frame (the WorldManager.UpdateGrafting / RunGraft role,
 * WorldManager.cs:84-103, MainWindow.cs:307-337, fed by requestChild, wavefront.cl:362-375):
 *  1. the stored child requests (ChildRequests, count in ChildRequestID; or, when the context's
 *     last frame was a direct frame traced with SVO_FRAME_KEEP_REQUESTS, the requests it kept, under the
 *     same limits: the same result without the 40-byte copies) are read in order;
 *     among those asking for a level (TreeSize >= 1: the cone needs detail below the block),
 *     the first request of every block address whose block is childless, at depth 1 <=
 *     d < max_depth, and can have children (an octant with both air and geometry,
 *     CanHaveChildren, TreeBuilder.cs:291-302) is kept;
 *  2. the first max_bricks of them, in request order, become graft records k (Grafting,
 *     GraftAddress = the block, Depth 0xFF) of one brick each: the block's 8 children at level
 *     d+1 from the terrain (GraftingBlocks[8k, 8k+8), MakeBlock rules, TreeBuilder.cs:74-93);
 *  3. the records are grafted as svo_graft does, collection age threshold prune_age;
 *  4. when free groups < gc_free_percent % of the pool: svo_reclaim (age > prune_age), then
 *     svo_compact. */
int svo_stream_update(svo_context *ctx, const svo_terrain *terrain, const svo_stream_options *options,
                      svo_stream_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* SVO_H */
