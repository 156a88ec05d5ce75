/*
 * svo_oracle.h -- CPU oracle for the wavefront sparse-voxel-octree cone tracer.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (SvoTracer.Kernel/wavefront.cl and the frame sequencing of
 * SvoTracer.Window/MainWindow.cs) used by tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py as the checker.  Nothing in the product
 * (svo-tracer_amd/) links, loads or calls it.
 *
 * It deliberately keeps the reference's own structure: every block step
 * re-walks the tree from the dense level (getBlockAddress, wavefront.cl:296-304)
 * instead of using a short stack, and the wavefront emulation reproduces the
 * host loop of MainWindow.cs:452-481 with order-preserving queue appends.  The
 * GPU product is structured differently (fused per-lane loop, LDS short stack),
 * so agreement between the two is a real check.
 *
 * Struct layouts are the natural OpenCL layouts of wavefront.cl:1-84 and are
 * declared here independently of include/svo.h.
 *
 * Pinning status of each function is recorded next to it; the pinning tests
 * live in tests/test_oracle_*.py.
 */
#ifndef SVO_ORACLE_H
#define SVO_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  uint32_t child;
  uint16_t chunk;
  int16_t normal_pitch;
  int16_t normal_yaw;
  uint8_t r, g, b, opacity, spare, gloss;
} orc_block;

typedef struct {
  uint16_t tick;
  uint16_t pad;
  uint32_t parent;
} orc_usage;

typedef struct {
  uint32_t address;
  uint16_t tick;
  uint8_t depth;
  uint8_t pad0;
  uint64_t location[3];
  uint8_t tree_size;
  uint8_t pad1[7];
} orc_child_request;

typedef struct {
  float origin[3];
  float facing[9];
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
} orc_trace_input;

/* Per-ray results of one primary frame. All arrays hold W*H entries; any may be NULL. */
typedef struct {
  uint32_t *hit;          /* block address re-walked by resolveRemainders, UINT32_MAX for background */
  uint32_t *visits;       /* number of base + block steps */
  uint64_t *hash;         /* FNV-1a-64 over the visit words ((depth<<32)|index) */
  float *rgba;            /* drawTrace output, 4 floats per ray */
  uint8_t *status;        /* 0 background, 1 material, |2 child requested */
  uint8_t *depth;         /* final Depths[] */
  uint64_t *location;     /* final Locations[], 3 per ray */
  uint64_t *visit_log;    /* optional: first log_stride visit words per ray */
  uint32_t log_stride;
  /* frame totals */
  uint64_t base_steps;
  uint64_t block_steps;
  uint64_t groups_touched; /* usage groups whose tick changed */
} orc_frame_out;

/* ---- individual helpers (exported for the pinning tests) ---- */
uint64_t orc_float_to_ulong(float x);
float orc_ulong_to_float(uint64_t x);
void orc_float_to_ulong_n(const float *x, uint64_t *out, size_t n);
void orc_ulong_to_float_n(const uint64_t *x, float *out, size_t n);
uint64_t orc_round_ulong(uint64_t value, uint32_t depth, int round_up);
uint32_t orc_pow_sum(uint32_t depth);
uint32_t orc_chunk_position(uint32_t depth, const uint64_t loc[3]);
uint32_t orc_base_location(uint32_t depth, const uint64_t loc[3]);
uint32_t orc_compare_positions(uint32_t depth, const uint64_t prev[3], const uint64_t loc[3],
                               const uint8_t sign[3]);
/* traverseChunk on a bare ray state; loc updated in place. */
void orc_traverse_chunk(uint32_t depth, uint64_t loc[3], const float dir[3], const float inv[3],
                        const uint8_t sign[3]);
/* init of one pixel: returns 1 (hit -> base queue) or 0 (background); writes dir, inv, loc, cone_fov. */
int orc_init_ray(const orc_trace_input *in, uint32_t x, uint32_t y, float dir[3], float inv[3],
                 uint64_t loc[3], float *cone_fov);
void orc_camera_lut(const orc_trace_input *in, float *cos_u, float *sin_u, float *cos_v, float *sin_v);

/* ---- whole-frame drivers ---- */

/* Per-ray independent tracing (the result every correct schedule must give).
 * `threads` > 1 uses OpenMP.  Usages/child requests updated as the reference frame
 * would leave them; child requests are stored in ascending ray-id order.
 * Returns 0 on success. */
int orc_trace_frame(const orc_trace_input *in, const uint16_t *bases, uint32_t n_bases,
                    const orc_block *blocks, uint32_t n_blocks, orc_usage *usages, uint32_t n_usages,
                    orc_child_request *requests, uint32_t request_capacity, uint32_t *request_count,
                    orc_frame_out *out, int threads);

/* Several cameras in one frame: outputs concatenated camera by camera (arrays sized for the
 * sum of W*H), cone aperture apertures[c] pixels (NULL: 1), child requests in ascending
 * concatenated index; all cameras must share tick and base depth. */
int orc_trace_batch(const orc_trace_input *ins, const float *apertures, uint32_t n_cams, const uint16_t *bases,
                    uint32_t n_bases, const orc_block *blocks, uint32_t n_blocks, orc_usage *usages,
                    uint32_t n_usages, orc_child_request *requests, uint32_t request_capacity,
                    uint32_t *request_count, orc_frame_out *out, int threads);

/* Trace only the rays listed in `ray_ids` (a bounded sample, used for the CPU baseline).
 * Usages are not touched.  hit/visits are written at the sample position. */
int orc_trace_rays(const orc_trace_input *in, const uint16_t *bases, uint32_t n_bases,
                   const orc_block *blocks, uint32_t n_blocks, const uint32_t *ray_ids, uint32_t n_rays,
                   uint32_t *hit, uint32_t *visits, int threads);

/* Exact single-threaded emulation of MainWindow.RunWavefront (MainWindow.cs:356-426) with
 * order-preserving queue appends.  Queue buffers must hold W*H entries.  Also fills `out`. */
int orc_wavefront_frame(const orc_trace_input *in, const uint16_t *bases, uint32_t n_bases,
                        const orc_block *blocks, uint32_t n_blocks, orc_usage *usages, uint32_t n_usages,
                        orc_child_request *requests, uint32_t request_capacity, uint32_t *request_count,
                        uint32_t *background_queue, uint32_t *background_count,
                        uint32_t *material_queue, uint32_t *material_count,
                        uint32_t *iterations, orc_frame_out *out);

/* ---- differentiable compositing (no reference counterpart) ----
 * A cone does not end at a material hit: the hit block (the block resolveRemainders would
 * read) becomes a layer and the cone continues past the cell, up to max_layers layers or
 * until it leaves the volume.  layers: max_layers u32 per cone (UINT32_MAX after the last);
 * bg: the sky/sun colour of the cone's direction (3 floats in [0,1]). */
int orc_composite_layers(const orc_trace_input *ins, const float *apertures, uint32_t n_views, const uint16_t *bases,
                         uint32_t n_bases, const orc_block *blocks, uint32_t n_blocks, uint32_t max_layers,
                         uint32_t *layers, float *bg, int threads);
double orc_composite_loss(const uint32_t *layers, const float *bg, const uint32_t *rays, uint32_t n_rays,
                          uint32_t max_layers, const double *params, uint32_t n_blocks, const double *target,
                          double *image, double *grads);

#ifdef __cplusplus
}
#endif
#endif
