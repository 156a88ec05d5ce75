// Internal declarations shared by the host sources of libsvo.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "svo.h"

static_assert(sizeof(svo_block) == 16, "Block is 16 bytes (wavefront.cl:1-12)");
static_assert(offsetof(svo_block, chunk) == 4 && offsetof(svo_block, normal_pitch) == 6 &&
                  offsetof(svo_block, normal_yaw) == 8 && offsetof(svo_block, colour_r) == 10 &&
                  offsetof(svo_block, opacity) == 13 && offsetof(svo_block, gloss) == 15,
              "Block field offsets");
static_assert(sizeof(svo_usage) == 8 && offsetof(svo_usage, parent) == 4, "Usage is 8 bytes (wavefront.cl:14-17)");
static_assert(sizeof(svo_parent) == 8, "Parent is 8 bytes");
static_assert(sizeof(svo_pruning) == 20 && offsetof(svo_pruning, address) == 4 && offsetof(svo_pruning, chunk) == 8 &&
                  offsetof(svo_pruning, colour_address) == 12 && offsetof(svo_pruning, child_address) == 16,
              "Pruning natural layout (wavefront.cl:24-36)");
static_assert(sizeof(svo_grafting) == 16 && offsetof(svo_grafting, depth) == 8 &&
                  offsetof(svo_grafting, graft_address) == 12,
              "Grafting natural layout (wavefront.cl:38-43)");
static_assert(sizeof(svo_child_request) == 40 && offsetof(svo_child_request, location) == 8 &&
                  offsetof(svo_child_request, tree_size) == 32,
              "ChildRequest natural layout (wavefront.cl:45-51)");
static_assert(sizeof(svo_trace_input) == 104 && offsetof(svo_trace_input, facing) == 12 &&
                  offsetof(svo_trace_input, fov) == 48 && offsetof(svo_trace_input, dof) == 56 &&
                  offsetof(svo_trace_input, screen_size) == 64 && offsetof(svo_trace_input, max_opacity) == 72 &&
                  offsetof(svo_trace_input, base_depth) == 73 && offsetof(svo_trace_input, tick) == 74 &&
                  offsetof(svo_trace_input, max_child_request_id) == 76 &&
                  offsetof(svo_trace_input, fov_multiplier) == 80 &&
                  offsetof(svo_trace_input, weighting_constant) == 92 &&
                  offsetof(svo_trace_input, ambient_light_level) == 96,
              "TraceInput natural layout (wavefront.cl:53-75)");
static_assert(sizeof(svo_update_input_data) == 20 && offsetof(svo_update_input_data, tick) == 2 &&
                  offsetof(svo_update_input_data, max_child_request_id) == 4 &&
                  offsetof(svo_update_input_data, memory_size) == 8 && offsetof(svo_update_input_data, offset) == 12 &&
                  offsetof(svo_update_input_data, graft_size) == 16,
              "UpdateInputData natural layout (wavefront.cl:77-84)");

namespace svo {

// Sets the thread-local error text and returns `code`.
int fail(int code, const std::string &msg);

inline uint32_t pow_sum(uint32_t depth) {
  uint32_t s = 0, p = 1;
  for (uint32_t i = 1; i <= depth; i++) {
    p *= 8u;
    s += p;
  }
  return s;
}

// Morton interleave of the low 10 bits of x,y,z: x bit i -> 3i, y -> 3i+1, z -> 3i+2
// (TreeBuilder.Interleave, TreeBuilder.cs:316-326).
inline uint32_t interleave3(uint32_t x, uint32_t y, uint32_t z) {
  uint32_t out = 0;
  for (int i = 0; i < 10; i++) {
    out |= ((x >> i) & 1u) << (3 * i);
    out |= ((y >> i) & 1u) << (3 * i + 1);
    out |= ((z >> i) & 1u) << (3 * i + 2);
  }
  return out;
}

struct Tree {
  uint32_t base_depth = 0;
  uint32_t max_depth = 0;
  uint32_t block_count = 0;
  std::vector<uint16_t> bases;
  std::vector<svo_block> blocks;          // trimmed to block_count
  std::vector<uint32_t> group_parent;     // parent block address per allocated group, UINT32_MAX otherwise
};

void camera_lut(const svo_trace_input &in, float *cos_u, float *sin_u, float *cos_v, float *sin_v);
// Entry k of the camera table laid out as cos_u[W], sin_u[W], cos_v[H], sin_v[H].
float camera_lut_at(const svo_trace_input &in, size_t k);

}  // namespace svo

struct svo_tree {
  svo::Tree t;
};
