// Device-side fixed-point and traversal helpers for gfx950.
//
// All floating-point code here is compiled with -ffp-contract=off and correctly
// rounded division/sqrt, so results are bit-identical to the defined semantics of
// the path (see the oracle's header for the list).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace svo {
namespace dev {

constexpr uint64_t kU64Max = ~0ull;
constexpr uint32_t kU32Max = 0xffffffffu;
constexpr uint32_t kStepCap = 1u << 20;     // defined termination bound per cone
constexpr uint32_t kMaxDepth = 60u;         // deepest level a cone may descend to
constexpr uint64_t kFnvOffset = 0xcbf29ce484222325ull;
constexpr uint64_t kFnvPrime = 0x100000001b3ull;

// floatToUlong, wavefront.cl:200-207 (saturating, NaN -> 0 where the reference is undefined).
// Branch-free: the conversion always runs on an in-range value (0 for NaN and overflow) and
// the saturated cases are selected afterwards, so lanes never split on it.
__device__ __forceinline__ uint64_t float_to_ulong(float x) {
  const float a = fabsf(x) * 18446744073709551616.0f;
  const bool in_range = a < 18446744073709551616.0f;  // false for NaN and >= 2^64
  const bool saturate = (x >= 0.999999911f) | (a >= 18446744073709551616.0f);
  const uint64_t v = (uint64_t)(in_range ? a : 0.0f);
  return saturate ? kU64Max : v;
}

// (uchar)f of the reference's float-to-uchar conversions: truncation, saturated to [0, 255]
// (NaN -> 0) where the reference's out-of-range conversion is undefined.
__device__ __forceinline__ uint32_t to_uchar(float f) {
  if (!(f > 0.0f)) return 0u;
  if (f >= 255.0f) return 255u;
  return (uint32_t)f;
}

// ulongToFloat, wavefront.cl:209: exact scaling by 2^-64.
__device__ __forceinline__ float ulong_to_float(uint64_t x) { return (float)x * 5.42101086242752217e-20f; }

// roundUlong, wavefront.cl:211-221, rearranged: with low = the bits below the level-(depth+1)
// cell, rounding up is (cell size - low) = (~v & lowmask) + 1, rounding down is low + 1,
// except in the first cell where the distance is v itself.
__device__ __forceinline__ uint64_t round_ulong(uint64_t v, uint32_t depth, bool up) {
  const uint64_t lowmask = kU64Max >> (depth + 1);
  if (!up && v <= lowmask) return v;
  return ((up ? ~v : v) & lowmask) + 1;
}

// chunkPosition, wavefront.cl:428-430.
__device__ __forceinline__ uint32_t chunk_position(uint32_t depth, uint64_t x, uint64_t y, uint64_t z) {
  const uint32_t s = 63u - depth;
  return (uint32_t)((x >> s) & 1u) | ((uint32_t)((y >> s) & 1u) << 1) | ((uint32_t)((z >> s) & 1u) << 2);
}

__device__ __forceinline__ uint32_t spread3(uint32_t v) {  // 10 bits -> every third bit
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// baseLocation, wavefront.cl:405-411: Morton code of the top `depth` bits (depth 1..10).
__device__ __forceinline__ uint32_t morton_top(uint32_t depth, uint64_t x, uint64_t y, uint64_t z) {
  const uint32_t s = 64u - depth;
  return spread3((uint32_t)(x >> s)) | (spread3((uint32_t)(y >> s)) << 1) | (spread3((uint32_t)(z >> s)) << 2);
}

__device__ __forceinline__ uint32_t spread3_7(uint32_t v) {  // 7 bits -> every third bit
  v = (v | (v << 8)) & 0x0000f00fu;
  v = (v | (v << 4)) & 0x000c30c3u;
  v = (v | (v << 2)) & 0x00249249u;
  return v;
}

// Morton code of the top `depth` bits for base levels (depth 1..7).
__device__ __forceinline__ uint32_t morton_top_base(uint32_t depth, uint64_t x, uint64_t y, uint64_t z) {
  const uint32_t s = 64u - depth;
  return spread3_7((uint32_t)(x >> s)) | (spread3_7((uint32_t)(y >> s)) << 1) | (spread3_7((uint32_t)(z >> s)) << 2);
}

// Offset of level d in Bases: powSum(d-1) = 8 + 64 + ... + 8^(d-1) = (8^d - 1)/7 - 1 (1 <= d <= 10).
__device__ __forceinline__ uint32_t base_offset(uint32_t d) { return ((1u << (3u * d)) - 1u) / 7u - 1u; }

// comparePositions, wavefront.cl:413-426; signs: bit a set when direction axis a >= 0.
// Level L counts as crossed when, on some axis, bit 63-L held the far value of the
// direction before the step and lost it; the reference pops one level per crossed
// level, starting at `depth`, and stops at level 1.  Stated in closed form: count the
// consecutive crossed levels from `depth` upwards.  Levels up to 31 live in the high
// words of the positions.
__device__ __forceinline__ uint32_t crossed_hi(uint64_t p, uint64_t n, bool pos) {
  const uint32_t ph = (uint32_t)(p >> 32), nh = (uint32_t)(n >> 32);
  return (ph ^ nh) & (pos ? ph : ~ph);
}
__device__ __forceinline__ uint64_t crossed_64(uint64_t p, uint64_t n, bool pos) { return (p ^ n) & (pos ? p : ~p); }

__device__ __forceinline__ uint32_t compare_positions(uint32_t depth, uint64_t px, uint64_t py, uint64_t pz,
                                                      uint64_t x, uint64_t y, uint64_t z, uint32_t signs) {
  const bool sx = signs & 1u, sy = signs & 2u, sz = signs & 4u;
  uint32_t k;
  if (depth <= 31u) {
    const uint32_t c = crossed_hi(px, x, sx) | crossed_hi(py, y, sy) | crossed_hi(pz, z, sz);
    const uint32_t t = ~(c >> (31u - depth));  // bit j: level depth-j not crossed
    k = t ? (uint32_t)__builtin_ctz(t) : 32u;
  } else {
    const uint64_t c = crossed_64(px, x, sx) | crossed_64(py, y, sy) | crossed_64(pz, z, sz);
    k = (uint32_t)__builtin_ctzll(~(c >> (63u - depth)));
  }
  return depth > k ? depth - k : 1u;
}

// leaving, wavefront.cl:433-437: a position sits on the face it is moving towards.
__device__ __forceinline__ bool leaving(uint64_t x, uint64_t y, uint64_t z, uint32_t signs) {
  return x == ((signs & 1u) ? kU64Max : 0ull) || y == ((signs & 2u) ? kU64Max : 0ull) ||
         z == ((signs & 4u) ? kU64Max : 0ull);
}

// One axis of traverseChunk's move: advance by n in the sign direction; landing exactly
// on the wrapped-around face (0 moving up, ULONG_MAX moving down) means the cube was
// left, so the position is pinned to the face it moved towards.
__device__ __forceinline__ uint64_t advance_axis(uint64_t v, uint64_t n, bool pos) {
  const uint64_t back = pos ? 0ull : kU64Max;  // the face opposite the motion
  v = pos ? v + n : v - n;
  return v == back ? ~back : v;
}

// traverseChunk, wavefront.cl:440-497 (the trailing setConeDepth is evaluated lazily
// where the cone is consumed).  Branch-free across the dominant axis m; ulongToFloat(d_m)
// reuses the conversion already made for the comparison (both are (float)d_m scaled by
// 2^-64), and only the two minor axes are converted back.
__device__ __forceinline__ void traverse_chunk(uint32_t depth, uint64_t &x, uint64_t &y, uint64_t &z, float dx,
                                               float dy, float dz, float ix, float iy, float iz, uint32_t signs) {
  const bool sx = signs & 1u, sy = signs & 2u, sz = signs & 4u;
  const uint64_t d0 = round_ulong(x, depth, sx), d1 = round_ulong(y, depth, sy), d2 = round_ulong(z, depth, sz);
  const float f0 = (float)d0, f1 = (float)d1, f2 = (float)d2;
  const float a0 = fabsf(f0 * ix), a1 = fabsf(f1 * iy), a2 = fabsf(f2 * iz);
  const int m = (a0 <= a1 && a0 <= a2) ? 0 : ((a1 <= a0 && a1 <= a2) ? 1 : 2);
  const float inv_m = m == 0 ? ix : (m == 1 ? iy : iz);
  const float ud = (m == 0 ? f0 : (m == 1 ? f1 : f2)) * 5.42101086242752217e-20f;
  const uint64_t dm = m == 0 ? d0 : (m == 1 ? d1 : d2);
  // minor axes: first = (m == 0 ? y : x), second = (m == 2 ? y : z)
  const uint64_t na = float_to_ulong(((m == 0 ? dy : dx) * inv_m) * ud);
  const uint64_t nb = float_to_ulong(((m == 2 ? dy : dz) * inv_m) * ud);
  const uint64_t n0 = m == 0 ? dm : na;
  const uint64_t n1 = m == 1 ? dm : (m == 0 ? na : nb);
  const uint64_t n2 = m == 2 ? dm : nb;
  x = advance_axis(x, n0, sx);
  y = advance_axis(y, n1, sy);
  z = advance_axis(z, n2, sz);
}

// ---- sign-canonical positions (the fused trace kernel) ----
// A cone's position is kept per axis as c = x for an axis it moves along towards 0 and
// c = ~x (= ULONG_MAX - x) for an axis it moves along towards ULONG_MAX, so that on every
// axis the cone moves towards c = 0.  The map is an involution (x = c ^ mask) and turns
// each sign-dependent rule of the step into one rule:
//   roundUlong up from x = (~x & lowmask) + 1 = (c & lowmask) + 1; roundUlong down from x
//   is the same except in the first cell (x <= lowmask), where it is x itself;
//   x + n (up) is ~(c - n) and x - n (down) is c - n, and landing on the wrapped-around
//   face (x = 0 up, x = ULONG_MAX down) is c - n = ULONG_MAX in both cases;
//   leaving (x on the face it moves towards) is c = 0;
//   comparePositions' "bit held the far value and lost it" is (pc ^ c) & ~pc;
//   chunkPosition's bit of an up axis is the complement of c's bit.
__device__ __forceinline__ uint64_t sign_mask64(uint32_t signs, uint32_t axis) {
  return (uint64_t)0 - (uint64_t)((signs >> axis) & 1u);
}
__device__ __forceinline__ uint32_t sign_mask32(uint32_t signs, uint32_t axis) { return 0u - ((signs >> axis) & 1u); }

__device__ __forceinline__ uint64_t round_canon(uint64_t c, uint32_t depth, bool pos) {
  const uint64_t lowmask = kU64Max >> (depth + 1);
  return (!pos && c <= lowmask) ? c : (c & lowmask) + 1;
}

__device__ __forceinline__ uint64_t advance_canon(uint64_t c, uint64_t n) {
  const uint64_t v = c - n;
  return v == kU64Max ? 0ull : v;
}

// traverse_chunk on sign-canonical positions (same floating-point operations in the same
// order; only the integer bookkeeping of the signs is gone).
__device__ __forceinline__ void traverse_canon(uint32_t depth, uint64_t &cx, uint64_t &cy, uint64_t &cz, float dx,
                                               float dy, float dz, float ix, float iy, float iz, uint32_t signs) {
  const uint64_t d0 = round_canon(cx, depth, signs & 1u), d1 = round_canon(cy, depth, signs & 2u),
                 d2 = round_canon(cz, depth, signs & 4u);
  const float f0 = (float)d0, f1 = (float)d1, f2 = (float)d2;
  const float a0 = fabsf(f0 * ix), a1 = fabsf(f1 * iy), a2 = fabsf(f2 * iz);
  const int m = (a0 <= a1 && a0 <= a2) ? 0 : ((a1 <= a0 && a1 <= a2) ? 1 : 2);
  const float inv_m = m == 0 ? ix : (m == 1 ? iy : iz);
  const float ud = (m == 0 ? f0 : (m == 1 ? f1 : f2)) * 5.42101086242752217e-20f;
  const uint64_t dm = m == 0 ? d0 : (m == 1 ? d1 : d2);
  const uint64_t na = float_to_ulong(((m == 0 ? dy : dx) * inv_m) * ud);
  const uint64_t nb = float_to_ulong(((m == 2 ? dy : dz) * inv_m) * ud);
  const uint64_t n0 = m == 0 ? dm : na;
  const uint64_t n1 = m == 1 ? dm : (m == 0 ? na : nb);
  const uint64_t n2 = m == 2 ? dm : nb;
  cx = advance_canon(cx, n0);
  cy = advance_canon(cy, n1);
  cz = advance_canon(cz, n2);
}

__device__ __forceinline__ bool leaving_canon(uint64_t cx, uint64_t cy, uint64_t cz) {
  return cx == 0ull || cy == 0ull || cz == 0ull;
}

// compare_positions on sign-canonical positions.
__device__ __forceinline__ uint32_t compare_canon(uint32_t depth, uint64_t px, uint64_t py, uint64_t pz, uint64_t x,
                                                  uint64_t y, uint64_t z) {
  uint32_t k;
  if (depth <= 31u) {
    const uint32_t pxh = (uint32_t)(px >> 32), pyh = (uint32_t)(py >> 32), pzh = (uint32_t)(pz >> 32);
    const uint32_t c = ((pxh ^ (uint32_t)(x >> 32)) & ~pxh) | ((pyh ^ (uint32_t)(y >> 32)) & ~pyh) |
                       ((pzh ^ (uint32_t)(z >> 32)) & ~pzh);
    const uint32_t t = ~(c >> (31u - depth));
    k = t ? (uint32_t)__builtin_ctz(t) : 32u;
  } else {
    const uint64_t c = ((px ^ x) & ~px) | ((py ^ y) & ~py) | ((pz ^ z) & ~pz);
    k = (uint32_t)__builtin_ctzll(~(c >> (63u - depth)));
  }
  return depth > k ? depth - k : 1u;
}

// Morton code of the top `depth` bits (depth <= 10) from the high words of the positions.
__device__ __forceinline__ uint32_t morton_top_hi(uint32_t depth, uint32_t xh, uint32_t yh, uint32_t zh) {
  const uint32_t s = 32u - depth;
  return spread3(xh >> s) | (spread3(yh >> s) << 1) | (spread3(zh >> s) << 2);
}
__device__ __forceinline__ uint32_t morton_top_base_hi(uint32_t depth, uint32_t xh, uint32_t yh, uint32_t zh) {
  const uint32_t s = 32u - depth;
  return spread3_7(xh >> s) | (spread3_7(yh >> s) << 1) | (spread3_7(zh >> s) << 2);
}

// coneSize(fast_length(Origin - loc)), wavefront.cl:387-399, returning the cone itself.
__device__ __forceinline__ float cone_at(float ox, float oy, float oz, uint64_t x, uint64_t y, uint64_t z, float fov,
                                         float dof0, float dof1) {
  const float mx = ox - ulong_to_float(x), my = oy - ulong_to_float(y), mz = oz - ulong_to_float(z);
  const float m = fabsf(sqrtf((mx * mx + my * my) + mz * mz));
  const float eye = fov * m;
  const float f = fabsf(dof0 * (dof1 - m));
  return eye < f ? f : eye;
}

// 2^-k (0 <= k <= 126) exactly.
__device__ __forceinline__ float exp2_neg(uint32_t k) { return __uint_as_float((127u - k) << 23); }

// traceIntoVolume, wavefront.cl:639-1088 (see the oracle for the per-axis statement).
__device__ __forceinline__ bool trace_into_volume(const float o[3], const float d[3], const float inv[3], float fov,
                                                  uint64_t &x, uint64_t &y, uint64_t &z) {
  bool lo[3], hi[3];
  bool out = false;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    lo[a] = o[a] < 0.0f;
    hi[a] = o[a] > 1.0f;
    out |= lo[a] | hi[a];
  }
  if (!out) {
    x = float_to_ulong(o[0]);
    y = float_to_ulong(o[1]);
    z = float_to_ulong(o[2]);
    return true;
  }
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if ((lo[a] | hi[a]) && d[a] == 0.0f) return false;
    const bool pos = d[a] >= 0.0f;
    if (lo[a] && !pos) return false;
    if (hi[a] && pos) return false;
  }
  const bool c0 = lo[0] | hi[0], c1 = lo[1] | hi[1], c2 = lo[2] | hi[2];
  const float m0 = c0 ? fabsf(((lo[0] ? 0.0f : 1.0f) - o[0]) * inv[0]) : 0.0f;
  const float m1 = c1 ? fabsf(((lo[1] ? 0.0f : 1.0f) - o[1]) * inv[1]) : 0.0f;
  const float m2 = c2 ? fabsf(((lo[2] ? 0.0f : 1.0f) - o[2]) * inv[2]) : 0.0f;
  // entry plane: first candidate (x,y,z order, the last candidate excluded) whose slab
  // distance is >= every other candidate's, else the last candidate
  const int last = c2 ? 2 : (c1 ? 1 : 0);
  const bool ok0 = (!c1 || m0 >= m1) && (!c2 || m0 >= m2);
  const bool ok1 = (!c0 || m1 >= m0) && (!c2 || m1 >= m2);
  const int pick = (c0 && last != 0 && ok0) ? 0 : ((c1 && last != 1 && ok1) ? 1 : last);
  const float m = pick == 0 ? m0 : (pick == 1 ? m1 : m2);
  float lf[3];
#pragma unroll
  for (int a = 0; a < 3; a++) lf[a] = (a == pick) ? (lo[a] ? 0.0f : 1.0f) : o[a] + d[a] * m;
  x = float_to_ulong(lf[0]);
  y = float_to_ulong(lf[1]);
  z = float_to_ulong(lf[2]);
  const float eye = fov * m;
  const float f = fabsf(0.0f * (0.0f - m));
  const float c = eye < f ? f : eye;
#pragma unroll
  for (int a = 0; a < 3; a++)
    if (lf[a] < -c || lf[a] > 1.0f + c) return false;
  return true;
}

// A 16-byte Block record fetched with one 128-bit load; out-of-range addresses read
// as an all-zero record (defined; the reference reads past its buffer).
__device__ __forceinline__ uint4 load_block(const uint4 *__restrict__ blocks, uint32_t n, uint32_t a) {
  return a < n ? blocks[a] : make_uint4(0u, 0u, 0u, 0u);
}
// The first 8 bytes of a Block record (Child @0, Chunk @4, NormalPitch @6): what a
// traversal step reads of a block.
__device__ __forceinline__ uint2 load_node(const uint4 *__restrict__ blocks, uint32_t n, uint32_t a) {
  return a < n ? reinterpret_cast<const uint2 *>(blocks)[2u * (size_t)a] : make_uint2(0u, 0u);
}
__device__ __forceinline__ uint32_t block_child(uint4 q) { return q.x; }
__device__ __forceinline__ uint32_t block_chunk(uint4 q) { return q.y & 0xffffu; }
__device__ __forceinline__ uint32_t block_r(uint4 q) { return (q.z >> 16) & 0xffu; }
__device__ __forceinline__ uint32_t block_g(uint4 q) { return q.z >> 24; }
__device__ __forceinline__ uint32_t block_b(uint4 q) { return q.w & 0xffu; }
__device__ __forceinline__ int16_t block_pitch(uint4 q) { return (int16_t)(q.y >> 16); }
__device__ __forceinline__ int16_t block_yaw(uint4 q) { return (int16_t)(q.z & 0xffffu); }

}  // namespace dev
}  // namespace svo
