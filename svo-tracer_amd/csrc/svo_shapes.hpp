// Sphere, box and sphere-plus-noise geometries shared by the host scene builder and the
// device TreeBuilder, so that both classify every cell and make every block identically.
// All arithmetic is IEEE basic operations and correctly rounded square roots (no
// contraction, no library transcendentals).
#pragma once

#include <math.h>
#include <stdint.h>

#include "svo_terrain.hpp"

namespace svo {
namespace shapes {

// Float bounds of a cell, as BoundingVolume / Location.CoordinateRange compute them
// (exact for levels <= 24).
struct Box {
  float min[3], max[3];
};

SVO_HD inline Box cell_box(uint32_t level, uint32_t x, uint32_t y, uint32_t z) {
  Box b;
  const float scale = ldexpf(1.0f, -(int)level);
  const uint32_t v[3] = {x, y, z};
  for (int a = 0; a < 3; a++) {
    b.min[a] = (float)v[a] * scale;
    b.max[a] = (float)(v[a] + 1) * scale;
  }
  return b;
}

// ---- SphereDefinition (SphereDefinition.cs), combined by TreeBuilder's rules, in float.
// Sphere i: centre s[4i..4i+2], radius s[4i+3].

SVO_HD inline bool sphere_within(const float *q, const Box &b) {
  for (int a = 0; a < 3; a++)
    if (!(b.min[a] < q[a] + q[3] && b.max[a] > q[a] - q[3])) return false;
  return true;
}

SVO_HD inline bool sphere_geo(const float *q, const Box &b) {
  float d2 = q[3] * q[3];
  for (int a = 0; a < 3; a++) {
    if (q[a] < b.min[a]) {
      const float t = q[a] - b.min[a];
      d2 -= t * t;
    } else if (q[a] > b.max[a]) {
      const float t = q[a] - b.max[a];
      d2 -= t * t;
    }
  }
  return d2 > 0;
}

SVO_HD inline bool sphere_air(const float *q, const Box &b) {
  for (int k = 0; k < 8; k++) {
    const float x = ((k & 1) ? b.max[0] : b.min[0]) - q[0];
    const float y = ((k & 2) ? b.max[1] : b.min[1]) - q[1];
    const float z = ((k & 4) ? b.max[2] : b.min[2]) - q[2];
    if (sqrtf((x * x + y * y) + z * z) > q[3]) return true;
  }
  return false;
}

SVO_HD inline void spheres_classify(const float *s, uint32_t n, const Box &b, bool *air, bool *geo) {
  bool any_air = true, any_geo = false;
  for (uint32_t i = 0; i < n; i++) {
    const float *q = s + 4 * i;
    if (!sphere_within(q, b)) continue;
    if (!sphere_air(q, b)) any_air = false;
    if (sphere_geo(q, b)) any_geo = true;
  }
  *air = any_air;
  *geo = any_geo;
}

SVO_HD inline void spheres_attributes(const float *s, uint32_t n, const Box &b, int16_t *pitch, int16_t *yaw,
                                      uint8_t rgb[3]) {
  *pitch = 0;
  *yaw = 0;
  rgb[0] = rgb[1] = rgb[2] = 0;
  bool coloured = false;
  for (uint32_t i = 0; i < n; i++) {
    const float *q = s + 4 * i;
    if (!sphere_within(q, b) || !sphere_geo(q, b)) continue;
    const float cx = ((b.max[0] - b.min[0]) / 2) + b.min[0];
    const float cy = ((b.max[1] - b.min[1]) / 2) + b.min[1];
    const float cz = ((b.max[2] - b.min[2]) / 2) + b.min[2];
    terrain::normal_angle(cx - q[0], cy - q[1], cz - q[2], pitch, yaw);
    if (!coloured) {
      rgb[0] = 255;  // SphereDefinition.Colour
      rgb[1] = 0;
      rgb[2] = 0;
      coloured = true;
    }
  }
}

// ---- axis-aligned boxes (the CubeDefinition role): geometry where the cell overlaps a box,
// air unless the cell lies inside one.  Box i: min c[6i..6i+2], max c[6i+3..6i+5].

SVO_HD inline void cubes_classify(const float *c, uint32_t n, const Box &v, bool *air, bool *geo) {
  bool any_geo = false, inside = false;
  for (uint32_t i = 0; i < n; i++) {
    const float *b = c + 6 * i;
    bool ov = true, in = true;
    for (int k = 0; k < 3; k++) {
      if (!(v.min[k] < b[3 + k] && v.max[k] > b[k])) ov = false;
      if (!(v.min[k] >= b[k] && v.max[k] <= b[3 + k])) in = false;
    }
    any_geo |= ov;
    inside |= in;
  }
  *geo = any_geo;
  *air = !inside;
}

// ---- unit sphere with value-noise fBm displacement in world space [-s, s]^3:
// d(p) = |p| - (1 + A fbm(p)); a cell is classified conservatively from the SDF at its
// centre and the Lipschitz bound `lip` of d.
struct Noise {
  uint64_t seed;
  uint32_t octaves;
  double lac, gain, amp, freq, s, lip;
};

SVO_HD inline double value_noise3(uint64_t seed, double x, double y, double z) {
  const double fx = floor(x), fy = floor(y), fz = floor(z);
  const int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
  const double tx = terrain::smooth(x - fx), ty = terrain::smooth(y - fy), tz = terrain::smooth(z - fz);
  double r = 0;
  for (int k = 0; k < 2; k++) {
    const double c00 = terrain::lattice(seed, ix, iy, iz + k), c01 = terrain::lattice(seed, ix + 1, iy, iz + k);
    const double c10 = terrain::lattice(seed, ix, iy + 1, iz + k), c11 = terrain::lattice(seed, ix + 1, iy + 1, iz + k);
    const double wy0 = c00 + (c01 - c00) * tx;
    const double wy1 = c10 + (c11 - c10) * tx;
    const double wz = wy0 + (wy1 - wy0) * ty;
    r += (k ? tz : 1.0 - tz) * wz;
  }
  return r;
}

SVO_HD inline double noise_fbm(const Noise &n, double x, double y, double z) {
  double sum = 0, norm = 0, a = 1, f = n.freq;
  for (uint32_t o = 0; o < n.octaves; o++) {
    sum += a * value_noise3(n.seed + o, x * f, y * f, z * f);
    norm += a;
    a *= n.gain;
    f *= n.lac;
  }
  return sum / norm;
}

SVO_HD inline double noise_sdf(const Noise &n, double x, double y, double z) {
  return sqrt(x * x + y * y + z * z) - (1.0 + n.amp * noise_fbm(n, x, y, z));
}

SVO_HD inline void noise_centre(const Noise &n, uint32_t level, uint32_t x, uint32_t y, uint32_t z, double p[3],
                                double *half) {
  const double size = 2.0 * n.s * ldexp(1.0, -(int)level);
  const uint32_t v[3] = {x, y, z};
  for (int a = 0; a < 3; a++) p[a] = ((double)v[a] + 0.5) * size - n.s;
  *half = 0.5 * size;
}

SVO_HD inline void noise_classify(const Noise &n, uint32_t level, uint32_t x, uint32_t y, uint32_t z, bool *air,
                                  bool *geo) {
  double p[3], h;
  noise_centre(n, level, x, y, z, p, &h);
  const double d = noise_sdf(n, p[0], p[1], p[2]);
  const double bound = n.lip * h * 1.7320508075688772;  // sqrt(3), correctly rounded
  *air = d > -bound;
  *geo = d < bound;
}

SVO_HD inline uint8_t channel(uint64_t h, int shift, int base) { return (uint8_t)(base + (int)((h >> shift) & 63u)); }

SVO_HD inline void noise_attributes(const Noise &n, uint32_t level, uint32_t x, uint32_t y, uint32_t z,
                                    int16_t *pitch, int16_t *yaw, uint8_t rgb[3]) {
  double p[3], h;
  noise_centre(n, level, x, y, z, p, &h);
  const double gx = noise_sdf(n, p[0] + h, p[1], p[2]) - noise_sdf(n, p[0] - h, p[1], p[2]);
  const double gy = noise_sdf(n, p[0], p[1] + h, p[2]) - noise_sdf(n, p[0], p[1] - h, p[2]);
  const double gz = noise_sdf(n, p[0], p[1], p[2] + h) - noise_sdf(n, p[0], p[1], p[2] - h);
  terrain::normal_angle(gx, gy, gz, pitch, yaw);
  const uint64_t hs =
      terrain::mix64(n.seed ^ ((uint64_t)level << 58) ^ ((uint64_t)x << 38) ^ ((uint64_t)y << 19) ^ z);
  rgb[0] = channel(hs, 0, 150);
  rgb[1] = channel(hs, 8, 90);
  rgb[2] = channel(hs, 16, 60);
}

}  // namespace shapes
}  // namespace svo
