// fBm heightfield terrain shared by the host scene builder and the device brick generator,
// so that a block generated on demand equals the block the builder would have made for the
// same cell.  All arithmetic is IEEE double/float basic operations (no contraction, no
// library transcendentals), which makes host and device results identical.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define SVO_HD __host__ __device__
#else
#define SVO_HD
#endif

namespace svo {
namespace terrain {

struct Spec {
  uint64_t seed;
  uint32_t octaves;
  uint32_t lres;  // samples per side = 2^lres
  double lacunarity, gain, frequency;
  double z_min, z_max;
};

SVO_HD inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Lattice value in [-1, 1) from integer coordinates and a seed (SplitMix64 finaliser).
SVO_HD inline double lattice(uint64_t seed, int64_t x, int64_t y, int64_t z) {
  uint64_t h = mix64(seed + 0x9e3779b97f4a7c15ull);
  h = mix64(h ^ (uint64_t)x * 0xd1b54a32d192ed03ull);
  h = mix64(h ^ (uint64_t)y * 0xabc98388fb8fac03ull);
  h = mix64(h ^ (uint64_t)z * 0x8cb92ba72f3d8dd7ull);
  return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

SVO_HD inline double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

SVO_HD inline double value_noise2(uint64_t seed, double x, double y) {
  const double fx = floor(x), fy = floor(y);
  const int64_t ix = (int64_t)fx, iy = (int64_t)fy;
  const double tx = smooth(x - fx), ty = smooth(y - fy);
  const double a = lattice(seed, ix, iy, 0), b = lattice(seed, ix + 1, iy, 0);
  const double c = lattice(seed, ix, iy + 1, 0), d = lattice(seed, ix + 1, iy + 1, 0);
  const double ab = a + (b - a) * tx, cd = c + (d - c) * tx;
  return ab + (cd - ab) * ty;
}

// Height of grid sample (i, j): octaves of value noise, normalised, mapped to [z_min, z_max].
SVO_HD inline float height(const Spec &s, uint32_t i, uint32_t j) {
  const double res = (double)(1ull << s.lres);
  const double x = ((double)i + 0.5) / res, y = ((double)j + 0.5) / res;
  double sum = 0.0, norm = 0.0, a = 1.0, f = s.frequency;
  for (uint32_t o = 0; o < s.octaves; o++) {
    sum += a * value_noise2(s.seed + o, x * f, y * f);
    norm += a;
    a *= s.gain;
    f *= s.lacunarity;
  }
  double v = 0.5 + 0.5 * (sum / norm) * 1.6;
  v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
  return (float)(s.z_min + (s.z_max - s.z_min) * v);
}

// Cell classification from the height range [lo, hi] of its footprint: geometry iff the
// cell's bottom is below the highest column, air iff its top is above the lowest.
SVO_HD inline void classify(uint32_t level, uint32_t z, float lo, float hi, bool *air, bool *geo) {
  const double scale = 1.0 / (double)(1ull << level);
  const double z0 = (double)z * scale, z1 = ((double)z + 1.0) * scale;
  *geo = z0 < (double)hi;
  *air = z1 > (double)lo;
}

// atan(x) by argument halving, atan(x) = 2 atan(x / (1 + sqrt(1 + x^2))), applied until
// |x| <= 0.2, then the Maclaurin series x - x^3/3 + x^5/5 - ... to x^21.
SVO_HD inline double atan_series(double x) {
  double k = 1.0;
  for (int i = 0; i < 4 && (x > 0.2 || x < -0.2); i++) {
    x = x / (1.0 + sqrt(1.0 + x * x));
    k *= 2.0;
  }
  const double x2 = x * x;
  double t = x, sum = 0.0;
  for (int n = 1; n <= 21; n += 2) {
    sum += ((n & 2) ? -t : t) / (double)n;
    t *= x2;
  }
  return k * sum;
}

constexpr double kPi = 3.14159265358979323846;

SVO_HD inline double atan2_d(double y, double x) {
  if (x == 0.0 && y == 0.0) return 0.0;
  if (fabs(y) <= fabs(x)) {
    const double a = atan_series(y / x);
    if (x > 0.0) return a;
    return y >= 0.0 ? a + kPi : a - kPi;
  }
  const double a = atan_series(x / y);
  return y > 0.0 ? kPi / 2.0 - a : -kPi / 2.0 - a;
}

SVO_HD inline double asin_d(double z) {
  z = z < -1.0 ? -1.0 : (z > 1.0 ? 1.0 : z);
  return atan2_d(z, sqrt((1.0 - z) * (1.0 + z)));
}

// NormalAngle (TreeBuilder.cs:328-334) with the decoder's axis convention
// (wavefront.cl:235-245): n = (cosYaw cosPitch, sinYaw cosPitch, sinPitch).
SVO_HD inline void normal_angle(double nx, double ny, double nz, int16_t *pitch, int16_t *yaw) {
  const double len = sqrt(nx * nx + ny * ny + nz * nz);
  if (!(len > 0.0)) {
    *pitch = 0;
    *yaw = 0;
    return;
  }
  nx /= len;
  ny /= len;
  nz /= len;
  *pitch = (int16_t)(32767.0 * (asin_d(nz) / kPi));
  *yaw = (int16_t)(32767.0 * (atan2_d(ny, nx) / kPi));
}

// Normal, colour of the cell at (level, x, y, z) given its footprint maximum `hi` and the
// finest grid (`top`, 2^lres per side).
SVO_HD inline void attributes(const Spec &s, const float *top, uint32_t level, uint32_t x, uint32_t y, uint32_t z,
                              float hi, int16_t *pitch, int16_t *yaw, uint8_t rgb[3]) {
  const int64_t res = (int64_t)1 << s.lres;
  auto sample = [&](int64_t i, int64_t j) -> double {
    i = i < 0 ? 0 : (i > res - 1 ? res - 1 : i);
    j = j < 0 ? 0 : (j > res - 1 ? res - 1 : j);
    return (double)top[j * res + i];
  };
  // sample-grid coordinates of the footprint centre and a half-width of >= 1 sample
  const double scale = (double)res / (double)(1ull << level);
  const double cx = ((double)x + 0.5) * scale, cy = ((double)y + 0.5) * scale;
  const double hw = 0.5 * scale > 1.0 ? 0.5 * scale : 1.0;
  const int64_t x0 = (int64_t)floor(cx - hw), x1 = (int64_t)floor(cx + hw);
  const int64_t y0 = (int64_t)floor(cy - hw), y1 = (int64_t)floor(cy + hw);
  const int64_t ci = (int64_t)floor(cx), cj = (int64_t)floor(cy);
  const double dx = (sample(x1, cj) - sample(x0, cj)) / ((double)(x1 - x0) / (double)res);
  const double dy = (sample(ci, y1) - sample(ci, y0)) / ((double)(y1 - y0) / (double)res);
  normal_angle(-dx, -dy, 1.0, pitch, yaw);
  double t = ((double)hi - 0.25) / 0.3;
  t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
  const uint64_t hs = mix64(s.seed ^ ((uint64_t)level << 58) ^ ((uint64_t)x << 38) ^ ((uint64_t)y << 19) ^ z);
  const int jitter = (int)(hs & 15u);
  const double c0 = 60 + 150 * t + jitter, c1 = 110 + 100 * t + jitter, c2 = 50 + 190 * t * t + jitter;
  rgb[0] = (uint8_t)(c0 < 255.0 ? c0 : 255.0);
  rgb[1] = (uint8_t)(c1 < 255.0 ? c1 : 255.0);
  rgb[2] = (uint8_t)(c2 < 255.0 ? c2 : 255.0);
}

}  // namespace terrain
}  // namespace svo
