// Host-side scene construction: a TreeBuilder-compatible octree builder
// (TreeBuilder.cs:102-256) over pluggable geometry, .oct persistence
// (OctreeSerializer.cs:13-45) and the camera table used by ray generation.
//
// Allocation order follows TreeBuilder exactly: dense Bases for levels 1..B,
// dense blocks at level B+2 only under level-(B+1) octants holding both air and
// geometry, a FIFO of free groups (sorted empty level-(B+1) cells, then groups
// 8^(B+1) .. maxSize/8-1), children assigned depth-first in x-outer/z-inner
// order of the dense level, and the final trimmed BlockCount rule of
// BuildBlocks (TreeBuilder.cs:254).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <thread>

#include "svo_internal.hpp"
#include "svo_shapes.hpp"
#include "svo_terrain.hpp"

namespace svo {
shapes::Noise noise_spec(const svo_geometry &g);

namespace {

struct Cell {
  uint32_t level, x, y, z;
};

// Combined geometry interface (the TreeBuilder.ContainsAir/ContainsGeo/GetNormal/
// GetColour roles, TreeBuilder.cs:30-66).
class Geometry {
 public:
  virtual ~Geometry() {}
  virtual void classify(const Cell &c, bool *air, bool *geo) const = 0;
  virtual void attributes(const Cell &c, int16_t *pitch, int16_t *yaw, uint8_t rgb[3]) const = 0;
};

// SphereDefinition (SphereDefinition.cs) combined by TreeBuilder's rules, in float (the
// shared rules of svo_shapes.hpp, which the device builder uses too).
class Spheres : public Geometry {
 public:
  std::vector<float> s;  // centre and radius per sphere
  explicit Spheres(const float *p, uint32_t n) : s(p, p + 4 * (size_t)n) {}
  void classify(const Cell &c, bool *a, bool *g) const override {
    shapes::spheres_classify(s.data(), (uint32_t)(s.size() / 4), shapes::cell_box(c.level, c.x, c.y, c.z), a, g);
  }
  void attributes(const Cell &c, int16_t *pitch, int16_t *yaw, uint8_t rgb[3]) const override {
    shapes::spheres_attributes(s.data(), (uint32_t)(s.size() / 4), shapes::cell_box(c.level, c.x, c.y, c.z), pitch,
                               yaw, rgb);
  }
};

// Axis-aligned boxes (the CubeDefinition role): geometry where the cell overlaps a box,
// air unless the cell lies inside one.
class Cubes : public Geometry {
 public:
  std::vector<float> b;  // min and max per box
  Cubes(const float *p, uint32_t n) : b(p, p + 6 * (size_t)n) {}
  void classify(const Cell &c, bool *a, bool *g) const override {
    shapes::cubes_classify(b.data(), (uint32_t)(b.size() / 6), shapes::cell_box(c.level, c.x, c.y, c.z), a, g);
  }
  void attributes(const Cell &, int16_t *pitch, int16_t *yaw, uint8_t rgb[3]) const override {
    *pitch = 0;
    *yaw = 0;
    rgb[0] = rgb[1] = rgb[2] = 255;
  }
};

// Unit-sphere SDF with value-noise fBm displacement (svo_shapes.hpp).
class SphereNoise : public Geometry {
 public:
  shapes::Noise n;
  explicit SphereNoise(const svo_geometry &g) : n(noise_spec(g)) {}
  void classify(const Cell &c, bool *a, bool *g) const override { shapes::noise_classify(n, c.level, c.x, c.y, c.z, a, g); }
  void attributes(const Cell &c, int16_t *pitch, int16_t *yaw, uint8_t rgb[3]) const override {
    shapes::noise_attributes(n, c.level, c.x, c.y, c.z, pitch, yaw, rgb);
  }
};

}  // namespace

// Parameters of an SVO_GEOMETRY_SPHERE_NOISE description (defaults applied), with the
// Lipschitz bound of its SDF; the device builder uses the same function.
shapes::Noise noise_spec(const svo_geometry &g) {
  shapes::Noise n;
  n.seed = g.seed;
  n.octaves = g.octaves ? g.octaves : 3;
  n.lac = g.lacunarity > 0 ? g.lacunarity : 2.0;
  n.gain = g.gain > 0 ? g.gain : 0.5;
  n.amp = g.amplitude;
  n.freq = g.frequency > 0 ? g.frequency : 4.0;
  n.s = g.half_extent > 0 ? g.half_extent : 1.25;
  double sum_a = 0, sum_af = 0, a = 1, f = n.freq;
  for (uint32_t o = 0; o < n.octaves; o++) {
    sum_a += a;
    sum_af += a * f;
    a *= n.gain;
    f *= n.lac;
  }
  // |d/dt smoothstep| <= 1.5, lattice range 2, three axes.
  const double noise_lip = 3.0 * std::sqrt(3.0) * sum_af / sum_a;
  n.lip = 1.0 + std::fabs(n.amp) * noise_lip;
  return n;
}

// Terrain parameters of an SVO_GEOMETRY_FBM_HEIGHTFIELD description (defaults applied); the
// device terrain of the streaming path uses the same function.
terrain::Spec terrain_spec(const svo_geometry &g) {
  terrain::Spec s;
  uint32_t res = g.resolution ? g.resolution : 4096, lres = 0;
  while ((1u << lres) < res) lres++;
  s.seed = g.seed;
  s.octaves = g.octaves ? g.octaves : 6;
  s.lres = lres;
  s.lacunarity = g.lacunarity > 0 ? g.lacunarity : 2.0;
  s.gain = g.gain > 0 ? g.gain : 0.5;
  s.frequency = g.frequency > 0 ? g.frequency : 4.0;
  s.z_min = g.z_min;
  s.z_max = g.z_max;
  if (!(s.z_max > s.z_min)) {
    s.z_min = 0.25;
    s.z_max = 0.55;
  }
  return s;
}

namespace {

// fBm heightfield on an R x R grid (R a power of two) covering the unit square;
// each grid cell is a column of constant height in [z_min, z_max].  Classification
// is exact through a min/max pyramid: geometry iff the cell's bottom is below the
// highest column of its footprint, air iff its top is above the lowest column.
class Heightfield : public Geometry {
 public:
  uint32_t res = 0, lres = 0;
  terrain::Spec spec;
  std::vector<std::vector<float>> hmin, hmax;  // per pyramid level L (2^L x 2^L)
  explicit Heightfield(const svo_geometry &g) {
    res = g.resolution ? g.resolution : 4096;
    while ((1u << lres) < res) lres++;
    res = 1u << lres;
    spec = terrain_spec(g);
    hmin.resize(lres + 1);
    hmax.resize(lres + 1);
    std::vector<float> &top = hmin[lres];
    top.resize((size_t)res * res);
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; t++) {
      pool.emplace_back([&, t]() {
        for (uint32_t j = t; j < res; j += nt)
          for (uint32_t i = 0; i < res; i++) top[(size_t)j * res + i] = terrain::height(spec, i, j);
      });
    }
    for (auto &th : pool) th.join();
    hmax[lres] = top;
    for (int L = (int)lres - 1; L >= 0; L--) {
      uint32_t n = 1u << L;
      hmin[L].resize((size_t)n * n);
      hmax[L].resize((size_t)n * n);
      const auto &mn = hmin[L + 1], &mx = hmax[L + 1];
      for (uint32_t j = 0; j < n; j++)
        for (uint32_t i = 0; i < n; i++) {
          size_t a = (size_t)(2 * j) * (2 * n) + 2 * i, b = a + 1, c = a + 2 * n, d = c + 1;
          hmin[L][(size_t)j * n + i] = std::min(std::min(mn[a], mn[b]), std::min(mn[c], mn[d]));
          hmax[L][(size_t)j * n + i] = std::max(std::max(mx[a], mx[b]), std::max(mx[c], mx[d]));
        }
    }
  }
  void footprint(const Cell &c, float *lo, float *hi) const {
    if (c.level <= lres) {
      size_t k = (size_t)c.y * (1u << c.level) + c.x;
      *lo = hmin[c.level][k];
      *hi = hmax[c.level][k];
    } else {
      uint32_t sh = c.level - lres;
      size_t k = (size_t)(c.y >> sh) * res + (c.x >> sh);
      *lo = *hi = hmin[lres][k];
    }
  }
  void classify(const Cell &c, bool *a, bool *g) const override {
    float lo, hi;
    footprint(c, &lo, &hi);
    terrain::classify(c.level, c.z, lo, hi, a, g);
  }
  void attributes(const Cell &c, int16_t *pitch, int16_t *yaw, uint8_t rgb[3]) const override {
    float lo, hi;
    footprint(c, &lo, &hi);
    terrain::attributes(spec, hmin[lres].data(), c.level, c.x, c.y, c.z, hi, pitch, yaw, rgb);
  }
};

// CanHaveChildren, TreeBuilder.cs:291-302.
inline bool can_have_children(uint16_t chunk) {
  for (int k = 0; k < 8; k++)
    if (((chunk >> (2 * k)) & 3u) == 3u) return true;
  return false;
}

// The builder proper (TreeBuilder.BuildTree and helpers).
class Builder {
 public:
  const Geometry &geo;
  Tree &t;
  uint32_t B, max_depth;
  std::vector<uint32_t> empties;
  size_t empty_next = 0;
  uint64_t range_next = 0, range_end = 0;
  bool exhausted = false;

  Builder(const Geometry &g, Tree &tree, uint32_t b, uint32_t md) : geo(g), t(tree), B(b), max_depth(md) {}

  // MakeChunk, TreeBuilder.cs:264-284.
  uint16_t make_chunk(const Cell &c) const {
    uint16_t chunk = 0;
    for (uint32_t k = 0; k < 8; k++) {
      Cell o{c.level + 1, 2 * c.x + (k & 1u), 2 * c.y + ((k >> 1) & 1u), 2 * c.z + ((k >> 2) & 1u)};
      bool a, g;
      geo.classify(o, &a, &g);
      chunk |= (uint16_t)((a ? 1u : 0u) << (2 * k));
      chunk |= (uint16_t)((g ? 1u : 0u) << (2 * k + 1));
    }
    return chunk;
  }

  // MakeBlock, TreeBuilder.cs:74-93 (ColourB takes colour[1] and ColourG colour[2], as there).
  svo_block make_block(const Cell &c) const {
    svo_block b;
    std::memset(&b, 0, sizeof b);
    b.chunk = make_chunk(c);
    b.child = UINT32_MAX;
    uint8_t rgb[3];
    geo.attributes(c, &b.normal_pitch, &b.normal_yaw, rgb);
    b.colour_r = rgb[0];
    b.colour_b = rgb[1];
    b.colour_g = rgb[2];
    b.opacity = 255;
    return b;
  }

  bool any_free() const { return empty_next < empties.size() || range_next < range_end; }
  uint32_t dequeue() {
    if (empty_next < empties.size()) return empties[empty_next++];
    return (uint32_t)range_next++;
  }

  void ensure(size_t n) {
    if (t.blocks.size() < n) {
      if (t.blocks.capacity() < n) t.blocks.reserve(std::max(n, 2 * t.blocks.capacity()));
      svo_block z;
      std::memset(&z, 0, sizeof z);
      t.blocks.resize(n, z);
    }
    size_t groups = (n + 7) / 8;
    if (t.group_parent.size() < groups) {
      if (t.group_parent.capacity() < groups) t.group_parent.reserve(std::max(groups, 2 * t.group_parent.capacity()));
      t.group_parent.resize(groups, UINT32_MAX);
    }
  }

  uint32_t alloc_group(uint32_t parent_address) {
    uint32_t g = dequeue();
    uint32_t a = g << 3;
    ensure((size_t)a + 8);
    t.group_parent[g] = parent_address;
    return a;
  }

  // BuildBlocks, TreeBuilder.cs:237-256.
  void build_blocks(uint32_t address, const Cell &parent, uint32_t depth) {
    for (uint32_t i = 0; i < 8; i++) {
      Cell c{depth, 2 * parent.x + (i & 1u), 2 * parent.y + ((i >> 1) & 1u), 2 * parent.z + ((i >> 2) & 1u)};
      svo_block blk = make_block(c);
      if (depth < max_depth && can_have_children(blk.chunk) && any_free()) {
        uint32_t na = alloc_group(address + i);
        blk.child = na;
        build_blocks(na, c, depth + 1);
      }
      t.blocks[address + i] = blk;
      if (t.block_count < address + i) t.block_count = address + i + 1;
    }
  }

  int run(uint64_t max_size) {
    const uint64_t dense_min = 1ull << (3 * B + 6);
    if (max_size < dense_min) max_size = dense_min;
    t.base_depth = B;
    t.max_depth = max_depth;
    t.block_count = 0;
    t.bases.assign(pow_sum(B), 0);
    // BuildBaseChunks, TreeBuilder.cs:152-172
    for (uint32_t d = 1; d <= B; d++) {
      uint32_t n = 1u << d;
      for (uint32_t z = 0; z < n; z++)
        for (uint32_t y = 0; y < n; y++)
          for (uint32_t x = 0; x < n; x++) t.bases[pow_sum(d - 1) + interleave3(x, y, z)] = make_chunk(Cell{d, x, y, z});
    }
    // PopulateInitialBlocks, TreeBuilder.cs:180-227
    const uint32_t bd = B + 2, scan = B + 1, sw = 1u << scan;
    const uint32_t base_start = pow_sum(B - 1);
    ensure((size_t)1 << (3 * bd));
    std::vector<uint8_t> children_needed((size_t)1 << (3 * bd), 0);
    for (uint32_t z = 0; z < sw; z++)
      for (uint32_t y = 0; y < sw; y++)
        for (uint32_t x = 0; x < sw; x++) {
          uint32_t bp = interleave3(x, y, z);
          if (((t.bases[base_start + (bp >> 3)] >> ((bp & 7u) * 2)) & 3u) == 3u) {
            for (uint32_t k = 0; k < 8; k++) {
              uint32_t x2 = k & 1u, y2 = (k >> 1) & 1u, z2 = (k >> 2) & 1u;
              uint32_t la = (bp << 3) + interleave3(x2, y2, z2);
              svo_block blk = make_block(Cell{bd, 2 * x + x2, 2 * y + y2, 2 * z + z2});
              t.blocks[la] = blk;
              children_needed[la] = can_have_children(blk.chunk) ? 1 : 0;
              if (t.block_count < la + 1) t.block_count = la + 1;
            }
          } else {
            empties.push_back(bp);
          }
        }
    std::sort(empties.begin(), empties.end());
    const uint64_t dense_groups = 1ull << (3 * (B + 1));
    range_next = dense_groups;
    range_end = (max_size >> 3) > dense_groups ? (max_size >> 3) : dense_groups;
    // BuildTree main loop, TreeBuilder.cs:120-142
    if (max_depth > bd) {
      uint32_t n = 1u << bd;
      for (uint32_t x = 0; x < n; x++)
        for (uint32_t y = 0; y < n; y++)
          for (uint32_t z = 0; z < n; z++) {
            uint32_t ia = interleave3(x, y, z);
            if (!children_needed[ia]) continue;
            if (!any_free()) return fail(SVO_ERR_CAPACITY, "svo_tree_build: free address queue exhausted");
            uint32_t a = alloc_group(ia);
            t.blocks[ia].child = a;
            build_blocks(a, Cell{bd, x, y, z}, bd + 1);
          }
    }
    // Trim, TreeBuilder.cs:145
    t.blocks.resize(t.block_count);
    t.group_parent.resize((t.block_count + 7) / 8, UINT32_MAX);
    return SVO_OK;
  }
};

// The builder geometry of a description (what TreeBuilder samples); `who` names the caller in
// error messages.
int make_geometry(const svo_geometry &d, const char *who, std::unique_ptr<Geometry> *out) {
  try {
    switch (d.kind) {
      case SVO_GEOMETRY_SPHERES:
        if (!d.shapes && d.count) return fail(SVO_ERR_INVALID_ARGUMENT, std::string(who) + ": no spheres");
        out->reset(new Spheres(d.shapes, d.count));
        return SVO_OK;
      case SVO_GEOMETRY_CUBES:
        if (!d.shapes && d.count) return fail(SVO_ERR_INVALID_ARGUMENT, std::string(who) + ": no cubes");
        out->reset(new Cubes(d.shapes, d.count));
        return SVO_OK;
      case SVO_GEOMETRY_SPHERE_NOISE:
        out->reset(new SphereNoise(d));
        return SVO_OK;
      case SVO_GEOMETRY_FBM_HEIGHTFIELD:
        out->reset(new Heightfield(d));
        return SVO_OK;
      default:
        return fail(SVO_ERR_INVALID_ARGUMENT, std::string(who) + ": unknown geometry kind");
    }
  } catch (const std::bad_alloc &) {
    return fail(SVO_ERR_OUT_OF_MEMORY, std::string(who) + ": out of host memory");
  }
}

}  // namespace

void camera_lut(const svo_trace_input &in, float *cos_u, float *sin_u, float *cos_v, float *sin_v) {
  const uint32_t W = in.screen_size[0], H = in.screen_size[1];
  for (uint32_t x = 0; x < W; x++) {
    cos_u[x] = camera_lut_at(in, x);
    sin_u[x] = camera_lut_at(in, (size_t)W + x);
  }
  for (uint32_t y = 0; y < H; y++) {
    cos_v[y] = camera_lut_at(in, 2 * (size_t)W + y);
    sin_v[y] = camera_lut_at(in, 2 * (size_t)W + H + y);
  }
}

float camera_lut_at(const svo_trace_input &in, size_t k) {
  const uint32_t W = in.screen_size[0], H = in.screen_size[1];
  const bool column = k < 2 * (size_t)W;
  const size_t j = column ? k : k - 2 * (size_t)W;
  const uint32_t n = column ? W : H;
  const bool sine = j >= n;
  const uint32_t i = (uint32_t)(sine ? j - n : j);
  // u = FoV.x ((W/2 - x) / W), v = FoV.y ((H/2 - y) / H), setupInitialTrace (wavefront.cl:563-595)
  const float size = (float)n;
  const float t = in.fov[column ? 0 : 1] * ((size / 2.0f - (float)i) / size);
  return (float)(sine ? std::sin((double)t) : std::cos((double)t));
}

}  // namespace svo

using svo::fail;

// A builder geometry kept for svo_geometry_blocks.
struct svo_geometry_host {
  std::unique_ptr<svo::Geometry> g;
};

extern "C" {

int svo_tree_build(const svo_geometry *geometry, uint32_t base_depth, uint32_t max_depth, uint32_t max_size,
                   svo_tree **out) {
  if (!geometry || !out) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_tree_build: null argument");
  if (base_depth < 1 || base_depth > 7) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_tree_build: base_depth must be 1..7");
  if (max_depth > 20) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_tree_build: max_depth must be <= 20");
  std::unique_ptr<svo::Geometry> g;
  int grc = svo::make_geometry(*geometry, "svo_tree_build", &g);
  if (grc) return grc;
  try {
    std::unique_ptr<svo_tree> tree(new svo_tree());
    svo::Builder b(*g, tree->t, base_depth, max_depth);
    int rc = b.run(max_size);
    if (rc != SVO_OK) return rc;
    *out = tree.release();
    return SVO_OK;
  } catch (const std::bad_alloc &) {
    return fail(SVO_ERR_OUT_OF_MEMORY, "svo_tree_build: out of host memory");
  }
}

int svo_geometry_create(const svo_geometry *geometry, svo_geometry_host **out) {
  if (!geometry || !out) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_geometry_create: null argument");
  *out = nullptr;
  std::unique_ptr<svo::Geometry> g;
  int rc = svo::make_geometry(*geometry, "svo_geometry_create", &g);
  if (rc) return rc;
  svo_geometry_host *h = new (std::nothrow) svo_geometry_host();
  if (!h) return fail(SVO_ERR_OUT_OF_MEMORY, "svo_geometry_create: out of host memory");
  h->g = std::move(g);
  *out = h;
  return SVO_OK;
}

int svo_geometry_destroy(svo_geometry_host *geometry) {
  delete geometry;
  return SVO_OK;
}

int svo_geometry_blocks(const svo_geometry_host *geometry, const uint32_t *cells, uint32_t n, svo_block *out) {
  if (!geometry || (n && (!cells || !out))) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_geometry_blocks: null argument");
  svo::Tree scratch;
  const svo::Builder b(*geometry->g, scratch, 1, 1);
  for (uint32_t k = 0; k < n; k++) {
    const uint32_t level = cells[4 * (size_t)k];
    if (level > 30) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_geometry_blocks: level must be <= 30");
    out[k] = b.make_block(svo::Cell{level, cells[4 * (size_t)k + 1], cells[4 * (size_t)k + 2], cells[4 * (size_t)k + 3]});
  }
  return SVO_OK;
}

int svo_tree_info_get(const svo_tree *tree, svo_tree_info *info) {
  if (!tree || !info) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_tree_info_get: null argument");
  info->base_depth = tree->t.base_depth;
  info->max_depth = tree->t.max_depth;
  info->block_count = tree->t.block_count;
  info->base_count = (uint32_t)tree->t.bases.size();
  info->group_count = (tree->t.block_count + 7) / 8;
  uint32_t alloc = 0;
  for (uint32_t p : tree->t.group_parent) alloc += p != UINT32_MAX;
  info->allocated_groups = alloc;
  return SVO_OK;
}

int svo_tree_copy_bases(const svo_tree *tree, uint16_t *dst, size_t count) {
  if (!tree || !dst) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_tree_copy_bases: null argument");
  if (count < tree->t.bases.size()) return fail(SVO_ERR_BUFFER_TOO_SMALL, "svo_tree_copy_bases: destination too small");
  std::memcpy(dst, tree->t.bases.data(), tree->t.bases.size() * sizeof(uint16_t));
  return SVO_OK;
}

int svo_tree_copy_blocks(const svo_tree *tree, svo_block *dst, size_t count) {
  if (!tree || !dst) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_tree_copy_blocks: null argument");
  if (count < tree->t.blocks.size()) return fail(SVO_ERR_BUFFER_TOO_SMALL, "svo_tree_copy_blocks: destination too small");
  std::memcpy(dst, tree->t.blocks.data(), tree->t.blocks.size() * sizeof(svo_block));
  return SVO_OK;
}

int svo_tree_initial_usages(const svo_tree *tree, uint16_t tick, svo_usage *dst, size_t count) {
  if (!tree || !dst) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_tree_initial_usages: null argument");
  const uint64_t dense_groups = 1ull << (3 * (tree->t.base_depth + 1));
  for (size_t g = 0; g < count; g++) {
    uint32_t parent = g < tree->t.group_parent.size() ? tree->t.group_parent[g] : UINT32_MAX;
    svo_usage u;
    u.pad = 0;
    if (g < dense_groups) {
      u.tick = 65535;
      u.parent = parent;
    } else if (parent != UINT32_MAX) {
      u.tick = tick;
      u.parent = parent;
    } else {
      u.tick = 0;
      u.parent = UINT32_MAX;
    }
    dst[g] = u;
  }
  return SVO_OK;
}

int svo_tree_save_oct(const svo_tree *tree, const char *path) {
  if (!tree || !path) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_tree_save_oct: null argument");
  FILE *f = std::fopen(path, "wb");
  if (!f) return fail(SVO_ERR_IO, std::string("svo_tree_save_oct: cannot open ") + path);
  uint8_t bd = (uint8_t)tree->t.base_depth;
  uint32_t bc = tree->t.block_count;
  bool ok = std::fwrite(&bd, 1, 1, f) == 1 && std::fwrite(&bc, 4, 1, f) == 1 &&
            std::fwrite(tree->t.bases.data(), 2, tree->t.bases.size(), f) == tree->t.bases.size() &&
            std::fwrite(tree->t.blocks.data(), 16, tree->t.blocks.size(), f) == tree->t.blocks.size();
  ok = (std::fclose(f) == 0) && ok;
  return ok ? SVO_OK : fail(SVO_ERR_IO, "svo_tree_save_oct: write failed");
}

int svo_tree_load_oct(const char *path, svo_tree **out) {
  if (!path || !out) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_tree_load_oct: null argument");
  FILE *f = std::fopen(path, "rb");
  if (!f) return fail(SVO_ERR_IO, std::string("svo_tree_load_oct: cannot open ") + path);
  std::unique_ptr<svo_tree> tree(new svo_tree());
  uint8_t bd = 0;
  uint32_t bc = 0;
  bool ok = std::fread(&bd, 1, 1, f) == 1 && std::fread(&bc, 4, 1, f) == 1 && bd >= 1 && bd <= 7;
  if (ok) {
    tree->t.base_depth = bd;
    tree->t.block_count = bc;
    tree->t.bases.resize(svo::pow_sum(bd));
    tree->t.blocks.resize(bc);
    ok = std::fread(tree->t.bases.data(), 2, tree->t.bases.size(), f) == tree->t.bases.size() &&
         std::fread(tree->t.blocks.data(), 16, bc, f) == bc;
  }
  std::fclose(f);
  if (!ok) return fail(SVO_ERR_IO, "svo_tree_load_oct: truncated or malformed file");
  // parents are reconstructed from the child links
  tree->t.group_parent.assign((bc + 7) / 8, UINT32_MAX);
  for (uint32_t a = 0; a < bc; a++) {
    uint32_t c = tree->t.blocks[a].child;
    if (c != UINT32_MAX && (c & 7u) == 0 && (c >> 3) < tree->t.group_parent.size() && tree->t.blocks[a].chunk != 0)
      tree->t.group_parent[c >> 3] = a;
  }
  *out = tree.release();
  return SVO_OK;
}

int svo_tree_free(svo_tree *tree) {
  delete tree;
  return SVO_OK;
}

int svo_camera_lut(const svo_trace_input *input, float *cos_u, float *sin_u, float *cos_v, float *sin_v) {
  if (!input || !cos_u || !sin_u || !cos_v || !sin_v) return fail(SVO_ERR_INVALID_ARGUMENT, "svo_camera_lut: null argument");
  svo::camera_lut(*input, cos_u, sin_u, cos_v, sin_v);
  return SVO_OK;
}

}  // extern "C"
