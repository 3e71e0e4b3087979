/*
 * shy_world.hpp -- C++ facade with the reference's World / Drop interface
 * (source/world.h:17-36, source/water.h:12-39) over the C ABI in shy.h.
 *
 * A headless host written against the reference swaps its includes for this
 * header and keeps its loop (SimpleHydrology.cpp:313-325):
 *
 *   shy::World::SEED = 42;
 *   shy::World::map.init(512, 512, shy::World::SEED);   // the reference's terrain, on the GPU
 *   for (;;) { shy::World::erode(512); shy::World::map.sync(); ... }
 *
 * The static parameters keep the reference's names and defaults; each call
 * into the engine forwards their current values.  Errors become shy::Error
 * exceptions on the host side of the boundary (the C ABI itself never throws).
 */
#ifndef SHY_WORLD_HPP
#define SHY_WORLD_HPP

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "shy.h"

namespace shy {

struct ivec2;
struct vec2 {
  float x = 0.0f, y = 0.0f;
  vec2() = default;
  vec2(float a, float b) : x(a), y(b) {}
  inline vec2(ivec2 v);  // implicit, as glm's conversion constructors
};
struct ivec2 {
  int x = 0, y = 0;
  ivec2() = default;
  explicit ivec2(int a) : x(a), y(a) {}
  ivec2(int a, int b) : x(a), y(b) {}
  ivec2(vec2 v) : x((int)v.x), y((int)v.y) {}  // implicit truncation toward zero, as glm
};
inline vec2::vec2(ivec2 v) : x((float)v.x), y((float)v.y) {}
inline vec2 operator+(vec2 a, vec2 b) { return vec2(a.x + b.x, a.y + b.y); }
inline vec2 operator-(vec2 a, vec2 b) { return vec2(a.x - b.x, a.y - b.y); }
inline ivec2 operator+(ivec2 a, ivec2 b) { return ivec2(a.x + b.x, a.y + b.y); }
inline ivec2 operator-(ivec2 a, ivec2 b) { return ivec2(a.x - b.x, a.y - b.y); }
inline ivec2 operator/(ivec2 a, int d) { return ivec2(a.x / d, a.y / d); }
inline bool operator==(ivec2 a, ivec2 b) { return a.x == b.x && a.y == b.y; }
struct vec3 {
  float x = 0.0f, y = 0.0f, z = 0.0f;
};

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& msg) : std::runtime_error(msg), code(code) {}
  int code;
};

inline void check(const shy_ctx* ctx, int rc) {
  if (rc != SHY_OK) throw Error(rc, shy_last_error(ctx));
}

/* quad::cell (source/cellpool.h:207-220) */
struct cell {
  float height, discharge, momentumx, momentumy;
  float discharge_track, momentumx_track, momentumy_track;
  float rootdensity;
};

/* mappool::sliceval (source/cellpool.h:44-47): a cell reference and its
   position in the slice, so `for (auto [cell, pos] : node.s)` binds `cell`
   to the cell itself, as the reference's loops do (world.h:56-61, 81-86). */
template <typename T>
struct sliceval {
  T& start;
  ivec2 pos;
};

/* mappool::slice (source/cellpool.h:73-99) over the host snapshot: the
   res.x by res.y cells from `org` of a grid with row stride `stride`,
   iterated x-major (pos.y fastest), as slice_iterator does (:49-71). */
template <typename T>
struct slice {
  T* root = nullptr;
  int stride = 0;
  ivec2 org = ivec2(0, 0);
  ivec2 res = ivec2(0, 0);

  size_t size() const { return (size_t)res.x * res.y; }
  bool oob(ivec2 p) const { return p.x < 0 || p.y < 0 || p.x >= res.x || p.y >= res.y; }
  T* get(ivec2 p) const {
    if (root == nullptr || oob(p)) return nullptr;
    return root + (size_t)(org.x + p.x) * stride + (org.y + p.y);
  }

  struct iterator {
    const slice* s;
    ivec2 pos;
    sliceval<T> operator*() const { return {*s->get(pos), pos}; }
    iterator& operator++() {
      if (++pos.y == s->res.y) {
        pos.y = 0;
        ++pos.x;
      }
      return *this;
    }
    bool operator!=(const iterator& o) const { return !(pos == o.pos); }
  };
  iterator begin() const { return iterator{this, res.x > 0 && res.y > 0 ? ivec2(0, 0) : ivec2(res.x, 0)}; }
  iterator end() const { return iterator{this, ivec2(res.x, 0)}; }
};

struct map;

/* quad::node (source/cellpool.h:222-250): one tile of the map, a view of the
   host snapshot.  The engine holds one seamless grid, so node::normal is the
   map's normal; the reference's own node::normal reads zero heights beyond
   its node (cellpool.h:237-240 through _normal), which equals the map's with
   the reference's single node (mapsize = 1, cellpool.h:171). */
struct node {
  ivec2 pos = ivec2(0, 0);  // absolute position of the node's first cell
  slice<cell> s;
  const map* owner = nullptr;

  cell* get(ivec2 p) { return s.get(p - pos); }
  bool oob(ivec2 p) const { return s.oob(p - pos); }
  float height(ivec2 p) {
    cell* c = get(p);
    return c == nullptr ? 0.0f : c->height;
  }
  inline float discharge(ivec2 p) const;  // erf(0.4 discharge), cellpool.h:242-244
  inline vec3 normal(ivec2 p) const;
};

/* quad::map (source/cellpool.h:319-449): the device terrain plus a host
   snapshot for the read accessors; sync() refreshes the snapshot and the
   node grid, upload() writes the snapshot back.  The nodes tile the map in
   tileres x tileres blocks (quad::tileres, cellpool.h:167-169), node (i, j)
   at index i*mapsize_y + j as map::get computes it (:421-426). */
struct map {
  shy_ctx* ctx = nullptr;
  int nx = 0, ny = 0;
  int tileres = 512;
  std::vector<cell> cells;
  std::vector<float> erfd, normals;
  std::vector<node> nodes;
  int mapsize_x = 0, mapsize_y = 0;

  // quad::map::init (cellpool.h:323-411): the OpenSimplex2 terrain sampled at
  // x / tileres (cellpool.h:360-404; the reference's tileres is 512,
  // :167-169), seed passed as the reference's int SEED
  void init(int sx, int sy, unsigned int seed);
  // the seeded Perlin fBm of the benchmark workloads instead
  void init_fbm(int sx, int sy, unsigned int seed);
  void release() {
    if (ctx) shy_destroy(ctx);
    ctx = nullptr;
  }
  void sync() {
    cells.resize((size_t)nx * ny);
    erfd.resize((size_t)nx * ny);
    normals.resize((size_t)3 * nx * ny);
    check(ctx, shy_download_cells(ctx, &cells[0].height));
    check(ctx, shy_download_field(ctx, SHY_ERF_DISCHARGE, erfd.data()));
    check(ctx, shy_download_normals(ctx, normals.data()));
    build_nodes();
  }
  void upload() { check(ctx, shy_upload_cells(ctx, &cells[0].height)); }
  void build_nodes() {
    if (tileres < 1) throw Error(SHY_ERR_ARG, "map: tileres must be >= 1");
    mapsize_x = (nx + tileres - 1) / tileres;
    mapsize_y = (ny + tileres - 1) / tileres;
    nodes.assign((size_t)mapsize_x * mapsize_y, node());
    for (int i = 0; i < mapsize_x; ++i)
      for (int j = 0; j < mapsize_y; ++j) {
        node& t = nodes[(size_t)i * mapsize_y + j];
        t.pos = ivec2(i * tileres, j * tileres);
        t.owner = this;
        t.s.root = cells.data();
        t.s.stride = ny;
        t.s.org = t.pos;
        t.s.res = ivec2(std::min(tileres, nx - t.pos.x), std::min(tileres, ny - t.pos.y));
      }
  }

  bool oob(ivec2 p) const { return p.x < 0 || p.y < 0 || p.x >= nx || p.y >= ny; }
  node* get(ivec2 p) {
    if (oob(p)) return nullptr;
    return &nodes[(size_t)(p.x / tileres) * mapsize_y + (p.y / tileres)];
  }
  cell* getCell(ivec2 p) { return oob(p) ? nullptr : get(p)->get(p); }
  float height(ivec2 p) const { return oob(p) ? 0.0f : cells[(size_t)p.x * ny + p.y].height; }
  /* node::discharge: erf(0.4 * discharge) (cellpool.h:242-244) */
  float discharge(ivec2 p) const { return oob(p) ? 0.0f : erfd[(size_t)p.x * ny + p.y]; }
  vec3 normal(ivec2 p) const {
    const float* n = &normals[3 * ((size_t)p.x * ny + p.y)];
    return vec3{n[0], n[1], n[2]};
  }
};

inline float node::discharge(ivec2 p) const { return owner->discharge(p); }
inline vec3 node::normal(ivec2 p) const { return owner->normal(p); }

struct Drop;

class World {
 public:
  inline static unsigned int SEED = 1;  // world.h:38
  inline static shy::map map;           // world.h:40
  inline static float lrate = 0.1f;     // world.h:42-44
  inline static float maxdiff = 0.01f;
  inline static float settling = 0.8f;

  static void erode(int cycles);  // world.h:54-88
  static void cascade(vec2 pos);  // world.h:90-168
  static shy_params params();
  static void push_params();
};

struct Drop {
  explicit Drop(vec2 p) : pos(p) {}  // water.h:14
  int age = 0;
  vec2 pos;
  vec2 speed = vec2(0.0f, 0.0f);
  float volume = 1.0f;
  float sediment = 0.0f;

  inline static float evapRate = 0.001f;  // water.h:43-50
  inline static float depositionRate = 0.1f;
  inline static float minVol = 0.01f;
  inline static float maxAge = 500.0f;
  inline static float entrainment = 10.0f;
  inline static float gravity = 1.0f;
  inline static float momentumTransfer = 1.0f;

  bool descend();  // water.h:58-156, on the live device map
};

inline shy_params World::params() {
  shy_params p;
  p.evapRate = Drop::evapRate;
  p.depositionRate = Drop::depositionRate;
  p.minVol = Drop::minVol;
  p.maxAge = Drop::maxAge;
  p.entrainment = Drop::entrainment;
  p.gravity = Drop::gravity;
  p.momentumTransfer = Drop::momentumTransfer;
  p.lrate = lrate;
  p.maxdiff = maxdiff;
  p.settling = settling;
  return p;
}

inline void World::push_params() {
  // forward the statics only when they changed: a new parameter set
  // invalidates captured graphs
  static shy_params last{};
  static const shy_ctx* last_ctx = nullptr;
  const shy_params p = params();
  const bool same = last_ctx == map.ctx && p.evapRate == last.evapRate &&
                    p.depositionRate == last.depositionRate && p.minVol == last.minVol &&
                    p.maxAge == last.maxAge && p.entrainment == last.entrainment &&
                    p.gravity == last.gravity && p.momentumTransfer == last.momentumTransfer &&
                    p.lrate == last.lrate && p.maxdiff == last.maxdiff && p.settling == last.settling;
  if (same) return;
  check(map.ctx, shy_set_params(map.ctx, &p));
  last = p;
  last_ctx = map.ctx;
}

inline void map::init(int sx, int sy, unsigned int seed) {
  release();
  nx = sx;
  ny = sy;
  const shy_params p = World::params();
  check(nullptr, shy_create(&p, nx, ny, seed, &ctx));
  check(ctx, shy_generate_reference(ctx, (int)seed, tileres));
  sync();
}

inline void map::init_fbm(int sx, int sy, unsigned int seed) {
  release();
  nx = sx;
  ny = sy;
  const shy_params p = World::params();
  check(nullptr, shy_create(&p, nx, ny, seed, &ctx));
  check(ctx, shy_generate_fbm(ctx, seed, 8, 0.0f, 2.0f, 0.5f));
  sync();
}

inline void World::erode(int cycles) {
  push_params();
  check(map.ctx, shy_erode(map.ctx, cycles));
}

inline void World::cascade(vec2 pos) {
  push_params();
  check(map.ctx, shy_cascade(map.ctx, pos.x, pos.y));
}

inline bool Drop::descend() {
  World::push_params();
  shy_drop d;
  d.age = age;
  d.pos[0] = pos.x;
  d.pos[1] = pos.y;
  d.speed[0] = speed.x;
  d.speed[1] = speed.y;
  d.volume = volume;
  d.sediment = sediment;
  int alive = 0;
  check(World::map.ctx, shy_descend(World::map.ctx, &d, &alive));
  age = d.age;
  pos = vec2(d.pos[0], d.pos[1]);
  speed = vec2(d.speed[0], d.speed[1]);
  volume = d.volume;
  sediment = d.sediment;
  return alive != 0;
}

}  // namespace shy

#endif
