/*
 * shy.h -- C ABI of the MI355X particle hydraulic-erosion engine.
 *
 * The reference (weigert/SimpleHydrology) has no FFI; its boundary is the
 * static C++ API of World and Drop.  Each entry point below names the
 * reference symbol it stands in for.  include/shy_world.hpp wraps these in a
 * C++ World/Drop facade with the reference's names.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative shy_status otherwise;
 *    shy_last_error(ctx) gives a message.  No exception crosses the boundary.
 *  - Maps are x-major: cell (x, y) lives at x*ny + y, the reference's
 *    math::flatten (source/include/math.h:11-14).
 *  - One context owns one terrain's device buffers and one HIP stream.  A
 *    context is not thread-safe; distinct contexts are independent.
 *  - Calls are asynchronous on the context's stream unless they return host
 *    data (download/stats/trace), which synchronise.
 */
#ifndef SHY_H
#define SHY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHY_ABI_VERSION 1

typedef enum shy_status {
  SHY_OK = 0,
  SHY_ERR_ARG = -1,      /* invalid argument (size, pointer, parameter) */
  SHY_ERR_HIP = -2,      /* HIP runtime error */
  SHY_ERR_NODEV = -3,    /* no GPU visible */
  SHY_ERR_NONFINITE = -4,/* input map holds NaN or Inf */
  SHY_ERR_STATE = -5,    /* call not valid in the context's current state */
  SHY_ERR_IO = -6        /* file could not be read or written, or is not a map dump */
} shy_status;

/* Parameter set: the reference's static members, same names and defaults.
   Drop::{evapRate..momentumTransfer}  source/water.h:27-33, 43-50
   World::{lrate, maxdiff, settling}   source/world.h:26-29, 42-44 */
typedef struct shy_params {
  float evapRate;         /* 0.001 */
  float depositionRate;   /* 0.1   */
  float minVol;           /* 0.01  */
  float maxAge;           /* 500   */
  float entrainment;      /* 10    */
  float gravity;          /* 1     */
  float momentumTransfer; /* 1     */
  float lrate;            /* 0.1   */
  float maxdiff;          /* 0.01  */
  float settling;         /* 0.8   */
} shy_params;

/* Plant parameters: the Plant statics of source/vegetation.h:37-41, same
   names and defaults. */
typedef struct shy_plant_params {
  float maxSize;       /* 1.5  */
  float growRate;      /* 0.05 */
  float maxSteep;      /* 0.8  */
  float maxDischarge;  /* 0.3  */
  float maxTreeHeight; /* 0.8  */
} shy_plant_params;

/* Cell fields, in the reference's quad::cell order (source/cellpool.h:207-220). */
typedef enum shy_field {
  SHY_HEIGHT = 0,
  SHY_DISCHARGE = 1,
  SHY_MOMENTUMX = 2,
  SHY_MOMENTUMY = 3,
  SHY_DISCHARGE_TRACK = 4,
  SHY_MOMENTUMX_TRACK = 5,
  SHY_MOMENTUMY_TRACK = 6,
  SHY_ROOTDENSITY = 7,
  SHY_ERF_DISCHARGE = 8 /* erf(0.4*discharge), the value node::discharge returns */
} shy_field;

/* Per-step state of one traced drop (Drop fields, source/water.h:18-23). */
typedef struct shy_drop_step {
  int32_t age;    /* age before the step */
  int32_t status; /* 0 continues, 1 out-of-bounds exit */
  float pos[2];   /* after the step */
  float speed[2]; /* after the step */
  float volume;   /* after evaporation */
  float sediment; /* after evaporation */
  float erosion;  /* effD*cdiff taken from the start cell */
  float h2;       /* downstream height used for the equilibrium */
} shy_drop_step;

/* A drop, as Drop holds it (source/water.h:18-23). */
typedef struct shy_drop {
  int32_t age;
  float pos[2];
  float speed[2];
  float volume;
  float sediment;
} shy_drop;

/* Counters summed over the cycles run since the last reset. */
typedef struct shy_stats {
  int64_t steps;      /* descend() calls that reach the physics update */
  int64_t walked;     /* spawns that walked */
  int64_t skipped;    /* spawns rejected by height < 0.1 (world.h:71-72) */
  int64_t exits;      /* out-of-bounds exits */
  int64_t ends_age;   /* ended by age > maxAge */
  int64_t ends_vol;   /* ended by volume < minVol */
  int64_t export40;   /* sediment carried off the map, 2^-40 units */
  int64_t gain40;     /* evaporation sediment gain, 2^-40 units */
  int64_t walk_dh;    /* net height committed by the walk, 2^-20 units */
  int64_t casc_moves; /* non-zero cascade transfers */
} shy_stats;

/* One terrain of an ensemble: its size, its drops per cycle (World::erode's
   argument), its World::SEED and its parameter set. */
typedef struct shy_terrain_desc {
  int nx, ny;
  int drops;
  uint64_t seed;
  shy_params params;
} shy_terrain_desc;

/* Launch shape (see shy_autotune):
   walk_block  threads per block of the walk kernels (64..1024, multiple of 64)
   walk_grid   0: one lane per drop; > 0: a persistent grid of that many blocks
               walking the wave's drops grid-stride
   tile        cascade/blend stencil tile: 0 = 32x8, 1 = 64x4, 2 = 16x16 cells */
typedef struct shy_tuning {
  int walk_block;
  int walk_grid;
  int tile;
} shy_tuning;

typedef struct shy_ctx shy_ctx;

/* Defaults of the reference statics (water.h:43-50, world.h:42-44). */
void shy_default_params(shy_params* p);

/* Number of visible GPUs (0 when none). */
int shy_device_count(void);

/* Create a terrain of nx by ny cells on the current device, all fields zero.
   Stands in for cellpool.reserve + quad::map::init's allocation
   (SimpleHydrology.cpp:36-38, source/cellpool.h:323-340); World::SEED
   (source/world.h:21) is `seed`. */
int shy_create(const shy_params* p, int nx, int ny, uint64_t seed, shy_ctx** out);
/* An ensemble of terrains of any sizes and parameter sets in one context;
   every cycle runs all of them in one launch per kernel.  No reference
   counterpart: the reference holds one static World (world.h:17-44).
   Terrain t spawns from the hash of (seed_t, t, cycle).  The single-terrain
   entry points below are valid only when nter == 1 (else SHY_ERR_STATE); the
   shy_terrain_* ones take a terrain index. */
int shy_ensemble_create(const shy_terrain_desc* terrains, int nter, shy_ctx** out);
int shy_ensemble_count(const shy_ctx* ctx, int* nter);
int shy_terrain_size(const shy_ctx* ctx, int terrain, int* nx, int* ny, int* drops);
int shy_terrain_set_params(shy_ctx* ctx, int terrain, const shy_params* p);
int shy_terrain_set_drops(shy_ctx* ctx, int terrain, int drops);
int shy_terrain_upload_field(shy_ctx* ctx, int terrain, int field, const float* src);
int shy_terrain_download_field(shy_ctx* ctx, int terrain, int field, float* dst);
int shy_terrain_generate_fbm(shy_ctx* ctx, int terrain, uint64_t seed, int octaves,
                             float base_freq, float lacunarity, float gain);
int shy_terrain_generate_reference(shy_ctx* ctx, int terrain, int seed, int tileres);
/* Device-to-device field transfer (no reference counterpart; for the
   surrogate): export writes terrain t's (height, discharge, momentumx,
   momentumy) as 4 floats per cell, x-major, into a caller-owned device
   buffer; import_flow reads (discharge, momentumx, momentumy) as 3 floats per
   cell and refreshes erf(0.4 discharge).  Both synchronise. */
int shy_terrain_export(shy_ctx* ctx, int terrain, float* dst4_device);
int shy_terrain_import_flow(shy_ctx* ctx, int terrain, const float* src3_device);
/* particle-steps per terrain since the last shy_reset_stats */
int shy_terrain_steps(shy_ctx* ctx, int64_t* steps);
/* The last cycle's walk of terrain t, one entry per drop (no reference
   counterpart; for measuring the walk): its particle-steps and how it ended
   (0 maxAge reached, water.h:74-76; 1 volume below minVol, water.h:78-82;
   2 left the map, water.h:139-142; 3 not walked, spawn height below 0.1,
   world.h:71-72), in walk-thread order.  Writes at most `capacity` entries
   and the number written to *count; *count = 0 before the first cycle or
   after the parameters or drop counts changed. */
int shy_terrain_drop_log(shy_ctx* ctx, int terrain, int32_t* steps, int32_t* ends, int capacity,
                         int* count);
/* `cycles` World::erode calls of every terrain (each with its own drops);
   use_graph=1 replays two-cycle hipGraphs. */
int shy_ensemble_erode(shy_ctx* ctx, int cycles, int use_graph);
int shy_destroy(shy_ctx* ctx);
const char* shy_last_error(const shy_ctx* ctx); /* never NULL */
int shy_set_params(shy_ctx* ctx, const shy_params* p);
int shy_get_params(const shy_ctx* ctx, shy_params* p);
int shy_size(const shy_ctx* ctx, int* nx, int* ny);

/* Whole-map transfer in the reference's AoS quad::cell layout: 8 floats per
   cell, x-major (source/cellpool.h:207-220).  Upload rejects NaN/Inf. */
int shy_upload_cells(shy_ctx* ctx, const float* cells8);
int shy_download_cells(shy_ctx* ctx, float* cells8);
/* One field as a dense x-major float array (shy_field). */
int shy_upload_field(shy_ctx* ctx, int field, const float* src);
int shy_download_field(shy_ctx* ctx, int field, float* dst);

/* The same transfer in the reference's multi-tile pool order
   (source/cellpool.h:171-176, 323-340, 421-426): a square map of side
   mapsize*tileres is mapsize^2 nodes, node (i, j) at index i*mapsize + j
   holding cells [i*tileres, (i+1)*tileres) x [j*tileres, (j+1)*tileres)
   x-major.  The engine itself holds one seamless grid, so drops cross node
   borders exactly as map.get routes them. */
int shy_upload_cells_tiled(shy_ctx* ctx, int tileres, const float* pool8);
int shy_download_cells_tiled(shy_ctx* ctx, int tileres, float* pool8);

/* quad::map::normal (source/cellpool.h:181-204, 445-447) of every cell,
   computed on the device from the current heights: 3 floats (x, y, z) per
   cell, x-major. */
int shy_download_normals(shy_ctx* ctx, float* xyz);

/* Fill the height field with the seeded 2D Perlin-gradient fBm (normalised
   to [0,1]); stands in for quad::map::init's noise stack
   (source/cellpool.h:349-409).  base_freq <= 0 means 1/max(nx, ny). */
int shy_generate_fbm(shy_ctx* ctx, uint64_t seed, int octaves, float base_freq,
                     float lacunarity, float gain);

/* quad::map::init's own terrain (source/cellpool.h:349-409): eight layers of
   FastNoiseLite OpenSimplex2 3D FBm noise (frequencies 1..128, weights
   0.6^(k+1)) sampled at (x/tileres, y/tileres, seed % 10000), normalised as
   (h - min)/(max - min) with min <= 0 <= max.  The reference's world is
   512x512 with tileres 512 (cellpool.h:167-176).  Writes heights only. */
int shy_generate_reference(shy_ctx* ctx, int seed, int tileres);

/* Render-facing outputs of the reference's frame loop, x-major, 4 floats per
   cell (SimpleHydrology.cpp:341-354): dischargeMap = (water_rgb, erf(0.4
   discharge)) and momentumMap = (0.5(1+erf(mx)), 0.5(1+erf(my)), 0.5, 1).
   Either output may be NULL. */
int shy_download_textures(shy_ctx* ctx, const float* water_rgb, float* discharge_rgba,
                          float* momentum_rgba);
/* updatenode's per-vertex mesh data (source/cellpool.h:286-304), 12 floats
   per cell, x-major: position (x, 80 h, y), normal, tangent T-P, bitangent
   B-P (T = the +x neighbour, B = the +y neighbour; heights off the map read
   0, as quad::node::height does). */
int shy_download_vertices(shy_ctx* ctx, float* vertex12);

/* Map dumps: this header, then nx*ny cells of 8 floats in the quad::cell
   order (source/cellpool.h:207-220), x-major, host byte order.  Load needs a
   context of the same size; it restores the cells, the parameters and the
   spawn stream (seed, cycle), so a loaded map continues bit for bit. */
#define SHY_MAP_MAGIC "SHYMAP01"
#define SHY_MAP_VERSION 1u
typedef struct shy_map_header {
  char magic[8];    /* SHY_MAP_MAGIC, no terminator */
  uint32_t version; /* SHY_MAP_VERSION */
  int32_t nx, ny;
  uint32_t cycle;   /* the next World::erode call's cycle number */
  uint64_t seed;
  shy_params params;
} shy_map_header;   /* 72 bytes */
int shy_save_map(shy_ctx* ctx, const char* path);
int shy_load_map(shy_ctx* ctx, const char* path);

/* Vegetation (source/vegetation.h:44-186), the other writer of the erosion's
   inputs: plants raise the root density around them (Plant::root), and the
   walk's deposition rate falls with it (water.h:86-87).  The plant list lives
   on the device.  One shy_vegetation_grow call is one Vegetation::grow(): a
   spawn attempt at a random cell, then every plant grows, may die and may
   seed a child within 4 cells; the children are visited in the same call, as
   the reference's loop visits plants appended during it.  Each generation
   decides against the map and root density as they stood when it began;
   root changes are summed exactly in units of 0.2 and applied per
   generation; draws come from counter hashes of (seed, step, plant, purpose),
   not rand().  Plants beyond `capacity` are dropped and counted. */
void shy_default_plant_params(shy_plant_params* p);
int shy_vegetation_enable(shy_ctx* ctx, int capacity, const shy_plant_params* p); /* NULL = defaults */
int shy_vegetation_grow(shy_ctx* ctx);
int shy_vegetation_count(shy_ctx* ctx, int* plants, int64_t* dropped);
/* the plant list in order: x, y, size per plant */
int shy_vegetation_download(shy_ctx* ctx, float* xys);
int shy_vegetation_upload(shy_ctx* ctx, const float* xys, int n);

/* World::erode(cycles) (source/world.h:54-88): one cycle of `drops` spawns.
   Wave-synchronous: every drop reads the map as it stood at the start of the
   cycle.  A counting pass walks every drop and logs each step's start cell;
   the log is partitioned by cell tile and counted per cell, so the commit
   pass can scale each step's erosion such that the k drops through a cell
   commit what k drops in sequence would.  Every commit is quantised to an
   integer (heights 2^-20, tracks 2^-24) and summed per cell in integers, with
   no per-step atomic on the map, so the result does not depend on drop order
   or launch shape; the cascade then runs as a weighted stencil over the cells
   drops arrived at, and the discharge/momentum fields are blended. */
int shy_erode(shy_ctx* ctx, int drops);
/* `cycles` consecutive World::erode(drops) calls; use_graph=1 replays them
   from a captured hipGraph. */
int shy_erode_cycles(shy_ctx* ctx, int cycles, int drops, int use_graph);

/* One drop traced on the frozen current map (no map update): Drop(pos) then
   Drop::descend() until it returns false (source/water.h:14, 58-156).
   Writes up to max_steps states and the number of steps taken. */
int shy_trace_drop(shy_ctx* ctx, float x, float y, int max_steps, shy_drop_step* out,
                   int* n_steps);

/* World::cascade(pos) (source/world.h:90-168) at one site, on the live map. */
int shy_cascade(shy_ctx* ctx, float x, float y);
/* One Drop::descend() call (source/water.h:58-156) on the live map, with the
   reference's immediate writes and inline cascade; *alive gets its result. */
int shy_descend(shy_ctx* ctx, shy_drop* drop, int* alive);

/* Spawn stream state: World::SEED and the number of cycles run. */
int shy_set_seed(shy_ctx* ctx, uint64_t seed, uint32_t cycle);
int shy_set_cycle(shy_ctx* ctx, uint32_t cycle);
int shy_get_cycle(const shy_ctx* ctx, uint32_t* cycle);

int shy_get_stats(shy_ctx* ctx, shy_stats* out);
int shy_reset_stats(shy_ctx* ctx);
int shy_synchronize(shy_ctx* ctx);

/* Launch shape of the drop walk: threads per block (64..1024, multiple of
   64).  0 restores the default. */
int shy_set_walk_block(shy_ctx* ctx, int threads);
int shy_set_tuning(shy_ctx* ctx, const shy_tuning* t);
int shy_get_tuning(const shy_ctx* ctx, shy_tuning* t);
/* Drop order of the count walk (no reference counterpart).  on = 1: each
   cycle assigns drops to walk threads by the position of their spawn cell,
   so that each XCD walks drops of one region and consecutive blocks walk
   neighbouring drops; on = 0: thread g walks drop g; on = -1 (the default):
   ordered when a cycle has at least 262144 drops (more walk threads than
   the GPU runs at once, where the order shortens the walk), else not.
   Results are bit-identical either way. */
int shy_set_drop_order(shy_ctx* ctx, int on);
/* Autotuner (no reference counterpart): times `cycles` cycles of every
   candidate launch shape (4 walk block sizes x one-lane-per-drop or
   persistent x 3 stencil tiles) from a copy of the live state, restores the
   state after each, keeps the fastest and returns it with its ms per cycle.
   Results are bit-identical under every shape. */
int shy_autotune(shy_ctx* ctx, int cycles, shy_tuning* chosen, float* best_ms);

/* Measurement (no reference counterpart): run `cycles` cycles eagerly with
   HIP events on the context's stream between kernels.  drops >= 0 sets the
   drops of a single-terrain context first; drops < 0 keeps every terrain's
   own.  ms[0] = wall time of the whole sequence; ms[1..5] = summed durations
   of the walk-record pack, the count walk, the partition of the step log by
   cell tile with the per-tile counts, the commit replay and the per-tile
   reduce + cascade/blend stencil, in milliseconds. */
int shy_time_cycles(shy_ctx* ctx, int cycles, int drops, float ms[6]);

#ifdef __cplusplus
}
#endif
#endif
