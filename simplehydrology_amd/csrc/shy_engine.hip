// shy_engine.hip -- C ABI (include/shy.h) over the gfx950 kernels.
//
// A context holds an ensemble of terrains (a single-terrain context is an
// ensemble of one) in HBM as concatenated structure-of-arrays, each terrain
// x-major in its own slice:
//   h[2]        float   heights, ping-pong (the cascade stencil reads a halo)
//   fld         float4  {momentumx, momentumy, discharge, erf(0.4 discharge)}
//   root        float   root density
//   visit[2]    uint8   cascade sites, ping-pong
//   first       int32   drops whose first step starts at the cell; cleared by
//                       tile_count_kernel
//   dh          int32   summed walk commits per cell, 2^-20 (tile_reduce_kernel)
//   trk[3]      int64   summed tracks per cell, 2^-24 (tile_reduce_kernel; stay
//                       readable until the next cycle, as the reference's
//                       *_track fields do)
//   deposit     int32   terminal deposits, 2^-20; cleared by tile_reduce_kernel
// and per wave (sized by total drops x (max maxAge + 1) log slots):
//   slog, strk  int2x2  [step][drop] step log of the count walk: {start cell,
//                       c_eq} (what the replay reads) and {momentum tracks}
//   qb          int32   [drop][step] height commit of each step (replay)
//   ent         int32 + {int32x3}  the valid slots grouped by cell tile
//                       (partition): start cells, then {slot, tracks}
//   vq          int32   [terrain][step] quantised volume track of a step
//   dlog, wsteps int4, int32 [drop] how each walk ended, its steps
//   hist        int32   per terrain a [tile][block] count matrix of the
//                       partition, then exclusive offsets over all of them
//   tstart      int32   [tile] first partition entry of every cell tile
//   pack        float4x2 per cell: the count walk's 32-byte record of the
//                       cycle's start (height, normal, flow), by pack_kernel
//   sr          float   burst scale per cell (tile_count_kernel)
//   tcount, tsum int32, int64x4  per-part tile counts and sums when a tile is
//                       split over kCountSplit / kReduceSplit blocks
// A device table of Terrain entries maps drops, cells and cascade tiles to
// their terrain, so every kernel covers the whole ensemble in one launch.
// Parity p = cycle & 1 selects h and visit; the cascade/blend kernel of cycle
// p clears visit[p^1] for the next cycle, so a cycle needs no memset and two
// cycles form a replayable hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "shy_kernels.hpp"

using namespace shy;

struct shy_ctx {
  int device = 0;
  std::vector<Terrain> ter;  // host copy of the table
  Terrain* d_ter = nullptr;
  int* gflag = nullptr;      // [terrain] a gravity term out of the walk's division range (pack)
  size_t n = 0;              // cells, all terrains
  int total_drops = 0;
  int total_tiles = 0;
  uint32_t cycle = 0;        // host mirror of state->cycle
  float* h[2] = {nullptr, nullptr};
  float4* fld = nullptr;
  float* root = nullptr;
  uint8_t* visit[2] = {nullptr, nullptr};
  int32_t* first = nullptr;
  float* sr = nullptr;
  int32_t* dh = nullptr;
  long long* trk = nullptr;  // 3*n: d, mx, my
  int32_t* deposit = nullptr;
  int2* slog = nullptr;      // [step][drop] step log of the count walk: {start cell, c_eq}
  int2* strk = nullptr;      // [step][drop] the steps' momentum tracks
  int32_t* qb = nullptr;     // [drop][step] height commits of the replay
  size_t qb_cap = 0;         // ints qb holds
  int qstride = 0;           // qb's stride per drop (log rows rounded up to kReplayDepth)
  int4* ent = nullptr;       // partition entries in tile order: wave_cap start cells, then
                             // wave_cap {slot, tracks} (16 B per log slot in all)
  int32_t* vq = nullptr;     // [terrain][row] quantised volume track per step
  size_t vq_cap = 0;
  int vq_rows = 0;
  bool vq_dirty = true;      // parameters changed since vq was built
  size_t wave_cap = 0;       // log slots the wave buffers can hold
  int32_t* hist = nullptr;   // per-terrain [tile][block] partition counts, then offsets
  int32_t* hbsum = nullptr;  // block sums of the hist scan
  size_t hist_cap = 0;       // ints hist holds (hbsum holds hist_cap / 1024 + 1)
  int32_t* tstart = nullptr; // [tile] first partition entry, [ntiles] = all entries
  int tstart_cap = 0;        // tiles tstart holds (plus one)
  int part_blocks = 0;       // partition geometry of the current wave: blocks over all terrains
  size_t part_entries = 0;   // count-matrix entries over all terrains
  int part_rows = 0;         // log rows
  int part_window = 0;       // most tiles of a terrain
  int part_maxw = 0;         // widest partition block (drops)
  bool walk_counts = false;  // the walk builds the partition's tile counts
  int ntiles = 0;            // cell tiles of kTileCells cells
  int32_t* tcount = nullptr;    // per-part tile counts (kCountSplit parts per tile)
  long long* tsum = nullptr;    // per-part tile sums (kReduceSplit parts per tile)
  bool tile_split = false;      // tiles split over several blocks (the part buffers are small)
  int4* dlog = nullptr;  // [drop] how each walk ended
  bool dlog_valid = false;  // dlog holds a cycle of the current wave geometry
  int32_t* dperm = nullptr;  // [drop] the terrain's drop each walk thread takes (order_kernel)
  long long pack_recs = 0;   // packed walk records over all terrains (pack_records)
  int drop_order = -1;       // assign drops to walk threads by spawn position: 1, 0, -1 auto
  int order_chunks = 0;      // blocks of order_kernel
  int32_t* wsteps = nullptr;  // [drop] steps of each walk
  float4* pack = nullptr;  // per cell, 2 x float4: the count walk's record
  size_t dlog_cap = 0;
  long long* tsteps = nullptr;
  DeviceState* state = nullptr;
  double* erf_a = nullptr;
  float2* part = nullptr;  // min/max partials
  uint8_t* perms = nullptr;
  shy_drop_step* trace_buf = nullptr;
  int trace_cap = 0;
  int* trace_n = nullptr;
  void* drop_buf = nullptr;
  hipStream_t stream = nullptr;
  shy_tuning tune{256, 0, 0};
  // vegetation (single terrain)
  int veg_cap = 0;
  int32_t *veg_px = nullptr, *veg_py = nullptr, *veg_alive = nullptr, *veg_cflag = nullptr;
  int32_t *veg_cx = nullptr, *veg_cy = nullptr, *veg_units = nullptr, *veg_st = nullptr;
  float* veg_size = nullptr;
  uint32_t veg_step = 0;
  shy_plant_params veg_P{};
  hipGraphExec_t graph = nullptr;
  std::string err;
};

namespace {

constexpr int kPartBlocks = 1024;
constexpr float kMaxAgeLimit = 1.0e6f;
constexpr int kTileY[3] = {32, 64, 16};
constexpr int kTileX[3] = {8, 4, 16};
constexpr int kWaveDrops = 64;  // drop ranges are padded to whole wave64s
// drop ranges of an ensemble are padded to 256 drops: walk blocks of 64, 128
// or 256 threads then never span terrains
constexpr int kDropAlign = 256;
// drops from which the automatic drop order is on: more walk threads than
// the GPU runs at once, so the walk's blocks run in sequence and neighbouring
// blocks share the L2 when their drops start near each other
constexpr int kOrderMinDrops = 262144;

bool ordered(const shy_ctx* c) {
  return c->drop_order < 0 ? c->total_drops >= kOrderMinDrops : c->drop_order != 0;
}
constexpr int kReplayBlock = 256;

std::string g_noctx_err;

int fail(shy_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg; else g_noctx_err = msg;
  return code;
}

#define HIPCHK(ctx, call)                                                               \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail((ctx), SHY_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define RETURN_IF(r)   \
  do {                 \
    int r_ = (r);      \
    if (r_) return r_; \
  } while (0)

void erf_coefficients(double* a) {
  // 2/sqrt(pi) (-1)^k / (k! (2k+1)), the series of Abramowitz & Stegun 7.1.5
  const double two_over_sqrtpi = 1.1283791670955126;
  double f = 1.0;
  for (int k = 0; k < kErfTerms; ++k) {
    if (k > 0) f = f * (double)k;
    const double v = two_over_sqrtpi / (f * (double)(2 * k + 1));
    a[k] = (k & 1) ? -v : v;
  }
}

void fbm_perm(uint64_t seed, int octave, uint8_t* perm) {
  for (int i = 0; i < 256; ++i) perm[i] = (uint8_t)i;
  const uint64_t key = mix64(mix64(seed ^ 0x5045524c494e4f49ULL) + (uint64_t)octave);
  for (int i = 255; i > 0; --i) {
    const uint64_t r = mix64(key ^ (uint64_t)i);
    const int j = (int)((r >> 32) % (uint64_t)(i + 1));
    const uint8_t t = perm[i];
    perm[i] = perm[j];
    perm[j] = t;
  }
}

bool params_ok(const shy_params* p) {
  const float v[] = {p->evapRate, p->depositionRate, p->minVol, p->maxAge, p->entrainment,
                     p->gravity, p->momentumTransfer, p->lrate, p->maxdiff, p->settling};
  for (float x : v)
    if (!std::isfinite(x)) return false;
  // maxAge bounds every walk loop (one age per step); evapRate < 1 keeps the
  // volume positive
  return p->evapRate < 1.0f && p->maxAge >= 0.0f && p->maxAge <= kMaxAgeLimit;
}

const char* kParamsMsg = "parameter not finite, evapRate >= 1 or maxAge outside [0, 1e6]";

int parity(const shy_ctx* c) { return (int)(c->cycle & 1u); }

void drop_graph(shy_ctx* c) {
  if (c->graph) hipGraphExecDestroy(c->graph);
  c->graph = nullptr;
}

// Cell tiles the partition's LDS histogram can hold (128 KiB).
constexpr int kPartMaxTiles = 32768;

// Offsets of every terrain's cells, drops and cascade tiles; uploads the table.
// Whether every volume a drop carries into a step (vol_0 = 1, vol_k+1 =
// evap_vol(vol_k), while age <= maxAge and vol >= minVol: water.h:74-82, 136)
// lies where the walk's reciprocal-form division is exact.
bool volumes_in_division_range(const shy_params& P) {
  float vol = 1.0f;
  for (int age = 0;; ++age) {
    if ((float)age > P.maxAge || vol < P.minVol) return true;
    if (!markstein_den_ok(vol)) return false;
    vol = evap_vol(vol, P.evapRate);
  }
}

int layout(shy_ctx* c) {
  long long cells = 0, drops = 0, tiles = 0, recs = 0;
  const int ty = kTileY[c->tune.tile], tx = kTileX[c->tune.tile];
  for (Terrain& t : c->ter) {
    t.vol_ok = volumes_in_division_range(t.P) ? 1 : 0;
    t.cell_off = cells;
    t.pack_off = recs;
    recs += pack_records(t.nx, t.ny);
    t.drop_off = (int)drops;
    t.tiles_y = (t.ny + ty - 1) / ty;
    t.tile_off = (int)tiles;
    cells += (long long)t.nx * t.ny;
    // whole waves per terrain: the walk reads its terrain entry once per wave
    const long long align = c->ter.size() > 1 ? kDropAlign : kWaveDrops;
    drops += ((long long)t.drops + align - 1) / align * align;
    tiles += (long long)t.tiles_y * ((t.nx + tx - 1) / tx);
    if (drops > INT32_MAX || tiles > INT32_MAX)
      return fail(c, SHY_ERR_ARG, "ensemble too large: drops and cascade tiles must stay below 2^31");
  }
  if ((cells + kTileCells - 1) / kTileCells > kPartMaxTiles)
    return fail(c, SHY_ERR_ARG, "context too large: at most 2^27 cells over all terrains");
  c->total_drops = (int)drops;
  c->total_tiles = (int)tiles;
  c->ntiles = (int)((cells + kTileCells - 1) / kTileCells);
  c->vq_dirty = true;
  c->dlog_valid = false;
  drop_graph(c);
  if (c->d_ter) {
    HIPCHK(c, hipMemcpyAsync(c->d_ter, c->ter.data(), c->ter.size() * sizeof(Terrain),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return SHY_OK;
}

size_t wave_steps(const shy_ctx* c) {
  float age = 0.0f;
  for (const Terrain& t : c->ter) age = std::max(age, t.P.maxAge);
  return (size_t)c->total_drops * ((size_t)age + 1);
}

// Partition blocks: at most kPartMaxBlocks, each a whole number of waves of
// drops over all log rows (the [tile][block] matrix stays small).
constexpr int kPartMaxBlocks = 512;
// blocks per tile in the tile kernels: the largest tiles hold several times
// the mean's entries; the counts split 4 ways, the sums (whose parts are
// 32 B per cell) 2 ways (1024^2 cycle: 1.185 ms at 4 and 4, 1.171 at 2 and 2,
// 1.165 at 4 and 2)
constexpr int kCountSplit = 4;
constexpr int kReduceSplit = 2;
constexpr int kTileSplitMaxTiles = 2048;
// partition blocks up to which each takes a CU's whole LDS (the scatter's
// rounds of 8192 entries): no more blocks than CUs
constexpr int kBigRoundBlocks = 256;

// The partition geometry: per terrain its tiles, its blocks of dwidth drops
// (dwidth = walk_block when the walk counts, else enough drops for at most
// kPartMaxBlocks blocks per terrain) and the offset of its count matrix.
// Fills the fields into t (when given); returns the count-matrix entries
// over all terrains and sets blocks (over all terrains) and window (most
// tiles of a terrain).
size_t part_geometry(const shy_ctx* c, int walk_block, std::vector<Terrain>* t, int& blocks,
                     int& window, int* chunks = nullptr) {
  size_t off = 0;
  int nchunk = 0;
  blocks = 0;
  window = 1;
  for (size_t k = 0; k < c->ter.size(); ++k) {
    const Terrain& T = c->ter[k];
    const long long end = (k + 1 < c->ter.size()) ? c->ter[k + 1].drop_off : c->total_drops;
    const int padded = (int)(end - T.drop_off);
    int dwidth = walk_block;
    if (dwidth == 0) {
      const int per = (padded + kPartMaxBlocks - 1) / kPartMaxBlocks;
      dwidth = std::max(kWaveDrops, (per + kWaveDrops - 1) / kWaveDrops * kWaveDrops);
    }
    const int nb = (padded + dwidth - 1) / dwidth;
    const int lo = (int)(T.cell_off >> kTileBits);
    const int tn = (int)((T.cell_off + (long long)T.nx * T.ny - 1) >> kTileBits) - lo + 1;
    if (t) {
      Terrain& o = (*t)[k];
      o.part_off = (long long)off;
      o.part_first = blocks;
      o.part_blocks = nb;
      o.part_dwidth = dwidth;
      o.tile_lo = lo;
      o.tile_n = tn;
      o.order_first = nchunk;
    }
    off += (size_t)tn * nb;
    blocks += nb;
    window = std::max(window, tn);
    nchunk += (padded + kOrderChunk - 1) / kOrderChunk;
  }
  if (chunks) *chunks = nchunk;
  return off;
}

// Size the logs, the partition and the per-drop buffers for the current
// drops (never inside a capture).
int ensure_wave(shy_ctx* c) {
  const size_t need = wave_steps(c);
  const size_t need_d = (size_t)std::max(c->total_drops, 1);
  if (need > (size_t)INT32_MAX - 16384)  // slot indices plus a round stay in int
    return fail(c, SHY_ERR_ARG, "drops x (maxAge+1) exceeds 2^31 log slots per wave");
  const int drops = std::max(c->total_drops, 1);
  const int rows = (int)(need / (size_t)drops);
  // The walk itself counts steps per tile (an LDS histogram per block) when
  // its blocks can be the partition blocks: one drop per thread (no
  // persistent grid), blocks inside one terrain (one terrain, or a block
  // size dividing the padded drop ranges), every terrain's tiles in its LDS
  // and count matrices of modest size.
  const int bs = c->tune.walk_block;
  int nb = 0, window = 0;
  const bool fit = c->ter.size() == 1 || kDropAlign % bs == 0;
  const bool walk_counts = c->total_drops > 0 && c->tune.walk_grid == 0 && fit &&
                           part_geometry(c, bs, nullptr, nb, window) <= ((size_t)1 << 22) &&
                           window <= kFuseMaxTiles;
  std::vector<Terrain> geo = c->ter;
  const size_t entries = part_geometry(c, walk_counts ? bs : 0, &geo, nb, window, &c->order_chunks);
  if (memcmp(geo.data(), c->ter.data(), geo.size() * sizeof(Terrain)) != 0) {
    drop_graph(c);
    c->ter = geo;
    if (c->d_ter) {
      HIPCHK(c, hipMemcpyAsync(c->d_ter, c->ter.data(), c->ter.size() * sizeof(Terrain),
                               hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
    }
  }
  if (walk_counts != c->walk_counts || nb != c->part_blocks || rows != c->part_rows ||
      entries != c->part_entries)
    drop_graph(c);  // a captured graph holds the old geometry
  c->walk_counts = walk_counts;
  c->part_blocks = nb;
  c->part_entries = entries;
  c->part_rows = rows;
  c->part_window = window;
  c->part_maxw = 0;
  for (const Terrain& t : c->ter) c->part_maxw = std::max(c->part_maxw, t.part_dwidth);
  const size_t hist_need = entries + 1;
  const int qstride = (rows + kReplayDepth - 1) / kReplayDepth * kReplayDepth;
  const size_t qb_need = (size_t)c->total_drops * qstride;
  if (qstride != c->qstride) drop_graph(c);
  c->qstride = qstride;
  {
    // the volume-track table: rows = max maxAge + 1 over the terrains
    float age = 0.0f;
    for (const Terrain& t : c->ter) age = std::max(age, t.P.maxAge);
    const int rows = (int)age + 1;
    const size_t vq_need = (size_t)rows * c->ter.size();
    if (vq_need > c->vq_cap) {
      drop_graph(c);
      HIPCHK(c, hipStreamSynchronize(c->stream));
      if (c->vq) hipFree(c->vq);
      c->vq = nullptr;
      c->vq_cap = 0;
      HIPCHK(c, hipMalloc(&c->vq, vq_need * sizeof(int32_t)));
      c->vq_cap = vq_need;
      c->vq_dirty = true;
    }
    if (c->vq_dirty || rows != c->vq_rows) {
      if (rows != c->vq_rows) drop_graph(c);
      c->vq_rows = rows;
      const int nter = (int)c->ter.size();
      volq_kernel<<<(nter + 63) / 64, 64, 0, c->stream>>>(c->d_ter, nter, rows, c->vq);
      HIPCHK(c, hipGetLastError());
      c->vq_dirty = false;
    }
  }
  {
    // split each tile's entries over several blocks while the per-part
    // buffers stay modest (up to 2048 tiles: 640 MiB of partial sums)
    const bool split = c->ntiles <= kTileSplitMaxTiles;
    if (split != c->tile_split || (split && !c->tcount)) {
      drop_graph(c);
      HIPCHK(c, hipStreamSynchronize(c->stream));
      if (c->tcount) hipFree(c->tcount);
      if (c->tsum) hipFree(c->tsum);
      c->tcount = nullptr;
      c->tsum = nullptr;
      c->tile_split = false;
      if (split) {
        const size_t cells = (size_t)c->ntiles * kTileCells;
        HIPCHK(c, hipMalloc(&c->tcount, cells * kCountSplit * sizeof(int32_t)));
        HIPCHK(c, hipMalloc(&c->tsum, 4 * cells * kReduceSplit * sizeof(long long)));
        c->tile_split = true;
      }
    }
  }
  if (c->ntiles > c->tstart_cap) {
    drop_graph(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->tstart) hipFree(c->tstart);
    c->tstart = nullptr;
    c->tstart_cap = 0;
    HIPCHK(c, hipMalloc(&c->tstart, ((size_t)c->ntiles + 1) * sizeof(int32_t)));
    c->tstart_cap = c->ntiles;
  }
  if (need <= c->wave_cap && need_d <= c->dlog_cap && hist_need <= c->hist_cap &&
      qb_need <= c->qb_cap)
    return SHY_OK;
  drop_graph(c);  // a captured graph holds the old buffers
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (need > c->wave_cap) {
    void* old[] = {c->slog, c->strk, c->ent};
    for (void* q : old)
      if (q) hipFree(q);
    c->slog = c->strk = nullptr;
    c->ent = nullptr;
    c->wave_cap = 0;
    HIPCHK(c, hipMalloc(&c->slog, need * sizeof(int2)));
    HIPCHK(c, hipMalloc(&c->strk, need * sizeof(int2)));
    HIPCHK(c, hipMalloc(&c->ent, need * sizeof(int4)));
    c->wave_cap = need;
  }
  if (qb_need > c->qb_cap) {
    if (c->qb) hipFree(c->qb);
    c->qb = nullptr;
    c->qb_cap = 0;
    HIPCHK(c, hipMalloc(&c->qb, qb_need * sizeof(int32_t)));
    c->qb_cap = qb_need;
  }
  if (need_d > c->dlog_cap) {
    if (c->dlog) hipFree(c->dlog);
    if (c->wsteps) hipFree(c->wsteps);
    if (c->dperm) hipFree(c->dperm);
    c->dlog = nullptr;
    c->wsteps = nullptr;
    c->dperm = nullptr;
    c->dlog_cap = 0;
    HIPCHK(c, hipMalloc(&c->dlog, need_d * sizeof(int4)));
    HIPCHK(c, hipMalloc(&c->wsteps, need_d * sizeof(int32_t)));
    HIPCHK(c, hipMalloc(&c->dperm, need_d * sizeof(int32_t)));
    c->dlog_cap = need_d;
  }
  if (hist_need > c->hist_cap) {
    if (c->hist) hipFree(c->hist);
    if (c->hbsum) hipFree(c->hbsum);
    c->hist = c->hbsum = nullptr;
    c->hist_cap = 0;
    HIPCHK(c, hipMalloc(&c->hist, hist_need * sizeof(int32_t)));
    HIPCHK(c, hipMalloc(&c->hbsum, (hist_need / kScanBlock + 1) * sizeof(int32_t)));
    c->hist_cap = hist_need;
  }
  return SHY_OK;
}

WaveArgs wave_args(shy_ctx* c, int p) {
  WaveArgs w;
  w.ter = c->d_ter;
  w.nter = (int)c->ter.size();
  w.total_drops = c->total_drops;
  w.h = c->h[p];
  w.fld = c->fld;
  w.root = c->root;
  w.first = c->first;
  w.slog = c->slog;
  w.strk = c->strk;
  w.qb = c->qb;
  w.sr = c->sr;
  w.qstride = c->qstride;
  w.dlog = c->dlog;
  w.steps = c->wsteps;
  w.hist = c->hist;
  w.perm = ordered(c) ? c->dperm : nullptr;
  w.pack = c->pack;
  w.deposit = c->deposit;
  w.visit = c->visit[p];
  w.state = c->state;
  w.gflag = c->gflag;
  w.tsteps = c->tsteps;
  return w;
}

int launch_cascade(shy_ctx* c, const CascArgs& a) {
  switch (c->tune.tile) {
    case 1: cascade_blend_kernel<64, 4><<<c->total_tiles, 256, 0, c->stream>>>(a); break;
    case 2: cascade_blend_kernel<16, 16><<<c->total_tiles, 256, 0, c->stream>>>(a); break;
    default: cascade_blend_kernel<32, 8><<<c->total_tiles, 256, 0, c->stream>>>(a); break;
  }
  HIPCHK(c, hipGetLastError());
  return SHY_OK;
}

// One cycle of every terrain.  ev, when given, receives events after the
// walk-record pack, the count walk, the partition + tile counts and the
// commit replay.  ensure_wave(c) must have sized the wave buffers.
int launch_cycle(shy_ctx* c, int p, hipEvent_t* ev = nullptr) {
  WaveArgs w = wave_args(c, p);
  const int bs = c->tune.walk_block;
  const int need = (c->total_drops + bs - 1) / bs;
  const int grid = c->tune.walk_grid > 0 ? std::min(c->tune.walk_grid, need) : need;
  const size_t n = c->n;
  const int nter = (int)c->ter.size();
  const bool walk = c->total_drops > 0;
  if (walk) {
    const size_t nrec = (size_t)c->pack_recs;
    pack_kernel<<<(unsigned)((nrec + 255) / 256), 256, 0, c->stream>>>(c->d_ter, nter, nrec,
                                                                        c->h[p], c->fld, c->pack,
                                                                        c->gflag);
    if (ordered(c))
      order_kernel<<<c->order_chunks, 1024, 0, c->stream>>>(c->d_ter, nter, c->total_drops, bs, c->state,
                                                 c->dperm);
    HIPCHK(c, hipGetLastError());
  }
  if (ev) HIPCHK(c, hipEventRecord(ev[0], c->stream));
  if (walk) {
    // the record prefetch pays where every drop runs at once (the walk is
    // bound by its chain's latency): the same waves the drop order leaves off
    const bool pf = !ordered(c);
    if (c->walk_counts)
      (pf ? walk_kernel<true, true> : walk_kernel<true, false>)<<<grid, bs, 0, c->stream>>>(w);
    else
      (pf ? walk_kernel<false, true> : walk_kernel<false, false>)<<<grid, bs, 0, c->stream>>>(w);
    HIPCHK(c, hipGetLastError());
    c->dlog_valid = true;
  }
  if (ev) HIPCHK(c, hipEventRecord(ev[1], c->stream));
  // partition of the valid log slots by cell tile (empty without drops)
  PartArgs pa;
  pa.slog = c->slog;
  pa.strk = c->strk;
  pa.steps = c->wsteps;
  pa.ter = c->d_ter;
  pa.nter = nter;
  pa.hist = c->hist;
  pa.ecell = reinterpret_cast<int32_t*>(c->ent);
  pa.erest = reinterpret_cast<EntRest*>(reinterpret_cast<int32_t*>(c->ent) + c->wave_cap);
  pa.rows = walk ? c->part_rows : 0;
  pa.drops = std::max(c->total_drops, 1);
  pa.tb = kTileBits;
  const size_t nh = c->part_entries;
  const int nbh = (int)((nh + kScanBlock - 1) / kScanBlock);
  const int nb = c->part_blocks;
  if (nh > 0) {
    if (!c->walk_counts) {
      if (c->part_window <= 4096)
        part_hist_kernel<4096><<<nb, 1024, 0, c->stream>>>(pa);
      else
        part_hist_kernel<kPartMaxTiles><<<nb, 1024, 0, c->stream>>>(pa);
    }
    scan_sums_kernel<<<nbh, kScanBlock, 0, c->stream>>>(c->hist, nh, c->hbsum);
    scan_top_kernel<<<1, kScanBlock, 0, c->stream>>>(c->hbsum, nbh);
    scan_apply_kernel<<<nbh, kScanBlock, 0, c->stream>>>(c->hist, c->hbsum, nh);
    const int mw = c->part_maxw;
    if (c->part_window <= 1024 && nb <= kBigRoundBlocks) {
      if (mw <= scatter_lds_drops<1024, 8192>())
        part_scatter_staged_kernel<1024, 8192, true><<<nb, 1024, 0, c->stream>>>(pa);
      else
        part_scatter_staged_kernel<1024, 8192, false><<<nb, 1024, 0, c->stream>>>(pa);
    } else if (c->part_window <= 1024) {
      if (mw <= scatter_lds_drops<1024, 4096>())
        part_scatter_staged_kernel<1024, 4096, true><<<nb, 1024, 0, c->stream>>>(pa);
      else
        part_scatter_staged_kernel<1024, 4096, false><<<nb, 1024, 0, c->stream>>>(pa);
    } else if (c->part_window <= 4096) {
      if (mw <= scatter_lds_drops<4096, 4096>())
        part_scatter_staged_kernel<4096, 4096, true><<<nb, 1024, 0, c->stream>>>(pa);
      else
        part_scatter_staged_kernel<4096, 4096, false><<<nb, 1024, 0, c->stream>>>(pa);
    } else
      part_scatter_kernel<kPartMaxTiles><<<nb, 1024, 0, c->stream>>>(pa);
  }
  tile_start_kernel<<<(c->ntiles + 256) / 256, 256, 0, c->stream>>>(
      c->d_ter, nter, c->ntiles, nh > 0 ? c->hist : nullptr, (long long)nh, c->tstart);
  if (c->tile_split) {
    tile_count_kernel<kCountSplit><<<c->ntiles * kCountSplit, 1024, 0, c->stream>>>(
        c->tstart, pa.ecell, n, c->sr, c->first, c->visit[p], c->fld, c->root, c->d_ter, nter,
        c->tcount);
    tile_count_finish_kernel<kCountSplit><<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(
        c->tcount, n, c->sr, c->first, c->visit[p], c->fld, c->root, c->d_ter, nter);
  } else {
    tile_count_kernel<1><<<c->ntiles, 1024, 0, c->stream>>>(c->tstart, pa.ecell, n, c->sr,
                                                            c->first, c->visit[p], c->fld,
                                                            c->root, c->d_ter, nter, nullptr);
  }
  HIPCHK(c, hipGetLastError());
  if (ev) HIPCHK(c, hipEventRecord(ev[2], c->stream));
  if (walk) {
    bool root = false;
    for (const Terrain& T : c->ter) root = root || T.has_root;
    const unsigned nb = (unsigned)((c->total_drops + kReplayBlock - 1) / kReplayBlock);
    if (root)
      replay_kernel<true><<<nb, kReplayBlock, 0, c->stream>>>(w);
    else
      replay_kernel<false><<<nb, kReplayBlock, 0, c->stream>>>(w);
    HIPCHK(c, hipGetLastError());
  }
  if (ev) HIPCHK(c, hipEventRecord(ev[3], c->stream));
  if (c->tile_split) {
    tile_reduce_kernel<kReduceSplit><<<c->ntiles * kReduceSplit, 1024, 0, c->stream>>>(
        c->tstart, pa.ecell, pa.erest, c->qb, c->qstride, c->vq, c->vq_rows, pa.drops, c->d_ter, nter,
        c->deposit, n, c->dh, c->trk, c->trk + n, c->trk + 2 * n, c->tsum);
    tile_reduce_finish_kernel<kReduceSplit><<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(
        c->tsum, c->deposit, n, c->dh, c->trk, c->trk + n, c->trk + 2 * n);
  } else {
    tile_reduce_kernel<1><<<c->ntiles, 1024, 0, c->stream>>>(
        c->tstart, pa.ecell, pa.erest, c->qb, c->qstride, c->vq, c->vq_rows, pa.drops, c->d_ter, nter,
        c->deposit, n, c->dh, c->trk, c->trk + n, c->trk + 2 * n, nullptr);
  }
  HIPCHK(c, hipGetLastError());
  CascArgs a;
  const int q = p ^ 1;
  a.ter = c->d_ter;
  a.nter = nter;
  a.h_in = c->h[p];
  a.h_out = c->h[q];
  a.dh = c->dh;
  a.visit = c->visit[p];
  a.td = c->trk;
  a.tmx = c->trk + n;
  a.tmy = c->trk + 2 * n;
  a.visit_clear = c->visit[q];
  a.fld = c->fld;
  a.erf_a = c->erf_a;
  a.state = c->state;
  a.gflag = c->gflag;
  return launch_cascade(c, a);
}

int refresh_erf(shy_ctx* c) {
  const int bs = 256;
  const size_t grid = (c->n + bs - 1) / bs;
  refresh_erf_kernel<<<grid, bs, 0, c->stream>>>(c->fld, c->n, c->erf_a);
  HIPCHK(c, hipGetLastError());
  return SHY_OK;
}

bool all_finite(const float* v, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(v[i])) return false;
  return true;
}

int check_terrain(shy_ctx* c, int t) {
  if (!c) return SHY_ERR_ARG;
  if (t < 0 || t >= (int)c->ter.size()) return fail(c, SHY_ERR_ARG, "terrain index out of range");
  return SHY_OK;
}

int need_single(shy_ctx* c, const char* what) {
  if (!c) return SHY_ERR_ARG;
  if (c->ter.size() != 1)
    return fail(c, SHY_ERR_STATE, std::string(what) + ": only for single-terrain contexts");
  return SHY_OK;
}

// One cycle eagerly (wave buffers sized first).
int erode_once(shy_ctx* c, hipEvent_t* ev = nullptr) {
  RETURN_IF(ensure_wave(c));
  RETURN_IF(launch_cycle(c, parity(c), ev));
  c->cycle += 1;
  return SHY_OK;
}

int erode_many(shy_ctx* c, int cycles, bool use_graph) {
  if (!use_graph) {
    for (int k = 0; k < cycles; ++k) RETURN_IF(erode_once(c));
    return SHY_OK;
  }
  if (cycles > 0 && parity(c) == 1) {
    RETURN_IF(erode_once(c));
    --cycles;
  }
  if (cycles >= 2) {
    RETURN_IF(ensure_wave(c));
    if (!c->graph) {
      hipGraph_t g;
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int r0 = launch_cycle(c, 0);
      int r1 = r0 ? r0 : launch_cycle(c, 1);
      hipError_t e = hipStreamEndCapture(c->stream, &g);
      if (r1) return r1;
      if (e != hipSuccess)
        return fail(c, SHY_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      if (e != hipSuccess) {
        c->graph = nullptr;
        return fail(c, SHY_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      }
    }
    for (int k = 0; k + 1 < cycles; k += 2) {
      HIPCHK(c, hipGraphLaunch(c->graph, c->stream));
      c->cycle += 2;
    }
  }
  if (cycles & 1) return erode_once(c);
  return SHY_OK;
}

int time_cycles(shy_ctx* c, int cycles, float* ms) {
  RETURN_IF(ensure_wave(c));
  constexpr int kSlots = 5;  // pack, count walk, partition + tile counts, replay, reduce + stencil
  const size_t ne = kSlots * (size_t)cycles + 1;
  std::vector<hipEvent_t> ev(ne, nullptr);
  auto cleanup = [&]() {
    for (hipEvent_t e : ev)
      if (e) hipEventDestroy(e);
  };
  for (auto& e : ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      cleanup();
      return fail(c, SHY_ERR_HIP, "shy_time_cycles: hipEventCreate failed");
    }
  }
  hipError_t e = hipEventRecord(ev[0], c->stream);
  for (int k = 0; k < cycles && e == hipSuccess; ++k) {
    int r = launch_cycle(c, parity(c), &ev[kSlots * k + 1]);
    if (r) {
      cleanup();
      return r;
    }
    c->cycle += 1;
    e = hipEventRecord(ev[kSlots * k + kSlots], c->stream);
  }
  if (e == hipSuccess) e = hipEventSynchronize(ev.back());
  float part[kSlots] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  for (int k = 0; k < cycles && e == hipSuccess; ++k) {
    for (int s = 0; s < kSlots && e == hipSuccess; ++s) {
      float t = 0.0f;
      e = hipEventElapsedTime(&t, ev[kSlots * k + s], ev[kSlots * k + s + 1]);
      part[s] += t;
    }
  }
  float tot = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&tot, ev[0], ev.back());
  cleanup();
  if (e != hipSuccess) return fail(c, SHY_ERR_HIP, std::string("shy_time_cycles: ") + hipGetErrorString(e));
  ms[0] = tot;
  for (int s = 0; s < kSlots; ++s) ms[1 + s] = part[s];
  return SHY_OK;
}

// device scratch of `bytes`, a kernel filling it, one copy to the host
template <class Launch>
int download_via(shy_ctx* c, const char* what, size_t bytes, void* host, Launch launch) {
  void* d = nullptr;
  HIPCHK(c, hipMalloc(&d, bytes));
  launch(d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(host, d, bytes, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(d);
  if (e != hipSuccess) return fail(c, SHY_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return SHY_OK;
}

// node-major pool order of a mapsize x mapsize node grid (cellpool.h:323-340,
// 421-426): node (i, j) is index i*mapsize + j, its tileres^2 cells x-major
int check_tiled(shy_ctx* c, int tileres, const char* what) {
  const Terrain& T = c->ter[0];
  if (tileres < 1 || T.nx != T.ny || T.nx % tileres)
    return fail(c, SHY_ERR_ARG, std::string(what) +
                                    ": the map must be square with a side a multiple of tileres");
  return SHY_OK;
}

size_t tiled_index(int x, int y, int tileres, int mapsize) {
  const int i = x / tileres, j = y / tileres;
  const size_t node = (size_t)i * mapsize + j;
  return node * (size_t)tileres * tileres + (size_t)(x % tileres) * tileres + (y % tileres);
}

}  // namespace

extern "C" {

void shy_default_params(shy_params* p) {
  if (!p) return;
  p->evapRate = 0.001f;
  p->depositionRate = 0.1f;
  p->minVol = 0.01f;
  p->maxAge = 500.0f;
  p->entrainment = 10.0f;
  p->gravity = 1.0f;
  p->momentumTransfer = 1.0f;
  p->lrate = 0.1f;
  p->maxdiff = 0.01f;
  p->settling = 0.8f;
}

int shy_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char* shy_last_error(const shy_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_noctx_err.c_str();
}

int shy_destroy(shy_ctx* c) {
  if (!c) return SHY_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  drop_graph(c);
  void* ptrs[] = {c->h[0], c->h[1], c->fld, c->root, c->visit[0], c->visit[1], c->first,
                  c->sr, c->dh, c->trk, c->deposit, c->slog, c->strk, c->qb, c->ent, c->gflag,
                  c->vq, c->tcount, c->tsum, c->hist, c->hbsum, c->tstart, c->dlog, c->wsteps, c->dperm, c->tsteps, c->veg_px, c->veg_py, c->veg_alive, c->veg_cflag,
                  c->veg_cx, c->veg_cy, c->veg_units, c->veg_st, c->veg_size, c->pack, c->d_ter, c->state, c->erf_a, c->part, c->perms, c->trace_buf,
                  c->trace_n, c->drop_buf};
  for (void* q : ptrs)
    if (q) hipFree(q);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return SHY_OK;
}

int shy_ensemble_create(const shy_terrain_desc* desc, int nter, shy_ctx** out) {
  if (!out) return fail(nullptr, SHY_ERR_ARG, "shy_ensemble_create: out is NULL");
  *out = nullptr;
  if (!desc || nter < 1) return fail(nullptr, SHY_ERR_ARG, "shy_ensemble_create: need >= 1 terrain");
  long long cells = 0, recs = 0;  // cells and packed walk records
  for (int t = 0; t < nter; ++t) {
    const shy_terrain_desc& d = desc[t];
    if (d.nx < 1 || d.ny < 1 || (long long)d.nx * d.ny > (1LL << 31) - 1 || d.drops < 0)
      return fail(nullptr, SHY_ERR_ARG,
                  "shy_create: nx, ny must be >= 1 with nx*ny < 2^31, drops >= 0");
    if (!params_ok(&d.params)) return fail(nullptr, SHY_ERR_ARG, std::string("shy_create: ") + kParamsMsg);
    cells += (long long)d.nx * d.ny;
    recs += pack_records(d.nx, d.ny);
  }
  if (cells > (1LL << 31) - 1 || recs > (1LL << 31) - 1)
    return fail(nullptr, SHY_ERR_ARG, "shy_ensemble_create: total cells must stay below 2^31");
  if (shy_device_count() < 1) return fail(nullptr, SHY_ERR_NODEV, "shy_create: no GPU visible");
  shy_ctx* c = new shy_ctx();
  c->ter.resize(nter);
  for (int t = 0; t < nter; ++t) {
    Terrain& T = c->ter[t];
    std::memset(&T, 0, sizeof(T));
    T.nx = desc[t].nx;
    T.ny = desc[t].ny;
    T.drops = desc[t].drops;
    T.seed = desc[t].seed;
    T.id = (unsigned)t;
    T.has_root = 0;
    T.P = desc[t].params;
  }
  c->n = (size_t)cells;
  const size_t n = c->n;
  auto bail = [&](hipError_t e, const char* what) {
    g_noctx_err = std::string("shy_create: ") + what + ": " + hipGetErrorString(e);
    shy_destroy(c);
    return SHY_ERR_HIP;
  };
  hipError_t e;
  if ((e = hipGetDevice(&c->device)) != hipSuccess) return bail(e, "hipGetDevice");
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
    return bail(e, "hipStreamCreate");
  if (layout(c)) {
    g_noctx_err = c->err;
    shy_destroy(c);
    return SHY_ERR_ARG;
  }
  for (int k = 0; k < 2; ++k) {
    if ((e = hipMalloc(&c->h[k], n * sizeof(float))) != hipSuccess) return bail(e, "hipMalloc h");
    if ((e = hipMalloc(&c->visit[k], n)) != hipSuccess) return bail(e, "hipMalloc visit");
  }
  if ((e = hipMalloc(&c->fld, n * sizeof(float4))) != hipSuccess) return bail(e, "hipMalloc fld");
  if ((e = hipMalloc(&c->root, n * sizeof(float))) != hipSuccess) return bail(e, "hipMalloc root");
  if ((e = hipMalloc(&c->first, n * sizeof(int32_t))) != hipSuccess) return bail(e, "hipMalloc first");
  if ((e = hipMalloc(&c->sr, n * sizeof(float))) != hipSuccess) return bail(e, "hipMalloc sr");
  c->pack_recs = recs;
  if ((e = hipMalloc(&c->pack, 2 * (size_t)recs * sizeof(float4))) != hipSuccess)
    return bail(e, "hipMalloc pack");
  if ((e = hipMalloc(&c->dh, n * sizeof(int32_t))) != hipSuccess) return bail(e, "hipMalloc dh");
  if ((e = hipMalloc(&c->trk, 3 * n * sizeof(long long))) != hipSuccess) return bail(e, "hipMalloc trk");
  if ((e = hipMalloc(&c->deposit, n * sizeof(int32_t))) != hipSuccess) return bail(e, "hipMalloc deposit");
  if ((e = hipMalloc(&c->tsteps, nter * sizeof(long long))) != hipSuccess) return bail(e, "hipMalloc tsteps");
  if ((e = hipMalloc(&c->d_ter, nter * sizeof(Terrain))) != hipSuccess) return bail(e, "hipMalloc terrains");
  if ((e = hipMalloc(&c->gflag, nter * sizeof(int))) != hipSuccess) return bail(e, "hipMalloc gflag");
  if ((e = hipMalloc(&c->state, sizeof(DeviceState))) != hipSuccess) return bail(e, "hipMalloc state");
  if ((e = hipMalloc(&c->erf_a, kErfTerms * sizeof(double))) != hipSuccess) return bail(e, "hipMalloc erf");
  if ((e = hipMalloc(&c->part, kPartBlocks * sizeof(float2))) != hipSuccess) return bail(e, "hipMalloc part");
  if ((e = hipMalloc(&c->perms, 16 * 256)) != hipSuccess) return bail(e, "hipMalloc perms");
  if ((e = hipMalloc(&c->trace_n, sizeof(int))) != hipSuccess) return bail(e, "hipMalloc trace_n");
  for (int k = 0; k < 2; ++k) {
    hipMemsetAsync(c->h[k], 0, n * sizeof(float), c->stream);
    hipMemsetAsync(c->visit[k], 0, n, c->stream);
  }
  hipMemsetAsync(c->fld, 0, n * sizeof(float4), c->stream);
  hipMemsetAsync(c->root, 0, n * sizeof(float), c->stream);
  hipMemsetAsync(c->first, 0, n * sizeof(int32_t), c->stream);
  hipMemsetAsync(c->sr, 0, n * sizeof(float), c->stream);
  hipMemsetAsync(c->dh, 0, n * sizeof(int32_t), c->stream);
  hipMemsetAsync(c->trk, 0, 3 * n * sizeof(long long), c->stream);
  hipMemsetAsync(c->deposit, 0, n * sizeof(int32_t), c->stream);
  hipMemsetAsync(c->tsteps, 0, nter * sizeof(long long), c->stream);
  hipMemsetAsync(c->gflag, 0, nter * sizeof(int), c->stream);
  DeviceState st{};
  st.cycle = 0;
  double coef[kErfTerms];
  erf_coefficients(coef);
  hipMemcpyAsync(c->state, &st, sizeof(st), hipMemcpyHostToDevice, c->stream);
  hipMemcpyAsync(c->erf_a, coef, sizeof(coef), hipMemcpyHostToDevice, c->stream);
  hipMemcpyAsync(c->d_ter, c->ter.data(), nter * sizeof(Terrain), hipMemcpyHostToDevice, c->stream);
  if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return bail(e, "init");
  *out = c;
  return SHY_OK;
}

int shy_create(const shy_params* p, int nx, int ny, uint64_t seed, shy_ctx** out) {
  shy_terrain_desc d;
  d.nx = nx;
  d.ny = ny;
  d.drops = 0;
  d.seed = seed;
  if (p) d.params = *p; else shy_default_params(&d.params);
  return shy_ensemble_create(&d, 1, out);
}

int shy_ensemble_count(const shy_ctx* c, int* nter) {
  if (!c || !nter) return SHY_ERR_ARG;
  *nter = (int)c->ter.size();
  return SHY_OK;
}

int shy_terrain_size(const shy_ctx* c, int t, int* nx, int* ny, int* drops) {
  if (!c || !nx || !ny || !drops || t < 0 || t >= (int)c->ter.size()) return SHY_ERR_ARG;
  *nx = c->ter[t].nx;
  *ny = c->ter[t].ny;
  *drops = c->ter[t].drops;
  return SHY_OK;
}

int shy_size(const shy_ctx* c, int* nx, int* ny) {
  if (!c || !nx || !ny) return SHY_ERR_ARG;
  *nx = c->ter[0].nx;
  *ny = c->ter[0].ny;
  return SHY_OK;
}

int shy_terrain_set_params(shy_ctx* c, int t, const shy_params* p) {
  RETURN_IF(check_terrain(c, t));
  if (!p) return fail(c, SHY_ERR_ARG, "shy_set_params: NULL argument");
  if (!params_ok(p)) return fail(c, SHY_ERR_ARG, std::string("shy_set_params: ") + kParamsMsg);
  hipSetDevice(c->device);
  c->ter[t].P = *p;
  return layout(c);
}

int shy_set_params(shy_ctx* c, const shy_params* p) {
  if (!c) return SHY_ERR_ARG;
  if (!p) return fail(c, SHY_ERR_ARG, "shy_set_params: NULL argument");
  if (!params_ok(p)) return fail(c, SHY_ERR_ARG, std::string("shy_set_params: ") + kParamsMsg);
  for (Terrain& t : c->ter) t.P = *p;
  hipSetDevice(c->device);
  return layout(c);
}

int shy_get_params(const shy_ctx* c, shy_params* p) {
  if (!c || !p) return SHY_ERR_ARG;
  *p = c->ter[0].P;
  return SHY_OK;
}

int shy_terrain_set_drops(shy_ctx* c, int t, int drops) {
  RETURN_IF(check_terrain(c, t));
  if (drops < 0) return fail(c, SHY_ERR_ARG, "drops must be >= 0");
  if (c->ter[t].drops == drops) return SHY_OK;
  hipSetDevice(c->device);
  c->ter[t].drops = drops;
  return layout(c);
}

int shy_set_cycle(shy_ctx* c, uint32_t cycle) {
  if (!c) return SHY_ERR_ARG;
  hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if ((cycle & 1u) != (c->cycle & 1u)) {
    // parity p: h[p] live, visit[p] the zero accumulator.  Move to q = p^1.
    const int p = parity(c), q = p ^ 1;
    HIPCHK(c, hipMemcpyAsync(c->h[q], c->h[p], c->n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->visit[q], 0, c->n, c->stream));
  }
  c->cycle = cycle;
  HIPCHK(c, hipMemcpyAsync(&c->state->cycle, &cycle, sizeof(cycle), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graph(c);
  return SHY_OK;
}

int shy_set_seed(shy_ctx* c, uint64_t seed, uint32_t cycle) {
  RETURN_IF(need_single(c, "shy_set_seed"));
  c->ter[0].seed = seed;
  RETURN_IF(layout(c));
  return shy_set_cycle(c, cycle);
}

int shy_get_cycle(const shy_ctx* c, uint32_t* cycle) {
  if (!c || !cycle) return SHY_ERR_ARG;
  *cycle = c->cycle;
  return SHY_OK;
}

int shy_terrain_upload_field(shy_ctx* c, int t, int field, const float* src) {
  RETURN_IF(check_terrain(c, t));
  if (!src) return fail(c, SHY_ERR_ARG, "shy_upload_field: NULL argument");
  if (field < 0 || field > SHY_ROOTDENSITY)
    return fail(c, SHY_ERR_ARG, "shy_upload_field: field must be a quad::cell member (0..7)");
  Terrain& T = c->ter[t];
  const size_t n = (size_t)T.nx * T.ny, o = (size_t)T.cell_off;
  if (!all_finite(src, n)) return fail(c, SHY_ERR_NONFINITE, "shy_upload_field: NaN or Inf in input");
  hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int p = parity(c);
  switch (field) {
    case SHY_HEIGHT:
      HIPCHK(c, hipMemcpy(c->h[p] + o, src, n * sizeof(float), hipMemcpyHostToDevice));
      break;
    case SHY_DISCHARGE:
    case SHY_MOMENTUMX:
    case SHY_MOMENTUMY: {
      std::vector<float4> f(n);
      HIPCHK(c, hipMemcpy(f.data(), c->fld + o, n * sizeof(float4), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; ++i) {
        if (field == SHY_DISCHARGE) f[i].z = src[i];
        else if (field == SHY_MOMENTUMX) f[i].x = src[i];
        else f[i].y = src[i];
      }
      HIPCHK(c, hipMemcpy(c->fld + o, f.data(), n * sizeof(float4), hipMemcpyHostToDevice));
      if (field == SHY_DISCHARGE) RETURN_IF(refresh_erf(c));
      break;
    }
    case SHY_DISCHARGE_TRACK:
    case SHY_MOMENTUMX_TRACK:
    case SHY_MOMENTUMY_TRACK: {
      // tracks are per-cycle scratch (world.h:56-61): they stay visible until
      // the next cycle overwrites them, as the reference's do
      std::vector<long long> q(n);
      for (size_t i = 0; i < n; ++i) q[i] = (long long)std::llrint((double)src[i] * 16777216.0);
      HIPCHK(c, hipMemcpy(c->trk + (size_t)(field - SHY_DISCHARGE_TRACK) * c->n + o, q.data(),
                          n * sizeof(long long), hipMemcpyHostToDevice));
      break;
    }
    case SHY_ROOTDENSITY: {
      bool any = false;
      for (size_t i = 0; i < n && !any; ++i) any = src[i] != 0.0f;
      HIPCHK(c, hipMemcpy(c->root + o, src, n * sizeof(float), hipMemcpyHostToDevice));
      if ((int)any != T.has_root) {
        T.has_root = any ? 1 : 0;
        RETURN_IF(layout(c));
      }
      break;
    }
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SHY_OK;
}

int shy_terrain_download_field(shy_ctx* c, int t, int field, float* dst) {
  RETURN_IF(check_terrain(c, t));
  if (!dst) return fail(c, SHY_ERR_ARG, "shy_download_field: NULL argument");
  if (field < 0 || field > SHY_ERF_DISCHARGE)
    return fail(c, SHY_ERR_ARG, "shy_download_field: unknown field");
  const Terrain& T = c->ter[t];
  const size_t n = (size_t)T.nx * T.ny, o = (size_t)T.cell_off;
  hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int p = parity(c);
  switch (field) {
    case SHY_HEIGHT:
      HIPCHK(c, hipMemcpy(dst, c->h[p] + o, n * sizeof(float), hipMemcpyDeviceToHost));
      break;
    case SHY_DISCHARGE:
    case SHY_MOMENTUMX:
    case SHY_MOMENTUMY:
    case SHY_ERF_DISCHARGE: {
      std::vector<float4> f(n);
      HIPCHK(c, hipMemcpy(f.data(), c->fld + o, n * sizeof(float4), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; ++i)
        dst[i] = field == SHY_DISCHARGE ? f[i].z
                 : field == SHY_MOMENTUMX ? f[i].x
                 : field == SHY_MOMENTUMY ? f[i].y
                                          : f[i].w;
      break;
    }
    case SHY_DISCHARGE_TRACK:
    case SHY_MOMENTUMX_TRACK:
    case SHY_MOMENTUMY_TRACK: {
      std::vector<long long> q(n);
      HIPCHK(c, hipMemcpy(q.data(), c->trk + (size_t)(field - SHY_DISCHARGE_TRACK) * c->n + o,
                          n * sizeof(long long), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; ++i) dst[i] = (float)((double)q[i] * kTQInv);
      break;
    }
    case SHY_ROOTDENSITY:
      HIPCHK(c, hipMemcpy(dst, c->root + o, n * sizeof(float), hipMemcpyDeviceToHost));
      break;
  }
  return SHY_OK;
}

int shy_terrain_export(shy_ctx* c, int t, float* dst4) {
  RETURN_IF(check_terrain(c, t));
  if (!dst4) return fail(c, SHY_ERR_ARG, "shy_terrain_export: NULL argument");
  const Terrain& T = c->ter[t];
  const size_t n = (size_t)T.nx * T.ny;
  hipSetDevice(c->device);
  export_fields_kernel<<<(n + 255) / 256, 256, 0, c->stream>>>(c->h[parity(c)] + T.cell_off,
                                                               c->fld + T.cell_off, n, dst4);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SHY_OK;
}

int shy_terrain_import_flow(shy_ctx* c, int t, const float* src3) {
  RETURN_IF(check_terrain(c, t));
  if (!src3) return fail(c, SHY_ERR_ARG, "shy_terrain_import_flow: NULL argument");
  const Terrain& T = c->ter[t];
  const size_t n = (size_t)T.nx * T.ny;
  hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  import_flow_kernel<<<(n + 255) / 256, 256, 0, c->stream>>>(src3, c->fld + T.cell_off, n,
                                                             c->erf_a);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SHY_OK;
}

int shy_upload_field(shy_ctx* c, int field, const float* src) {
  RETURN_IF(need_single(c, "shy_upload_field"));
  return shy_terrain_upload_field(c, 0, field, src);
}

int shy_download_field(shy_ctx* c, int field, float* dst) {
  RETURN_IF(need_single(c, "shy_download_field"));
  return shy_terrain_download_field(c, 0, field, dst);
}

int shy_download_normals(shy_ctx* c, float* xyz) {
  RETURN_IF(need_single(c, "shy_download_normals"));
  if (!xyz) return fail(c, SHY_ERR_ARG, "shy_download_normals: NULL argument");
  hipSetDevice(c->device);
  const Terrain& T = c->ter[0];
  float* d = nullptr;
  HIPCHK(c, hipMalloc((void**)&d, 3 * c->n * sizeof(float)));
  const size_t grid = (c->n + 255) / 256;
  normals_kernel<<<grid, 256, 0, c->stream>>>(c->h[parity(c)], T.nx, T.ny, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(xyz, d, 3 * c->n * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(d);
  if (e != hipSuccess) return fail(c, SHY_ERR_HIP, std::string("shy_download_normals: ") + hipGetErrorString(e));
  return SHY_OK;
}

int shy_download_textures(shy_ctx* c, const float* water_rgb, float* discharge_rgba,
                          float* momentum_rgba) {
  RETURN_IF(need_single(c, "shy_download_textures"));
  if (!discharge_rgba && !momentum_rgba) return SHY_OK;
  if (discharge_rgba && !water_rgb)
    return fail(c, SHY_ERR_ARG, "shy_download_textures: water_rgb is NULL");
  hipSetDevice(c->device);
  const size_t n = c->n, grid = (n + 255) / 256;
  const float wr = water_rgb ? water_rgb[0] : 0.0f, wg = water_rgb ? water_rgb[1] : 0.0f,
              wb = water_rgb ? water_rgb[2] : 0.0f;
  if (discharge_rgba)
    RETURN_IF(download_via(c, "shy_download_textures", n * sizeof(float4), discharge_rgba,
                           [&](void* d) {
                             textures_kernel<<<grid, 256, 0, c->stream>>>(
                                 c->fld, n, wr, wg, wb, c->erf_a, (float4*)d, nullptr);
                           }));
  if (momentum_rgba)
    RETURN_IF(download_via(c, "shy_download_textures", n * sizeof(float4), momentum_rgba,
                           [&](void* d) {
                             textures_kernel<<<grid, 256, 0, c->stream>>>(
                                 c->fld, n, 0.0f, 0.0f, 0.0f, c->erf_a, nullptr, (float4*)d);
                           }));
  return SHY_OK;
}

int shy_download_vertices(shy_ctx* c, float* vertex12) {
  RETURN_IF(need_single(c, "shy_download_vertices"));
  if (!vertex12) return fail(c, SHY_ERR_ARG, "shy_download_vertices: NULL argument");
  hipSetDevice(c->device);
  const Terrain& T = c->ter[0];
  const size_t grid = (c->n + 255) / 256;
  return download_via(c, "shy_download_vertices", 12 * c->n * sizeof(float), vertex12,
                      [&](void* d) {
                        vertices_kernel<<<grid, 256, 0, c->stream>>>(c->h[parity(c)], T.nx, T.ny,
                                                                    (float*)d);
                      });
}

int shy_save_map(shy_ctx* c, const char* path) {
  RETURN_IF(need_single(c, "shy_save_map"));
  if (!path) return fail(c, SHY_ERR_ARG, "shy_save_map: NULL path");
  shy_map_header hd;
  std::memset(&hd, 0, sizeof(hd));
  std::memcpy(hd.magic, SHY_MAP_MAGIC, 8);
  hd.version = SHY_MAP_VERSION;
  hd.nx = c->ter[0].nx;
  hd.ny = c->ter[0].ny;
  hd.cycle = c->cycle;
  hd.seed = c->ter[0].seed;
  hd.params = c->ter[0].P;
  std::vector<float> cells(8 * c->n);
  RETURN_IF(shy_download_cells(c, cells.data()));
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(c, SHY_ERR_IO, std::string("shy_save_map: cannot open ") + path);
  const bool ok = std::fwrite(&hd, sizeof(hd), 1, f) == 1 &&
                  std::fwrite(cells.data(), sizeof(float), cells.size(), f) == cells.size();
  const bool closed = std::fclose(f) == 0;
  if (!ok || !closed) return fail(c, SHY_ERR_IO, std::string("shy_save_map: write failed: ") + path);
  return SHY_OK;
}

int shy_load_map(shy_ctx* c, const char* path) {
  RETURN_IF(need_single(c, "shy_load_map"));
  if (!path) return fail(c, SHY_ERR_ARG, "shy_load_map: NULL path");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(c, SHY_ERR_IO, std::string("shy_load_map: cannot open ") + path);
  shy_map_header hd;
  std::vector<float> cells(8 * c->n);
  bool ok = std::fread(&hd, sizeof(hd), 1, f) == 1;
  if (ok && (std::memcmp(hd.magic, SHY_MAP_MAGIC, 8) != 0 || hd.version != SHY_MAP_VERSION)) {
    std::fclose(f);
    return fail(c, SHY_ERR_IO, std::string("shy_load_map: not a map dump: ") + path);
  }
  if (ok && (hd.nx != c->ter[0].nx || hd.ny != c->ter[0].ny)) {
    std::fclose(f);
    return fail(c, SHY_ERR_ARG, "shy_load_map: map size differs from the context's");
  }
  ok = ok && std::fread(cells.data(), sizeof(float), cells.size(), f) == cells.size();
  std::fclose(f);
  if (!ok) return fail(c, SHY_ERR_IO, std::string("shy_load_map: short file: ") + path);
  RETURN_IF(shy_set_params(c, &hd.params));
  RETURN_IF(shy_upload_cells(c, cells.data()));
  return shy_set_seed(c, hd.seed, hd.cycle);
}

int shy_download_cells_tiled(shy_ctx* c, int tileres, float* pool8) {
  RETURN_IF(need_single(c, "shy_download_cells_tiled"));
  if (!pool8) return fail(c, SHY_ERR_ARG, "shy_download_cells_tiled: NULL argument");
  RETURN_IF(check_tiled(c, tileres, "shy_download_cells_tiled"));
  const int n = c->ter[0].nx, mapsize = n / tileres;
  std::vector<float> cells(8 * c->n);
  RETURN_IF(shy_download_cells(c, cells.data()));
  for (int x = 0; x < n; ++x)
    for (int y = 0; y < n; ++y)
      std::memcpy(pool8 + 8 * tiled_index(x, y, tileres, mapsize),
                  cells.data() + 8 * ((size_t)x * n + y), 8 * sizeof(float));
  return SHY_OK;
}

int shy_upload_cells_tiled(shy_ctx* c, int tileres, const float* pool8) {
  RETURN_IF(need_single(c, "shy_upload_cells_tiled"));
  if (!pool8) return fail(c, SHY_ERR_ARG, "shy_upload_cells_tiled: NULL argument");
  RETURN_IF(check_tiled(c, tileres, "shy_upload_cells_tiled"));
  const int n = c->ter[0].nx, mapsize = n / tileres;
  std::vector<float> cells(8 * c->n);
  for (int x = 0; x < n; ++x)
    for (int y = 0; y < n; ++y)
      std::memcpy(cells.data() + 8 * ((size_t)x * n + y),
                  pool8 + 8 * tiled_index(x, y, tileres, mapsize), 8 * sizeof(float));
  return shy_upload_cells(c, cells.data());
}

void shy_default_plant_params(shy_plant_params* p) {
  if (!p) return;
  p->maxSize = 1.5f;
  p->growRate = 0.05f;
  p->maxSteep = 0.8f;
  p->maxDischarge = 0.3f;
  p->maxTreeHeight = 0.8f;
}

int shy_vegetation_enable(shy_ctx* c, int capacity, const shy_plant_params* p) {
  RETURN_IF(need_single(c, "shy_vegetation_enable"));
  if (capacity < 1) return fail(c, SHY_ERR_ARG, "shy_vegetation_enable: capacity must be >= 1");
  shy_plant_params P;
  shy_default_plant_params(&P);
  if (p) P = *p;
  const float v[] = {P.maxSize, P.growRate, P.maxSteep, P.maxDischarge, P.maxTreeHeight};
  for (float x : v)
    if (!std::isfinite(x)) return fail(c, SHY_ERR_ARG, "shy_vegetation_enable: non-finite parameter");
  hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (capacity != c->veg_cap) {
    void* old[] = {c->veg_px, c->veg_py, c->veg_alive, c->veg_cflag, c->veg_cx, c->veg_cy,
                   c->veg_size};
    for (void* q : old)
      if (q) hipFree(q);
    c->veg_px = c->veg_py = c->veg_alive = c->veg_cflag = c->veg_cx = c->veg_cy = nullptr;
    c->veg_size = nullptr;
    c->veg_cap = 0;
    const size_t b = (size_t)capacity * sizeof(int32_t);
    HIPCHK(c, hipMalloc(&c->veg_px, b));
    HIPCHK(c, hipMalloc(&c->veg_py, b));
    HIPCHK(c, hipMalloc(&c->veg_alive, b));
    HIPCHK(c, hipMalloc(&c->veg_cflag, b));
    HIPCHK(c, hipMalloc(&c->veg_cx, b));
    HIPCHK(c, hipMalloc(&c->veg_cy, b));
    HIPCHK(c, hipMalloc(&c->veg_size, (size_t)capacity * sizeof(float)));
    c->veg_cap = capacity;
  }
  if (!c->veg_units) HIPCHK(c, hipMalloc(&c->veg_units, c->n * sizeof(int32_t)));
  if (!c->veg_st) HIPCHK(c, hipMalloc(&c->veg_st, 4 * sizeof(int32_t)));
  HIPCHK(c, hipMemsetAsync(c->veg_units, 0, c->n * sizeof(int32_t), c->stream));
  HIPCHK(c, hipMemsetAsync(c->veg_st, 0, 4 * sizeof(int32_t), c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->veg_P = P;
  c->veg_step = 0;
  if (!c->ter[0].has_root) {  // the walk now reads the root density
    c->ter[0].has_root = 1;
    RETURN_IF(layout(c));
  }
  return SHY_OK;
}

static VegArgs veg_args(shy_ctx* c) {
  VegArgs a;
  a.h = c->h[parity(c)];
  a.fld = c->fld;
  a.root = c->root;
  a.units = c->veg_units;
  a.nx = c->ter[0].nx;
  a.ny = c->ter[0].ny;
  a.px = c->veg_px;
  a.py = c->veg_py;
  a.size = c->veg_size;
  a.alive = c->veg_alive;
  a.cflag = c->veg_cflag;
  a.cx = c->veg_cx;
  a.cy = c->veg_cy;
  a.st = c->veg_st;
  a.cap = c->veg_cap;
  a.P = c->veg_P;
  return a;
}

static int veg_state(shy_ctx* c, int32_t st[4]) {
  HIPCHK(c, hipMemcpyAsync(st, c->veg_st, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SHY_OK;
}

int shy_vegetation_grow(shy_ctx* c) {
  RETURN_IF(need_single(c, "shy_vegetation_grow"));
  if (!c->veg_cap) return fail(c, SHY_ERR_STATE, "shy_vegetation_grow: vegetation not enabled");
  hipSetDevice(c->device);
  const VegArgs a = veg_args(c);
  const uint64_t key = cycle_key(c->ter[0].seed, 0xffffffffu, c->veg_step);
  const size_t grid = (c->n + 255) / 256;
  veg_spawn_kernel<<<1, 64, 0, c->stream>>>(a, key);
  veg_apply_kernel<<<grid, 256, 0, c->stream>>>(c->root, c->veg_units, c->n);
  HIPCHK(c, hipGetLastError());
  int32_t st[4];
  RETURN_IF(veg_state(c, st));
  int g0 = 0, g1 = st[0];
  for (int gen = 0; gen < 1024 && g1 > g0; ++gen) {
    veg_gen_kernel<<<(g1 - g0 + 255) / 256, 256, 0, c->stream>>>(a, key, g0, g1);
    veg_append_kernel<<<1, kScanBlock, 0, c->stream>>>(a, g0, g1);
    veg_apply_kernel<<<grid, 256, 0, c->stream>>>(c->root, c->veg_units, c->n);
    HIPCHK(c, hipGetLastError());
    RETURN_IF(veg_state(c, st));
    g0 = g1;
    g1 = st[0];
  }
  veg_compact_kernel<<<1, kScanBlock, 0, c->stream>>>(a);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->veg_step++;
  return SHY_OK;
}

int shy_vegetation_count(shy_ctx* c, int* plants, int64_t* dropped) {
  RETURN_IF(need_single(c, "shy_vegetation_count"));
  if (!c->veg_cap) return fail(c, SHY_ERR_STATE, "shy_vegetation_count: vegetation not enabled");
  int32_t st[4];
  RETURN_IF(veg_state(c, st));
  if (plants) *plants = st[0];
  if (dropped) *dropped = st[2];
  return SHY_OK;
}

int shy_vegetation_download(shy_ctx* c, float* xys) {
  int n = 0;
  RETURN_IF(shy_vegetation_count(c, &n, nullptr));
  if (!xys) return fail(c, SHY_ERR_ARG, "shy_vegetation_download: NULL argument");
  std::vector<int32_t> x(n), y(n);
  std::vector<float> sz(n);
  if (n) {
    HIPCHK(c, hipMemcpy(x.data(), c->veg_px, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(y.data(), c->veg_py, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(sz.data(), c->veg_size, n * sizeof(float), hipMemcpyDeviceToHost));
  }
  for (int k = 0; k < n; ++k) {
    xys[3 * k] = (float)x[k];
    xys[3 * k + 1] = (float)y[k];
    xys[3 * k + 2] = sz[k];
  }
  return SHY_OK;
}

int shy_vegetation_upload(shy_ctx* c, const float* xys, int n) {
  RETURN_IF(need_single(c, "shy_vegetation_upload"));
  if (!c->veg_cap) return fail(c, SHY_ERR_STATE, "shy_vegetation_upload: vegetation not enabled");
  if (n < 0 || n > c->veg_cap || (n && !xys))
    return fail(c, SHY_ERR_ARG, "shy_vegetation_upload: bad plant count or NULL list");
  const int nx = c->ter[0].nx, ny = c->ter[0].ny;
  std::vector<int32_t> x(n), y(n), alive(n, 1);
  std::vector<float> sz(n);
  for (int k = 0; k < n; ++k) {
    x[k] = (int32_t)xys[3 * k];
    y[k] = (int32_t)xys[3 * k + 1];
    sz[k] = xys[3 * k + 2];
    if (x[k] < 0 || y[k] < 0 || x[k] >= nx || y[k] >= ny || !std::isfinite(sz[k]))
      return fail(c, SHY_ERR_ARG, "shy_vegetation_upload: plant off the map or size not finite");
  }
  hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (n) {
    HIPCHK(c, hipMemcpy(c->veg_px, x.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->veg_py, y.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->veg_size, sz.data(), n * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->veg_alive, alive.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  const int32_t st[4] = {n, 0, 0, 0};
  HIPCHK(c, hipMemcpy(c->veg_st, st, sizeof(st), hipMemcpyHostToDevice));
  return SHY_OK;
}

int shy_upload_cells(shy_ctx* c, const float* cells8) {
  RETURN_IF(need_single(c, "shy_upload_cells"));
  if (!cells8) return fail(c, SHY_ERR_ARG, "shy_upload_cells: NULL argument");
  std::vector<float> f(c->n);
  for (int field = 0; field < 8; ++field) {
    for (size_t i = 0; i < c->n; ++i) f[i] = cells8[i * 8 + field];
    RETURN_IF(shy_upload_field(c, field, f.data()));
  }
  return SHY_OK;
}

int shy_download_cells(shy_ctx* c, float* cells8) {
  RETURN_IF(need_single(c, "shy_download_cells"));
  if (!cells8) return fail(c, SHY_ERR_ARG, "shy_download_cells: NULL argument");
  std::vector<float> f(c->n);
  for (int field = 0; field < 8; ++field) {
    RETURN_IF(shy_download_field(c, field, f.data()));
    for (size_t i = 0; i < c->n; ++i) cells8[i * 8 + field] = f[i];
  }
  return SHY_OK;
}

int shy_terrain_generate_fbm(shy_ctx* c, int t, uint64_t seed, int octaves, float base_freq,
                             float lacunarity, float gain) {
  RETURN_IF(check_terrain(c, t));
  if (octaves < 1 || octaves > 16) return fail(c, SHY_ERR_ARG, "shy_generate_fbm: octaves must be 1..16");
  if (!std::isfinite(base_freq) || !std::isfinite(lacunarity) || !std::isfinite(gain))
    return fail(c, SHY_ERR_ARG, "shy_generate_fbm: non-finite argument");
  const Terrain& T = c->ter[t];
  if (base_freq <= 0.0f) base_freq = 1.0f / (float)(T.nx > T.ny ? T.nx : T.ny);
  hipSetDevice(c->device);
  std::vector<uint8_t> perms(16 * 256);
  for (int o = 0; o < octaves; ++o) fbm_perm(seed, o, perms.data() + o * 256);
  HIPCHK(c, hipMemcpyAsync(c->perms, perms.data(), octaves * 256, hipMemcpyHostToDevice, c->stream));
  const size_t n = (size_t)T.nx * T.ny;
  float* out = c->h[parity(c)] + T.cell_off;
  const int bs = 256;
  const size_t grid = (n + bs - 1) / bs;
  fbm_kernel<<<grid, bs, 0, c->stream>>>(out, T.nx, T.ny, c->perms, octaves, base_freq, lacunarity,
                                         gain);
  HIPCHK(c, hipGetLastError());
  minmax_partial_kernel<<<kPartBlocks, 256, 0, c->stream>>>(out, n, c->part);
  HIPCHK(c, hipGetLastError());
  minmax_fold_kernel<<<1, 256, 0, c->stream>>>(c->part, kPartBlocks, 0);
  HIPCHK(c, hipGetLastError());
  normalize_kernel<<<grid, bs, 0, c->stream>>>(out, n, c->part);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SHY_OK;
}

int shy_terrain_generate_reference(shy_ctx* c, int t, int seed, int tileres) {
  RETURN_IF(check_terrain(c, t));
  if (tileres < 1) return fail(c, SHY_ERR_ARG, "shy_generate_reference: tileres must be >= 1");
  const Terrain& T = c->ter[t];
  hipSetDevice(c->device);
  const size_t n = (size_t)T.nx * T.ny;
  float* out = c->h[parity(c)] + T.cell_off;
  const int bs = 256;
  const size_t grid = (n + bs - 1) / bs;
  fnl_world_kernel<<<grid, bs, 0, c->stream>>>(out, T.nx, T.ny, seed, (float)tileres);
  HIPCHK(c, hipGetLastError());
  minmax_partial_kernel<<<kPartBlocks, 256, 0, c->stream>>>(out, n, c->part);
  HIPCHK(c, hipGetLastError());
  minmax_fold_kernel<<<1, 256, 0, c->stream>>>(c->part, kPartBlocks, 1);
  HIPCHK(c, hipGetLastError());
  normalize_kernel<<<grid, bs, 0, c->stream>>>(out, n, c->part);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SHY_OK;
}

int shy_generate_reference(shy_ctx* c, int seed, int tileres) {
  RETURN_IF(need_single(c, "shy_generate_reference"));
  return shy_terrain_generate_reference(c, 0, seed, tileres);
}

int shy_generate_fbm(shy_ctx* c, uint64_t seed, int octaves, float base_freq, float lacunarity,
                     float gain) {
  RETURN_IF(need_single(c, "shy_generate_fbm"));
  return shy_terrain_generate_fbm(c, 0, seed, octaves, base_freq, lacunarity, gain);
}

int shy_erode(shy_ctx* c, int drops) {
  RETURN_IF(need_single(c, "shy_erode"));
  if (drops < 0) return fail(c, SHY_ERR_ARG, "shy_erode: drops must be >= 0");
  hipSetDevice(c->device);
  RETURN_IF(shy_terrain_set_drops(c, 0, drops));
  return erode_once(c);
}

int shy_erode_cycles(shy_ctx* c, int cycles, int drops, int use_graph) {
  RETURN_IF(need_single(c, "shy_erode_cycles"));
  if (cycles < 0 || drops < 0) return fail(c, SHY_ERR_ARG, "shy_erode_cycles: negative count");
  hipSetDevice(c->device);
  RETURN_IF(shy_terrain_set_drops(c, 0, drops));
  return erode_many(c, cycles, use_graph != 0);
}

int shy_ensemble_erode(shy_ctx* c, int cycles, int use_graph) {
  if (!c) return SHY_ERR_ARG;
  if (cycles < 0) return fail(c, SHY_ERR_ARG, "shy_ensemble_erode: negative count");
  hipSetDevice(c->device);
  return erode_many(c, cycles, use_graph != 0);
}

int shy_trace_drop(shy_ctx* c, float x, float y, int max_steps, shy_drop_step* out, int* n_steps) {
  RETURN_IF(need_single(c, "shy_trace_drop"));
  if (!out || !n_steps) return fail(c, SHY_ERR_ARG, "shy_trace_drop: NULL argument");
  if (max_steps < 1) return fail(c, SHY_ERR_ARG, "shy_trace_drop: max_steps must be >= 1");
  const Terrain& T = c->ter[0];
  if (!std::isfinite(x) || !std::isfinite(y) || std::fabs(x) > 1e9f || std::fabs(y) > 1e9f) {
    *n_steps = 0;
    return SHY_OK;
  }
  const int ix = (int)x, iy = (int)y;
  if (ix < 0 || iy < 0 || ix >= T.nx || iy >= T.ny) {
    *n_steps = 0;
    return SHY_OK;  // Drop::descend returns false on a NULL node (water.h:62-64)
  }
  hipSetDevice(c->device);
  if (c->trace_cap < max_steps) {
    if (c->trace_buf) hipFree(c->trace_buf);
    c->trace_buf = nullptr;
    c->trace_cap = 0;
    HIPCHK(c, hipMalloc(&c->trace_buf, (size_t)max_steps * sizeof(shy_drop_step)));
    c->trace_cap = max_steps;
  }
  WalkArgs a;
  std::memset(&a, 0, sizeof(a));
  a.h = c->h[parity(c)];
  a.fld = c->fld;
  a.root = T.has_root ? c->root : nullptr;
  a.state = c->state;
  a.seed = T.seed;
  a.nx = T.nx;
  a.ny = T.ny;
  a.drops = 1;
  a.P = T.P;
  trace_kernel<<<1, 64, 0, c->stream>>>(a, x, y, c->trace_buf, max_steps, c->trace_n);
  HIPCHK(c, hipGetLastError());
  int n = 0;
  HIPCHK(c, hipMemcpyAsync(&n, c->trace_n, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int m = n < max_steps ? n : max_steps;
  HIPCHK(c, hipMemcpy(out, c->trace_buf, (size_t)m * sizeof(shy_drop_step), hipMemcpyDeviceToHost));
  *n_steps = n;
  return SHY_OK;
}

int shy_cascade(shy_ctx* c, float x, float y) {
  RETURN_IF(need_single(c, "shy_cascade"));
  const Terrain& T = c->ter[0];
  if (!std::isfinite(x) || !std::isfinite(y) || std::fabs(x) > 1e9f || std::fabs(y) > 1e9f)
    return fail(c, SHY_ERR_ARG, "shy_cascade: position off the map");
  const int ix = (int)x, iy = (int)y;
  if (ix < 0 || iy < 0 || ix >= T.nx || iy >= T.ny)
    return fail(c, SHY_ERR_ARG, "shy_cascade: position off the map");
  hipSetDevice(c->device);
  cascade_live_kernel<<<1, 64, 0, c->stream>>>(c->h[parity(c)], T.nx, T.ny, T.P, x, y);
  HIPCHK(c, hipGetLastError());
  return SHY_OK;
}

int shy_descend(shy_ctx* c, shy_drop* drop, int* alive) {
  RETURN_IF(need_single(c, "shy_descend"));
  if (!drop || !alive) return fail(c, SHY_ERR_ARG, "shy_descend: NULL argument");
  if (!std::isfinite(drop->pos[0]) || !std::isfinite(drop->pos[1]) ||
      std::fabs(drop->pos[0]) > 1e9f || std::fabs(drop->pos[1]) > 1e9f) {
    *alive = 0;
    return SHY_OK;
  }
  hipSetDevice(c->device);
  if (!c->drop_buf) HIPCHK(c, hipMalloc(&c->drop_buf, sizeof(shy_drop) + sizeof(int)));
  const Terrain& T = c->ter[0];
  LiveArgs a;
  a.h = c->h[parity(c)];
  a.fld = c->fld;
  a.root = c->root;
  a.td = c->trk;
  a.tmx = c->trk + c->n;
  a.tmy = c->trk + 2 * c->n;
  a.erf_a = c->erf_a;
  a.nx = T.nx;
  a.ny = T.ny;
  a.P = T.P;
  shy_drop* dd = (shy_drop*)c->drop_buf;
  int* da = (int*)((char*)c->drop_buf + sizeof(shy_drop));
  HIPCHK(c, hipMemcpyAsync(dd, drop, sizeof(shy_drop), hipMemcpyHostToDevice, c->stream));
  descend_live_kernel<<<1, 64, 0, c->stream>>>(a, dd, da);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(drop, dd, sizeof(shy_drop), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(alive, da, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SHY_OK;
}

int shy_get_stats(shy_ctx* c, shy_stats* out) {
  if (!c || !out) return SHY_ERR_ARG;
  hipSetDevice(c->device);
  DeviceState st;
  HIPCHK(c, hipMemcpyAsync(&st, c->state, sizeof(st), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  out->steps = st.stats[ST_STEPS];
  out->walked = st.stats[ST_WALKED];
  out->skipped = st.stats[ST_SKIPPED];
  out->exits = st.stats[ST_EXITS];
  out->ends_age = st.stats[ST_ENDS_AGE];
  out->ends_vol = st.stats[ST_ENDS_VOL];
  out->export40 = st.stats[ST_EXPORT40];
  out->gain40 = st.stats[ST_GAIN40];
  out->walk_dh = st.stats[ST_WALK_DH];
  out->casc_moves = st.stats[ST_CASC_MOVES];
  return SHY_OK;
}

int shy_terrain_steps(shy_ctx* c, int64_t* steps) {
  if (!c || !steps) return SHY_ERR_ARG;
  hipSetDevice(c->device);
  HIPCHK(c, hipMemcpyAsync(steps, c->tsteps, c->ter.size() * sizeof(long long),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SHY_OK;
}

int shy_terrain_drop_log(shy_ctx* c, int t, int32_t* steps, int32_t* ends, int capacity,
                         int* count) {
  RETURN_IF(check_terrain(c, t));
  if (!count || capacity < 0 || (capacity > 0 && (!steps || !ends)))
    return fail(c, SHY_ERR_ARG, "shy_terrain_drop_log: bad output arguments");
  *count = 0;
  if (!c->dlog_valid || !c->dlog) return SHY_OK;
  hipSetDevice(c->device);
  const Terrain& T = c->ter[t];
  const int end = t + 1 < (int)c->ter.size() ? c->ter[t + 1].drop_off : c->total_drops;
  const int padded = end - T.drop_off;
  std::vector<int4> h((size_t)std::max(padded, 0));
  if (padded > 0) {
    HIPCHK(c, hipMemcpyAsync(h.data(), c->dlog + T.drop_off, (size_t)padded * sizeof(int4),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  int k = 0;
  for (const int4& d : h) {
    if (d.y == kEndPad) continue;
    if (k >= capacity) break;
    steps[k] = d.x;
    ends[k] = d.y;
    ++k;
  }
  *count = k;
  return SHY_OK;
}

int shy_reset_stats(shy_ctx* c) {
  if (!c) return SHY_ERR_ARG;
  hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemsetAsync(c->state->stats, 0, sizeof(c->state->stats), c->stream));
  HIPCHK(c, hipMemsetAsync(c->tsteps, 0, c->ter.size() * sizeof(long long), c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SHY_OK;
}

int shy_synchronize(shy_ctx* c) {
  if (!c) return SHY_ERR_ARG;
  hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SHY_OK;
}

int shy_set_tuning(shy_ctx* c, const shy_tuning* t) {
  if (!c || !t) return SHY_ERR_ARG;
  shy_tuning v = *t;
  if (v.walk_block == 0) v.walk_block = 256;
  if (v.walk_block < 64 || v.walk_block > 1024 || v.walk_block % 64)
    return fail(c, SHY_ERR_ARG, "walk_block must be a multiple of 64 in 64..1024");
  if (v.walk_grid < 0) return fail(c, SHY_ERR_ARG, "walk_grid must be >= 0");
  if (v.tile < 0 || v.tile > 2) return fail(c, SHY_ERR_ARG, "tile must be 0 (32x8), 1 (64x4) or 2 (16x16)");
  hipSetDevice(c->device);
  c->tune = v;
  return layout(c);
}

int shy_get_tuning(const shy_ctx* c, shy_tuning* t) {
  if (!c || !t) return SHY_ERR_ARG;
  *t = c->tune;
  return SHY_OK;
}

int shy_set_drop_order(shy_ctx* c, int on) {
  if (!c) return SHY_ERR_ARG;
  if (on < -1 || on > 1) return fail(c, SHY_ERR_ARG, "drop order must be 1, 0 or -1 (automatic)");
  if (c->drop_order != on) drop_graph(c);
  c->drop_order = on;
  return SHY_OK;
}

int shy_set_walk_block(shy_ctx* c, int threads) {
  if (!c) return SHY_ERR_ARG;
  shy_tuning t = c->tune;
  t.walk_block = threads;
  return shy_set_tuning(c, &t);
}

int shy_time_cycles(shy_ctx* c, int cycles, int drops, float ms[6]) {
  if (!c || !ms || cycles < 1) return fail(c, SHY_ERR_ARG, "shy_time_cycles: bad argument");
  hipSetDevice(c->device);
  if (drops >= 0) {
    RETURN_IF(need_single(c, "shy_time_cycles with drops >= 0"));
    RETURN_IF(shy_terrain_set_drops(c, 0, drops));
  }
  return time_cycles(c, cycles, ms);
}

int shy_autotune(shy_ctx* c, int cycles, shy_tuning* chosen, float* best_ms) {
  if (!c || cycles < 1) return fail(c, SHY_ERR_ARG, "shy_autotune: bad argument");
  hipSetDevice(c->device);
  RETURN_IF(ensure_wave(c));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // Candidates are timed from a copy of the live state, which is restored
  // after each one, so tuning never changes the trajectory.
  const size_t n = c->n;
  const int p0 = parity(c);
  const uint32_t cyc0 = c->cycle;
  std::vector<float> h0(n);
  std::vector<float4> f0(n);
  std::vector<uint8_t> v0(n);
  std::vector<long long> t0(3 * n), ts0(c->ter.size());
  DeviceState s0;
  HIPCHK(c, hipMemcpy(h0.data(), c->h[p0], n * sizeof(float), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(f0.data(), c->fld, n * sizeof(float4), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(v0.data(), c->visit[p0], n, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(t0.data(), c->trk, 3 * n * sizeof(long long), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(ts0.data(), c->tsteps, ts0.size() * sizeof(long long), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(&s0, c->state, sizeof(s0), hipMemcpyDeviceToHost));
  auto restore = [&]() -> int {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->cycle = cyc0;
    HIPCHK(c, hipMemcpy(c->h[p0], h0.data(), n * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->fld, f0.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->visit[p0], v0.data(), n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(c->visit[p0 ^ 1], 0, n));
    HIPCHK(c, hipMemcpy(c->trk, t0.data(), 3 * n * sizeof(long long), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->tsteps, ts0.data(), ts0.size() * sizeof(long long), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->state, &s0, sizeof(s0), hipMemcpyHostToDevice));
    return SHY_OK;
  };
  int cus = 256;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess)
    cus = 256;
  std::vector<shy_tuning> cand;
  for (int tile = 0; tile < 3; ++tile)
    for (int bs : {64, 128, 256, 512})
      for (int per_cu : {0, 8})  // 0: one lane per drop; 8: persistent, 8 blocks' worth per CU
        cand.push_back(shy_tuning{bs, per_cu ? per_cu * cus * 256 / bs : 0, tile});
  float best = 1e30f;
  shy_tuning bt = c->tune;
  for (const shy_tuning& t : cand) {
    c->tune = t;
    RETURN_IF(layout(c));
    float ms[6];
    RETURN_IF(time_cycles(c, 1, ms));  // warm-up
    RETURN_IF(restore());
    RETURN_IF(time_cycles(c, cycles, ms));
    RETURN_IF(restore());
    if (ms[0] < best) {
      best = ms[0];
      bt = t;
    }
  }
  c->tune = bt;
  RETURN_IF(layout(c));
  if (chosen) *chosen = bt;
  if (best_ms) *best_ms = best / (float)cycles;
  return SHY_OK;
}

}  // extern "C"
