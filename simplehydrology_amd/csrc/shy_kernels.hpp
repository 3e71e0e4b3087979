// shy_kernels.hpp -- kernels of the erosion hot path (gfx950, wave64).
//
// One cycle (= one World::erode call, source/world.h:54-88) is one wave of
// drops that all read the map as it stood when the cycle began:
//
//  pack_kernel           one 32-byte record per cell of what a walk step reads
//                        there, the cell's normal precomputed, in 2x2 cell
//                        blocks (four records per 128-byte line)
//  order_kernel          (many drops) which drop each walk thread takes
//  walk_kernel           one drop per lane walks its trajectory; for every
//                        step it logs, coalesced as [step][drop], 16 bytes
//                        {start cell, c_eq} + {tracks} of what the
//                        snapshot fixes (and, where its blocks are the
//                        partition blocks, counts the step's cell tile in
//                        LDS); per drop it logs how the walk ended, counts
//                        its first step's cell and marks its end cell.  The
//                        cascade sites (arrival cells, water.h:151) follow
//                        per cell: a cell is a site iff more steps start there
//                        than first steps, or a walk ended there.
//  part_*_kernel, scan   group the valid log slots by cell tile (4096 cells):
//                        per-block LDS histograms over the tiles of the
//                        block's terrain (built by the walk itself where its
//                        blocks are the partition blocks), one exclusive scan
//                        of every terrain's [tile][block] matrix, then an
//                        LDS-staged scatter of {slot, cell, tracks} into tile
//                        order.  No global atomic per step: the device's
//                        atomic rate, not its bandwidth, bounded a walk that
//                        took one.
//  tile_count_kernel     per tile, the steps starting at each cell (LDS) ->
//                        cascade sites and the burst scale
//                        s = (1-(1-gamma)^k)/(gamma k), gamma = min(1, effD (1 +
//                        entrainment erf(0.4 d))): the k drops through one cell
//                        on the same snapshot commit together what k drops in
//                        sequence would (water.h:127-132 as a linear feedback),
//                        which keeps the wave stable where crowded cells would
//                        overshoot.
//  replay_kernel         replays each drop's logged steps: the only chain left
//                        is the sediment recurrence (arithmetic); the log loads
//                        stream and the per-cell gathers (burst scale, root
//                        density where a terrain has one) do not depend on the
//                        previous step, so they are issued ahead.  Each step's erosion/deposition effD*cdiff*s is
//                        quantised once to int32 (2^-20, moved between map and
//                        drop as that integer) and stored drop-major.
//  tile_reduce_kernel    per tile, sums in LDS the height commits (gathered by
//                        slot) and the discharge/momentum tracks (int32, 2^-24)
//                        of its steps in integers (so the result does not
//                        depend on the order drops ran), plus the terminal
//                        deposits: no read-modify-write of the map per step.
//  cascade_blend_kernel  2D-tiled stencil: applies the walk's commits, runs the
//                        reference cascade (world.h:90-168) of every cell a drop
//                        arrived at (water.h:151) as a weighted Jacobi sweep
//                        (LDS tile, 2-cell height halo, 3-cell mark halo),
//                        applies the quantised transfers, blends
//                        discharge/momentum (world.h:80-86), refreshes
//                        erf(0.4 d) and clears the other parity's marks.
#pragma once

#include "shy_device.hpp"

namespace shy {

// Streaming accesses: the step log is written once by the walk and read once
// by the partition and once by the replay, far more bytes than the L2 holds.
// Its stores and loads carry the non-temporal hint, which keeps the lines
// that are reused (the walk records, the burst scales) in the L2 instead
// (1024^2 cycle: 1.23 -> 1.19 ms, the walk 4% and the partition 16% faster).  The partition entries and the
// replay's commits are read again by gathers, and keep the default policy.
typedef int shy_v2i __attribute__((ext_vector_type(2)));
__device__ inline int2 ld_stream(const int2* p) {
  const shy_v2i v = __builtin_nontemporal_load(reinterpret_cast<const shy_v2i*>(p));
  return make_int2(v.x, v.y);
}
__device__ inline int ld_stream(const int* p) { return __builtin_nontemporal_load(p); }
__device__ inline void st_stream(int2* p, int2 v) {
  shy_v2i w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<shy_v2i*>(p));
}

enum StatSlot {
  ST_STEPS = 0, ST_WALKED, ST_SKIPPED, ST_EXITS, ST_ENDS_AGE, ST_ENDS_VOL, ST_EXPORT40,
  ST_GAIN40, ST_WALK_DH, ST_CASC_MOVES, ST_COUNT
};

struct DeviceState {
  long long stats[ST_COUNT];
  unsigned int cycle;
  unsigned int pad;
};

// One terrain of an ensemble: its slice of the concatenated cell arrays,
// its drops in the concatenated wave, its cascade tiles and its parameters.
// A single-terrain context is an ensemble of one.
struct Terrain {
  long long cell_off;
  long long pack_off;  // the terrain's first packed walk record (pack_rec)
  int nx, ny;
  int drop_off, drops;
  int tile_off, tiles_y;
  // the partition: the terrain's cell tiles [tile_lo, tile_lo + tile_n),
  // its blocks [part_first, part_first + part_blocks) of part_dwidth drops,
  // and its tile_n x part_blocks count matrix at part_off
  long long part_off;
  int part_first, part_blocks, part_dwidth;
  int tile_lo, tile_n;
  int order_first;  // the drop order's first chunk
  unsigned long long seed;
  unsigned int id;  // terrain number in the spawn hash
  int has_root;
  // every volume a drop of this terrain can carry into a step lies where the
  // walk's reciprocal-form division is exact (markstein_den_ok; host)
  int vol_ok;
  shy_params P;
};

// Whole-ensemble arguments of the walk kernels (concatenated arrays).
// The partition's cell tiles: 2^kTileBits consecutive cell indices over all
// terrains.
constexpr int kTileBits = 12;             // cells per tile = 4096
constexpr int kTileCells = 1 << kTileBits;
constexpr int kFuseMaxTiles = 4096;       // most tiles of a terrain whose walk blocks count in LDS

struct WaveArgs {
  const Terrain* __restrict__ ter;
  int nter;
  // [terrain] set by pack_kernel when some cell's gravity term lies outside
  // the range of the walk's reciprocal-form division this cycle (cleared by
  // the stencil at the end of the cycle)
  const int* __restrict__ gflag;
  int total_drops;
  const float* __restrict__ h;
  const float4* __restrict__ fld;
  const float* __restrict__ root;
  int32_t* first;     // drops whose first step starts at the cell
  int2* slog;         // [step][drop] {start cell over all terrains, c_eq bits}
  int2* strk;         // [step][drop] {qt(vol sx), qt(vol sy)}: the step's momentum tracks
  int32_t* qb;        // [drop][step] height commit of the step, 2^-20 (replay)
  int qstride;        // qb's stride per drop: log rows rounded up to kReplayDepth
  const float* __restrict__ sr;   // burst scale per cell (tile_count_kernel)
  int4* dlog;         // [drop] {steps, end kind, end cell, unused}
  int32_t* steps;     // [drop] steps of the walk (the log rows it filled)
  int32_t* hist;      // per-terrain [tile][walk block] step counts, when the walk counts them
  const int32_t* __restrict__ perm;  // [drop] the terrain's drop a thread walks (order_kernel), or null
  const float4* __restrict__ pack;  // per cell 32 B: what one walk step reads (pack_kernel), 2x2 cell blocks (pack_rec)
  int32_t* deposit;
  uint8_t* visit;
  DeviceState* state;
  long long* tsteps;  // particle-steps per terrain
};

// how a drop's walk ended (dlog.y)
enum DropEnd { kEndAge = 0, kEndVol = 1, kEndExit = 2, kEndSkipped = 3, kEndPad = 4 };

// largest k in [0, nter) with key(k) <= v (keys ascending, key(0) <= v)
template <class Key>
__device__ inline int find_terrain(int nter, long long v, Key key) {
  int lo = 0, hi = nter - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (key(mid) <= v) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// One terrain's view, as one drop walks it (pointers offset to the terrain).
struct WalkArgs {
  const float* __restrict__ h;
  const float4* __restrict__ fld;  // {momentumx, momentumy, discharge, erf(0.4*discharge)}
  const float* __restrict__ root;  // nullptr when the map has no root density
  int2* slog;                      // [step][drop] step log (count pass): {cell, c_eq}
  int2* strk;                      // [step][drop] its tracks
  int cell_off;                    // the terrain's first cell over all terrains
  const float4* __restrict__ pack; // packed cells, or nullptr (read h, fld, root)
  uint8_t* visit;                  // cascade sites
  DeviceState* state;
  uint64_t seed;
  uint32_t terrain;
  int nx, ny;
  int drops;
  shy_params P;
};

struct CascArgs {
  const Terrain* __restrict__ ter;
  int nter;
  const float* __restrict__ h_in;
  float* __restrict__ h_out;
  const int32_t* __restrict__ dh;     // walk commits per cell, 2^-20
  const uint8_t* __restrict__ visit;
  const long long* __restrict__ td;   // tracks per cell, 2^-24
  const long long* __restrict__ tmx;
  const long long* __restrict__ tmy;
  uint8_t* visit_clear;
  float4* fld;
  const double* __restrict__ erf_a;
  DeviceState* state;
  int* gflag;  // cleared for the next cycle (pack_kernel sets it)
};

// ------------------------------------------------------------ reductions
template <int NC>
__device__ inline void block_add_stats(long long (&v)[NC], long long* out) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] += __shfl_xor(v[c], o, 64);
  }
  __shared__ long long red[16][NC];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < NC; ++c) red[w][c] = v[c];
  }
  __syncthreads();
  if (threadIdx.x < NC) {
    long long s = 0;
    for (int k = 0; k < nw; ++k) s += red[k][threadIdx.x];
    if (s != 0)
      __hip_atomic_fetch_add(out + threadIdx.x, s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ------------------------------------------------------------------ walk
// One cache line per step: {h, h(x-1), h(x+1), h(y-1)}, {h(y+1), mx, my, d},
// {erf(0.4 d), root, 0, 0}; neighbours off the map repeat the cell's own
// height, as load_unpacked's clamped reads do (the normal never uses them).
__device__ inline CellView load_packed(const float4* __restrict__ pk, int i) {
  const float4 a = pk[2 * (size_t)i], b = pk[2 * (size_t)i + 1];
  CellView v;
  v.h = a.x; v.n0 = a.y; v.n2 = a.z; v.erfd = a.w;
  v.mx = b.x; v.my = b.y; v.d = b.z; v.rf = b.w;
  v.xm = v.xp = v.ym = v.yp = 0.0f;  // not read: the record holds the normal
  v.root = 0.0f;                      // not read by the count walk (the replay reads it)
  return v;
}

// The same fields from the map arrays (the single-drop trace reads the live
// map, which has no packed records).
__device__ inline CellView load_unpacked(const WalkArgs& a, int x, int y, int i) {
  CellView c;
  const int ny = a.ny;
  c.h = a.h[i];
  c.xm = a.h[x > 0 ? i - ny : i];
  c.xp = a.h[x + 1 < a.nx ? i + ny : i];
  c.ym = a.h[y > 0 ? i - 1 : i];
  c.yp = a.h[y + 1 < ny ? i + 1 : i];
  const float4 f = a.fld[i];
  c.mx = f.x;
  c.my = f.y;
  c.d = f.z;
  c.erfd = f.w;
  c.root = a.root ? a.root[i] : 0.0f;
  return c;
}

struct LaneStats {
  long long steps = 0, exits = 0, ends_age = 0, ends_vol = 0, export40 = 0, gain40 = 0,
            walk_dh = 0;
};

__device__ inline void atomic_add_i32(int32_t* p, int32_t v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void atomic_add_i64(long long* p, long long v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

enum WalkMode { kTrace = 0, kCount = 1 };

// One drop from spawn to termination: Drop(pos) then while(drop.descend())
// (source/world.h:74-76, source/water.h:58-156) on the cycle's snapshot.
// kTrace records per-step states of one drop (unscaled, as Drop::descend);
// kCount writes the step log that the partition and the replay read and
// calls step_hook(start cell) for every step.
// Returns the steps taken; *end and *end_cell say how and where it ended.
struct NoStepHook {
  __device__ void operator()(int) const {}
};

template <int MODE, class Hook = NoStepHook, bool PREFETCH = false, bool SAFE = false>
__device__ int walk_drop(const WalkArgs& a, int lane_drop, float x0, float y0, int* end,
                         int* end_cell, shy_drop_step* tr, int max_tr, int* tr_n,
                         Hook step_hook = Hook(), uint32_t* pf_sink = nullptr) {
  constexpr bool TRACE = MODE == kTrace;
  constexpr bool COUNT = MODE == kCount;
  const shy_params& P = a.P;
  const int nx = a.nx, ny = a.ny;
  int ix = (int)x0, iy = (int)y0;
  int i0 = ix * ny + iy;
  float px = x0, py = y0, sx = 0.0f, sy = 0.0f, vol = 1.0f, sed = 0.0f;
  int age = 0, steps = 0;
  int nrec = 0;
  // the log of step k is stored during step k+1, after that step's record
  // load is issued: vector memory operations complete in issue order, so the
  // load's wait then does not include the stores' acknowledgements.  The
  // stores are unconditional (a conditional store makes every later wait a
  // full one): step 0 first stores a placeholder into its own slot, which its
  // real log overwrites one step later or, for a drop that takes no step,
  // which no reader looks at (its step count is 0)
  size_t pk = (size_t)lane_drop;
  int2 ps = make_int2(a.cell_off + i0, 0), pt = make_int2(0, 0);
  // the count walk reads the cycle's packed records, the trace the map arrays
  auto load_cell = [&](int x, int y, int i) {
    if constexpr (COUNT) return load_packed(a.pack, pack_rec(x, y, a.ny));
    else return load_unpacked(a, x, y, i);
  };
  CellView c = load_cell(ix, iy, i0);
  // PREFETCH (count walk): at the top of each step, a load of the record of
  // the cell the drop would reach at its current speed.  When the new speed
  // keeps the drop on that cell, the step's real record load, issued after
  // the motion update, finds its line already on the way: a shorter step
  // where the walk is bound by its chain's latency, extra traffic where it is
  // bound by bandwidth (waves of many more drops than the GPU runs at once).
  // The prefetched word is folded into pf_sink at the end of the step, after
  // the wait for the step's record load (issued after it), so it never adds
  // a wait, and it keeps the load alive.
  uint32_t pf_acc = 0;
  float pf = 0.0f;
  // 1 / vol for the gravity term, and whether vol lies where the
  // reciprocal-form division is exact, computed a step ahead, off the chain
  // that leads from a record to the next one
  float yv = 1.0f / vol;
  // The first record is waited for here, before the loop: left pending, it
  // makes the wait at the top of the loop count the previous step's log
  // stores as well, on every step.
  if constexpr (COUNT) __builtin_amdgcn_s_waitcnt(0);
  for (;;) {
    if constexpr (COUNT && PREFETCH) {
      int qx = (int)(px + sx), qy = (int)(py + sy);
      qx = qx < 0 ? 0 : (qx >= nx ? nx - 1 : qx);
      qy = qy < 0 ? 0 : (qy >= ny ? ny - 1 : qy);
      pf = a.pack[2 * (size_t)pack_rec(qx, qy, ny)].x;
    }
    float n0, n1, n2;
    if constexpr (COUNT) {  // the packed record holds the cell's gravity terms (water.h:70, 95)
      n0 = c.n0;
      n2 = c.n2;
    } else {
      normal_from(c, ix, iy, nx, ny, n0, n1, n2);  // water.h:70
    }
    (void)n1;
    if ((float)age > P.maxAge || vol < P.minVol) {  // water.h:74-82
      *end = (float)age > P.maxAge ? kEndAge : kEndVol;
      *end_cell = i0;
      break;
    }
    const size_t klog = (size_t)age * a.drops + lane_drop;  // step k = age of this drop
    ++steps;
    float effD = P.depositionRate * (1.0f - c.root);  // water.h:86-87
    if (effD < 0) effD = 0;
    if constexpr (COUNT) {  // water.h:95-111
      // the record holds the gravity terms k n (x, z).  SAFE: every gravity
      // term of the terrain and every volume lies where the reciprocal-form
      // division is exact; else each division checks its operands
      if constexpr (SAFE) {
        sx = sx + markstein_div(n0, vol, yv);
        sy = sy + markstein_div(n2, vol, yv);
      } else {
        sx = sx + div_rcp(n0, vol, yv);
        sy = sy + div_rcp(n2, vol, yv);
      }
      step_coupling(P, vol, c.d, c.mx, c.my, c.rf, sx, sy, px, py);
    } else {
      step_motion(P, n0, n2, vol, c.d, c.mx, c.my, sx, sy, px, py);
    }
    // ivec2(pos) (truncation toward zero) and the exit test: (int)p lies in
    // [0, n) iff -1 < p < n, and a NaN coordinate fails the test (a speed made
    // NaN by a volume that underflowed to 0 with minVol <= 0: the reference's
    // x86 build sends it off the map too); the exit test does not wait for the
    // conversion, and off the map the cell index is not used
    int jx, jy;
    bool ex;
    if constexpr (COUNT) {
      jx = (int)px;
      jy = (int)py;
      ex = !(px > -1.0f && px < (float)nx && py > -1.0f && py < (float)ny);
    } else {
      jx = cell_coord(px);
      jy = cell_coord(py);
      ex = off_map(jx, jy, nx, ny);
    }
    const int j = jx * ny + jy;
    // off the map the record is not used; loading the current cell's keeps
    // the load unconditional, so its wait is counted exactly
    const CellView cn = ex ? (COUNT ? load_cell(ix, iy, i0) : c) : load_cell(jx, jy, j);
    if (COUNT) {
      st_stream(&a.slog[pk], ps);
      st_stream(&a.strk[pk], pt);
    }
    const float h2 = ex ? (float)((double)c.h - 0.002) : cn.h;  // water.h:120-124
    float c_eq = (1.0f + P.entrainment * c.erfd) * (c.h - h2);  // water.h:127
    if (c_eq < 0) c_eq = 0;
    if (COUNT) {
      // everything of the step that the snapshot fixes; the tracks
      // (water.h:115-117) use the volume before this step's evaporation
      pk = klog;
      ps = make_int2(a.cell_off + i0, __float_as_int(c_eq));
      pt = make_int2(quant_t(vol * sx), quant_t(vol * sy));
      step_hook(a.cell_off + i0);
    }
    if (TRACE) {
      const float ero = effD * (c_eq - sed);  // water.h:128-132
      sed += ero;
      sed = evap_sed(sed, P.evapRate);  // water.h:135
      const float v1 = evap_vol(vol, P.evapRate);
      if (nrec < max_tr) {
        shy_drop_step r;
        r.age = age; r.status = ex ? 1 : 0;
        r.pos[0] = px; r.pos[1] = py; r.speed[0] = sx; r.speed[1] = sy;
        r.volume = v1; r.sediment = sed; r.erosion = ero; r.h2 = h2;
        tr[nrec] = r;
      }
      ++nrec;
      if (ex || nrec >= max_tr) {
        *end = ex ? kEndExit : kEndAge;
        *end_cell = i0;
        break;
      }
    }
    vol = evap_vol(vol, P.evapRate);  // water.h:136
    if (COUNT) yv = 1.0f / vol;
    if (ex) {  // water.h:139-142
      *end = kEndExit;
      *end_cell = i0;
      break;
    }
    ++age;
    c = cn;
    ix = jx; iy = jy; i0 = j;
    if constexpr (COUNT && PREFETCH) pf_acc ^= __float_as_uint(pf);
  }
  if (COUNT) {
    st_stream(&a.slog[pk], ps);
    st_stream(&a.strk[pk], pt);
    if (pf_sink) *pf_sink = pf_acc;
  }
  if (TRACE) *tr_n = nrec;
  return steps;
}

// The count walk's 32-byte per-cell record for this cycle (load_packed),
// every cell of every terrain: {h, normal x, normal z, erf(0.4 d)},
// {momentum x, momentum y, discharge, 1/|momentum|}.  The normal
// (quad::_normal, cellpool.h:181-204) and the momentum's inverse length
// (water.h:101-104) depend only on the cell's snapshot, so they are computed
// here once per cell, with the walk's own arithmetic, instead of at every
// step: a shorter chain per step.
// One thread per record, numbered over all terrains in record order
// (Terrain::pack_off), so that four consecutive threads write one 128-byte
// line; padding records (odd nx or ny) are skipped and never read.
__global__ void pack_kernel(const Terrain* __restrict__ ter, int nter, size_t nrec,
                            const float* __restrict__ h, const float4* __restrict__ fld,
                            float4* pack, int* gflag) {
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nrec) return;
  const int t = find_terrain(nter, (long long)g, [&](int q) { return ter[q].pack_off; });
  const Terrain& T = ter[t];
  const int nx = T.nx, ny = T.ny, nyb = (ny + 1) >> 1;
  const int rr = (int)(g - (size_t)T.pack_off), b = rr >> 2, q = rr & 3;
  const int bx = b / nyb, by = b - bx * nyb;
  const int x = 2 * bx + (q >> 1), y = 2 * by + (q & 1);  // pack_rec(x, y, ny) = rr
  if (x >= nx || y >= ny) return;
  const size_t i = (size_t)x * ny + y;
  const float* hh = h + T.cell_off;
  const float hc = hh[i];
  const float xm = hh[x > 0 ? i - ny : i], xp = hh[x + 1 < nx ? i + ny : i];
  const float ym = hh[y > 0 ? i - 1 : i], yp = hh[y + 1 < ny ? i + 1 : i];
  const float4 f = fld[(size_t)T.cell_off + i];
  CellView c;
  c.h = hc; c.xm = xm; c.xp = xp; c.ym = ym; c.yp = yp;
  float n0, n1, n2;
  normal_from(c, x, y, nx, ny, n0, n1, n2);  // water.h:70, once per cell instead of per step
  // the gravity terms k n (water.h:95's numerators; the terrain is flagged
  // where one lies outside the range of the walk's reciprocal-form division),
  // and 1/|momentum| for the coupling term (read only where |momentum| > 0)
  const float k = kLod * T.P.gravity;
  const float a0 = k * n0, a2 = k * n2;
  if (!(markstein_num_ok(a0) && markstein_num_ok(a2))) gflag[t] = 1;
  const float rf = f.x * f.x + f.y * f.y > 0.0f ? inv_len(f.x, f.y) : 0.0f;
  pack[2 * g] = make_float4(hc, a0, a2, f.w);
  pack[2 * g + 1] = make_float4(f.x, f.y, f.z, rf);
}

// Count pass over every drop of every terrain; grid-stride, so the grid may
// be smaller than the wave (a persistent launch).
// With COUNT_TILES the grid is one block per partition block (one drop per
// thread, no grid stride): the block counts its steps per tile of its
// terrain in LDS and writes its column of the terrain's [tile][block]
// matrix, which part_hist_kernel would otherwise build from the log.
template <bool COUNT_TILES, bool PREFETCH>
__global__ void __launch_bounds__(1024) walk_kernel(WaveArgs w) {
  __shared__ int th[COUNT_TILES ? kFuseMaxTiles : 1];
  int wlo = 0, wn = 0;
  long long hbase = 0;
  int hstride = 0;
  if (COUNT_TILES) {
    // the block's drops lie in one terrain: the partition blocks are the walk
    // blocks (part_dwidth = blockDim, which divides every padded drop range)
    const Terrain& T = w.ter[find_terrain(
        w.nter, (long long)blockIdx.x * blockDim.x,
        [&](int q) { return (long long)w.ter[q].drop_off; })];
    wlo = T.tile_lo;
    wn = T.tile_n;
    hbase = T.part_off + ((int)blockIdx.x - T.part_first);
    hstride = T.part_blocks;
    for (int j = threadIdx.x; j < wn; j += blockDim.x) th[j] = 0;
    __syncthreads();
  }
  auto count_step = [&](int cell) {
    if (COUNT_TILES) atomicAdd(&th[(cell >> kTileBits) - wlo], 1);
  };
  const unsigned int cycle = w.state->cycle;
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < w.total_drops;
       g += gridDim.x * blockDim.x) {
    // drop ranges are padded to whole waves, so the wave shares one terrain
    // and its entry lives in scalar registers
    const int g0 = __builtin_amdgcn_readfirstlane(g);
    const int t = __builtin_amdgcn_readfirstlane(
        find_terrain(w.nter, g0, [&](int k) { return (long long)w.ter[k].drop_off; }));
    const Terrain T = w.ter[t];
    int4 dl = make_int4(0, kEndPad, 0, 0);
    const int loc = w.perm ? w.perm[g] : g - T.drop_off;  // the drop this thread walks
    if (loc < T.drops) {  // not a padding lane
      WalkArgs a{};
      a.h = w.h + T.cell_off;
      a.fld = w.fld + T.cell_off;
      a.root = T.has_root ? w.root + T.cell_off : nullptr;
      a.cell_off = (int)T.cell_off;
      a.slog = w.slog;
      a.strk = w.strk;
      a.pack = w.pack ? w.pack + 2 * T.pack_off : nullptr;
      a.nx = T.nx;
      a.ny = T.ny;
      a.drops = w.total_drops;  // log stride
      a.P = T.P;
      const uint32_t ncell = (uint32_t)T.nx * (uint32_t)T.ny;
      const uint32_t cid =
          spawn_cell(cycle_key(T.seed, T.id, cycle), (uint32_t)loc, ncell);
      dl.y = kEndSkipped;
      if ((double)a.h[cid] >= 0.1) {  // world.h:71-72
        int end = 0, end_cell = 0;
        uint32_t sink = 0;
        const float x0 = (float)(cid / (uint32_t)T.ny), y0 = (float)(cid % (uint32_t)T.ny);
        // uniform over the wave (one terrain per wave): the step loop without
        // the division's operand checks wherever they cannot fail
        if (T.vol_ok && !w.gflag[t])
          dl.x = walk_drop<kCount, decltype(count_step), PREFETCH, true>(
              a, g, x0, y0, &end, &end_cell, nullptr, 0, nullptr, count_step, &sink);
        else
          dl.x = walk_drop<kCount, decltype(count_step), PREFETCH, false>(
              a, g, x0, y0, &end, &end_cell, nullptr, 0, nullptr, count_step, &sink);
        dl.w = (int)sink;  // not read: it keeps the record prefetches
        dl.y = end;
        dl.z = end_cell;
        if (dl.x > 0) {
          // arrivals = steps past the first; a walk that ended after a step
          // ended on its last arrival (water.h:151)
          __hip_atomic_fetch_add(&w.first[T.cell_off + cid], 1, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
          if (end != kEndExit) w.visit[T.cell_off + end_cell] = 1;
        }
      }
    }
    w.dlog[g] = dl;
    w.steps[g] = dl.x;
  }
  if (COUNT_TILES) {
    __syncthreads();
    for (int r = threadIdx.x; r < wn; r += blockDim.x) w.hist[hbase + (long long)r * hstride] = th[r];
  }
}

// The commits of every drop from its step log.  The arithmetic is the
// walk's, step for step: cdiff = c_eq - sediment, one quantised amount
// q = quant_h(effD cdiff s) moves from the map into the drop, then
// evaporation (water.h:128-136); the terminal deposit or the export ends the
// drop (water.h:74-82, 139-142).  -q is written drop-major, four steps per
// 16-byte store: a drop's consecutive steps often start in one cell tile, so
// tile_reduce_kernel's gathers of them share lines.
// Steps of the log in flight per lane: while step k computes, the log row
// of step k+D is loaded and the per-cell gather of step k+D-L is issued with
// the cell of a row loaded L steps ago, so no step waits on a load issued
// shortly before it.  One wave per SIMD is all a 65536-drop wave fills, so
// the ring is the latency hiding.
constexpr int kReplayDepth = 32;
constexpr int kReplayLag = 16;

// ROOT: some terrain has a root density (a terrain without one holds zeros,
// so the same gather serves every terrain); without, each step gathers only
// the 4-byte burst scale.
template <bool ROOT>
__global__ void __launch_bounds__(256) replay_kernel(WaveArgs w) {
  constexpr int D = kReplayDepth, L = kReplayLag;
  static_assert(L >= 1 && L < D, "gathers run D - L steps ahead of use");
  LaneStats S;
  long long walked = 0, skipped = 0;
  const size_t stride = (size_t)w.total_drops;
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < w.total_drops;
       g += gridDim.x * blockDim.x) {
    const int g0 = __builtin_amdgcn_readfirstlane(g);
    const int t = __builtin_amdgcn_readfirstlane(
        find_terrain(w.nter, g0, [&](int k) { return (long long)w.ter[k].drop_off; }));
    const Terrain T = w.ter[t];
    const int4 dl = w.dlog[g];
    const int n = dl.y <= kEndExit ? dl.x : 0;
    if (dl.y == kEndSkipped) skipped += 1;
    if (dl.y <= kEndExit) {
      walked += 1;
      const float* __restrict__ sr = w.sr;
      const float* __restrict__ rmap = w.root;
      const float dep = T.P.depositionRate, e = T.P.evapRate;
      const double ce = 1.0 - (double)e, ye = 1.0 / ce;  // evaporation divisor, its reciprocal
      float sed = 0.0f, vol = 1.0f;
      double gain = 0.0;
      long long wdh = 0;
      if (n > 0) {
        // ring slot u holds step k = k0 + u.  Every load is unconditional (the
        // row index clamped to the drop's last step): a conditional load
        // makes each later wait a full one, which would serialise the ring.
        const int last = n - 1;
        auto slot = [&](int k) { return (size_t)(k < last ? k : last) * stride + g; };
        float cq[D], sc[D], rt[D];  // c_eq; burst scale and root density of the start cell
        int cl[D];
#pragma unroll
        for (int u = 0; u < D; ++u) {
          const int2 l = ld_stream(&w.slog[slot(u)]);
          cl[u] = l.x;
          cq[u] = __int_as_float(l.y);
        }
#pragma unroll
        for (int u = 0; u < D; ++u) {
          if (u < D - L) {
            sc[u] = sr[cl[u]];  // burst scale and root density of the start cell
            rt[u] = ROOT ? rmap[cl[u]] : 0.0f;
          } else {
            sc[u] = 0.0f;  // gathered in the loop's first L steps
            rt[u] = 0.0f;
          }
        }
        static_assert(D % 4 == 0, "the ring stores its commits as int4s");
        int4* qdst = reinterpret_cast<int4*>(w.qb + (size_t)g * w.qstride);
        for (int k0 = 0; k0 < n; k0 += D) {
          int qv[D];
#pragma unroll
          for (int u = 0; u < D; ++u) {
            const int k = k0 + u;
            qv[u] = 0;
            if (k < n) {
              float effD = dep * (1.0f - rt[u]);  // water.h:86-87
              if (effD < 0) effD = 0;
              const float cdiff = cq[u] - sed;
              const int32_t q = quant_h((effD * cdiff) * sc[u]);
              qv[u] = -q;
              wdh -= q;
              sed += (float)q * kHQInv;
              const float s1 = evap_sed_rcp(sed, ce, ye);  // = evap_sed(sed, e)
              gain += (double)s1 - (double)sed;
              sed = s1;
              vol = evap_vol(vol, e);
            }
            // slot u takes step k + D; the gather of step k + D - L (slot v,
            // whose row was loaded L steps ago) goes out now
            const int2 l = ld_stream(&w.slog[slot(k + D)]);
            cl[u] = l.x;
            cq[u] = __int_as_float(l.y);
            const int v = (u + D - L) % D;  // constant once unrolled
            sc[v] = sr[cl[v]];
            rt[v] = ROOT ? rmap[cl[v]] : 0.0f;
          }
#pragma unroll
          for (int u = 0; u < D; u += 4)
            qdst[(k0 + u) >> 2] = make_int4(qv[u], qv[u + 1], qv[u + 2], qv[u + 3]);
        }
      }
      S.steps += n;
      if (dl.y == kEndExit) {
        S.exits++;
        S.export40 += quant_40((double)sed);
      } else {
        const int32_t q = quant_h(sed);
        atomic_add_i32(&w.deposit[T.cell_off + dl.z], q);
        wdh += q;
        if (dl.y == kEndAge) S.ends_age++; else S.ends_vol++;
      }
      S.walk_dh += wdh;
      S.gain40 += quant_40(gain);
    }
    if (w.tsteps) {
      // the wave shares one terrain (drop ranges are whole waves): one atomic
      // per wave and terrain instead of one per drop on a shared counter
      long long mine = n;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) mine += __shfl_xor(mine, o, 64);
      if ((threadIdx.x & 63) == 0 && mine)
        __hip_atomic_fetch_add(&w.tsteps[t], mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  long long v[ST_CASC_MOVES] = {S.steps, walked, skipped, S.exits, S.ends_age,
                                S.ends_vol, S.export40, S.gain40, S.walk_dh};
  block_add_stats<ST_CASC_MOVES>(v, w.state->stats);
}

// r^k by binary exponentiation, least significant bit first.
__device__ inline float powi_f32(float r, int k) {
  float res = 1.0f, b = r;
  while (k > 0) {
    if (k & 1) res = res * b;
    b = b * b;
    k >>= 1;
  }
  return res;
}

// Burst scale of a cell whose count pass saw k walk steps start there.  One
// step moves the cell a fraction g = effD (1 + entrainment erf(0.4 d)) of its
// distance to equilibrium (water.h:127-132), k steps in sequence
// R = 1 - (1-g)^k of it; the k wave commits are scaled to move it R in total.
// A wave updates a cell and its downstream neighbour from one snapshot (a
// Jacobi step), which is stable only while R <= 1, so for g > 1 the k commits
// together move the cell exactly to the snapshot's equilibrium (R = 1).
__device__ inline float burst_scale(int k, float effD, float entrainment, float erfd) {
  float g = effD * (1.0f + entrainment * erfd);
  if (!(g > 0.0f)) return 1.0f;
  if (g > 1.0f) return 1.0f / ((float)(k < 1 ? 1 : k) * g);
  if (k <= 1) return 1.0f;
  return (1.0f - powi_f32(1.0f - g, k)) / (g * (float)k);
}

// ----------------------------------------------------- bucket offsets
constexpr int kScanBlock = 1024;

// exclusive scan of v over the block's kScanBlock threads (Hillis-Steele in
// LDS); total receives the block sum
__device__ inline int block_exclusive_scan(int v, int* tmp, int& total) {
  const int t = threadIdx.x;
  tmp[t] = v;
  __syncthreads();
  for (int o = 1; o < kScanBlock; o <<= 1) {
    const int add = t >= o ? tmp[t - o] : 0;
    __syncthreads();
    tmp[t] += add;
    __syncthreads();
  }
  total = tmp[kScanBlock - 1];
  const int incl = tmp[t];
  __syncthreads();
  return incl - v;
}

// per-block sums of the step counts
__global__ void __launch_bounds__(kScanBlock) scan_sums_kernel(const int32_t* __restrict__ count,
                                                               size_t n, int32_t* bsum) {
  __shared__ int tmp[kScanBlock];
  const size_t i = (size_t)blockIdx.x * kScanBlock + threadIdx.x;
  int total;
  block_exclusive_scan(i < n ? count[i] : 0, tmp, total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// exclusive scan of the block sums in one block (each thread a run of them)
__global__ void __launch_bounds__(kScanBlock) scan_top_kernel(int32_t* bsum, int nb) {
  __shared__ int tmp[kScanBlock];
  const int per = (nb + kScanBlock - 1) / kScanBlock;
  const int b0 = threadIdx.x * per;
  int s = 0;
  for (int b = b0; b < b0 + per && b < nb; ++b) s += bsum[b];
  int total;
  int run = block_exclusive_scan(s, tmp, total);
  for (int b = b0; b < b0 + per && b < nb; ++b) {
    const int v = bsum[b];
    bsum[b] = run;
    run += v;
  }
}

// in-place exclusive scan of v[0, n) from the per-block sums of
// scan_sums_kernel after scan_top_kernel; v[n] receives the total
__global__ void __launch_bounds__(kScanBlock) scan_apply_kernel(int32_t* v,
                                                                const int32_t* __restrict__ bsum,
                                                                size_t n) {
  __shared__ int tmp[kScanBlock];
  const size_t i = (size_t)blockIdx.x * kScanBlock + threadIdx.x;
  const int x = i < n ? v[i] : 0;
  int total;
  const int ex = block_exclusive_scan(x, tmp, total);
  if (i < n) v[i] = bsum[blockIdx.x] + ex;
  if (i + 1 == n) v[n] = bsum[blockIdx.x] + ex + x;
}

// ------------------------------------------------- partition by cell tile
// The step log is [step][drop]; the per-cell sums need it grouped by cell.
// Cells are cut into tiles of 2^tb consecutive indices (over all terrains).
// Each block owns dwidth consecutive drops over all rows of the log (slot s
// of a block: row s / width, drop g0 + s % width), and its drops belong to
// one terrain (dwidth divides every terrain's padded drop range), so its
// steps fall in that terrain's tiles.  Every terrain owns a [tile][block]
// count matrix of its tiles by its blocks, stored one after the other in
// terrain order: part_hist_kernel (or the walk) fills the block's column in
// LDS, the scan turns all matrices into exclusive offsets at once, and the
// scatter places every valid slot at its tile's next free entry (an LDS
// cursor per tile), one 16-byte entry per step.  A tile shared by two
// terrains ends the first terrain's matrix and opens the next one's, so its
// entries stay contiguous.  Order inside a tile does not matter: every
// consumer sums integers.
// The rest of a partition entry beside its start cell: the log slot and the
// step's quantised momentum tracks.
struct EntRest {
  int slot, tx, ty;
};

struct PartArgs {
  const int2* __restrict__ slog;      // [slot] {start cell, c_eq}
  const int2* __restrict__ strk;      // [slot] {qt(vol sx), qt(vol sy)}
  const int32_t* __restrict__ steps;  // [drop] valid rows of the log
  const Terrain* __restrict__ ter;
  int nter;
  int32_t* hist;   // per terrain [tile][block] counts, then exclusive offsets; [total] = entries
  // [entry] in tile order: the start cell (all tile_count_kernel reads) and
  // {slot, qt(vol sx), qt(vol sy)}
  int32_t* ecell;
  EntRest* erest;
  int rows;        // log rows (0: no walk this cycle)
  int drops;       // log stride = all drops (padded)
  int tb;          // log2 cells per tile
};

// A partition block: its drops [g0, g0 + width), its slots (width x rows),
// its column of its terrain's count matrix (base, stride = the terrain's
// blocks) and its terrain's tiles [wlo, wlo + wn).
struct PartBlock {
  int g0, width, nslots;
  long long base;
  int stride;
  int wlo, wn;
};

__device__ inline PartBlock part_block(const PartArgs& a) {
  const int b = blockIdx.x;
  const int t =
      find_terrain(a.nter, (long long)b, [&](int q) { return (long long)a.ter[q].part_first; });
  const Terrain& T = a.ter[t];
  const int j = b - T.part_first;
  // the terrain's whole padded drop range: the drop order may place real
  // drops on any of its threads (padding threads log no steps)
  const int end = t + 1 < a.nter ? a.ter[t + 1].drop_off : a.drops;
  const int left = end - T.drop_off - j * T.part_dwidth;
  PartBlock p;
  p.g0 = T.drop_off + j * T.part_dwidth;
  p.width = left < T.part_dwidth ? (left > 0 ? left : 0) : T.part_dwidth;
  p.nslots = p.width * a.rows;
  p.base = T.part_off + j;
  p.stride = T.part_blocks;
  p.wlo = T.tile_lo;
  p.wn = T.tile_n;
  return p;
}

// slot s of the block -> log index k, or -1 when the drop took fewer steps
__device__ inline int part_slot(const PartArgs& a, int g0, int width, int s) {
  const int row = s / width, g = g0 + (s - row * width);
  return row < a.steps[g] ? row * a.drops + g : -1;
}

// MAXT bounds the tiles of any terrain (the host checks).
template <int MAXT>
__global__ void __launch_bounds__(1024) part_hist_kernel(PartArgs a) {
  __shared__ int h[MAXT];
  const PartBlock col = part_block(a);
  const int g0 = col.g0, width = col.width, ns = col.nslots;
  for (int t = threadIdx.x; t < col.wn; t += blockDim.x) h[t] = 0;
  __syncthreads();
  for (int s = (int)threadIdx.x; s < ns; s += 4 * blockDim.x) {
    int c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int x = s + u * (int)blockDim.x;
      const int k = x < ns ? part_slot(a, g0, width, x) : -1;
      c[u] = k >= 0 ? a.slog[k].x : -1;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (c[u] >= 0) atomicAdd(&h[(c[u] >> a.tb) - col.wlo], 1);
  }
  __syncthreads();
  for (int r = threadIdx.x; r < col.wn; r += blockDim.x)
    a.hist[col.base + (long long)r * col.stride] = h[r];
}

// Scatter without staging (terrains of any size): one 16-byte entry per step at
// its tile's LDS cursor.
template <int MAXT>
__global__ void __launch_bounds__(1024) part_scatter_kernel(PartArgs a) {
  __shared__ int cur[MAXT];
  const PartBlock col = part_block(a);
  const int g0 = col.g0, width = col.width, ns = col.nslots;
  for (int r = threadIdx.x; r < col.wn; r += blockDim.x)
    cur[r] = a.hist[col.base + (long long)r * col.stride];
  __syncthreads();
  for (int s = (int)threadIdx.x; s < ns; s += blockDim.x) {
    const int k = part_slot(a, g0, width, s);
    if (k >= 0) {
      const int c = a.slog[k].x;
      const int2 tk = a.strk[k];
      const int e = atomicAdd(&cur[(c >> a.tb) - col.wlo], 1);
      a.ecell[e] = c;
      a.erest[e] = EntRest{k, tk.x, tk.y};
    }
  }
}

// First entry of every tile in the scanned count matrices: tile t starts at
// its row in the terrain holding its first cell (a tile shared by
// neighbouring terrains continues in the next terrain's first row, which
// follows in memory; a terrain without blocks has an empty matrix whose
// offset is the next one's); start[ntiles] = all entries.  Without hist
// (no entries) every start is 0.
__global__ void tile_start_kernel(const Terrain* __restrict__ ter, int nter, int ntiles,
                                  const int32_t* __restrict__ hist, long long total_pos,
                                  int32_t* start) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  if (!hist) {
    start[t] = 0;
    return;
  }
  if (t == ntiles) {
    start[t] = hist[total_pos];
    return;
  }
  const int q = find_terrain(nter, (long long)t << kTileBits,
                             [&](int k) { return ter[k].cell_off; });
  const Terrain& T = ter[q];
  start[t] = hist[T.part_off + (long long)(t - T.tile_lo) * T.part_blocks];
}

// Exclusive scan of v[0, m) into out[0, m) by a block of 1024 threads,
// m <= 4096 (each thread a run of up to four); wsum holds 16 ints.  Every
// thread of the block must call it.
__device__ inline void block_scan_runs(const int* v, int* out, int m, int* wsum) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int per = (m + 1023) >> 10;
  int loc[4];
  int sum = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int i = tid * per + j;
    loc[j] = j < per && i < m ? v[i] : 0;
    sum += loc[j];
  }
  int x = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wv] = x;
  __syncthreads();
  if (tid < 64) {
    const int w = tid < 16 ? wsum[tid] : 0;
    int z = w;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      const int y = __shfl_up(z, o, 64);
      if (lane >= o) z += y;
    }
    if (tid < 16) wsum[tid] = z - w;
  }
  __syncthreads();
  int run = wsum[wv] + x - sum;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int i = tid * per + j;
    if (j < per && i < m) out[i] = run;
    run += loc[j];
  }
}

// ---------------------------------------------------------------- drop order
// Each step of the count walk waits on its cell's 32-byte record, which
// mostly misses the XCD's L2.  When a cycle has more drops than the GPU runs
// at once, the walk's blocks run in sequence, and blocks that walk drops
// starting near each other on one XCD share its L2.  Drops are independent
// within a cycle, every log is indexed by the thread that walked it and
// every per-cell result is an integer sum, so any assignment of drops to
// threads gives the same results.  order_kernel assigns them spatially, per
// chunk of kOrderChunk threads of a terrain (one block each): the chunk's
// drops sorted by the Morton code of their spawn cell on a grid of at most
// 64 x 64 bins, and the sorted list laid onto the chunk's walk blocks class
// by class, class x holding the blocks b with b % 8 == x (the XCD a
// workgroup runs on: workgroups are dispatched round-robin over the eight
// XCDs).  Every chunk puts the same eighth of its terrain on the same XCD,
// and the lanes of a wave start near each other.
// perm[g] = the terrain's drop walked by thread g (padding threads receive
// numbers >= drops).  A chunk's partial last block of bs keeps its
// positions, so the order is a permutation for every bs; the classes match
// the XCDs when bs divides the terrain's drop offset and the walk grid is
// one block per bs drops.
constexpr int kOrderChunk = 65536;

// position of sorted entry s of a chunk of nbl blocks of bs whose first
// block is global block b0: class-major, blocks of a class in order
__device__ inline int order_position(int s, int bs, int nbl, int b0) {
  const int k = s / bs, o = s - k * bs;
  int cum = 0;
  for (int x = 0; x < 8; ++x) {
    const int d = (x - b0) & 7;
    const int n = d < nbl ? (nbl - 1 - d) / 8 + 1 : 0;
    if (k < cum + n) return (d + 8 * (k - cum)) * bs + o;
    cum += n;
  }
  return s;  // the partial last block keeps its positions
}

__device__ inline int morton_bin(uint32_t cid, int ny, int sx, int sy) {
  uint32_t bx = (cid / (uint32_t)ny) >> sx, by = (cid % (uint32_t)ny) >> sy;
  uint32_t m = 0;
#pragma unroll
  for (int i = 0; i < 6; ++i) m |= ((bx >> i) & 1u) << (2 * i) | ((by >> i) & 1u) << (2 * i + 1);
  return (int)m;
}

// One block of 1024 threads per chunk; Terrain::order_first numbers the
// chunks of the terrains.  The second pass recomputes each drop's bin (a
// global round trip per drop costs more at one block per CU).
__global__ void __launch_bounds__(1024) order_kernel(const Terrain* __restrict__ ter, int nter,
                                                     int total_drops, int bs,
                                                     const DeviceState* __restrict__ state,
                                                     int32_t* perm) {
  __shared__ int cnt[4096];
  __shared__ int off[4096];
  __shared__ int wsum[16];
  const int t = find_terrain(nter, (long long)blockIdx.x,
                             [&](int q) { return (long long)ter[q].order_first; });
  const Terrain& T = ter[t];
  const int end = t + 1 < nter ? ter[t + 1].drop_off : total_drops;
  const int P = end - T.drop_off;  // the terrain's padded drops
  const int c0 = ((int)blockIdx.x - T.order_first) * kOrderChunk;
  const int c1 = c0 + kOrderChunk < P ? c0 + kOrderChunk : P;
  const int r1 = c1 < T.drops ? c1 : T.drops;  // real drops of the chunk: [c0, r1)
  const int nbl = (c1 - c0) / bs, b0 = (T.drop_off + c0) / bs;
  const uint32_t ncell = (uint32_t)T.nx * (uint32_t)T.ny;
  const uint64_t key = cycle_key(T.seed, T.id, state->cycle);
  int sx = 0, sy = 0;
  while (((T.nx - 1) >> sx) >= 64) ++sx;
  while (((T.ny - 1) >> sy) >= 64) ++sy;
  for (int b = threadIdx.x; b < 4096; b += 1024) cnt[b] = 0;
  __syncthreads();
  for (int i = c0 + (int)threadIdx.x; i < r1; i += 1024)
    atomicAdd(&cnt[morton_bin(spawn_cell(key, (uint32_t)i, ncell), T.ny, sx, sy)], 1);
  __syncthreads();
  block_scan_runs(cnt, off, 4096, wsum);
  __syncthreads();
  for (int i = c0 + (int)threadIdx.x; i < r1; i += 1024) {
    const int s = atomicAdd(&off[morton_bin(spawn_cell(key, (uint32_t)i, ncell), T.ny, sx, sy)], 1);
    perm[T.drop_off + c0 + order_position(s, bs, nbl, b0)] = i;
  }
  for (int s = (r1 > c0 ? r1 : c0) - c0 + (int)threadIdx.x; s < c1 - c0; s += 1024)
    perm[T.drop_off + c0 + order_position(s, bs, nbl, b0)] = c0 + s;
}

// The scatter with its writes grouped: each round of RS slots is ranked per
// tile in LDS, staged in tile order and written out so that the round's
// entries of one tile go to consecutive addresses.  LDS holds the block's
// terrain's tiles (at most MAXT, which the host checks) and RS entries.
template <int MAXT, int RS>
constexpr int scatter_lds_drops() {
  return (MAXT <= 1024 && RS <= 4096) ? 1008 : 4096;
}

// LDSW: the block's step counts are staged in LDS (every block width at most
// scatter_lds_drops, which the host checks); else they are read from memory.
// The choice is made at compile time: a pointer that may address either
// makes every read of it a flat access, whose wait covers all memory loads
// in flight, and so would serialise the next round's loads.
template <int MAXT, int RS, bool LDSW>
// (rounds of 4096: two blocks of 1024 threads per CU, at most 64 VGPRs;
// rounds of 8192 fill the LDS of a CU with one block)
__global__ void __launch_bounds__(1024, RS <= 4096 ? 8 : 4) part_scatter_staged_kernel(PartArgs a) {
  constexpr int PER = RS / 1024;
  // drops whose step counts the block keeps in LDS: with 1024 tiles and 4096
  // staged entries the block stays within 80 KB, two blocks per CU
  constexpr int kMaxW = LDSW ? scatter_lds_drops<MAXT, RS>() : 1;
  __shared__ int cur[MAXT];  // next entry of each tile of the block's terrain
  __shared__ int cnt[MAXT];  // the round's entries per tile
  __shared__ int off[MAXT];  // their exclusive offsets in the stage
  __shared__ int4 stage[RS];
  __shared__ int wst[kMaxW];  // steps of the block's drops
  __shared__ int wsum[16];
  const int tid = threadIdx.x;
  const PartBlock col = part_block(a);
  const int g0 = col.g0, width = col.width, ns = col.nslots;
  if (ns == 0) return;  // uniform over the block
  const int wlo = col.wlo, wn = col.wn;
  if (LDSW)
    for (int d = tid; d < width; d += 1024) wst[d] = a.steps[g0 + d];
  for (int t = tid; t < wn; t += 1024) cur[t] = a.hist[col.base + (long long)t * col.stride];
  __syncthreads();
  // A thread's slots are tid, tid + 1024, tid + 2048, ... (round r0 takes
  // r0 + u * 1024 + tid for u < PER): slot x is row x / width, drop x %
  // width, kept as a cursor advanced by 1024 slots at a time.  A slot is
  // valid when its row lies below its drop's step count (steps <= rows, so
  // that also bounds x < ns).
  const int qa = 1024 / width, ra = 1024 - qa * width;
  int crow = tid / width, cd = tid - crow * width;
  auto next_slot = [&]() {
    const int st = LDSW ? wst[cd] : a.steps[g0 + cd];
    const int k = crow < st ? crow * a.drops + g0 + cd : -1;
    cd += ra;
    crow += qa;
    if (cd >= width) {
      cd -= width;
      ++crow;
    }
    return k;
  };
  // The next round's log records are loaded while this round is ranked and
  // staged, and taken into registers before this round's entries are
  // written out: the wait for them then does not cover those stores.  The
  // loads are unconditional (an invalid slot reads record 0, unused); the
  // slots are found first, then their loads are issued together.
  int kq[PER], lc[PER];
  int2 lt[PER];
  auto load_round = [&]() {
#pragma unroll
    for (int u = 0; u < PER; ++u) kq[u] = next_slot();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int k = kq[u] >= 0 ? kq[u] : 0;
      lc[u] = ld_stream(&a.slog[k].x);
      lt[u] = ld_stream(&a.strk[k]);
    }
  };
  int4 ev[PER];
  int rk[PER];
  auto take_round = [&]() {
#pragma unroll
    for (int u = 0; u < PER; ++u) ev[u] = make_int4(kq[u], lc[u], lt[u].x, lt[u].y);
  };
  // tile of staged record u (-1: no step)
  auto tile_of = [&](int u) { return ev[u].x >= 0 ? (ev[u].y >> a.tb) - wlo : -1; };
  load_round();
  take_round();
  for (int r0 = 0; r0 < ns; r0 += RS) {
    for (int t = tid; t < wn; t += 1024) cnt[t] = 0;
    load_round();  // round r0 + RS
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int t = tile_of(u);
      rk[u] = t >= 0 ? atomicAdd(&cnt[t], 1) : 0;
    }
    __syncthreads();
    block_scan_runs(cnt, off, wn, wsum);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int t = tile_of(u);
      if (t >= 0) stage[off[t] + rk[u]] = ev[u];
    }
    __syncthreads();
    take_round();
    const int tot = off[wn - 1] + cnt[wn - 1];
    for (int i = tid; i < tot; i += 1024) {
      const int4 v = stage[i];
      const int t = (v.y >> a.tb) - wlo;
      const int e = cur[t] + (i - off[t]);
      a.ecell[e] = v.y;
      a.erest[e] = EntRest{v.x, v.z, v.w};
    }
    __syncthreads();
    for (int t = tid; t < wn; t += 1024) cur[t] += cnt[t];
  }
}


// Per tile: the walk steps starting at each cell (LDS histogram of the
// tile's entries) -> burst scale, the replay's 4-byte gather per step;
// cascade sites where more steps
// start than first steps (an arrival, water.h:151); clears the first-step
// counts.  With S > 1 each tile's entries are split over S blocks (block
// t*S + s takes the s-th part), which write their counts to part[block] and
// tile_count_finish_kernel adds them: the largest tiles hold several times
// the mean's entries, and one block per tile would wait for them.

// entry range of part s of tile t
template <int S>
__device__ inline void tile_part(const int32_t* __restrict__ start, int t, int s, int& a, int& b) {
  const int e0 = start[t], e1 = start[t + 1];
  if (S == 1) {
    a = e0;
    b = e1;
  } else {
    const long long len = e1 - e0;
    a = e0 + (int)(len * s / S);
    b = e0 + (int)(len * (s + 1) / S);
  }
}

// burst scale and cascade site of cell i, k steps starting there
__device__ inline void cell_count_finish(size_t i, int k, float* sr, int32_t* first,
                                         uint8_t* visit, const float4* __restrict__ fld,
                                         const float* __restrict__ root,
                                         const Terrain* __restrict__ ter, int nter) {
  const int t = find_terrain(nter, (long long)i, [&](int q) { return ter[q].cell_off; });
  const shy_params& P = ter[t].P;
  const float rd = ter[t].has_root ? root[i] : 0.0f;
  float effD = P.depositionRate * (1.0f - rd);
  if (effD < 0) effD = 0;
  sr[i] = burst_scale(k, effD, P.entrainment, fld[i].w);
  const int f = first[i];
  if (k > f) visit[i] = 1;
  if (f) first[i] = 0;
}

template <int S>
__global__ void __launch_bounds__(1024) tile_count_kernel(
    const int32_t* __restrict__ start, const int32_t* __restrict__ ecell, size_t n,
    float* sr, int32_t* first, uint8_t* visit, const float4* __restrict__ fld,
    const float* __restrict__ root, const Terrain* __restrict__ ter, int nter, int32_t* part) {
  __shared__ int cnt[kTileCells];
  for (int j = threadIdx.x; j < kTileCells; j += blockDim.x) cnt[j] = 0;
  __syncthreads();
  const int t = blockIdx.x / S, s = blockIdx.x % S;
  const int base = t << kTileBits;
  int e0, e1;
  tile_part<S>(start, t, s, e0, e1);
  // four entries per thread in flight
  for (int e = e0 + (int)threadIdx.x; e < e1; e += 4 * blockDim.x) {
    int c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int x = e + u * (int)blockDim.x;
      c[u] = x < e1 ? ecell[x] - base : -1;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (c[u] >= 0) atomicAdd(&cnt[c[u]], 1);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < kTileCells; j += blockDim.x) {
    if (S > 1) {
      part[(size_t)blockIdx.x * kTileCells + j] = cnt[j];
    } else {
      const size_t i = (size_t)base + j;
      if (i >= n) break;
      cell_count_finish(i, cnt[j], sr, first, visit, fld, root, ter, nter);
    }
  }
}

// the S parts' counts of every cell (fixed order), then as tile_count_kernel
template <int S>
__global__ void tile_count_finish_kernel(const int32_t* __restrict__ part, size_t n, float* sr,
                                         int32_t* first, uint8_t* visit,
                                         const float4* __restrict__ fld,
                                         const float* __restrict__ root,
                                         const Terrain* __restrict__ ter, int nter) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t t = i >> kTileBits, j = i & (kTileCells - 1);
  int k = 0;
#pragma unroll
  for (int s = 0; s < S; ++s) k += part[(t * S + s) * kTileCells + j];
  cell_count_finish(i, k, sr, first, visit, fld, root, ter, nter);
}

// quant_t of the volume a drop of terrain t carries into step k, for every
// terrain and step: vol_0 = 1, vol_k+1 = evap_vol(vol_k) (water.h:136), the
// same rounding as the walk.  [terrain][rows]
__global__ void volq_kernel(const Terrain* __restrict__ ter, int nter, int rows, int32_t* vq) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nter) return;
  const float e = ter[t].P.evapRate;
  float vol = 1.0f;
  for (int k = 0; k < rows; ++k) {
    vq[(size_t)t * rows + k] = quant_t(vol);
    vol = evap_vol(vol, e);
  }
}

// Per tile: the height commits (gathered from the replay's drop-major
// output) and the tracks of every step starting at each cell, summed in
// LDS in integers (so the result does not depend on the order drops ran),
// plus the terminal deposits; clears the deposits.  The volume track of a
// step is vq[terrain][step row].  With S > 1 the tile's entries are split as
// in tile_count_kernel and tile_reduce_finish_kernel adds the parts.
template <int S>
__global__ void __launch_bounds__(1024) tile_reduce_kernel(
    const int32_t* __restrict__ start, const int32_t* __restrict__ ecell,
    const EntRest* __restrict__ erest,
    const int32_t* __restrict__ qb, int qstride, const int32_t* __restrict__ vq, int rows,
    int drops, const Terrain* __restrict__ ter, int nter, int32_t* deposit, size_t n,
    int32_t* dh, long long* td, long long* tmx, long long* tmy, long long* part) {
  __shared__ unsigned long long acc[4][kTileCells];
  for (int j = threadIdx.x; j < kTileCells; j += blockDim.x) {
    acc[0][j] = 0;
    acc[1][j] = 0;
    acc[2][j] = 0;
    acc[3][j] = 0;
  }
  __syncthreads();
  const int tile = blockIdx.x / S, s = blockIdx.x % S;
  const int base = tile << kTileBits;
  // terrains of the tile's first and last cell: one terrain in nearly every tile
  auto key = [&](int q) { return ter[q].cell_off; };
  const int t_lo = find_terrain(nter, (long long)base, key);
  const long long last = (long long)base + kTileCells - 1 < (long long)n ? base + kTileCells - 1
                                                                          : (long long)n - 1;
  const int t_hi = find_terrain(nter, last, key);
  int e0, e1;
  tile_part<S>(start, tile, s, e0, e1);
  // four entries per thread in flight: the entry loads, then their gathers
  for (int e = e0 + (int)threadIdx.x; e < e1; e += 4 * blockDim.x) {
    int4 p[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int x = e + u * (int)blockDim.x;
      if (x < e1) {
        const EntRest r = erest[x];
        p[u] = make_int4(r.slot, ecell[x], r.tx, r.ty);
      } else {
        p[u] = make_int4(-1, 0, 0, 0);
      }
    }
    int q[4], v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      q[u] = 0;
      v[u] = 0;
      if (p[u].x >= 0) {
        const int t = t_lo == t_hi ? t_lo : find_terrain(nter, (long long)p[u].y, key);
        const int row = p[u].x / drops;
        q[u] = qb[(size_t)(p[u].x - row * drops) * qstride + row];
        v[u] = vq[(size_t)t * rows + row];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (p[u].x >= 0) {
        const int c = p[u].y - base;
        atomicAdd(&acc[0][c], (unsigned long long)(long long)q[u]);
        atomicAdd(&acc[1][c], (unsigned long long)(long long)v[u]);
        atomicAdd(&acc[2][c], (unsigned long long)(long long)p[u].z);
        atomicAdd(&acc[3][c], (unsigned long long)(long long)p[u].w);
      }
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < kTileCells; j += blockDim.x) {
    if (S > 1) {
      long long* o = part + (size_t)blockIdx.x * 4 * kTileCells + j;
#pragma unroll
      for (int f = 0; f < 4; ++f) o[(size_t)f * kTileCells] = (long long)acc[f][j];
    } else {
      const size_t i = (size_t)base + j;
      if (i >= n) break;
      dh[i] = (int32_t)((long long)acc[0][j] + deposit[i]);
      td[i] = (long long)acc[1][j];
      tmx[i] = (long long)acc[2][j];
      tmy[i] = (long long)acc[3][j];
      deposit[i] = 0;
    }
  }
}

// the S parts' sums of every cell (fixed order), then as tile_reduce_kernel
template <int S>
__global__ void tile_reduce_finish_kernel(const long long* __restrict__ part, int32_t* deposit,
                                          size_t n, int32_t* dh, long long* td, long long* tmx,
                                          long long* tmy) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t t = i >> kTileBits, j = i & (kTileCells - 1);
  long long a[4] = {0, 0, 0, 0};
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const long long* o = part + (t * S + s) * 4 * kTileCells + j;
#pragma unroll
    for (int f = 0; f < 4; ++f)
      a[f] = (long long)((unsigned long long)a[f] + (unsigned long long)o[(size_t)f * kTileCells]);
  }
  dh[i] = (int32_t)(a[0] + deposit[i]);
  td[i] = a[1];
  tmx[i] = a[2];
  tmy[i] = a[3];
  deposit[i] = 0;
}

__global__ void trace_kernel(WalkArgs a, float x0, float y0, shy_drop_step* out, int max_steps,
                             int* n_out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int end = 0, end_cell = 0;
  walk_drop<kTrace>(a, 0, x0, y0, &end, &end_cell, out, max_steps, n_out);
}

// --------------------------------------------------------------- cascade
// Neighbour offsets in the reference's enumeration order (world.h:94-103);
// slot 8 is the centre.
__constant__ int kOffX[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
__constant__ int kOffY[8] = {-1, 0, 1, -1, 1, -1, 0, 1};

// World::cascade (world.h:116-166) of one site against the heights read by
// get(dx, dy) (offsets from the centre): in-bounds neighbours sorted
// ascending by height (stable, as libstdc++'s insertion sort for <= 16
// elements), the centre re-read after every transfer.  emit(slot, dq) gets
// each quantised transfer; returns the number of transfers.
template <class Get, class Emit>
__device__ inline int cascade_site(int cx, int cy, int nx, int ny, const shy_params& P,
                                   float hc, Get get, Emit emit) {
  float sh[8], sd[8];
  int sl[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int ox = k < 3 ? -1 : (k < 5 ? 0 : 1);
    const int oy = (k == 0 || k == 3 || k == 5) ? -1 : ((k == 1 || k == 6) ? 0 : 1);
    const bool in = !off_map(cx + ox, cy + oy, nx, ny);
    const float vx = (float)ox, vy = (float)oy;
    sh[k] = in ? get(ox, oy) : __builtin_inff();
    sd[k] = __builtin_sqrtf(vx * vx + vy * vy);  // length(vec2(nn))
    sl[k] = in ? k : -1;
  }
  // stable sort: odd-even transposition (adjacent exchanges on strict >),
  // absent neighbours carry +inf and settle at the end
#pragma unroll
  for (int r = 0; r < 8; ++r) {
#pragma unroll
    for (int i = (r & 1); i + 1 < 8; i += 2) {
      if (sh[i + 1] < sh[i]) {
        float t = sh[i]; sh[i] = sh[i + 1]; sh[i + 1] = t;
        t = sd[i]; sd[i] = sd[i + 1]; sd[i + 1] = t;
        int u = sl[i]; sl[i] = sl[i + 1]; sl[i + 1] = u;
      }
    }
  }
  int moves = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    if (sl[i] < 0) continue;
    const float diff = hc - sh[i];
    if (diff == 0) continue;
    float excess;
    if ((double)sh[i] > 0.1)
      excess = __builtin_fabsf(diff) - sd[i] * P.maxdiff * kLod;
    else
      excess = __builtin_fabsf(diff);
    if (excess <= 0) continue;
    const float transfer = P.settling * excess / 2.0f;
    ++moves;
    if (diff > 0) {
      hc -= transfer;
      emit(sl[i], transfer, true);
    } else {
      hc += transfer;
      emit(sl[i], transfer, false);
    }
  }
  return moves;
}

// One terrain's view of the cascade/blend stencil.
struct CascView {
  const float* __restrict__ h_in;
  float* __restrict__ h_out;
  const int32_t* __restrict__ dh;
  const uint8_t* __restrict__ visit;
  const long long* __restrict__ td;
  const long long* __restrict__ tmx;
  const long long* __restrict__ tmy;
  uint8_t* visit_clear;
  float4* fld;
  const double* __restrict__ erf_a;
  DeviceState* state;
  int nx, ny;
  shy_params P;
};

// The stencil is instantiated for the tile shapes the autotuner chooses
// from (extent along y, the contiguous axis, by extent along x; 256 threads).
template <int kCbY, int kCbX>
__global__ void __launch_bounds__(kCbY * kCbX) cascade_blend_kernel(CascArgs e) {
  constexpr int kCbThreads = kCbY * kCbX;
  __shared__ float Hs[kCbX + 4][kCbY + 4];   // post-walk heights, 2-cell halo
  __shared__ uint8_t Vs[kCbX + 6][kCbY + 6]; // cascade sites, 3-cell halo
  __shared__ uint8_t Wc[kCbX + 4][kCbY + 4]; // site windows covering a cell, 2-cell halo
  __shared__ int32_t Wd[(kCbX + 2) * (kCbY + 2)][9];
  if (blockIdx.x == 0 && e.gflag)
    for (int q = threadIdx.x; q < e.nter; q += kCbThreads) e.gflag[q] = 0;
  const int tt = find_terrain(e.nter, (long long)blockIdx.x,
                              [&](int q) { return (long long)e.ter[q].tile_off; });
  const Terrain& T = e.ter[tt];
  const int lt = (int)blockIdx.x - T.tile_off;
  CascView a;
  a.h_in = e.h_in + T.cell_off;
  a.h_out = e.h_out + T.cell_off;
  a.dh = e.dh + T.cell_off;
  a.visit = e.visit + T.cell_off;
  a.td = e.td + T.cell_off;
  a.tmx = e.tmx + T.cell_off;
  a.tmy = e.tmy + T.cell_off;
  a.visit_clear = e.visit_clear + T.cell_off;
  a.fld = e.fld + T.cell_off;
  a.erf_a = e.erf_a;
  a.state = e.state;
  a.nx = T.nx;
  a.ny = T.ny;
  a.P = T.P;
  const int tid = threadIdx.x;
  const int x0 = (lt / T.tiles_y) * kCbX, y0 = (lt % T.tiles_y) * kCbY;
  const int nx = a.nx, ny = a.ny;
  for (int k = tid; k < (kCbX + 4) * (kCbY + 4); k += kCbThreads) {
    const int lx = k / (kCbY + 4), ly = k % (kCbY + 4);
    const int gx = x0 + lx - 2, gy = y0 + ly - 2;
    float v = 0.0f;
    if (!off_map(gx, gy, nx, ny)) {
      const int i = gx * ny + gy;
      v = a.h_in[i] + (float)a.dh[i] * kHQInv;
    }
    Hs[lx][ly] = v;
  }
  for (int k = tid; k < (kCbX + 6) * (kCbY + 6); k += kCbThreads) {
    const int lx = k / (kCbY + 6), ly = k % (kCbY + 6);
    const int gx = x0 + lx - 3, gy = y0 + ly - 3;
    Vs[lx][ly] = off_map(gx, gy, nx, ny) ? 0 : a.visit[gx * ny + gy];
  }
  __syncthreads();
  for (int k = tid; k < (kCbX + 4) * (kCbY + 4); k += kCbThreads) {
    const int lx = k / (kCbY + 4), ly = k % (kCbY + 4);  // Vs coords lx+1, ly+1
    int cnt = 0;
#pragma unroll
    for (int dx = 0; dx < 3; ++dx)
#pragma unroll
      for (int dy = 0; dy < 3; ++dy) cnt += Vs[lx + dx][ly + dy];
    Wc[lx][ly] = (uint8_t)cnt;
  }
  __syncthreads();
  long long moves = 0;
  for (int k = tid; k < (kCbX + 2) * (kCbY + 2); k += kCbThreads) {
    const int lx = k / (kCbY + 2), ly = k % (kCbY + 2);
    int32_t* w = Wd[k];
#pragma unroll
    for (int s = 0; s < 9; ++s) w[s] = 0;
    if (!Vs[lx + 2][ly + 2]) continue;
    const int gx = x0 + lx - 1, gy = y0 + ly - 1;
    const int hx = lx + 1, hy = ly + 1;  // centre in Hs / Wc coordinates
    // window weight: the most windows covering any of its on-map cells
    int m = 1;
#pragma unroll
    for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy)
        if (!off_map(gx + dx, gy + dy, nx, ny) && Wc[hx + dx][hy + dy] > m) m = Wc[hx + dx][hy + dy];
    const float fm = (float)m;
    // halo windows are recomputed by the neighbouring block; count own ones
    const bool own = lx >= 1 && lx <= kCbX && ly >= 1 && ly <= kCbY;
    const int mv = cascade_site(
        gx, gy, nx, ny, a.P, Hs[hx][hy],
        [&](int ox, int oy) { return Hs[hx + ox][hy + oy]; },
        [&](int slot, float transfer, bool down) {
          const int32_t q = quant_h(transfer / fm);
          if (down) { w[8] -= q; w[slot] += q; } else { w[8] += q; w[slot] -= q; }
        });
    if (own) moves += mv;
  }
  __syncthreads();
  const int tx = tid / kCbY, ty = tid % kCbY;
  const int gx = x0 + tx, gy = y0 + ty;
  const bool mine = !off_map(gx, gy, nx, ny);
  if (mine) {
    const int i = gx * ny + gy;
    int32_t tot = 0;
    // window centred at (gx-ox, gy-oy) sees this cell at offset (ox, oy)
#pragma unroll
    for (int s = 0; s < 9; ++s) {
      const int ox = s == 8 ? 0 : (s < 3 ? -1 : (s < 5 ? 0 : 1));
      const int oy = s == 8 ? 0 : ((s == 0 || s == 3 || s == 5) ? -1 : ((s == 1 || s == 6) ? 0 : 1));
      const int wx = tx - ox + 1, wy = ty - oy + 1;
      if (Vs[wx + 2][wy + 2]) tot += Wd[wx * (kCbY + 2) + wy][s];
    }
    a.h_out[i] = Hs[tx + 2][ty + 2] + (float)tot * kHQInv;
    // world.h:81-86
    const float4 f = a.fld[i];
    const float tD = track_value(a.td[i]);
    const float tX = track_value(a.tmx[i]);
    const float tY = track_value(a.tmy[i]);
    const float keep = 1.0f - a.P.lrate;
    const float d = keep * f.z + a.P.lrate * tD;
    const float mx = keep * f.x + a.P.lrate * tX;
    const float my = keep * f.y + a.P.lrate * tY;
    a.fld[i] = make_float4(mx, my, d, erf_f32(a.erf_a, 0.4f * d));
    a.visit_clear[i] = 0;
  }
  long long v[1] = {moves};
  block_add_stats<1>(v, &a.state->stats[ST_CASC_MOVES]);
  if (blockIdx.x == 0 && tid == 0) a.state->cycle = a.state->cycle + 1u;
}

// ---------------------------------------------------------- field kernels
// terrain fields to a caller's device buffer: 4 floats per cell (h, d, mx, my)
__global__ void export_fields_kernel(const float* __restrict__ h, const float4* __restrict__ fld,
                                     size_t n, float* dst) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 f = fld[i];
  dst[4 * i] = h[i];
  dst[4 * i + 1] = f.z;
  dst[4 * i + 2] = f.x;
  dst[4 * i + 3] = f.y;
}

// discharge/momentum from a caller's device buffer (3 floats per cell:
// d, mx, my); refreshes erf(0.4 d)
__global__ void import_flow_kernel(const float* __restrict__ src, float4* fld, size_t n,
                                   const double* __restrict__ erf_a) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float d = src[3 * i];
  fld[i] = make_float4(src[3 * i + 1], src[3 * i + 2], d, erf_f32(erf_a, 0.4f * d));
}

// quad::map::normal (cellpool.h:181-204, 445-447) of every cell, xyz.
__global__ void normals_kernel(const float* __restrict__ h, int nx, int ny, float* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)nx * ny) return;
  const int x = (int)(i / (size_t)ny), y = (int)(i % (size_t)ny);
  CellView c;
  c.h = h[i];
  c.xm = h[x > 0 ? i - ny : i];
  c.xp = h[x + 1 < nx ? i + ny : i];
  c.ym = h[y > 0 ? i - 1 : i];
  c.yp = h[y + 1 < ny ? i + 1 : i];
  float n0, n1, n2;
  normal_from(c, x, y, nx, ny, n0, n1, n2);
  out[3 * i] = n0;
  out[3 * i + 1] = n1;
  out[3 * i + 2] = n2;
}

// The textures the reference's frame loop builds after each erode call
// (SimpleHydrology.cpp:341-354): dischargeMap = (water colour, erf(0.4 d)),
// momentumMap = (0.5 (1 + erf(mx)), 0.5 (1 + erf(my)), 0.5, 1).
__global__ void textures_kernel(const float4* __restrict__ fld, size_t n, float wr, float wg,
                                float wb, const double* __restrict__ erf_a, float4* dis,
                                float4* mom) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 f = fld[i];
  if (dis) dis[i] = make_float4(wr, wg, wb, f.w);
  if (mom)
    mom[i] = make_float4(0.5f * (1.0f + erf_f32(erf_a, f.x)), 0.5f * (1.0f + erf_f32(erf_a, f.y)),
                         0.5f, 1.0f);
}

// updatenode's vertex (cellpool.h:286-304): P = (x, 80 h(x,y), y), the
// normal, T - P and B - P with T = (x+1, 80 h(x+1,y), y), B = (x, 80 h(x,y+1),
// y+1); a height off the map reads 0 (quad::node::height).
__global__ void vertices_kernel(const float* __restrict__ h, int nx, int ny, float* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)nx * ny) return;
  const int x = (int)(i / (size_t)ny), y = (int)(i % (size_t)ny);
  CellView c;
  c.h = h[i];
  c.xm = h[x > 0 ? i - ny : i];
  c.xp = h[x + 1 < nx ? i + ny : i];
  c.ym = h[y > 0 ? i - 1 : i];
  c.yp = h[y + 1 < ny ? i + 1 : i];
  float n0, n1, n2;
  normal_from(c, x, y, nx, ny, n0, n1, n2);
  const float py = kMapScale * c.h;
  const float ty = kMapScale * (x + 1 < nx ? c.xp : 0.0f);
  const float by = kMapScale * (y + 1 < ny ? c.yp : 0.0f);
  float* o = out + 12 * i;
  o[0] = (float)x; o[1] = py; o[2] = (float)y;
  o[3] = n0; o[4] = n1; o[5] = n2;
  o[6] = (float)(x + 1) - (float)x; o[7] = ty - py; o[8] = (float)y - (float)y;
  o[9] = (float)x - (float)x; o[10] = by - py; o[11] = (float)(y + 1) - (float)y;
}

__global__ void refresh_erf_kernel(float4* fld, size_t n, const double* __restrict__ erf_a) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 f = fld[i];
  f.w = erf_f32(erf_a, 0.4f * f.z);
  fld[i] = f;
}

// ----------------------------------------------------------- vegetation
// Vegetation::grow (vegetation.h:120-186) on the device, generation by
// generation: the spawn attempt, then every plant of the current generation
// grows, may die (root -1) and may seed a child (root +1), all against the
// map and root density as they stood when the generation began; the
// children form the next generation of the same call (the reference's loop
// also visits plants appended during it).  Root-density changes are integer
// units of 0.2 (the stencil weights 1.0, 0.6, 0.4 are 5, 3, 2 units), summed
// order-free and applied once per generation.  Random draws replace rand():
// counter hashes of (seed, step, plant slot, purpose).

struct VegArgs {
  const float* __restrict__ h;
  const float4* __restrict__ fld;  // .w = erf(0.4 discharge) = map.discharge
  float* root;
  int32_t* units;                  // pending root-density change, 0.2 units
  int nx, ny;
  int32_t* px;
  int32_t* py;
  float* size;
  int32_t* alive;
  int32_t* cflag;
  int32_t* cx;
  int32_t* cy;
  int32_t* st;                     // {count, children of the last generation, dropped}
  int cap;
  shy_plant_params P;
};

enum VegDraw { kVegX = 0, kVegY = 1, kVegDie = 2, kVegTry = 3, kVegDx = 4, kVegDy = 5, kVegRoot = 6 };
constexpr uint32_t kVegSpawnSlot = 0xffffffffu;

__host__ __device__ inline uint32_t veg_draw(uint64_t key, uint32_t slot, int purpose, uint32_t m) {
  const uint64_t r = mix64(key ^ (((uint64_t)slot << 8) | (uint64_t)purpose));
  return (uint32_t)(((r >> 32) * (uint64_t)m) >> 32);
}

__device__ inline float veg_normal_y(const VegArgs& a, int x, int y) {
  const int ny = a.ny;
  const size_t i = (size_t)x * ny + y;
  CellView c;
  c.h = a.h[i];
  c.xm = a.h[x > 0 ? i - ny : i];
  c.xp = a.h[x + 1 < a.nx ? i + ny : i];
  c.ym = a.h[y > 0 ? i - 1 : i];
  c.yp = a.h[y + 1 < ny ? i + 1 : i];
  float n0, n1, n2;
  normal_from(c, x, y, a.nx, ny, n0, n1, n2);
  return n1;
}

// Plant::root (vegetation.h:87-117): weights 1.0 / 0.6 / 0.4 = 5 / 3 / 2 units
__device__ inline void veg_root(const VegArgs& a, int x, int y, int sign) {
#pragma unroll
  for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
      const int u = (dx == 0 && dy == 0) ? 5 : (dx == 0 || dy == 0) ? 3 : 2;
      const int X = x + dx, Y = y + dy;
      if (off_map(X, Y, a.nx, a.ny)) continue;
      __hip_atomic_fetch_add(&a.units[(size_t)X * a.ny + Y], sign * u, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Vegetation::grow's spawn attempt (vegetation.h:124-136, Plant::spawn :72-85)
__global__ void veg_spawn_kernel(VegArgs a, uint64_t key) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int x = (int)veg_draw(key, kVegSpawnSlot, kVegX, (uint32_t)a.nx);
  const int y = (int)veg_draw(key, kVegSpawnSlot, kVegY, (uint32_t)a.ny);
  const size_t i = (size_t)x * a.ny + y;
  if (a.fld[i].w >= a.P.maxDischarge) return;
  if (veg_normal_y(a, x, y) < a.P.maxSteep) return;
  if (a.h[i] >= a.P.maxTreeHeight) return;
  const int n = a.st[0];
  if (n >= a.cap) {
    a.st[2] += 1;
    return;
  }
  a.px[n] = x;
  a.py[n] = y;
  a.size[n] = 0.0f;
  a.alive[n] = 1;
  a.st[0] = n + 1;
  veg_root(a, x, y, 1);
}

// one generation [g0, g1): grow, die (Plant::die :63-70), seed a child (:144-184)
__global__ void veg_gen_kernel(VegArgs a, uint64_t key, int g0, int g1) {
  const int k = g0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g1) return;
  a.cflag[k] = 0;
  const shy_plant_params& P = a.P;
  float sz = a.size[k];
  sz += P.growRate * (P.maxSize - sz);  // Plant::grow, :59-61
  a.size[k] = sz;
  const int x = a.px[k], y = a.py[k];
  const size_t i = (size_t)x * a.ny + y;
  if (a.fld[i].w >= P.maxDischarge || a.h[i] >= P.maxTreeHeight ||
      veg_draw(key, (uint32_t)k, kVegDie, 1000) == 0) {
    a.alive[k] = 0;
    veg_root(a, x, y, -1);
    return;
  }
  if (veg_draw(key, (uint32_t)k, kVegTry, 20) != 0) return;
  const int nx = x + (int)veg_draw(key, (uint32_t)k, kVegDx, 9) - 4;
  const int ny = y + (int)veg_draw(key, (uint32_t)k, kVegDy, 9) - 4;
  if (off_map(nx, ny, a.nx, a.ny)) return;
  const size_t j = (size_t)nx * a.ny + ny;
  if (a.fld[j].w >= P.maxDischarge) return;
  const float r = (float)veg_draw(key, (uint32_t)k, kVegRoot, 1000);
  if ((double)r / 1000.0 <= (double)a.root[j]) return;
  if (veg_normal_y(a, nx, ny) <= P.maxSteep) return;
  a.cflag[k] = 1;
  a.cx[k] = nx;
  a.cy[k] = ny;
}

// exclusive block scan (1024 threads) of v; returns the block total
__device__ inline int veg_block_scan(int v, int* tmp, int& total) {
  return block_exclusive_scan(v, tmp, total);
}

// appends the children of [g0, g1) in parent order (one block)
__global__ void __launch_bounds__(kScanBlock) veg_append_kernel(VegArgs a, int g0, int g1) {
  __shared__ int tmp[kScanBlock];
  __shared__ int base;
  if (threadIdx.x == 0) base = a.st[0];
  __syncthreads();
  int added = 0, dropped = 0;
  for (int c0 = g0; c0 < g1; c0 += kScanBlock) {
    const int k = c0 + threadIdx.x;
    const int f = k < g1 ? a.cflag[k] : 0;
    int total;
    const int off = veg_block_scan(f, tmp, total);
    const int b = base;
    if (f) {
      const int slot = b + off;
      if (slot < a.cap) {
        a.px[slot] = a.cx[k];
        a.py[slot] = a.cy[k];
        a.size[slot] = 0.0f;
        a.alive[slot] = 1;
        veg_root(a, a.cx[k], a.cy[k], 1);
        added += 1;
      } else {
        dropped += 1;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) base = b + total;
    __syncthreads();
  }
  // totals (slots are contiguous from the old count, so the new count is
  // min(old + children, capacity))
  int tot_add, tot_drop;
  veg_block_scan(added, tmp, tot_add);
  veg_block_scan(dropped, tmp, tot_drop);
  if (threadIdx.x == 0) {
    a.st[0] += tot_add;
    a.st[1] = tot_add;
    a.st[2] += tot_drop;
  }
}

// root density += 0.2 x the generation's units; clears the units
__global__ void veg_apply_kernel(float* root, int32_t* units, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int u = units[i];
  if (u) {
    root[i] += 0.2f * (float)u;
    units[i] = 0;
  }
}

// removes dead plants, keeping the order (one block)
__global__ void __launch_bounds__(kScanBlock) veg_compact_kernel(VegArgs a) {
  __shared__ int tmp[kScanBlock];
  __shared__ int base;
  if (threadIdx.x == 0) base = 0;
  __syncthreads();
  const int n = a.st[0];
  for (int c0 = 0; c0 < n; c0 += kScanBlock) {
    const int k = c0 + threadIdx.x;
    const int f = k < n ? a.alive[k] : 0;
    int x = 0, y = 0;
    float sz = 0.0f;
    if (f) {
      x = a.px[k];
      y = a.py[k];
      sz = a.size[k];
    }
    int total;
    const int off = veg_block_scan(f, tmp, total);
    const int b = base;
    __syncthreads();
    if (f) {
      const int d = b + off;
      a.px[d] = x;
      a.py[d] = y;
      a.size[d] = sz;
      a.alive[d] = 1;
    }
    __syncthreads();
    if (threadIdx.x == 0) base = b + total;
    __syncthreads();
  }
  if (threadIdx.x == 0) a.st[0] = base;
}

// ------------------------------------------------------------------ fBm
__device__ inline float fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }

__device__ inline float grad2(int h, float x, float y) {
  switch (h & 7) {
    case 0: return x + y;
    case 1: return -x + y;
    case 2: return x - y;
    case 3: return -x - y;
    case 4: return x;
    case 5: return -x;
    case 6: return y;
    default: return -y;
  }
}

// Perlin gradient noise in 2D (Perlin 2002 fade curve, 8 gradient
// directions) over a per-octave 256-entry permutation.
__device__ inline float perlin2(const uint8_t* __restrict__ P, float x, float y) {
  const float xf = __builtin_floorf(x), yf = __builtin_floorf(y);
  const int xi = ((int)xf) & 255, yi = ((int)yf) & 255;
  const float fx = x - xf, fy = y - yf;
  const float u = fade(fx), v = fade(fy);
  const int xj = (xi + 1) & 255;
  const int aa = P[(P[xi] + yi) & 255], ab = P[(P[xi] + yi + 1) & 255];
  const int ba = P[(P[xj] + yi) & 255], bb = P[(P[xj] + yi + 1) & 255];
  const float n00 = grad2(aa, fx, fy);
  const float n10 = grad2(ba, fx - 1.0f, fy);
  const float n01 = grad2(ab, fx, fy - 1.0f);
  const float n11 = grad2(bb, fx - 1.0f, fy - 1.0f);
  const float x1 = n00 + u * (n10 - n00);
  const float x2 = n01 + u * (n11 - n01);
  return x1 + v * (x2 - x1);
}

__global__ void fbm_kernel(float* out, int nx, int ny, const uint8_t* __restrict__ perms,
                           int octaves, float base, float lac, float gain) {
  __shared__ uint8_t P[16 * 256];
  for (int k = threadIdx.x; k < octaves * 256; k += blockDim.x) P[k] = perms[k];
  __syncthreads();
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)nx * ny) return;
  const int x = (int)(i / (size_t)ny), y = (int)(i % (size_t)ny);
  float s = 0.0f, amp = 1.0f, f = base;
  for (int o = 0; o < octaves; ++o) {
    s = s + amp * perlin2(P + o * 256, (float)x * f, (float)y * f);
    f = f * lac;
    amp = amp * gain;
  }
  out[i] = s;
}

// ------------------------------------------- the reference's own terrain
// The OpenSimplex2 restatement below follows FastNoiseLite, which carries
// this notice:
//
//   MIT License
//
//   Copyright(c) 2020 Jordan Peck (jordan.me2@gmail.com)
//   Copyright(c) 2020 Contributors
//
//   Permission is hereby granted, free of charge, to any person obtaining a
//   copy of this software and associated documentation files(the "Software"),
//   to deal in the Software without restriction, including without limitation
//   the rights to use, copy, modify, merge, publish, distribute, sublicense,
//   and / or sell copies of the Software, and to permit persons to whom the
//   Software is furnished to do so, subject to the following conditions :
//
//   The above copyright notice and this permission notice shall be included in
//   all copies or substantial portions of the Software.
//
//   THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS OR
//   IMPLIED, INCLUDING BUT NOT LIMITED TO THE WARRANTIES OF MERCHANTABILITY,
//   FITNESS FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT.IN NO EVENT SHALL THE
//   AUTHORS OR COPYRIGHT HOLDERS BE LIABLE FOR ANY CLAIM, DAMAGES OR OTHER
//   LIABILITY, WHETHER IN AN ACTION OF CONTRACT, TORT OR OTHERWISE, ARISING
//   FROM, OUT OF OR IN CONNECTION WITH THE SOFTWARE OR THE USE OR OTHER
//   DEALINGS IN THE SOFTWARE.
//
// quad::map::init (cellpool.h:349-409): eight layers of FastNoiseLite noise
// (OpenSimplex2, 3D, FBm of 3 octaves, seed 1337, lacunarity 2, gain 0.5,
// bounding 1/1.75) at frequencies 1, 2, 4, ... and weights 0.6, 0.36, ...,
// sampled at (x/tileres, y/tileres, SEED % 10000); then (h - min)/(max - min)
// with min <= 0 <= max.  The noise follows FastNoiseLite.h:321-339 (GetNoise),
// 687-727 (the DefaultOpenSimplex2 rotation), 866-885 (FBm) and 1054-1153
// (SingleOpenSimplex2), every float operation in the same order.

constexpr int32_t kOsPrimeX = 501125321;
constexpr int32_t kOsPrimeY = 1136930381;
constexpr int32_t kOsPrimeZ = 1720413743;

// integer products wrap mod 2^32 (the reference's signed overflow)
__device__ inline int32_t wrap_mul(int32_t a, int32_t b) {
  return (int32_t)((uint32_t)a * (uint32_t)b);
}
__device__ inline int32_t wrap_add(int32_t a, int32_t b) {
  return (int32_t)((uint32_t)a + (uint32_t)b);
}

// The 3D gradient set of FastNoiseLite.h:2529-2547: 64 entries = the 12
// cube-edge directions five times over, then edges 8, 1, 9, 3.  Edge e has
// its zero in component e/4; the two non-zero components are + unless bit 0
// (first) or bit 1 (second) of e%4 is set.
__device__ inline float os2_grad(int32_t seed, int32_t i, int32_t j, int32_t k, float x, float y,
                                 float z) {
  int32_t h = wrap_mul(seed ^ i ^ j ^ k, 0x27d4eb2d);  // FastNoiseLite.h:500-506
  h ^= h >> 15;                                       // arithmetic shift
  const int entry = (h & (63 << 2)) >> 2;             // :542-554
  const int e = entry < 60 ? entry % 12 : (0x3918 >> (4 * (entry - 60))) & 15;
  const float a = (e & 1) ? -1.0f : 1.0f, b = (e & 2) ? -1.0f : 1.0f;
  const int zero = e >> 2;
  const float gx = zero == 0 ? 0.0f : a;
  const float gy = zero == 0 ? a : (zero == 1 ? 0.0f : b);
  const float gz = zero == 2 ? 0.0f : b;
  return x * gx + y * gy + z * gz;
}

__device__ inline int os2_round(float f) { return f >= 0 ? (int)(f + 0.5f) : (int)(f - 0.5f); }

// SingleOpenSimplex2 (FastNoiseLite.h:1054-1153): two offset rotated cube grids
__device__ inline float os2_single3(int32_t seed, float x, float y, float z) {
  int32_t i = os2_round(x), j = os2_round(y), k = os2_round(z);
  float x0 = (float)(x - (float)i), y0 = (float)(y - (float)j), z0 = (float)(z - (float)k);
  int xs = (int)(-1.0f - x0) | 1, ys = (int)(-1.0f - y0) | 1, zs = (int)(-1.0f - z0) | 1;
  float ax0 = (float)xs * -x0, ay0 = (float)ys * -y0, az0 = (float)zs * -z0;
  i = wrap_mul(i, kOsPrimeX);
  j = wrap_mul(j, kOsPrimeY);
  k = wrap_mul(k, kOsPrimeZ);
  float value = 0.0f;
  float a = (0.6f - x0 * x0) - (y0 * y0 + z0 * z0);
  for (int l = 0;; ++l) {
    if (a > 0) value += (a * a) * (a * a) * os2_grad(seed, i, j, k, x0, y0, z0);
    float b = a + 1.0f;
    int32_t i1 = i, j1 = j, k1 = k;
    float x1 = x0, y1 = y0, z1 = z0;
    if (ax0 >= ay0 && ax0 >= az0) {
      x1 += (float)xs;
      b -= (float)(xs * 2) * x1;
      i1 = wrap_add(i1, -wrap_mul(xs, kOsPrimeX));
    } else if (ay0 > ax0 && ay0 >= az0) {
      y1 += (float)ys;
      b -= (float)(ys * 2) * y1;
      j1 = wrap_add(j1, -wrap_mul(ys, kOsPrimeY));
    } else {
      z1 += (float)zs;
      b -= (float)(zs * 2) * z1;
      k1 = wrap_add(k1, -wrap_mul(zs, kOsPrimeZ));
    }
    if (b > 0) value += (b * b) * (b * b) * os2_grad(seed, i1, j1, k1, x1, y1, z1);
    if (l == 1) break;
    ax0 = 0.5f - ax0;
    ay0 = 0.5f - ay0;
    az0 = 0.5f - az0;
    x0 = (float)xs * ax0;
    y0 = (float)ys * ay0;
    z0 = (float)zs * az0;
    a += (0.75f - ax0) - (ay0 + az0);
    i = wrap_add(i, (xs >> 1) & kOsPrimeX);
    j = wrap_add(j, (ys >> 1) & kOsPrimeY);
    k = wrap_add(k, (zs >> 1) & kOsPrimeZ);
    xs = -xs;
    ys = -ys;
    zs = -zs;
    seed = ~seed;
  }
  return value * 32.69428253173828125f;
}

// GetNoise(x, y, z) of a FastNoiseLite(1337) set to OpenSimplex2 + FBm with
// its default 3 octaves, at frequency f
__device__ inline float fnl_noise3(float f, float x, float y, float z) {
  x *= f;
  y *= f;
  z *= f;
  const float r3 = (float)(2.0 / 3.0);
  const float r = (x + y + z) * r3;  // rotation, not skew
  x = r - x;
  y = r - y;
  z = r - z;
  int32_t seed = 1337;
  float sum = 0.0f, amp = 1.0f / 1.75f;
  for (int o = 0; o < 3; ++o) {
    const float noise = os2_single3(seed++, x, y, z);
    sum += noise * amp;
    amp *= 1.0f + 0.0f * ((noise + 1.0f) * 0.5f - 1.0f);  // Lerp(1, .., weightedStrength 0)
    x *= 2.0f;
    y *= 2.0f;
    z *= 2.0f;
    amp *= 0.5f;
  }
  return sum;
}

__global__ void fnl_world_kernel(float* out, int nx, int ny, int seed, float tileres) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)nx * ny) return;
  const int x = (int)(i / (size_t)ny), y = (int)(i % (size_t)ny);
  const float px = (float)x / tileres, py = (float)y / tileres;
  const float pz = (float)(seed % 10000);
  float h = 0.0f, freq = 1.0f, scale = 0.6f;
  for (int o = 0; o < 8; ++o) {
    h += scale * fnl_noise3(freq, px, py, pz);
    freq *= 2.0f;
    scale = (float)((double)scale * 0.6);
  }
  out[i] = h;
}

// min/max of a float array: per-block partials then one block folds them.
__global__ void minmax_partial_kernel(const float* __restrict__ v, size_t n, float2* part) {
  __shared__ float smn[256], smx[256];
  float mn = __builtin_inff(), mx = -__builtin_inff();
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float t = v[i];
    mn = t < mn ? t : mn;
    mx = t > mx ? t : mx;
  }
  smn[threadIdx.x] = mn;
  smx[threadIdx.x] = mx;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smn[threadIdx.x] = smn[threadIdx.x + s] < smn[threadIdx.x] ? smn[threadIdx.x + s] : smn[threadIdx.x];
      smx[threadIdx.x] = smx[threadIdx.x + s] > smx[threadIdx.x] ? smx[threadIdx.x + s] : smx[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = make_float2(smn[0], smx[0]);
}

// fold the per-block partials into part[0] (one block); with_zero starts
// the fold at min = max = 0 (quad::map::init, cellpool.h:385-392)
__global__ void minmax_fold_kernel(float2* part, int nparts, int with_zero) {
  __shared__ float smn[256], smx[256];
  float mn = with_zero ? 0.0f : __builtin_inff(), mx = with_zero ? 0.0f : -__builtin_inff();
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
    const float2 p = part[k];
    mn = p.x < mn ? p.x : mn;
    mx = p.y > mx ? p.y : mx;
  }
  smn[threadIdx.x] = mn;
  smx[threadIdx.x] = mx;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smn[threadIdx.x] = smn[threadIdx.x + s] < smn[threadIdx.x] ? smn[threadIdx.x + s] : smn[threadIdx.x];
      smx[threadIdx.x] = smx[threadIdx.x + s] > smx[threadIdx.x] ? smx[threadIdx.x + s] : smx[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[0] = make_float2(smn[0], smx[0]);
}

__global__ void normalize_kernel(float* v, size_t n, const float2* part) {
  const float2 p = part[0];
  const float range = p.y - p.x;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  v[i] = range > 0.0f ? (v[i] - p.x) / range : 0.0f;
}

// ------------------------------------------ single-site reference kernels
// World::cascade(pos) on the live map (world.h:90-168).
__device__ void cascade_live(float* h, int nx, int ny, const shy_params& P, float fx, float fy) {
  const int cx = cell_coord(fx), cy = cell_coord(fy);
  if (off_map(cx, cy, nx, ny)) return;
  const int ci = cx * ny + cy;
  // neighbour heights are captured before the loop (world.h:116-125); the
  // centre is re-read from the map every iteration (world.h:138)
  float cap[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int x = cx + kOffX[k], y = cy + kOffY[k];
    cap[k] = off_map(x, y, nx, ny) ? 0.0f : h[x * ny + y];
  }
  cascade_site(
      cx, cy, nx, ny, P, h[ci],
      [&](int ox, int oy) {
        const int s = ox < 0 ? 1 + oy : (ox == 0 ? (oy < 0 ? 3 : 4) : 6 + oy);
        return cap[s];
      },
      [&](int slot, float transfer, bool down) {
        const int n = (cx + kOffX[slot]) * ny + (cy + kOffY[slot]);
        if (down) { h[ci] -= transfer; h[n] += transfer; } else { h[ci] += transfer; h[n] -= transfer; }
      });
}

__global__ void cascade_live_kernel(float* h, int nx, int ny, shy_params P, float fx, float fy) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  cascade_live(h, nx, ny, P, fx, fy);
}

struct LiveArgs {
  float* h;
  const float4* fld;
  const float* root;
  long long* td;  // the tracks visible until the next cycle
  long long* tmx;
  long long* tmy;
  const double* erf_a;
  int nx, ny;
  shy_params P;
};

// One Drop::descend() call on the live map (water.h:58-156), one thread:
// writes go straight to the map and the cascade runs inline, as in the
// reference.  Tracks accumulate in fixed point (int32 steps of 2^-24 summed
// into int64).
__global__ void descend_live_kernel(LiveArgs a, shy_drop* drop, int* alive) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  shy_drop D = *drop;
  const shy_params& P = a.P;
  const int nx = a.nx, ny = a.ny;
  const int ix = cell_coord(D.pos[0]), iy = cell_coord(D.pos[1]);
  *alive = 0;
  if (off_map(ix, iy, nx, ny)) return;  // water.h:62-68
  const int i0 = ix * ny + iy;
  CellView c;
  c.h = a.h[i0];
  c.xm = a.h[ix > 0 ? i0 - ny : i0];
  c.xp = a.h[ix + 1 < nx ? i0 + ny : i0];
  c.ym = a.h[iy > 0 ? i0 - 1 : i0];
  c.yp = a.h[iy + 1 < ny ? i0 + 1 : i0];
  float n0, n1, n2;
  normal_from(c, ix, iy, nx, ny, n0, n1, n2);
  (void)n1;
  if ((float)D.age > P.maxAge || D.volume < P.minVol) {  // water.h:74-82
    a.h[i0] += D.sediment;
    *drop = D;
    return;
  }
  const float4 f = a.fld[i0];
  float effD = P.depositionRate * (1.0f - a.root[i0]);
  if (effD < 0) effD = 0;
  float sx = D.speed[0], sy = D.speed[1], px = D.pos[0], py = D.pos[1];
  step_motion(P, n0, n2, D.volume, f.z, f.x, f.y, sx, sy, px, py);
  a.td[i0] += quant_t(D.volume);
  a.tmx[i0] += quant_t(D.volume * sx);
  a.tmy[i0] += quant_t(D.volume * sy);
  const int jx = cell_coord(px), jy = cell_coord(py);
  const bool ex = off_map(jx, jy, nx, ny);
  const float h2 = ex ? (float)((double)a.h[i0] - 0.002) : a.h[jx * ny + jy];
  const float E = erf_f32(a.erf_a, 0.4f * f.z);
  float c_eq = (1.0f + P.entrainment * E) * (a.h[i0] - h2);
  if (c_eq < 0) c_eq = 0;
  const float cdiff = (c_eq - D.sediment);
  D.sediment += effD * cdiff;
  a.h[i0] -= effD * cdiff;
  D.sediment = evap_sed(D.sediment, P.evapRate);
  D.volume = evap_vol(D.volume, P.evapRate);
  D.speed[0] = sx; D.speed[1] = sy; D.pos[0] = px; D.pos[1] = py;
  if (ex) {
    D.volume = 0.0f;
    *drop = D;
    return;
  }
  cascade_live(a.h, nx, ny, P, px, py);
  D.age += 1;
  *drop = D;
  *alive = 1;
}

}  // namespace shy
