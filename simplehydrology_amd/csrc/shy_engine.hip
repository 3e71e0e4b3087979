// shy_engine.hip -- C ABI (include/shy.h) over the gfx950 kernels.
//
// A context holds one terrain in HBM as structure-of-arrays, x-major:
//   h[2]        float   heights, ping-pong (the cascade stencil reads a halo)
//   fld         float4  {momentumx, momentumy, discharge, erf(0.4 discharge)}
//   root        float   root density
//   visit[2]    uint8   cascade sites, ping-pong
//   count       int32   walk steps per start cell (tickets); cleared by reduce
//   offset      int32   bucket start per cell (prefix sum of count)
//   scale       float   burst scale per cell
//   dh          int32   summed walk commits per cell, 2^-20 (written by reduce)
//   trk[3]      int64   summed tracks per cell, 2^-24 (written by reduce; stay
//                       readable until the next cycle, as the reference's
//                       *_track fields do)
//   deposit     int32   terminal deposits, 2^-20; cleared by reduce
// and per wave (sized by drops x (maxAge+1) steps):
//   ticket      int32   [step][drop] ticket log of the count pass
//   rec         int4    per-step commit records, bucketed by start cell
// Parity p = cycle & 1 selects h and visit; the cascade/blend kernel of cycle
// p clears visit[p^1] for the next cycle, so a cycle needs no memset and two
// cycles form a replayable hipGraph.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "shy_kernels.hpp"

using namespace shy;

struct shy_ctx {
  int device = 0;
  int nx = 0, ny = 0;
  size_t n = 0;
  shy_params P{};
  uint64_t seed = 1;
  uint32_t cycle = 0;  // host mirror of state->cycle
  float* h[2] = {nullptr, nullptr};
  float4* fld = nullptr;
  float* root = nullptr;
  bool has_root = false;
  uint8_t* visit[2] = {nullptr, nullptr};
  int32_t* count = nullptr;
  int32_t* offset = nullptr;
  float* scale = nullptr;
  int32_t* dh = nullptr;
  long long* trk = nullptr;  // 3*n: d, mx, my
  int32_t* deposit = nullptr;
  int32_t* bsum = nullptr;
  int32_t* ticket = nullptr;
  int4* rec = nullptr;
  size_t wave_cap = 0;  // steps the ticket log and buckets can hold
  DeviceState* state = nullptr;
  double* erf_a = nullptr;
  float2* part = nullptr;  // min/max partials
  uint8_t* perms = nullptr;
  shy_drop_step* trace_buf = nullptr;
  int trace_cap = 0;
  int* trace_n = nullptr;
  void* drop_buf = nullptr;
  hipStream_t stream = nullptr;
  int walk_block = 256;
  hipGraphExec_t graph = nullptr;
  int graph_drops = -1;
  std::string err;
};

namespace {

constexpr int kPartBlocks = 1024;
constexpr float kMaxAgeLimit = 1.0e6f;

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

int parity(const shy_ctx* c) { return (int)(c->cycle & 1u); }

WalkArgs walk_args(shy_ctx* c, int drops) {
  const int p = parity(c);
  WalkArgs a;
  a.h = c->h[p];
  a.fld = c->fld;
  a.root = c->has_root ? c->root : nullptr;
  a.count = c->count;
  a.ticket = c->ticket;
  a.offset = c->offset;
  a.scale = nullptr;
  a.rec = c->rec;
  a.deposit = c->deposit;
  a.visit = c->visit[p];
  a.state = c->state;
  a.seed = c->seed;
  a.terrain = 0;
  a.nx = c->nx;
  a.ny = c->ny;
  a.drops = drops;
  a.P = c->P;
  return a;
}

void drop_graph(shy_ctx* c);

size_t wave_steps(const shy_ctx* c, int drops) {
  return (size_t)drops * ((size_t)c->P.maxAge + 1);
}

// Size the ticket log and buckets for `drops` (never inside a capture).
int ensure_wave(shy_ctx* c, int drops) {
  const size_t need = wave_steps(c, drops);
  if (need > (size_t)INT32_MAX)
    return fail(c, SHY_ERR_ARG, "drops x (maxAge+1) exceeds 2^31 steps per wave");
  if (need <= c->wave_cap) return SHY_OK;
  drop_graph(c);  // a captured graph holds the old buffers
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->ticket) hipFree(c->ticket);
  if (c->rec) hipFree(c->rec);
  c->ticket = nullptr;
  c->rec = nullptr;
  c->wave_cap = 0;
  HIPCHK(c, hipMalloc(&c->ticket, need * sizeof(int32_t)));
  HIPCHK(c, hipMalloc(&c->rec, need * sizeof(int4)));
  c->wave_cap = need;
  return SHY_OK;
}

// ev, when given, receives events after the count walk, the scan and the
// commit walk.  The wave buffers must already hold wave_steps(c, drops).
int launch_cycle(shy_ctx* c, int drops, int p, hipEvent_t* ev = nullptr) {
  WalkArgs w = walk_args(c, drops);
  w.h = c->h[p];
  w.visit = c->visit[p];
  const int bs = c->walk_block;
  const int grid = (drops + bs - 1) / bs;
  const size_t n = c->n;
  const int nb = (int)((n + kScanBlock - 1) / kScanBlock);
  if (drops > 0) {
    walk_kernel<kCount><<<grid, bs, 0, c->stream>>>(w);
    HIPCHK(c, hipGetLastError());
  }
  if (ev) HIPCHK(c, hipEventRecord(ev[0], c->stream));
  if (drops > 0) {
    scan_sums_kernel<<<nb, kScanBlock, 0, c->stream>>>(c->count, n, c->bsum);
    scan_top_kernel<<<1, kScanBlock, 0, c->stream>>>(c->bsum, nb);
    scan_cells_kernel<<<nb, kScanBlock, 0, c->stream>>>(c->count, c->bsum, n, c->offset, c->scale,
                                                       c->fld, c->has_root ? c->root : nullptr,
                                                       c->P);
    HIPCHK(c, hipGetLastError());
  }
  if (ev) HIPCHK(c, hipEventRecord(ev[1], c->stream));
  if (drops > 0) {
    w.scale = c->scale;
    walk_kernel<kCommit><<<grid, bs, 0, c->stream>>>(w);
    HIPCHK(c, hipGetLastError());
  }
  if (ev) HIPCHK(c, hipEventRecord(ev[2], c->stream));
  reduce_kernel<<<(n + kReduceCells - 1) / kReduceCells, kReduceCells, 0, c->stream>>>(
      c->count, c->offset, c->rec, c->deposit, n, c->dh, c->trk, c->trk + n, c->trk + 2 * n);
  HIPCHK(c, hipGetLastError());
  CascArgs a;
  const int q = p ^ 1;
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
  a.nx = c->nx;
  a.ny = c->ny;
  a.P = c->P;
  dim3 cgrid((c->ny + kCbY - 1) / kCbY, (c->nx + kCbX - 1) / kCbX);
  cascade_blend_kernel<<<cgrid, kCbThreads, 0, c->stream>>>(a);
  HIPCHK(c, hipGetLastError());
  return SHY_OK;
}

void drop_graph(shy_ctx* c) {
  if (c->graph) hipGraphExecDestroy(c->graph);
  c->graph = nullptr;
  c->graph_drops = -1;
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
  void* ptrs[] = {c->h[0], c->h[1], c->fld, c->root, c->visit[0], c->visit[1], c->count,
                  c->offset, c->scale, c->dh, c->trk, c->deposit, c->bsum, c->ticket, c->rec,
                  c->state, c->erf_a, c->part, c->perms, c->trace_buf, c->trace_n,
                  c->drop_buf};
  for (void* q : ptrs)
    if (q) hipFree(q);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return SHY_OK;
}

int shy_create(const shy_params* p, int nx, int ny, uint64_t seed, shy_ctx** out) {
  if (!out) return fail(nullptr, SHY_ERR_ARG, "shy_create: out is NULL");
  *out = nullptr;
  if (nx < 1 || ny < 1 || (long long)nx * ny > (1LL << 31) - 1)
    return fail(nullptr, SHY_ERR_ARG, "shy_create: nx, ny must be >= 1 with nx*ny < 2^31");
  if (shy_device_count() < 1) return fail(nullptr, SHY_ERR_NODEV, "shy_create: no GPU visible");
  shy_ctx* c = new shy_ctx();
  if (p) {
    if (!params_ok(p)) {
      delete c;
      return fail(nullptr, SHY_ERR_ARG, "shy_create: parameter not finite, evapRate >= 1 or maxAge outside [0, 1e6]");
    }
    c->P = *p;
  } else {
    shy_default_params(&c->P);
  }
  c->nx = nx;
  c->ny = ny;
  c->n = (size_t)nx * ny;
  c->seed = seed;
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
  const size_t nb = (n + kScanBlock - 1) / kScanBlock;
  for (int k = 0; k < 2; ++k) {
    if ((e = hipMalloc(&c->h[k], n * sizeof(float))) != hipSuccess) return bail(e, "hipMalloc h");
    if ((e = hipMalloc(&c->visit[k], n)) != hipSuccess) return bail(e, "hipMalloc visit");
  }
  if ((e = hipMalloc(&c->fld, n * sizeof(float4))) != hipSuccess) return bail(e, "hipMalloc fld");
  if ((e = hipMalloc(&c->count, n * sizeof(int32_t))) != hipSuccess) return bail(e, "hipMalloc count");
  if ((e = hipMalloc(&c->offset, n * sizeof(int32_t))) != hipSuccess) return bail(e, "hipMalloc offset");
  if ((e = hipMalloc(&c->scale, n * sizeof(float))) != hipSuccess) return bail(e, "hipMalloc scale");
  if ((e = hipMalloc(&c->dh, n * sizeof(int32_t))) != hipSuccess) return bail(e, "hipMalloc dh");
  if ((e = hipMalloc(&c->trk, 3 * n * sizeof(long long))) != hipSuccess) return bail(e, "hipMalloc trk");
  if ((e = hipMalloc(&c->deposit, n * sizeof(int32_t))) != hipSuccess) return bail(e, "hipMalloc deposit");
  if ((e = hipMalloc(&c->bsum, nb * sizeof(int32_t))) != hipSuccess) return bail(e, "hipMalloc bsum");
  if ((e = hipMalloc(&c->root, n * sizeof(float))) != hipSuccess) return bail(e, "hipMalloc root");
  if ((e = hipMalloc(&c->state, sizeof(DeviceState))) != hipSuccess) return bail(e, "hipMalloc state");
  if ((e = hipMalloc(&c->erf_a, kErfTerms * sizeof(double))) != hipSuccess)
    return bail(e, "hipMalloc erf");
  if ((e = hipMalloc(&c->part, kPartBlocks * sizeof(float2))) != hipSuccess)
    return bail(e, "hipMalloc part");
  if ((e = hipMalloc(&c->perms, 16 * 256)) != hipSuccess) return bail(e, "hipMalloc perms");
  if ((e = hipMalloc(&c->trace_n, sizeof(int))) != hipSuccess) return bail(e, "hipMalloc trace_n");
  for (int k = 0; k < 2; ++k) {
    hipMemsetAsync(c->h[k], 0, n * sizeof(float), c->stream);
    hipMemsetAsync(c->visit[k], 0, n, c->stream);
  }
  hipMemsetAsync(c->fld, 0, n * sizeof(float4), c->stream);
  hipMemsetAsync(c->root, 0, n * sizeof(float), c->stream);
  hipMemsetAsync(c->count, 0, n * sizeof(int32_t), c->stream);
  hipMemsetAsync(c->offset, 0, n * sizeof(int32_t), c->stream);
  hipMemsetAsync(c->dh, 0, n * sizeof(int32_t), c->stream);
  hipMemsetAsync(c->trk, 0, 3 * n * sizeof(long long), c->stream);
  hipMemsetAsync(c->deposit, 0, n * sizeof(int32_t), c->stream);
  DeviceState st{};
  st.cycle = 0;
  double coef[kErfTerms];
  erf_coefficients(coef);
  hipMemcpyAsync(c->state, &st, sizeof(st), hipMemcpyHostToDevice, c->stream);
  hipMemcpyAsync(c->erf_a, coef, sizeof(coef), hipMemcpyHostToDevice, c->stream);
  if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return bail(e, "init");
  *out = c;
  return SHY_OK;
}

int shy_set_params(shy_ctx* c, const shy_params* p) {
  if (!c || !p) return fail(c, SHY_ERR_ARG, "shy_set_params: NULL argument");
  if (!params_ok(p)) return fail(c, SHY_ERR_ARG, "shy_set_params: parameter not finite, evapRate >= 1 or maxAge outside [0, 1e6]");
  c->P = *p;
  drop_graph(c);
  return SHY_OK;
}

int shy_get_params(const shy_ctx* c, shy_params* p) {
  if (!c || !p) return SHY_ERR_ARG;
  *p = c->P;
  return SHY_OK;
}

int shy_size(const shy_ctx* c, int* nx, int* ny) {
  if (!c || !nx || !ny) return SHY_ERR_ARG;
  *nx = c->nx;
  *ny = c->ny;
  return SHY_OK;
}

int shy_set_seed(shy_ctx* c, uint64_t seed, uint32_t cycle) {
  if (!c) return SHY_ERR_ARG;
  hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->seed = seed;
  // parity follows the cycle number; move the live buffers if it flips
  if ((cycle & 1u) != (c->cycle & 1u)) {
    // parity p: h[p] live, visit[p] the zero accumulator.  Move to q = p^1.
    const int p = parity(c), q = p ^ 1;
    const size_t n = c->n;
    HIPCHK(c, hipMemcpyAsync(c->h[q], c->h[p], n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->visit[q], 0, n, c->stream));
  }
  c->cycle = cycle;
  HIPCHK(c, hipMemcpyAsync(&c->state->cycle, &cycle, sizeof(cycle), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graph(c);
  return SHY_OK;
}

int shy_get_cycle(const shy_ctx* c, uint32_t* cycle) {
  if (!c || !cycle) return SHY_ERR_ARG;
  *cycle = c->cycle;
  return SHY_OK;
}

int shy_upload_field(shy_ctx* c, int field, const float* src) {
  if (!c || !src) return fail(c, SHY_ERR_ARG, "shy_upload_field: NULL argument");
  if (field < 0 || field > SHY_ROOTDENSITY)
    return fail(c, SHY_ERR_ARG, "shy_upload_field: field must be a quad::cell member (0..7)");
  if (!all_finite(src, c->n)) return fail(c, SHY_ERR_NONFINITE, "shy_upload_field: NaN or Inf in input");
  hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int p = parity(c);
  const size_t n = c->n;
  switch (field) {
    case SHY_HEIGHT:
      HIPCHK(c, hipMemcpy(c->h[p], src, n * sizeof(float), hipMemcpyHostToDevice));
      break;
    case SHY_DISCHARGE:
    case SHY_MOMENTUMX:
    case SHY_MOMENTUMY: {
      std::vector<float4> f(n);
      HIPCHK(c, hipMemcpy(f.data(), c->fld, n * sizeof(float4), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; ++i) {
        if (field == SHY_DISCHARGE) f[i].z = src[i];
        else if (field == SHY_MOMENTUMX) f[i].x = src[i];
        else f[i].y = src[i];
      }
      HIPCHK(c, hipMemcpy(c->fld, f.data(), n * sizeof(float4), hipMemcpyHostToDevice));
      if (field == SHY_DISCHARGE) {
        int r = refresh_erf(c);
        if (r) return r;
      }
      break;
    }
    case SHY_DISCHARGE_TRACK:
    case SHY_MOMENTUMX_TRACK:
    case SHY_MOMENTUMY_TRACK: {
      // tracks are per-cycle scratch (world.h:56-61): they stay visible until
      // the next cycle overwrites them, as the reference's do
      std::vector<long long> q(n);
      for (size_t i = 0; i < n; ++i) q[i] = (long long)std::llrint((double)src[i] * 16777216.0);
      HIPCHK(c, hipMemcpy(c->trk + (size_t)(field - SHY_DISCHARGE_TRACK) * n, q.data(),
                          n * sizeof(long long), hipMemcpyHostToDevice));
      break;
    }
    case SHY_ROOTDENSITY: {
      bool any = false;
      for (size_t i = 0; i < n && !any; ++i) any = src[i] != 0.0f;
      HIPCHK(c, hipMemcpy(c->root, src, n * sizeof(float), hipMemcpyHostToDevice));
      if (any != c->has_root) drop_graph(c);
      c->has_root = any;
      break;
    }
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SHY_OK;
}

int shy_download_field(shy_ctx* c, int field, float* dst) {
  if (!c || !dst) return fail(c, SHY_ERR_ARG, "shy_download_field: NULL argument");
  if (field < 0 || field > SHY_ERF_DISCHARGE)
    return fail(c, SHY_ERR_ARG, "shy_download_field: unknown field");
  hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int p = parity(c);
  const size_t n = c->n;
  switch (field) {
    case SHY_HEIGHT:
      HIPCHK(c, hipMemcpy(dst, c->h[p], n * sizeof(float), hipMemcpyDeviceToHost));
      break;
    case SHY_DISCHARGE:
    case SHY_MOMENTUMX:
    case SHY_MOMENTUMY:
    case SHY_ERF_DISCHARGE: {
      std::vector<float4> f(n);
      HIPCHK(c, hipMemcpy(f.data(), c->fld, n * sizeof(float4), hipMemcpyDeviceToHost));
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
      HIPCHK(c, hipMemcpy(q.data(), c->trk + (size_t)(field - SHY_DISCHARGE_TRACK) * n,
                          n * sizeof(long long), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; ++i) dst[i] = (float)((double)q[i] * kTQInv);
      break;
    }
    case SHY_ROOTDENSITY:
      HIPCHK(c, hipMemcpy(dst, c->root, n * sizeof(float), hipMemcpyDeviceToHost));
      break;
  }
  return SHY_OK;
}

int shy_download_normals(shy_ctx* c, float* xyz) {
  if (!c || !xyz) return fail(c, SHY_ERR_ARG, "shy_download_normals: NULL argument");
  hipSetDevice(c->device);
  float* d = nullptr;
  HIPCHK(c, hipMalloc((void**)&d, 3 * c->n * sizeof(float)));
  const size_t grid = (c->n + 255) / 256;
  normals_kernel<<<grid, 256, 0, c->stream>>>(c->h[parity(c)], c->nx, c->ny, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(xyz, d, 3 * c->n * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(d);
  if (e != hipSuccess) return fail(c, SHY_ERR_HIP, std::string("shy_download_normals: ") + hipGetErrorString(e));
  return SHY_OK;
}

int shy_upload_cells(shy_ctx* c, const float* cells8) {
  if (!c || !cells8) return fail(c, SHY_ERR_ARG, "shy_upload_cells: NULL argument");
  std::vector<float> f(c->n);
  for (int field = 0; field < 8; ++field) {
    for (size_t i = 0; i < c->n; ++i) f[i] = cells8[i * 8 + field];
    int r = shy_upload_field(c, field, f.data());
    if (r) return r;
  }
  return SHY_OK;
}

int shy_download_cells(shy_ctx* c, float* cells8) {
  if (!c || !cells8) return fail(c, SHY_ERR_ARG, "shy_download_cells: NULL argument");
  std::vector<float> f(c->n);
  for (int field = 0; field < 8; ++field) {
    int r = shy_download_field(c, field, f.data());
    if (r) return r;
    for (size_t i = 0; i < c->n; ++i) cells8[i * 8 + field] = f[i];
  }
  return SHY_OK;
}

int shy_generate_fbm(shy_ctx* c, uint64_t seed, int octaves, float base_freq, float lacunarity,
                     float gain) {
  if (!c) return SHY_ERR_ARG;
  if (octaves < 1 || octaves > 16) return fail(c, SHY_ERR_ARG, "shy_generate_fbm: octaves must be 1..16");
  if (!std::isfinite(base_freq) || !std::isfinite(lacunarity) || !std::isfinite(gain))
    return fail(c, SHY_ERR_ARG, "shy_generate_fbm: non-finite argument");
  if (base_freq <= 0.0f) base_freq = 1.0f / (float)(c->nx > c->ny ? c->nx : c->ny);
  hipSetDevice(c->device);
  std::vector<uint8_t> perms(16 * 256);
  for (int o = 0; o < octaves; ++o) fbm_perm(seed, o, perms.data() + o * 256);
  HIPCHK(c, hipMemcpyAsync(c->perms, perms.data(), octaves * 256, hipMemcpyHostToDevice, c->stream));
  float* out = c->h[parity(c)];
  const int bs = 256;
  const size_t grid = (c->n + bs - 1) / bs;
  fbm_kernel<<<grid, bs, 0, c->stream>>>(out, c->nx, c->ny, c->perms, octaves, base_freq,
                                         lacunarity, gain);
  HIPCHK(c, hipGetLastError());
  minmax_partial_kernel<<<kPartBlocks, 256, 0, c->stream>>>(out, c->n, c->part);
  HIPCHK(c, hipGetLastError());
  minmax_fold_kernel<<<1, 256, 0, c->stream>>>(c->part, kPartBlocks);
  HIPCHK(c, hipGetLastError());
  normalize_kernel<<<grid, bs, 0, c->stream>>>(out, c->n, c->part);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SHY_OK;
}

int shy_erode(shy_ctx* c, int drops) {
  if (!c) return SHY_ERR_ARG;
  if (drops < 0) return fail(c, SHY_ERR_ARG, "shy_erode: drops must be >= 0");
  hipSetDevice(c->device);
  int r = ensure_wave(c, drops);
  if (r) return r;
  r = launch_cycle(c, drops, parity(c));
  if (r) return r;
  c->cycle += 1;
  return SHY_OK;
}

int shy_erode_cycles(shy_ctx* c, int cycles, int drops, int use_graph) {
  if (!c) return SHY_ERR_ARG;
  if (cycles < 0 || drops < 0) return fail(c, SHY_ERR_ARG, "shy_erode_cycles: negative count");
  hipSetDevice(c->device);
  if (!use_graph) {
    for (int k = 0; k < cycles; ++k) {
      int r = shy_erode(c, drops);
      if (r) return r;
    }
    return SHY_OK;
  }
  if (cycles > 0 && parity(c) == 1) {
    int r = shy_erode(c, drops);
    if (r) return r;
    --cycles;
  }
  if (cycles >= 2) {
    int rw = ensure_wave(c, drops);
    if (rw) return rw;
    if (!c->graph || c->graph_drops != drops) {
      drop_graph(c);
      hipGraph_t g;
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int r0 = launch_cycle(c, drops, 0);
      int r1 = r0 ? r0 : launch_cycle(c, drops, 1);
      hipError_t e = hipStreamEndCapture(c->stream, &g);
      if (r1) return r1;
      if (e != hipSuccess) return fail(c, SHY_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      if (e != hipSuccess) {
        c->graph = nullptr;
        return fail(c, SHY_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      }
      c->graph_drops = drops;
    }
    for (int k = 0; k + 1 < cycles; k += 2) {
      HIPCHK(c, hipGraphLaunch(c->graph, c->stream));
      c->cycle += 2;
    }
  }
  if (cycles & 1) return shy_erode(c, drops);
  return SHY_OK;
}

int shy_trace_drop(shy_ctx* c, float x, float y, int max_steps, shy_drop_step* out, int* n_steps) {
  if (!c || !out || !n_steps) return fail(c, SHY_ERR_ARG, "shy_trace_drop: NULL argument");
  if (max_steps < 1) return fail(c, SHY_ERR_ARG, "shy_trace_drop: max_steps must be >= 1");
  if (!std::isfinite(x) || !std::isfinite(y) || std::fabs(x) > 1e9f || std::fabs(y) > 1e9f) {
    *n_steps = 0;
    return SHY_OK;
  }
  const int ix = (int)x, iy = (int)y;
  if (ix < 0 || iy < 0 || ix >= c->nx || iy >= c->ny) {
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
  WalkArgs a = walk_args(c, 1);
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
  if (!c) return SHY_ERR_ARG;
  const int ix = (int)x, iy = (int)y;
  if (!std::isfinite(x) || !std::isfinite(y) || ix < 0 || iy < 0 || ix >= c->nx || iy >= c->ny)
    return fail(c, SHY_ERR_ARG, "shy_cascade: position off the map");
  hipSetDevice(c->device);
  cascade_live_kernel<<<1, 64, 0, c->stream>>>(c->h[parity(c)], c->nx, c->ny, c->P, x, y);
  HIPCHK(c, hipGetLastError());
  return SHY_OK;
}

int shy_descend(shy_ctx* c, shy_drop* drop, int* alive) {
  if (!c || !drop || !alive) return fail(c, SHY_ERR_ARG, "shy_descend: NULL argument");
  if (!std::isfinite(drop->pos[0]) || !std::isfinite(drop->pos[1]) ||
      std::fabs(drop->pos[0]) > 1e9f || std::fabs(drop->pos[1]) > 1e9f) {
    *alive = 0;
    return SHY_OK;
  }
  hipSetDevice(c->device);
  if (!c->drop_buf) HIPCHK(c, hipMalloc(&c->drop_buf, sizeof(shy_drop) + sizeof(int)));
  LiveArgs a;
  const int p = parity(c);
  a.h = c->h[p];
  a.fld = c->fld;
  a.root = c->root;
  a.td = c->trk;
  a.tmx = c->trk + c->n;
  a.tmy = c->trk + 2 * c->n;
  a.erf_a = c->erf_a;
  a.nx = c->nx;
  a.ny = c->ny;
  a.P = c->P;
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

int shy_reset_stats(shy_ctx* c) {
  if (!c) return SHY_ERR_ARG;
  hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemsetAsync(c->state->stats, 0, sizeof(c->state->stats), c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SHY_OK;
}

int shy_synchronize(shy_ctx* c) {
  if (!c) return SHY_ERR_ARG;
  hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SHY_OK;
}

int shy_time_cycles(shy_ctx* c, int cycles, int drops, float* ms) {
  if (!c || !ms || cycles < 1 || drops < 0)
    return fail(c, SHY_ERR_ARG, "shy_time_cycles: bad argument");
  hipSetDevice(c->device);
  int rw = ensure_wave(c, drops);
  if (rw) return rw;
  // per cycle: start, after count walk, after scan, after commit walk; + end
  const size_t ne = 4 * (size_t)cycles + 1;
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
    int r = launch_cycle(c, drops, parity(c), &ev[4 * k + 1]);
    if (r) {
      cleanup();
      return r;
    }
    c->cycle += 1;
    e = hipEventRecord(ev[4 * k + 4], c->stream);
  }
  if (e == hipSuccess) e = hipEventSynchronize(ev.back());
  float part[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int k = 0; k < cycles && e == hipSuccess; ++k) {
    for (int s = 0; s < 4 && e == hipSuccess; ++s) {
      float t = 0.0f;
      e = hipEventElapsedTime(&t, ev[4 * k + s], ev[4 * k + s + 1]);
      part[s] += t;
    }
  }
  float tot = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&tot, ev[0], ev.back());
  cleanup();
  if (e != hipSuccess) return fail(c, SHY_ERR_HIP, std::string("shy_time_cycles: ") + hipGetErrorString(e));
  ms[0] = tot;
  for (int s = 0; s < 4; ++s) ms[1 + s] = part[s];
  return SHY_OK;
}

int shy_set_walk_block(shy_ctx* c, int threads) {
  if (!c) return SHY_ERR_ARG;
  if (threads == 0) threads = 256;
  if (threads < 64 || threads > 1024 || threads % 64)
    return fail(c, SHY_ERR_ARG, "shy_set_walk_block: threads must be a multiple of 64 in 64..1024");
  c->walk_block = threads;
  drop_graph(c);
  return SHY_OK;
}

}  // extern "C"
