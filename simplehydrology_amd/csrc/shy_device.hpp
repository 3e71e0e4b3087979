// shy_device.hpp -- device arithmetic of the erosion hot path (gfx950).
//
// Every float expression keeps the evaluation order of the reference source
// (source/water.h, source/world.h, source/cellpool.h) and of glm's generic
// vector code, so that each rounding happens where the reference's does.
// Build with -ffp-contract=off: a fused multiply-add would skip a rounding.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/shy.h"

namespace shy {

constexpr float kMapScale = 80.0f;  // quad::mapscale, cellpool.h:165
constexpr float kLod = 1.0f;        // quad::lodsize, cellpool.h:178
constexpr float kHQ = 1048576.0f;   // height commits: int32 in 2^-20 units
constexpr float kHQInv = 9.5367431640625e-07f;
constexpr float kTQ = 16777216.0f;  // track commits: int32 in 2^-24 units
constexpr double kTQInv = 5.9604644775390625e-08;
constexpr double kSQ40 = 1099511627776.0;  // ledger sums in 2^-40 units
constexpr int kErfTerms = 72;

// Record of cell (x, y) among a terrain's packed walk records: 2x2 blocks of
// cells, the blocks x-major over ceil(ny/2) block rows, so that one 128-byte
// line (four 32-byte records) holds a 2x2 neighbourhood: a drop's next cell
// lies in the line of its current one more often than with 4 cells along y
// (about 26% of steps against 10%, measured along drop paths at 1024^2).
// A terrain holds ceil(nx/2) * ceil(ny/2) * 4 records (pack_records).
__host__ __device__ inline int pack_rec(int x, int y, int ny) {
  return (((x >> 1) * ((ny + 1) >> 1) + (y >> 1)) << 2) + ((x & 1) << 1) + (y & 1);
}
__host__ __device__ inline long long pack_records(int nx, int ny) {
  return (long long)((nx + 1) >> 1) * ((ny + 1) >> 1) * 4;
}

// ---------------------------------------------------------------- hashing
// SplitMix64 output function (Steele, Lea, Flood, OOPSLA 2014).
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// Key of one cycle's spawn stream: (World::SEED, terrain, cycle).
__host__ __device__ inline uint64_t cycle_key(uint64_t seed, uint32_t terrain, uint32_t cycle) {
  uint64_t k = mix64(seed);
  k = mix64(k ^ (uint64_t)terrain);
  return mix64(k ^ ((uint64_t)cycle << 32));
}

// Spawn cell of drop `i`: uniform over [0, ncells) by multiply-shift of the
// hash's high word (Lemire 2019).  Replaces rand()%tileres (world.h:69).
__host__ __device__ inline uint32_t spawn_cell(uint64_t key, uint32_t i, uint32_t ncells) {
  uint64_t r = mix64(key ^ (uint64_t)i);
  return (uint32_t)(((r >> 32) * (uint64_t)ncells) >> 32);
}

// ---------------------------------------------------------- fixed point
__device__ inline int32_t quant_h(float v) {
  float s = v * kHQ;
  if (!(s == s)) s = 0.0f;
  if (s > 1073741824.0f) s = 1073741824.0f;
  if (s < -1073741824.0f) s = -1073741824.0f;
  return (int32_t)__builtin_rintf(s);
}
__device__ inline int32_t quant_t(float v) {
  float s = v * kTQ;
  if (!(s == s)) s = 0.0f;
  if (s > 2.0e9f) s = 2.0e9f;
  if (s < -2.0e9f) s = -2.0e9f;
  return (int32_t)__builtin_rintf(s);
}
__device__ inline long long quant_40(double v) {
  double s = v * kSQ40;
  if (!(s == s)) s = 0.0;
  if (s > 9.0e18) s = 9.0e18;
  if (s < -9.0e18) s = -9.0e18;
  return (long long)__builtin_rint(s);
}
__device__ inline float track_value(long long q) { return (float)((double)q * kTQInv); }

// ------------------------------------------------------------------- erf
// erf(x) = 2/sqrt(pi) sum_k (-1)^k x^(2k+1) / (k! (2k+1))  (Abramowitz &
// Stegun 7.1.5), Horner in x^2 in double, one rounding to float.  |x| >= 4
// returns +-1: erfc(4) = 1.5e-8 is below half an ulp of 1.0f.  The
// coefficient table is filled by the host (erf_coefficients in
// shy_engine.hip) before any kernel runs.
__device__ inline float erf_f32(const double* __restrict__ a, float x) {
  if (x == 0.0f) return x;
  double ax = __builtin_fabs((double)x);
  float r;
  if (ax >= 4.0) {
    r = 1.0f;
  } else {
    int n = ax < 0.5 ? 16 : (ax < 2.0 ? 36 : kErfTerms);
    double t = ax * ax;
    double p = a[n - 1];
    for (int k = n - 2; k >= 0; --k) p = p * t + a[k];
    r = (float)(ax * p);
  }
  return x < 0.0f ? -r : r;
}

// ------------------------------------------------------- cell neighbourhood
// What one walk step reads at its cell: the 5-point height stencil of
// quad::_normal (cellpool.h:181-204) and the frozen per-cell fields.
struct CellView {
  float h, xm, xp, ym, yp;  // h(x,y), h(x-1,y), h(x+1,y), h(x,y-1), h(x,y+1)
  float mx, my, d, erfd;    // momentum, discharge, erf(0.4 d)
  float root;               // rootdensity
  float n0, n2;             // from a packed record: the gravity terms lod*gravity*normal (x, z)
  float rf;                 // from a packed record: 1/|(mx, my)| when |(mx, my)| > 0, its sign
                            // set where a gravity term is outside markstein_num_ok
};

// quad::_normal (cellpool.h:181-204) from a loaded stencil.  Plane k is
// summed only when its diagonal corner is on the map.
__device__ inline void normal_from(const CellView& c, int x, int y, int nx, int ny, float& n0,
                                   float& n1, float& n2) {
  const float s0 = 1.0f, s1 = kMapScale, s2 = 1.0f;
  n0 = 0.0f; n1 = 0.0f; n2 = 0.0f;
  float ax, ay, az, bx, by, bz;
  const bool xp = x + 1 < nx, xm = x - 1 >= 0, yp = y + 1 < ny, ym = y - 1 >= 0;
  if (xp && yp) {
    ax = s0 * 0.0f; ay = s1 * (c.yp - c.h); az = s2 * 1.0f;
    bx = s0 * 1.0f; by = s1 * (c.xp - c.h); bz = s2 * 0.0f;
    n0 = n0 + (ay * bz - by * az); n1 = n1 + (az * bx - bz * ax); n2 = n2 + (ax * by - bx * ay);
  }
  if (xm && ym) {
    ax = s0 * 0.0f; ay = s1 * (c.ym - c.h); az = s2 * -1.0f;
    bx = s0 * -1.0f; by = s1 * (c.xm - c.h); bz = s2 * 0.0f;
    n0 = n0 + (ay * bz - by * az); n1 = n1 + (az * bx - bz * ax); n2 = n2 + (ax * by - bx * ay);
  }
  if (xp && ym) {
    ax = s0 * 1.0f; ay = s1 * (c.xp - c.h); az = s2 * 0.0f;
    bx = s0 * 0.0f; by = s1 * (c.ym - c.h); bz = s2 * -1.0f;
    n0 = n0 + (ay * bz - by * az); n1 = n1 + (az * bx - bz * ax); n2 = n2 + (ax * by - bx * ay);
  }
  if (xm && yp) {
    ax = s0 * -1.0f; ay = s1 * (c.xm - c.h); az = s2 * 0.0f;
    bx = s0 * 0.0f; by = s1 * (c.yp - c.h); bz = s2 * 1.0f;
    n0 = n0 + (ay * bz - by * az); n1 = n1 + (az * bx - bz * ax); n2 = n2 + (ax * by - bx * ay);
  }
  float dd = n0 * n0 + n1 * n1 + n2 * n2;
  if (__builtin_sqrtf(dd) > 0.0f) {
    float r = 1.0f / __builtin_sqrtf(dd);
    n0 = n0 * r; n1 = n1 * r; n2 = n2 * r;
  }
}

// 1/|f| of the momentum field, as the coupling below normalises it
__device__ inline float inv_len(float fx, float fy) {
  return 1.0f / __builtin_sqrtf(fx * fx + fy * fy);
}

// a / b, correctly rounded, given y = 1.0f / b (correctly rounded): q = a y is
// within an ulp of a/b, the remainder r = a - b q is exact in one FMA, and
// RN(q + r y) is the correctly rounded quotient (Markstein's theorem).  The
// theorem needs q, r and y in the normal range: with 2^-60 <= |a| <= 2^60 and
// 2^-60 <= b <= 2^60 they are (|q| <= 2^120, |r| >= 2^-84); markstein_ok
// tests that.  r = 0 means q is exact, which keeps the sign of a zero.
__host__ __device__ inline bool markstein_num_ok(float a) {
  const float aa = __builtin_fabsf(a);
  return (aa >= 0x1p-60f || aa == 0.0f) && aa <= 0x1p60f;
}
__host__ __device__ inline bool markstein_den_ok(float b) { return b >= 0x1p-60f && b <= 0x1p60f; }
__host__ __device__ inline float markstein_div(float a, float b, float y) {
  const float q = a * y;
  const float r = __builtin_fmaf(-q, b, a);
  return r == 0.0f ? q : __builtin_fmaf(r, y, q);
}
// Outside the theorem's range (a tiny slope, or a volume driven to the
// extremes by minVol <= 0 or a negative evapRate) the plain division runs.
__host__ __device__ inline float div_rcp(float a, float b, float y) {
  if (!(markstein_num_ok(a) && markstein_den_ok(b))) return a / b;
  return markstein_div(a, b, y);
}

// water.h:95-111 after the gravity term: momentum coupling, fixed step of
// length sqrt(2).  rf = inv_len(fx, fy), used only when |f| > 0.
__device__ inline void step_coupling(const shy_params& P, float vol, float d, float fx, float fy,
                                     float rf, float& sx, float& sy, float& px, float& py) {
  if (__builtin_sqrtf(fx * fx + fy * fy) > 0.0f && __builtin_sqrtf(sx * sx + sy * sy) > 0.0f) {
    float rs = 1.0f / __builtin_sqrtf(sx * sx + sy * sy);
    float dt = (fx * rf) * (sx * rs) + (fy * rf) * (sy * rs);
    float c = ((kLod * P.momentumTransfer) * dt) / (vol + d);
    sx = sx + c * fx;
    sy = sy + c * fy;
  }
  if (__builtin_sqrtf(sx * sx + sy * sy) > 0.0f) {
    const float k2 = kLod * __builtin_sqrtf(2.0f);  // lodsize*sqrt(2.0f)
    float r = 1.0f / __builtin_sqrtf(sx * sx + sy * sy);
    sx = k2 * (sx * r);
    sy = k2 * (sy * r);
  }
  px = px + sx;
  py = py + sy;
}

// water.h:95-111: gravity (k n) / vol, momentum coupling, fixed step.
__device__ inline void step_motion(const shy_params& P, float n0, float n2, float vol, float d,
                                   float fx, float fy, float& sx, float& sy, float& px,
                                   float& py) {
  const float k = kLod * P.gravity, yv = 1.0f / vol;
  sx = sx + div_rcp(k * n0, vol, yv);  // (k n0) / vol
  sy = sy + div_rcp(k * n2, vol, yv);
  step_coupling(P, vol, d, fx, fy, inv_len(fx, fy), sx, sy, px, py);
}

// water.h:135-136, evaluated in double as the reference's 1.0-evapRate is.
__host__ __device__ inline float evap_sed(float sed, float e) {
  return (float)((double)sed / (1.0 - (double)e));
}

// The same quotient with the divisor's reciprocal given: c = 1 - e and
// y = 1.0 / c (correctly rounded).  Markstein's theorem: with y within half
// an ulp of 1/c and q = RN(a y) within one ulp of a/c, the remainder
// r = a - c q is exact in one FMA and RN(q + r y) is the correctly rounded
// a/c (no overflow or underflow in double: a is a float and c = 1 - e with e
// a float below 1, so 2^-24 <= c <= 2^129).  Three dependent f64 operations
// instead of the division's sequence.
__host__ __device__ inline float evap_sed_rcp(float sed, double c, double y) {
  const double a = (double)sed;
  const double q = a * y;
  const double r = __builtin_fma(-q, c, a);
  return (float)__builtin_fma(r, y, q);
}
__host__ __device__ inline float evap_vol(float vol, float e) {
  return (float)((double)vol * (1.0 - (double)e));
}

// One coordinate of ivec2(pos): truncation toward zero, as glm's vec2 ->
// ivec2 conversion.  A position is finite and within maxAge * sqrt(2) of its
// spawn unless the speed became NaN (a volume that underflowed to 0 with
// minVol <= 0 makes the gravity term infinite); a NaN coordinate names no
// cell (-1, off the map), where the reference's x86 conversion (INT_MIN)
// sends it too.
__host__ __device__ inline int cell_coord(float p) { return p == p ? (int)p : -1; }

__device__ inline bool off_map(int x, int y, int nx, int ny) {
  return x < 0 || y < 0 || x >= nx || y >= ny;
}

}  // namespace shy
