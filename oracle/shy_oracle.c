/*
 * shy_oracle.c -- CPU oracle for the particle hydraulic-erosion hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): only tests/,
 * the smoke check and bench.py's CPU-baseline leg load it.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math), the
 * reference's -O2 (Makefile:3-4).  Float expressions are written in the
 * evaluation order of the reference's C++/glm source so that every rounding
 * happens where the reference's does; the glm forms restated here are
 * length(v) = sqrt(dot(v,v)), normalize(v) = v * (1/sqrt(dot(v,v))),
 * dot = left-to-right sum of products, cross = (a.y*b.z-b.y*a.z,
 * a.z*b.x-b.z*a.x, a.x*b.y-b.x*a.y), vec2->ivec2 = truncation toward zero.
 */
#include "shy_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAPSCALE 80.0f /* quad::mapscale, cellpool.h:165 */
#define LOD 1.0f       /* quad::lodsize, cellpool.h:178 (int 1 -> float) */
#define HQ 1048576.0f  /* height deltas: int32, 2^-20 units */
#define HQ_INV 9.5367431640625e-07f
#define TQ 16777216.0f /* track commits: int32 in 2^-24 units, summed in int64 */
#define TQ_INV 5.9604644775390625e-08
#define SQ40 1099511627776.0 /* ledger sums: 2^-40 units */

void orc_default_params(orc_params* p) {
  /* water.h:43-50 and world.h:42-44 */
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

/* ------------------------------------------------------------------ erf */
/* erf(x) = 2/sqrt(pi) * sum_k (-1)^k x^(2k+1) / (k! (2k+1))   (A&S 7.1.5),
   Horner in t = x^2, in double, rounded once to float.  |x| >= 4 gives +-1
   (erfc(4) = 1.5e-8 < 2^-25, so the float result is exactly 1). */
#define ERF_N 72
static double g_erf_a[ERF_N];
static int g_erf_init = 0;

static void erf_init(void) {
  const double two_over_sqrtpi = 1.1283791670955126;
  double f = 1.0;
  for (int k = 0; k < ERF_N; ++k) {
    if (k > 0) f = f * (double)k;
    double a = two_over_sqrtpi / (f * (double)(2 * k + 1));
    g_erf_a[k] = (k & 1) ? -a : a;
  }
  g_erf_init = 1;
}

float orc_erf(float x) {
  if (!g_erf_init) erf_init();
  if (x == 0.0f) return x;
  double ax = fabs((double)x);
  float r;
  if (ax >= 4.0) {
    r = 1.0f;
  } else {
    int n = ax < 0.5 ? 16 : (ax < 2.0 ? 36 : ERF_N);
    double t = ax * ax;
    double p = g_erf_a[n - 1];
    for (int k = n - 2; k >= 0; --k) p = p * t + g_erf_a[k];
    r = (float)(ax * p);
  }
  return x < 0.0f ? -r : r;
}

void orc_erf_many(const float* x, float* y, int64_t n) {
  for (int64_t i = 0; i < n; ++i) y[i] = orc_erf(x[i]);
}

/* ------------------------------------------------------------------ rng */
/* SplitMix64 output function (Steele, Lea, Flood 2014). */
uint64_t orc_mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

uint64_t orc_cycle_key(uint64_t seed, uint32_t terrain, uint32_t cycle) {
  uint64_t k = orc_mix64(seed);
  k = orc_mix64(k ^ (uint64_t)terrain);
  return orc_mix64(k ^ ((uint64_t)cycle << 32));
}

/* Uniform cell index in [0, ncells): Lemire multiply-shift on the high word. */
uint32_t orc_spawn_cell(uint64_t cycle_key, uint32_t drop, uint32_t ncells) {
  uint64_t r = orc_mix64(cycle_key ^ (uint64_t)drop);
  return (uint32_t)(((r >> 32) * (uint64_t)ncells) >> 32);
}

/* ------------------------------------------------------------------ fBm */
void orc_fbm_perm(uint64_t seed, int octave, uint8_t* perm) {
  for (int i = 0; i < 256; ++i) perm[i] = (uint8_t)i;
  uint64_t key = orc_mix64(orc_mix64(seed ^ 0x5045524c494e4f49ULL) + (uint64_t)octave);
  for (int i = 255; i > 0; --i) {
    uint64_t r = orc_mix64(key ^ (uint64_t)i);
    int j = (int)((r >> 32) % (uint64_t)(i + 1));
    uint8_t t = perm[i];
    perm[i] = perm[j];
    perm[j] = t;
  }
}

static float fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }

static float grad2(int h, float x, float y) {
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

static float perlin2(const uint8_t* P, float x, float y) {
  float xf = floorf(x), yf = floorf(y);
  int xi = ((int)xf) & 255, yi = ((int)yf) & 255;
  float fx = x - xf, fy = y - yf;
  float u = fade(fx), v = fade(fy);
  int xj = (xi + 1) & 255;
  int aa = P[(P[xi] + yi) & 255], ab = P[(P[xi] + yi + 1) & 255];
  int ba = P[(P[xj] + yi) & 255], bb = P[(P[xj] + yi + 1) & 255];
  float n00 = grad2(aa, fx, fy);
  float n10 = grad2(ba, fx - 1.0f, fy);
  float n01 = grad2(ab, fx, fy - 1.0f);
  float n11 = grad2(bb, fx - 1.0f, fy - 1.0f);
  float x1 = n00 + u * (n10 - n00);
  float x2 = n01 + u * (n11 - n01);
  return x1 + v * (x2 - x1);
}

void orc_fbm(uint64_t seed, int nx, int ny, int octaves, float base_freq,
             float lacunarity, float gain, float* out) {
  uint8_t perm[16][256];
  if (octaves > 16) octaves = 16;
  for (int o = 0; o < octaves; ++o) orc_fbm_perm(seed, o, perm[o]);
  float mn = INFINITY, mx = -INFINITY;
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y) {
      float s = 0.0f, amp = 1.0f, f = base_freq;
      for (int o = 0; o < octaves; ++o) {
        s = s + amp * perlin2(perm[o], (float)x * f, (float)y * f);
        f = f * lacunarity;
        amp = amp * gain;
      }
      out[(size_t)x * ny + y] = s;
      mn = s < mn ? s : mn;
      mx = s > mx ? s : mx;
    }
  float range = mx - mn;
  for (size_t i = 0; i < (size_t)nx * ny; ++i)
    out[i] = range > 0.0f ? (out[i] - mn) / range : 0.0f;
}

/* --------------------------------------------- the reference's terrain */
/* The OpenSimplex2 restatement below follows FastNoiseLite, which carries
   this notice:

     MIT License

     Copyright(c) 2020 Jordan Peck (jordan.me2@gmail.com)
     Copyright(c) 2020 Contributors

     Permission is hereby granted, free of charge, to any person obtaining a
     copy of this software and associated documentation files(the "Software"),
     to deal in the Software without restriction, including without limitation
     the rights to use, copy, modify, merge, publish, distribute, sublicense,
     and / or sell copies of the Software, and to permit persons to whom the
     Software is furnished to do so, subject to the following conditions :

     The above copyright notice and this permission notice shall be included in
     all copies or substantial portions of the Software.

     THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS OR
     IMPLIED, INCLUDING BUT NOT LIMITED TO THE WARRANTIES OF MERCHANTABILITY,
     FITNESS FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT.IN NO EVENT SHALL THE
     AUTHORS OR COPYRIGHT HOLDERS BE LIABLE FOR ANY CLAIM, DAMAGES OR OTHER
     LIABILITY, WHETHER IN AN ACTION OF CONTRACT, TORT OR OTHERWISE, ARISING
     FROM, OUT OF OR IN CONNECTION WITH THE SOFTWARE OR THE USE OR OTHER
     DEALINGS IN THE SOFTWARE. */
/* quad::map::init (cellpool.h:349-409) drives FastNoiseLite, a third-party
   header vendored in the reference (source/include/FastNoiseLite.h, MIT, no
   version number in the file).  Restated here: GetNoise(x, y, z)
   (FastNoiseLite.h:321-339) of a FastNoiseLite(1337) with NoiseType
   OpenSimplex2 and FractalType FBm at its defaults (3 octaves, lacunarity 2,
   gain 0.5, weighted strength 0, bounding 1/1.75; :114-129), which is the
   DefaultOpenSimplex2 rotation (:687-727), the FBm sum (:866-885) and
   SingleOpenSimplex2 (:1054-1153).  Signed overflow in the hash wraps mod
   2^32 (what the reference's compiled code does). */
#define OS_PX 501125321
#define OS_PY 1136930381
#define OS_PZ 1720413743

static int32_t os_mul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
static int32_t os_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }

/* FastNoiseLite.h:2529-2547: entries 0..59 run through the 12 cube edges
   five times; 60..63 repeat edges 8, 1, 9, 3 */
static const float OS_EDGE[12][3] = {
    {0, 1, 1}, {0, -1, 1}, {0, 1, -1}, {0, -1, -1}, {1, 0, 1}, {-1, 0, 1},
    {1, 0, -1}, {-1, 0, -1}, {1, 1, 0}, {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}};

int orc_os2_edge(int entry) {
  static const int tail[4] = {8, 1, 9, 3};
  return entry < 60 ? entry % 12 : tail[entry - 60];
}

static float os_grad(int32_t seed, int32_t i, int32_t j, int32_t k, float x, float y, float z) {
  int32_t h = os_mul(seed ^ i ^ j ^ k, 0x27d4eb2d); /* :500-506 */
  h ^= h >> 15;                                     /* :545 */
  const float* g = OS_EDGE[orc_os2_edge((h & 252) >> 2)];
  return x * g[0] + y * g[1] + z * g[2];
}

static int os_round(float f) { return f >= 0 ? (int)(f + 0.5f) : (int)(f - 0.5f); } /* :453 */

float orc_os2_single3(int32_t seed, float x, float y, float z) {
  int32_t i = os_round(x), j = os_round(y), k = os_round(z);
  float x0 = x - (float)i, y0 = y - (float)j, z0 = z - (float)k;
  int xs = (int)(-1.0f - x0) | 1, ys = (int)(-1.0f - y0) | 1, zs = (int)(-1.0f - z0) | 1;
  float ax0 = (float)xs * -x0, ay0 = (float)ys * -y0, az0 = (float)zs * -z0;
  i = os_mul(i, OS_PX);
  j = os_mul(j, OS_PY);
  k = os_mul(k, OS_PZ);
  float value = 0.0f;
  float a = (0.6f - x0 * x0) - (y0 * y0 + z0 * z0);
  for (int l = 0; l < 2; ++l) {
    if (a > 0) value += (a * a) * (a * a) * os_grad(seed, i, j, k, x0, y0, z0);
    float b = a + 1.0f;
    int32_t i1 = i, j1 = j, k1 = k;
    float x1 = x0, y1 = y0, z1 = z0;
    if (ax0 >= ay0 && ax0 >= az0) {
      x1 += (float)xs;
      b -= (float)(xs * 2) * x1;
      i1 = os_add(i1, -os_mul(xs, OS_PX));
    } else if (ay0 > ax0 && ay0 >= az0) {
      y1 += (float)ys;
      b -= (float)(ys * 2) * y1;
      j1 = os_add(j1, -os_mul(ys, OS_PY));
    } else {
      z1 += (float)zs;
      b -= (float)(zs * 2) * z1;
      k1 = os_add(k1, -os_mul(zs, OS_PZ));
    }
    if (b > 0) value += (b * b) * (b * b) * os_grad(seed, i1, j1, k1, x1, y1, z1);
    if (l == 1) break;
    ax0 = 0.5f - ax0;
    ay0 = 0.5f - ay0;
    az0 = 0.5f - az0;
    x0 = (float)xs * ax0;
    y0 = (float)ys * ay0;
    z0 = (float)zs * az0;
    a += (0.75f - ax0) - (ay0 + az0);
    i = os_add(i, (xs >> 1) & OS_PX);
    j = os_add(j, (ys >> 1) & OS_PY);
    k = os_add(k, (zs >> 1) & OS_PZ);
    xs = -xs;
    ys = -ys;
    zs = -zs;
    seed = ~seed;
  }
  return value * 32.69428253173828125f;
}

float orc_fnl_noise3(float freq, float x, float y, float z) {
  x *= freq;
  y *= freq;
  z *= freq;
  const float r3 = (float)(2.0 / 3.0);
  const float r = (x + y + z) * r3;
  x = r - x;
  y = r - y;
  z = r - z;
  int32_t seed = 1337;
  float sum = 0.0f, amp = 1 / 1.75f;
  for (int o = 0; o < 3; ++o) {
    const float noise = orc_os2_single3(seed++, x, y, z);
    sum += noise * amp;
    const float t = 0.0f, lo = 1.0f, hi = (noise + 1) * 0.5f; /* Lerp(1, hi, 0) */
    amp *= lo + t * (hi - lo);
    x *= 2.0f;
    y *= 2.0f;
    z *= 2.0f;
    amp *= 0.5f;
  }
  return sum;
}

void orc_fnl_world(int nx, int ny, int seed, int tileres, float* out) {
  /* cellpool.h:360-381: eight layers, frequency x2 and weight x0.6 (in
     double, stored to float) per layer */
  const float pz = (float)(seed % 10000);
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y) {
      const float px = (float)x / (float)tileres, py = (float)y / (float)tileres;
      float h = 0.0f, freq = 1.0f, scale = 0.6f;
      for (int o = 0; o < 8; ++o) {
        h += scale * orc_fnl_noise3(freq, px, py, pz);
        freq *= 2;
        scale = (float)(scale * 0.6);
      }
      out[(size_t)x * ny + y] = h;
    }
  /* cellpool.h:385-404: min and max start at 0 */
  float mn = 0.0f, mx = 0.0f;
  const size_t n = (size_t)nx * ny;
  for (size_t i = 0; i < n; ++i) {
    mn = mn < out[i] ? mn : out[i];
    mx = mx > out[i] ? mx : out[i];
  }
  for (size_t i = 0; i < n; ++i) out[i] = (out[i] - mn) / (mx - mn);
}

/* ------------------------------------------------------------ vegetation */
/* Vegetation::grow (vegetation.h:120-186) with the device's generation
   semantics (see include/shy.h): spawn attempt, then generations of plants
   that grow (:59-61), may die (:63-70) and may seed a child (:144-184), each
   generation deciding against the map and root density as they stood when it
   began; root changes in integer units of 0.2 (Plant::root's 1.0 / 0.6 / 0.4,
   :87-117) applied per generation; counter-hash draws. */
static uint32_t veg_draw(uint64_t key, uint32_t slot, int purpose, uint32_t m) {
  const uint64_t r = orc_mix64(key ^ (((uint64_t)slot << 8) | (uint64_t)purpose));
  return (uint32_t)(((r >> 32) * (uint64_t)m) >> 32);
}

static void veg_root(int32_t* units, int nx, int ny, int x, int y, int sign) {
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy) {
      const int u = (dx == 0 && dy == 0) ? 5 : (dx == 0 || dy == 0) ? 3 : 2;
      const int X = x + dx, Y = y + dy;
      if (X < 0 || Y < 0 || X >= nx || Y >= ny) continue;
      units[(size_t)X * ny + Y] += sign * u;
    }
}

static void veg_apply(float* root, int32_t* units, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (units[i]) {
      root[i] = root[i] + 0.2f * (float)units[i];
      units[i] = 0;
    }
}

int orc_veg_grow(const float* h, const float* d, float* root, int nx, int ny, int32_t* px,
                 int32_t* py, float* size, int count, int cap, const orc_plant_params* P,
                 uint64_t seed, uint32_t step, int64_t* dropped) {
  const size_t n = (size_t)nx * ny;
  int32_t* units = (int32_t*)calloc(n, sizeof(int32_t));
  int32_t* alive = (int32_t*)malloc((size_t)cap * sizeof(int32_t));
  int32_t* cflag = (int32_t*)malloc((size_t)cap * sizeof(int32_t));
  int32_t* cx = (int32_t*)malloc((size_t)cap * sizeof(int32_t));
  int32_t* cy = (int32_t*)malloc((size_t)cap * sizeof(int32_t));
  for (int k = 0; k < cap; ++k) alive[k] = 1;
  const uint64_t key = orc_cycle_key(seed, 0xffffffffu, step);
  float nrm[3];
  { /* spawn attempt (:124-136) */
    const int x = (int)veg_draw(key, 0xffffffffu, 0, (uint32_t)nx);
    const int y = (int)veg_draw(key, 0xffffffffu, 1, (uint32_t)ny);
    const size_t i = (size_t)x * ny + y;
    orc_normal(h, nx, ny, x, y, nrm);
    if (orc_erf(0.4f * d[i]) < P->maxDischarge && !(nrm[1] < P->maxSteep) &&
        h[i] < P->maxTreeHeight) {
      if (count < cap) {
        px[count] = x;
        py[count] = y;
        size[count] = 0.0f;
        alive[count] = 1;
        count++;
        veg_root(units, nx, ny, x, y, 1);
      } else {
        (*dropped)++;
      }
    }
  }
  veg_apply(root, units, n);
  int g0 = 0, g1 = count;
  for (int gen = 0; gen < 1024 && g1 > g0; ++gen) {
    for (int k = g0; k < g1; ++k) {
      cflag[k] = 0;
      size[k] += P->growRate * (P->maxSize - size[k]);
      const int x = px[k], y = py[k];
      const size_t i = (size_t)x * ny + y;
      if (orc_erf(0.4f * d[i]) >= P->maxDischarge || h[i] >= P->maxTreeHeight ||
          veg_draw(key, (uint32_t)k, 2, 1000) == 0) {
        alive[k] = 0;
        veg_root(units, nx, ny, x, y, -1);
        continue;
      }
      if (veg_draw(key, (uint32_t)k, 3, 20) != 0) continue;
      const int X = x + (int)veg_draw(key, (uint32_t)k, 4, 9) - 4;
      const int Y = y + (int)veg_draw(key, (uint32_t)k, 5, 9) - 4;
      if (X < 0 || Y < 0 || X >= nx || Y >= ny) continue;
      const size_t j = (size_t)X * ny + Y;
      if (orc_erf(0.4f * d[j]) >= P->maxDischarge) continue;
      const float r = (float)veg_draw(key, (uint32_t)k, 6, 1000);
      if ((double)r / 1000.0 <= (double)root[j]) continue;
      orc_normal(h, nx, ny, X, Y, nrm);
      if (nrm[1] <= P->maxSteep) continue;
      cflag[k] = 1;
      cx[k] = X;
      cy[k] = Y;
    }
    for (int k = g0; k < g1; ++k) {
      if (!cflag[k]) continue;
      if (count < cap) {
        px[count] = cx[k];
        py[count] = cy[k];
        size[count] = 0.0f;
        alive[count] = 1;
        count++;
        veg_root(units, nx, ny, cx[k], cy[k], 1);
      } else {
        (*dropped)++;
      }
    }
    veg_apply(root, units, n);
    g0 = g1;
    g1 = count;
  }
  int m = 0;
  for (int k = 0; k < count; ++k)
    if (alive[k]) {
      px[m] = px[k];
      py[m] = py[k];
      size[m] = size[k];
      m++;
    }
  free(units);
  free(alive);
  free(cflag);
  free(cx);
  free(cy);
  return m;
}

/* ------------------------------------------------------------ accessors */
/* ivec2(pos) of one coordinate: truncation toward zero (glm's conversion).
   A NaN coordinate (a speed made NaN by a volume that underflowed to 0 with
   minVol <= 0) is -1, off the map: (int)NaN is undefined in C, and the
   reference's x86 build gets INT_MIN, also off the map. */
static inline int cell_coord(float p) { return p == p ? (int)p : -1; }

static inline int oob(int nx, int ny, int x, int y) {
  /* quad::map::oob, cellpool.h:413-419 */
  return x < 0 || y < 0 || x >= nx || y >= ny;
}

/* quad::_normal, cellpool.h:181-204.  st = float stride between cells. */
static void normal_st(const float* h, int st, int nx, int ny, int x, int y, float* n) {
#define HT(X, Y) (oob(nx, ny, (X), (Y)) ? 0.0f : h[((size_t)(X) * ny + (Y)) * st])
  const float s0 = 1.0f, s1 = MAPSCALE, s2 = 1.0f; /* vec3(1.0, mapscale, 1.0) */
  float n0 = 0.0f, n1 = 0.0f, n2 = 0.0f;
  float ax, ay, az, bx, by, bz;
  const float hp = HT(x, y);
  if (!oob(nx, ny, x + 1, y + 1)) { /* :187-188 */
    ax = s0 * 0.0f; ay = s1 * (HT(x, y + 1) - hp); az = s2 * 1.0f;
    bx = s0 * 1.0f; by = s1 * (HT(x + 1, y) - hp); bz = s2 * 0.0f;
    n0 = n0 + (ay * bz - by * az); n1 = n1 + (az * bx - bz * ax); n2 = n2 + (ax * by - bx * ay);
  }
  if (!oob(nx, ny, x - 1, y - 1)) { /* :190-191 */
    ax = s0 * 0.0f; ay = s1 * (HT(x, y - 1) - hp); az = s2 * -1.0f;
    bx = s0 * -1.0f; by = s1 * (HT(x - 1, y) - hp); bz = s2 * 0.0f;
    n0 = n0 + (ay * bz - by * az); n1 = n1 + (az * bx - bz * ax); n2 = n2 + (ax * by - bx * ay);
  }
  if (!oob(nx, ny, x + 1, y - 1)) { /* :194-195 */
    ax = s0 * 1.0f; ay = s1 * (HT(x + 1, y) - hp); az = s2 * 0.0f;
    bx = s0 * 0.0f; by = s1 * (HT(x, y - 1) - hp); bz = s2 * -1.0f;
    n0 = n0 + (ay * bz - by * az); n1 = n1 + (az * bx - bz * ax); n2 = n2 + (ax * by - bx * ay);
  }
  if (!oob(nx, ny, x - 1, y + 1)) { /* :197-198 */
    ax = s0 * -1.0f; ay = s1 * (HT(x - 1, y) - hp); az = s2 * 0.0f;
    bx = s0 * 0.0f; by = s1 * (HT(x, y + 1) - hp); bz = s2 * 1.0f;
    n0 = n0 + (ay * bz - by * az); n1 = n1 + (az * bx - bz * ax); n2 = n2 + (ax * by - bx * ay);
  }
#undef HT
  float dd = n0 * n0 + n1 * n1 + n2 * n2; /* dot, left to right */
  if (sqrtf(dd) > 0.0f) {                 /* length(n) > 0 */
    float r = 1.0f / sqrtf(dd);           /* glm inversesqrt */
    n0 = n0 * r; n1 = n1 * r; n2 = n2 * r;
  }
  n[0] = n0; n[1] = n1; n[2] = n2;
}

void orc_normal(const float* h, int nx, int ny, int x, int y, float* n3) {
  normal_st(h, 1, nx, ny, x, y, n3);
}

/* ---------------------------------------------------- shared step physics */
/* water.h:86-111: forces, momentum coupling, fixed step of length sqrt(2).
   Returns nothing; updates speed and pos. */
static inline void step_motion(const orc_params* P, const float* n, float vol, float d,
                               float fx, float fy, float* sx, float* sy, float* px,
                               float* py) {
  float k = LOD * P->gravity;
  *sx = *sx + (k * n[0]) / vol; /* :95 speed += lodsize*gravity*vec2(n.x,n.z)/volume */
  *sy = *sy + (k * n[2]) / vol;
  /* :97-99 */
  if (sqrtf(fx * fx + fy * fy) > 0.0f && sqrtf(*sx * *sx + *sy * *sy) > 0.0f) {
    float rf = 1.0f / sqrtf(fx * fx + fy * fy);
    float rs = 1.0f / sqrtf(*sx * *sx + *sy * *sy);
    float dt = (fx * rf) * (*sx * rs) + (fy * rf) * (*sy * rs);
    float c = ((LOD * P->momentumTransfer) * dt) / (vol + d);
    *sx = *sx + c * fx;
    *sy = *sy + c * fy;
  }
  /* :108-109 */
  if (sqrtf(*sx * *sx + *sy * *sy) > 0.0f) {
    const float k2 = LOD * sqrtf(2.0f);
    float r = 1.0f / sqrtf(*sx * *sx + *sy * *sy);
    *sx = k2 * (*sx * r);
    *sy = k2 * (*sy * r);
  }
  *px = *px + *sx; /* :111 */
  *py = *py + *sy;
}

static inline float evap_sed(float sed, float e) { return (float)((double)sed / (1.0 - (double)e)); }
static inline float evap_vol(float vol, float e) { return (float)((double)vol * (1.0 - (double)e)); }

/* ------------------------------------------------ (a) sequential reference */
/* 0: glibc erff as the reference; 1: the project erf (to compare bit for bit
   with a device restatement of the same live semantics). */
static int g_seq_project_erf = 0;
void orc_set_seq_erf(int use_project) { g_seq_project_erf = use_project; }

enum { C_H = 0, C_D, C_MX, C_MY, C_DT, C_MXT, C_MYT, C_ROOT, C_N };

void orc_seq_cascade(float* cells, int nx, int ny, const orc_params* P, float fx, float fy) {
  /* World::cascade, world.h:90-168 */
  static const int off[8][2] = {{-1, -1}, {-1, 0}, {-1, 1}, {0, -1},
                                {0, 1},   {1, -1}, {1, 0},  {1, 1}};
  int px[8], py[8];
  float sh[8], sd[8];
  int num = 0;
  int cx = cell_coord(fx), cy = cell_coord(fy);
  for (int k = 0; k < 8; ++k) {
    int x = cx + 1 * off[k][0], y = cy + 1 * off[k][1];
    if (oob(nx, ny, x, y)) continue;
    float vx = (float)off[k][0], vy = (float)off[k][1];
    px[num] = x; py[num] = y;
    sh[num] = cells[((size_t)x * ny + y) * C_N + C_H];
    sd[num] = sqrtf(vx * vx + vy * vy); /* length(vec2(nn)) */
    ++num;
  }
  /* std::sort on <= 16 elements: libstdc++ insertion sort (stable for '<') */
  for (int i = 1; i < num; ++i) {
    int tx = px[i], ty = py[i];
    float th = sh[i], td = sd[i];
    int j = i;
    while (j > 0 && th < sh[j - 1]) {
      px[j] = px[j - 1]; py[j] = py[j - 1]; sh[j] = sh[j - 1]; sd[j] = sd[j - 1];
      --j;
    }
    px[j] = tx; py[j] = ty; sh[j] = th; sd[j] = td;
  }
  float* hc = &cells[((size_t)cx * ny + cy) * C_N + C_H];
  for (int i = 0; i < num; ++i) {
    float diff = *hc - sh[i];
    if (diff == 0) continue;
    float excess;
    if ((double)sh[i] > 0.1)
      excess = fabsf(diff) - sd[i] * P->maxdiff * LOD;
    else
      excess = fabsf(diff);
    if (excess <= 0) continue;
    float transfer = P->settling * excess / 2.0f;
    float* hn = &cells[((size_t)px[i] * ny + py[i]) * C_N + C_H];
    if (diff > 0) {
      *hc -= transfer;
      *hn += transfer;
    } else {
      *hc += transfer;
      *hn -= transfer;
    }
  }
}

/* Drop::descend, water.h:58-156, on the live map.  Returns 0 when the call
   ends before the physics update (NULL cell, age, volume), 1 for a step that
   continues, 2 for the out-of-bounds exit step.  *gain receives evaporation's
   sediment increase of the step (ledger only). */
static int seq_descend_impl(float* cells, int nx, int ny, const orc_params* P, int32_t* age,
                            float* pos, float* speed, float* volume, float* sediment,
                            double* gain) {
  int ix = cell_coord(pos[0]), iy = cell_coord(pos[1]);
  if (oob(nx, ny, ix, iy)) return 0; /* :62-68 */
  float* c = &cells[((size_t)ix * ny + iy) * C_N];
  float n[3];
  normal_st(cells, C_N, nx, ny, ix, iy, n); /* :70 */
  if ((float)*age > P->maxAge) {            /* :74-77 */
    c[C_H] += *sediment;
    return 0;
  }
  if (*volume < P->minVol) { /* :79-82 */
    c[C_H] += *sediment;
    return 0;
  }
  float effD = P->depositionRate * (1.0f - c[C_ROOT]); /* :86-87 */
  if (effD < 0) effD = 0;
  float sx = speed[0], sy = speed[1], px = pos[0], py = pos[1];
  step_motion(P, n, *volume, c[C_D], c[C_MX], c[C_MY], &sx, &sy, &px, &py);
  c[C_DT] += *volume; /* :115-117 */
  c[C_MXT] += *volume * sx;
  c[C_MYT] += *volume * sy;
  int jx = cell_coord(px), jy = cell_coord(py);
  float h2;
  if (oob(nx, ny, jx, jy)) /* :120-124 */
    h2 = (float)((double)c[C_H] - 0.002);
  else
    h2 = cells[((size_t)jx * ny + jy) * C_N + C_H];
  /* node::discharge, cellpool.h:242-244 */
  float E = g_seq_project_erf ? orc_erf(0.4f * c[C_D]) : erff(0.4f * c[C_D]);
  float c_eq = (1.0f + P->entrainment * E) * (c[C_H] - h2); /* :127 */
  if (c_eq < 0) c_eq = 0;
  float cdiff = (c_eq - *sediment);
  *sediment += effD * cdiff; /* :131 */
  c[C_H] -= effD * cdiff;    /* :132 */
  float s1 = evap_sed(*sediment, P->evapRate); /* :135 */
  if (gain) *gain += (double)s1 - (double)*sediment;
  *sediment = s1;
  *volume = evap_vol(*volume, P->evapRate); /* :136 */
  speed[0] = sx; speed[1] = sy; pos[0] = px; pos[1] = py;
  if (oob(nx, ny, jx, jy)) { /* :139-142 */
    *volume = 0.0f;
    return 2;
  }
  orc_seq_cascade(cells, nx, ny, P, px, py); /* :151 */
  *age += 1;                                 /* :153 */
  return 1;
}

int orc_seq_descend(float* cells, int nx, int ny, const orc_params* P, int32_t* age,
                    float* pos, float* speed, float* volume, float* sediment) {
  return seq_descend_impl(cells, nx, ny, P, age, pos, speed, volume, sediment, NULL) == 1;
}

void orc_seq_erode(float* cells, int nx, int ny, const orc_params* P, uint64_t seed,
                   uint32_t terrain, uint32_t cycle, int drops, orc_seq_ledger* led) {
  /* World::erode, world.h:54-88 (spawn stream: the counter hash, not rand()) */
  size_t ncell = (size_t)nx * ny;
  for (size_t i = 0; i < ncell; ++i) { /* :56-61 */
    cells[i * C_N + C_DT] = 0;
    cells[i * C_N + C_MXT] = 0;
    cells[i * C_N + C_MYT] = 0;
  }
  uint64_t key = orc_cycle_key(seed, terrain, cycle);
  for (int i = 0; i < drops; ++i) { /* :64-78 */
    uint32_t cid = orc_spawn_cell(key, (uint32_t)i, (uint32_t)ncell);
    int x = (int)(cid / (uint32_t)ny), y = (int)(cid % (uint32_t)ny);
    if ((double)cells[(size_t)cid * C_N + C_H] < 0.1) { /* :71-72 */
      if (led) led->skipped++;
      continue;
    }
    if (led) led->walked++;
    int32_t age = 0;
    float pos[2] = {(float)x, (float)y}, speed[2] = {0.0f, 0.0f};
    float vol = 1.0f, sed = 0.0f;
    double gain = 0.0;
    int64_t steps = 0;
    for (;;) { /* :76 while(drop.descend()); */
      int r = seq_descend_impl(cells, nx, ny, P, &age, pos, speed, &vol, &sed, &gain);
      if (r) ++steps;
      if (r == 2 && led) {
        led->exits++;
        led->export_ += (double)sed;
      }
      if (r != 1) break;
    }
    if (led) {
      led->steps += steps;
      led->gain += gain;
    }
  }
  for (size_t i = 0; i < ncell; ++i) { /* :81-86 */
    float* c = &cells[i * C_N];
    c[C_D] = (1.0f - P->lrate) * c[C_D] + P->lrate * c[C_DT];
    c[C_MX] = (1.0f - P->lrate) * c[C_MX] + P->lrate * c[C_MXT];
    c[C_MY] = (1.0f - P->lrate) * c[C_MY] + P->lrate * c[C_MYT];
  }
}

/* ---------------------------------------------------- fixed-point commits */
static inline int32_t qh(float v) {
  float s = v * HQ;
  if (!(s == s)) s = 0.0f;
  if (s > 1073741824.0f) s = 1073741824.0f;
  if (s < -1073741824.0f) s = -1073741824.0f;
  return (int32_t)rintf(s);
}
static inline int32_t qt(float v) {
  float s = v * TQ;
  if (!(s == s)) s = 0.0f;
  if (s > 2.0e9f) s = 2.0e9f;
  if (s < -2.0e9f) s = -2.0e9f;
  return (int32_t)rintf(s);
}
static inline int64_t q40(double v) {
  double s = v * SQ40;
  if (!(s == s)) s = 0.0;
  if (s > 9.0e18) s = 9.0e18;
  if (s < -9.0e18) s = -9.0e18;
  return (int64_t)rint(s);
}

/* ------------------------------------------------- (b) frozen single drop */
int orc_trace_frozen(const float* h, const float* d, const float* mx, const float* my,
                     const float* root, const float* erfd, int nx, int ny,
                     const orc_params* P, float x0, float y0, int max_steps,
                     int use_ref_erf, orc_step* out, int* end_reason) {
  int32_t age = 0;
  float px = x0, py = y0, sx = 0.0f, sy = 0.0f, vol = 1.0f, sed = 0.0f;
  int nst = 0;
  *end_reason = 0;
  for (;;) {
    int ix = cell_coord(px), iy = cell_coord(py);
    if (oob(nx, ny, ix, iy)) { *end_reason = 1; break; }
    size_t i0 = (size_t)ix * ny + iy;
    float n[3];
    normal_st(h, 1, nx, ny, ix, iy, n);
    if ((float)age > P->maxAge) { *end_reason = 2; break; }
    if (vol < P->minVol) { *end_reason = 3; break; }
    float effD = P->depositionRate * (1.0f - root[i0]);
    if (effD < 0) effD = 0;
    step_motion(P, n, vol, d[i0], mx[i0], my[i0], &sx, &sy, &px, &py);
    int jx = cell_coord(px), jy = cell_coord(py);
    float hc = h[i0];
    float h2 = oob(nx, ny, jx, jy) ? (float)((double)hc - 0.002) : h[(size_t)jx * ny + jy];
    float E = use_ref_erf ? erff(0.4f * d[i0]) : erfd[i0];
    float c_eq = (1.0f + P->entrainment * E) * (hc - h2);
    if (c_eq < 0) c_eq = 0;
    float cdiff = (c_eq - sed);
    float ero = effD * cdiff;
    sed += ero;
    sed = evap_sed(sed, P->evapRate);
    vol = evap_vol(vol, P->evapRate);
    int ex = oob(nx, ny, jx, jy);
    if (nst < max_steps) {
      orc_step* o = &out[nst];
      o->age = age; o->status = ex;
      o->pos[0] = px; o->pos[1] = py; o->speed[0] = sx; o->speed[1] = sy;
      o->volume = vol; o->sediment = sed; o->erosion = ero; o->h2 = h2;
    }
    ++nst;
    if (ex) { *end_reason = 4; break; }
    ++age;
    if (nst >= max_steps) { *end_reason = 5; break; }
  }
  return nst;
}

/* --------------------------------------------------- (c) wave-synchronous */
/*
 * One cycle = one wave.  Every drop reads the map as it stood when the cycle
 * began, so a drop cannot see the edits of any drop of the same cycle, its own
 * included.  Committing each drop's full erosion against that stale state
 * overshoots without bound: one drop moves a cell a fraction
 * g = effD*(1 + entrainment*erf(0.4 d)) (about 1.1 on rivers) of the way to
 * its equilibrium, so k >= 2 drops through a cell on the same snapshot
 * overcorrect, and drops pinned in pits re-erode the same cells up to maxAge
 * times.  The wave therefore runs in two passes over identical trajectories
 * (trajectories depend only on the snapshot):
 *
 *  pass A  counts k(c), the walk steps that start at cell c;
 *  scale   s(c) = (1 - (1-gamma)^k) / (gamma k), gamma = min(1, g(c)):
 *          k drops of identical inputs applied one after another under the
 *          linear feedback of water.h:127-132 take (1-(1-gamma)^k)/gamma times
 *          one drop's erosion, so the k wave commits sum to that burst total;
 *  pass B  commits erosion/deposition q = rint(2^20 effD*cdiff*s(c)) and moves
 *          the drop's sediment by the same q 2^-20, the full terminal deposit, the
 *          tracks, and marks arrival cells as cascade sites;
 *  cascade the reference window sequence (world.h:116-166) at every site,
 *          on the heights after pass B, each window's transfers divided by
 *          m(c) = max over its cells of the number of site windows covering
 *          that cell (a Jacobi sweep that cannot overshoot where windows
 *          overlap);
 *  blend   world.h:81-86.
 */
typedef struct wacc {
  int32_t* dh;
  int64_t *td, *tmx, *tmy;
  uint8_t* visit;
  int32_t* count;
  const float* scale; /* NULL in pass A */
} wacc;

static void wave_walk(const float* h, const float* d, const float* mx, const float* my,
                      const float* root, const float* erfd, int nx, int ny,
                      const orc_params* P, float x0, float y0, wacc* A, orc_stats* st) {
  const int commit = A->scale != NULL;
  int32_t age = 0;
  float px = x0, py = y0, sx = 0.0f, sy = 0.0f, vol = 1.0f, sed = 0.0f;
  double gain = 0.0;
  int64_t steps = 0;
  for (;;) {
    int ix = cell_coord(px), iy = cell_coord(py);
    size_t i0 = (size_t)ix * ny + iy; /* in bounds: spawn or checked below */
    float n[3];
    normal_st(h, 1, nx, ny, ix, iy, n);
    if ((float)age > P->maxAge || vol < P->minVol) {
      if (commit) {
        int32_t q = qh(sed);
        A->dh[i0] += q;
        st->walk_dh += q;
        if ((float)age > P->maxAge) st->ends_age++; else st->ends_vol++;
      }
      break;
    }
    ++steps;
    if (!commit) A->count[i0] += 1;
    float effD = P->depositionRate * (1.0f - root[i0]);
    if (effD < 0) effD = 0;
    step_motion(P, n, vol, d[i0], mx[i0], my[i0], &sx, &sy, &px, &py);
    if (commit) {
      A->td[i0] += qt(vol);
      A->tmx[i0] += qt(vol * sx);
      A->tmy[i0] += qt(vol * sy);
    }
    int jx = cell_coord(px), jy = cell_coord(py);
    int ex = oob(nx, ny, jx, jy);
    float hc = h[i0];
    float h2 = ex ? (float)((double)hc - 0.002) : h[(size_t)jx * ny + jy];
    float c_eq = (1.0f + P->entrainment * erfd[i0]) * (hc - h2);
    if (c_eq < 0) c_eq = 0;
    float cdiff = (c_eq - sed);
    if (commit) {
      /* one quantised amount moves between map and drop, so the ledger
         closes in fixed point */
      int32_t q = qh((effD * cdiff) * A->scale[i0]);
      sed += (float)q * HQ_INV;
      A->dh[i0] -= q;
      st->walk_dh -= q;
    } else {
      sed += effD * cdiff;
    }
    float s1 = evap_sed(sed, P->evapRate);
    gain += (double)s1 - (double)sed;
    sed = s1;
    vol = evap_vol(vol, P->evapRate);
    if (ex) {
      if (commit) {
        st->exits++;
        st->export40 += q40((double)sed);
      }
      break;
    }
    if (commit) A->visit[(size_t)jx * ny + jy] = 1;
    ++age;
  }
  if (commit) {
    st->steps += steps;
    st->gain40 += q40(gain);
  }
}

/* r^k by binary exponentiation, least significant bit first */
static float powi_f32(float r, int32_t k) {
  float res = 1.0f, b = r;
  while (k > 0) {
    if (k & 1) res = res * b;
    b = b * b;
    k >>= 1;
  }
  return res;
}

/* Burst scale of a cell that k count-pass steps start at: with g = effD (1 +
   entrainment erf(0.4 d)) the fraction of its distance to equilibrium one
   step moves it (water.h:127-132), k steps in sequence move it
   R = 1 - (1-g)^k; the k commits are scaled by R / (k g).  For g > 1, R is
   capped at 1: a wave updates a cell and its downstream neighbour from one
   snapshot, a Jacobi step that is stable only for R <= 1. */
float orc_burst_scale(int32_t k, float effD, float entrainment, float erfd) {
  float g = effD * (1.0f + entrainment * erfd);
  if (!(g > 0.0f)) return 1.0f;
  if (g > 1.0f) return 1.0f / ((float)(k < 1 ? 1 : k) * g);
  if (k <= 1) return 1.0f;
  return (1.0f - powi_f32(1.0f - g, k)) / (g * (float)k);
}

/* Cascade of one site on the post-walk heights H: the reference's sorted,
   centre-re-reading window (world.h:116-166); each transfer divided by the
   window weight m and quantised once, committed as +-q. */
static int64_t wave_cascade_site(const float* H, int nx, int ny, const orc_params* P,
                                 int cx, int cy, int m, int32_t* C) {
  static const int off[8][2] = {{-1, -1}, {-1, 0}, {-1, 1}, {0, -1},
                                {0, 1},   {1, -1}, {1, 0},  {1, 1}};
  int pidx[8];
  float sh[8], sd[8];
  int num = 0;
  for (int k = 0; k < 8; ++k) {
    int x = cx + off[k][0], y = cy + off[k][1];
    if (oob(nx, ny, x, y)) continue;
    float vx = (float)off[k][0], vy = (float)off[k][1];
    pidx[num] = x * ny + y;
    sh[num] = H[(size_t)x * ny + y];
    sd[num] = sqrtf(vx * vx + vy * vy);
    ++num;
  }
  for (int i = 1; i < num; ++i) {
    int ti = pidx[i];
    float th = sh[i], td = sd[i];
    int j = i;
    while (j > 0 && th < sh[j - 1]) {
      pidx[j] = pidx[j - 1]; sh[j] = sh[j - 1]; sd[j] = sd[j - 1];
      --j;
    }
    pidx[j] = ti; sh[j] = th; sd[j] = td;
  }
  size_t ci = (size_t)cx * ny + cy;
  float hc = H[ci];
  int64_t moves = 0;
  for (int i = 0; i < num; ++i) {
    float diff = hc - sh[i];
    if (diff == 0) continue;
    float excess;
    if ((double)sh[i] > 0.1)
      excess = fabsf(diff) - sd[i] * P->maxdiff * LOD;
    else
      excess = fabsf(diff);
    if (excess <= 0) continue;
    float transfer = P->settling * excess / 2.0f;
    int32_t q = qh(transfer / (float)m);
    ++moves;
    if (diff > 0) {
      hc -= transfer;
      C[ci] -= q;
      C[pidx[i]] += q;
    } else {
      hc += transfer;
      C[ci] += q;
      C[pidx[i]] -= q;
    }
  }
  return moves;
}

/* The cascade stage of a wave (world.h:90-168 at every site of the cycle):
   the window sequence of wave_cascade_site at each site, on the post-walk
   heights H, each window weighted by m = the largest number of site windows
   covering any cell of the window.  Adds the integer deltas to C (2^-20
   units); returns the number of non-zero transfers. */
int64_t orc_wave_cascade(const float* H, const uint8_t* visit, int nx, int ny,
                         const orc_params* P, int32_t* C) {
  size_t n = (size_t)nx * ny;
  int32_t* wc = (int32_t*)calloc(n, sizeof(int32_t));
  int64_t moves = 0;
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y) {
      if (!visit[(size_t)x * ny + y]) continue;
      for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b)
          if (!oob(nx, ny, x + a, y + b)) wc[(size_t)(x + a) * ny + (y + b)] += 1;
    }
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y) {
      if (!visit[(size_t)x * ny + y]) continue;
      int m = 1;
      for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b)
          if (!oob(nx, ny, x + a, y + b) && wc[(size_t)(x + a) * ny + (y + b)] > m)
            m = wc[(size_t)(x + a) * ny + (y + b)];
      moves += wave_cascade_site(H, nx, ny, P, x, y, m, C);
    }
  free(wc);
  return moves;
}

/* One wave of the drops [drop0, drop1) of a cycle on the current heights h:
   count walk, burst scale, scaled commits, the cascade over the wave's sites,
   and the heights updated in place.  Tracks are added to A's track sums (the
   caller blends once per cycle); dh/visit, when given, are accumulated. */
static void wave_pass(float* h, const float* d, const float* mx, const float* my,
                      const float* root, const float* erfd, int nx, int ny, const orc_params* P,
                      uint64_t key, int drop0, int drop1, wacc* A, orc_stats* local,
                      int32_t* dh_out, uint8_t* visit_out) {
  size_t n = (size_t)nx * ny;
  memset(A->dh, 0, n * sizeof(int32_t));
  memset(A->visit, 0, n);
  memset(A->count, 0, n * sizeof(int32_t));
  float* scale = (float*)malloc(n * sizeof(float));
  int32_t* C = (int32_t*)calloc(n, sizeof(int32_t));
  float* H = (float*)malloc(n * sizeof(float));

  /* pass A: visit counts (no commits) */
  A->scale = NULL;
  for (int i = drop0; i < drop1; ++i) {
    uint32_t cid = orc_spawn_cell(key, (uint32_t)i, (uint32_t)n);
    if ((double)h[cid] < 0.1) continue;
    wave_walk(h, d, mx, my, root, erfd, nx, ny, P, (float)(cid / (uint32_t)ny),
              (float)(cid % (uint32_t)ny), A, local);
  }
  for (size_t c = 0; c < n; ++c) {
    float effD = P->depositionRate * (1.0f - root[c]);
    if (effD < 0) effD = 0;
    scale[c] = orc_burst_scale(A->count[c], effD, P->entrainment, erfd[c]);
  }
  /* pass B: scaled commits */
  A->scale = scale;
  for (int i = drop0; i < drop1; ++i) {
    uint32_t cid = orc_spawn_cell(key, (uint32_t)i, (uint32_t)n);
    if ((double)h[cid] < 0.1) {
      local->skipped++;
      continue;
    }
    local->walked++;
    wave_walk(h, d, mx, my, root, erfd, nx, ny, P, (float)(cid / (uint32_t)ny),
              (float)(cid % (uint32_t)ny), A, local);
  }
  if (dh_out)
    for (size_t i = 0; i < n; ++i) dh_out[i] += A->dh[i];
  if (visit_out)
    for (size_t i = 0; i < n; ++i) visit_out[i] |= A->visit[i];

  /* apply the walk's height commits */
  for (size_t i = 0; i < n; ++i) H[i] = h[i] + (float)A->dh[i] * HQ_INV;

  /* cascade: weighted Jacobi sweep over the sites */
  local->casc_moves += orc_wave_cascade(H, A->visit, nx, ny, P, C);

  for (size_t i = 0; i < n; ++i) h[i] = H[i] + (float)C[i] * HQ_INV;
  free(scale); free(C); free(H);
}

void orc_wave_cycle_ex(float* h, float* d, float* mx, float* my, const float* root,
                       float* erfd, int nx, int ny, const orc_params* P, uint64_t seed,
                       uint32_t terrain, uint32_t cycle, int drops, orc_stats* st,
                       int32_t* dh_out, uint8_t* visit_out, int64_t* trk_out) {
  size_t n = (size_t)nx * ny;
  wacc A;
  A.dh = (int32_t*)calloc(n, sizeof(int32_t));
  A.td = (int64_t*)calloc(n, sizeof(int64_t));
  A.tmx = (int64_t*)calloc(n, sizeof(int64_t));
  A.tmy = (int64_t*)calloc(n, sizeof(int64_t));
  A.visit = (uint8_t*)calloc(n, 1);
  A.count = (int32_t*)calloc(n, sizeof(int32_t));
  orc_stats local;
  memset(&local, 0, sizeof(local));
  uint64_t key = orc_cycle_key(seed, terrain, cycle);
  if (dh_out) memset(dh_out, 0, n * sizeof(int32_t));
  if (visit_out) memset(visit_out, 0, n);

  wave_pass(h, d, mx, my, root, erfd, nx, ny, P, key, 0, drops, &A, &local, dh_out, visit_out);
  if (trk_out) { /* the cycle's summed tracks, 2^-24 units, before the blend */
    memcpy(trk_out, A.td, n * sizeof(int64_t));
    memcpy(trk_out + n, A.tmx, n * sizeof(int64_t));
    memcpy(trk_out + 2 * n, A.tmy, n * sizeof(int64_t));
  }

  /* blend (world.h:81-86), refresh erf(0.4 d) */
  for (size_t i = 0; i < n; ++i) {
    float t_d = (float)((double)A.td[i] * TQ_INV);
    float t_x = (float)((double)A.tmx[i] * TQ_INV);
    float t_y = (float)((double)A.tmy[i] * TQ_INV);
    d[i] = (1.0f - P->lrate) * d[i] + P->lrate * t_d;
    mx[i] = (1.0f - P->lrate) * mx[i] + P->lrate * t_x;
    my[i] = (1.0f - P->lrate) * my[i] + P->lrate * t_y;
    erfd[i] = orc_erf(0.4f * d[i]);
  }
  if (st) {
    st->steps += local.steps; st->walked += local.walked; st->skipped += local.skipped;
    st->exits += local.exits; st->ends_age += local.ends_age; st->ends_vol += local.ends_vol;
    st->export40 += local.export40; st->gain40 += local.gain40;
    st->walk_dh += local.walk_dh; st->casc_moves += local.casc_moves;
  }
  free(A.dh); free(A.td); free(A.tmx); free(A.tmy); free(A.visit); free(A.count);
}

void orc_wave_cycle(float* h, float* d, float* mx, float* my, const float* root,
                    float* erfd, int nx, int ny, const orc_params* P, uint64_t seed,
                    uint32_t terrain, uint32_t cycle, int drops, orc_stats* st,
                    int32_t* dh_out, uint8_t* visit_out) {
  orc_wave_cycle_ex(h, d, mx, my, root, erfd, nx, ny, P, seed, terrain, cycle, drops, st,
                    dh_out, visit_out, NULL);
}
