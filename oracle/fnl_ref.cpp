// Reference build of quad::map::init's terrain, for the oracle's checks only.
//
// Compiled by `make ref` against the reference's own vendored FastNoiseLite.h
// where it lies (source/include/FastNoiseLite.h); nothing of the reference is
// copied here.  The layer loop restates cellpool.h:349-404 around the real
// FastNoiseLite object (the rest of cellpool.h needs glm and OpenGL and
// cannot be built in this image).  Output goes to oracle/_ref/ only.
#include <cstddef>

#include "FastNoiseLite.h"

extern "C" {

// FastNoiseLite(1337), OpenSimplex2, FBm (defaults), at frequency freq
float fnl_ref_noise3(float freq, float x, float y, float z) {
  FastNoiseLite noise;
  noise.SetNoiseType(FastNoiseLite::NoiseType_OpenSimplex2);
  noise.SetFractalType(FastNoiseLite::FractalType_FBm);
  noise.SetFrequency(freq);
  return noise.GetNoise(x, y, z);
}

void fnl_ref_world(int nx, int ny, int seed, int tileres, float* out) {
  FastNoiseLite noise;
  noise.SetNoiseType(FastNoiseLite::NoiseType_OpenSimplex2);
  noise.SetFractalType(FastNoiseLite::FractalType_FBm);
  const size_t n = (size_t)nx * ny;
  for (size_t i = 0; i < n; ++i) out[i] = 0.0f;
  float frequency = 1.0f;
  float scale = 0.6f;
  for (int o = 0; o < 8; ++o) {
    noise.SetFrequency(frequency);
    for (int x = 0; x < nx; ++x)
      for (int y = 0; y < ny; ++y) {
        const float px = (float)x / (float)tileres, py = (float)y / (float)tileres;
        out[(size_t)x * ny + y] += scale * noise.GetNoise(px, py, (float)(seed % 10000));
      }
    frequency *= 2;
    scale *= 0.6;
  }
  float mn = 0.0f, mx = 0.0f;
  for (size_t i = 0; i < n; ++i) {
    mn = (mn < out[i]) ? mn : out[i];
    mx = (mx > out[i]) ? mx : out[i];
  }
  for (size_t i = 0; i < n; ++i) out[i] = (out[i] - mn) / (mx - mn);
}
}
