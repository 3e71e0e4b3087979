// world_demo.cpp -- the reference's headless erosion loop
// (SimpleHydrology.cpp:36-38, 313-325) written against the C++ facade.
//
//   world_demo [size] [cycles] [drops] [seed]
//
// Prints one JSON line: an FNV-1a hash of the final height and discharge
// bits, the particle-step count, and the steps of one Drop walked with
// Drop::descend on the live map afterwards.
#include <cstdint>
#include <cstdio>
#include <string>
#include <cstdlib>
#include <cstring>

#include "../../include/shy_world.hpp"

static uint64_t fnv(uint64_t h, float v) {
  uint32_t b;
  std::memcpy(&b, &v, 4);
  for (int k = 0; k < 4; ++k) {
    h ^= (b >> (8 * k)) & 0xffu;
    h *= 1099511628211ULL;
  }
  return h;
}

int main(int argc, char** argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 256;
  const int cycles = argc > 2 ? std::atoi(argv[2]) : 10;
  const int drops = argc > 3 ? std::atoi(argv[3]) : 4096;
  const unsigned seed = argc > 4 ? (unsigned)std::atoi(argv[4]) : 1u;
  const bool fbm = argc > 5 && std::string(argv[5]) == "fbm";
  try {
    shy::World::SEED = seed;
    if (fbm) shy::World::map.init_fbm(n, n, shy::World::SEED);
    else shy::World::map.init(n, n, shy::World::SEED);
    for (int i = 0; i < cycles; ++i) shy::World::erode(drops);
    shy::World::map.sync();
    uint64_t h = 1469598103934665603ULL;
    for (const shy::cell& c : shy::World::map.cells) {
      h = fnv(h, c.height);
      h = fnv(h, c.discharge);
    }
    shy_stats st;
    shy::check(shy::World::map.ctx, shy_get_stats(shy::World::map.ctx, &st));
    shy::Drop drop(shy::vec2(n / 2 + 0.5f, n / 3 + 0.5f));
    int steps = 0;
    while (drop.descend()) ++steps;
    std::printf("{\"size\": %d, \"cycles\": %d, \"drops\": %d, \"seed\": %u, \"hash\": \"%016llx\", "
                "\"particle_steps\": %lld, \"drop_steps\": %d, \"drop_volume\": %.9g}\n",
                n, cycles, drops, seed, (unsigned long long)h, (long long)st.steps, steps,
                drop.volume);
    shy::World::map.release();
  } catch (const shy::Error& e) {
    std::fprintf(stderr, "world_demo: %s (code %d)\n", e.what(), e.code);
    return 1;
  }
  return 0;
}
