// world_demo.cpp -- the reference's headless erosion loop
// (SimpleHydrology.cpp:36-38, 313-325) written against the C++ facade, with
// the map reads of its frame loop and of Vegetation through the node-level
// chain.
//
//   world_demo [size] [cycles] [drops] [seed] [reference|fbm] [tileres]
//
// Prints one JSON line:
//   hash         FNV-1a of the final (height, discharge) bits, x-major, each
//                cell read as World::map.get(p)->get(p) with p a vec2, as the
//                tree placement does (SimpleHydrology.cpp:331)
//   hash_nodes   the same bits in the order `for (auto& node : map.nodes)
//                for (auto [cell, pos] : node.s)` visits them (world.h:56-61)
//   hash_momentum FNV-1a of (momentumx, momentumy) read as
//                node = map.get(p); cell = node->get(p) (SimpleHydrology.cpp:349-352)
//   neighbours   in-bounds results of map.getCell(pos + vec2(dx, dy)) over
//                the 8 offsets of every cell (vegetation.h:91-115)
//   particle_steps, drop_steps, drop_volume  the step count, and one Drop
//                walked with Drop::descend on the live map afterwards
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

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
  const int tileres = argc > 6 ? std::atoi(argv[6]) : 512;
  try {
    using shy::World;
    World::SEED = seed;
    World::map.tileres = tileres;
    if (fbm) World::map.init_fbm(n, n, World::SEED);
    else World::map.init(n, n, World::SEED);
    for (int i = 0; i < cycles; ++i) World::erode(drops);
    World::map.sync();

    uint64_t h = 1469598103934665603ULL, hm = h, hn = h;
    long long neighbours = 0;
    for (int x = 0; x < n; ++x)
      for (int y = 0; y < n; ++y) {
        const shy::vec2 p(x + 0.5f, y + 0.25f);
        h = fnv(h, World::map.get(p)->get(p)->height);
        h = fnv(h, World::map.get(p)->get(p)->discharge);
        auto node = World::map.get(p);
        auto cell = node->get(p);
        hm = fnv(hm, cell->momentumx);
        hm = fnv(hm, cell->momentumy);
        for (int dx = -1; dx <= 1; ++dx)
          for (int dy = -1; dy <= 1; ++dy)
            if ((dx || dy) && World::map.getCell(p + shy::vec2((float)dx, (float)dy)) != nullptr)
              ++neighbours;
      }
    for (auto& node : World::map.nodes)
      for (auto [cell, pos] : node.s) {
        hn = fnv(hn, cell.height);
        hn = fnv(hn, cell.discharge);
        (void)pos;
      }

    shy_stats st;
    shy::check(World::map.ctx, shy_get_stats(World::map.ctx, &st));
    shy::Drop drop(shy::vec2(n / 2 + 0.5f, n / 3 + 0.5f));
    int steps = 0;
    while (drop.descend()) ++steps;
    std::printf("{\"size\": %d, \"cycles\": %d, \"drops\": %d, \"seed\": %u, \"tileres\": %d, "
                "\"nodes\": %zu, \"hash\": \"%016llx\", \"hash_nodes\": \"%016llx\", "
                "\"hash_momentum\": \"%016llx\", \"neighbours\": %lld, "
                "\"particle_steps\": %lld, \"drop_steps\": %d, \"drop_volume\": %.9g}\n",
                n, cycles, drops, seed, tileres, World::map.nodes.size(),
                (unsigned long long)h, (unsigned long long)hn, (unsigned long long)hm,
                neighbours, (long long)st.steps, steps, drop.volume);
    World::map.release();
  } catch (const shy::Error& e) {
    std::fprintf(stderr, "world_demo: %s (code %d)\n", e.what(), e.code);
    return 1;
  }
  return 0;
}
