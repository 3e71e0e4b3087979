// Host check of the walk records' 2x2-block layout (pack_rec, pack_records in
// simplehydrology_amd/csrc/shy_device.hpp): for each terrain size given as
// "nx ny" pairs on the command line, every cell maps to its own record below
// pack_records(nx, ny); the four cells of a 2x2 block share one 128-byte line
// (four 32-byte records); the records the pack kernel writes in record order
// map back to their cells.  Prints the cells checked and the mismatches;
// exits 1 on any mismatch.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../simplehydrology_amd/csrc/shy_device.hpp"

using shy::pack_rec;
using shy::pack_records;

int main(int argc, char** argv) {
  long long checked = 0, bad = 0;
  for (int a = 1; a + 1 < argc; a += 2) {
    const int nx = std::atoi(argv[a]), ny = std::atoi(argv[a + 1]);
    const long long nr = pack_records(nx, ny);
    std::vector<int> owner(nr, -1);
    for (int x = 0; x < nx; ++x)
      for (int y = 0; y < ny; ++y) {
        const int r = pack_rec(x, y, ny);
        ++checked;
        if (r < 0 || r >= nr || owner[r] != -1) { ++bad; continue; }
        owner[r] = x * ny + y;
        // one line per 2x2 block: the block's four records share r / 4
        if (r / 4 != pack_rec(x & ~1, y & ~1, ny) / 4) ++bad;
      }
    // record order back to cells, as pack_kernel numbers its threads
    const int nyb = (ny + 1) >> 1;
    for (int r = 0; r < nr; ++r) {
      const int b = r >> 2, q = r & 3, bx = b / nyb, by = b - bx * nyb;
      const int x = 2 * bx + (q >> 1), y = 2 * by + (q & 1);
      const bool cell = x < nx && y < ny;
      if (cell != (owner[r] != -1) || (cell && owner[r] != x * ny + y)) ++bad;
    }
  }
  std::printf("cells %lld mismatches %lld\n", checked, bad);
  return bad ? 1 : 0;
}
