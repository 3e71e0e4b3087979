"""The C++ facade's host-side map view (include/shy_world.hpp), CPU only: a
program written like the reference's call sites compiles against it and
reads and writes the host snapshot through the node-level chain.  No engine
call is made (the map is filled by hand, no context is created)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r'''
#include <cstdio>
#include "shy_world.hpp"
using shy::World;
using shy::vec2;
using shy::ivec2;
int main() {
  const int nx = 10, ny = 7;
  World::map.nx = nx; World::map.ny = ny; World::map.tileres = 4;
  World::map.cells.assign(nx * ny, shy::cell{});
  for (int i = 0; i < nx * ny; ++i) World::map.cells[i].height = 0.5f + i;
  World::map.erfd.assign(nx * ny, 0.25f);
  World::map.build_nodes();
  int bad = 0;
  // world.h:56-61 -- track reset through node slices, structured bindings by value
  for (auto& node : World::map.nodes)
    for (auto [cell, pos] : node.s) {
      cell.discharge_track = 7.0f;
      cell.momentumx_track = (float)(node.pos.x + pos.x);
      cell.momentumy_track = (float)(node.pos.y + pos.y);
    }
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y) {
      const shy::cell& c = World::map.cells[x * ny + y];
      if (c.discharge_track != 7.0f || c.momentumx_track != x || c.momentumy_track != y) ++bad;
    }
  // water.h:60-68 -- const glm::ivec2 ipos = pos; node = map.get(ipos); cell = node->get(ipos)
  const vec2 pos(3.7f, 5.2f);
  const ivec2 ipos = pos;
  shy::node* node = World::map.get(ipos);
  shy::cell* cell = node ? node->get(ipos) : nullptr;
  if (!cell || cell->height != 0.5f + 3 * ny + 5) ++bad;
  if (node->pos.x != 0 || node->pos.y != 4) ++bad;
  // SimpleHydrology.cpp:331 -- map.get(t.pos)->get(t.pos)->height with a vec2
  if (World::map.get(vec2(9.9f, 6.1f))->get(vec2(9.9f, 6.1f))->height != 0.5f + 9 * ny + 6) ++bad;
  // vegetation.h:91-115 -- getCell(pos + vec2(...)), nullptr off the map
  if (World::map.getCell(vec2(9.0f, 0.0f) + vec2(1, 0)) != nullptr) ++bad;
  if (World::map.getCell(vec2(0.0f, 0.0f) + vec2(-1, 0)) != nullptr) ++bad;
  if (World::map.getCell(vec2(8.5f, 5.5f) + vec2(1, 1)) != &World::map.cells[9 * ny + 6]) ++bad;
  // node-level accessors: height 0 off the node, discharge = erf field
  if (node->height(ivec2(9, 9)) != 0.0f || node->discharge(ipos) != 0.25f) ++bad;
  if (World::map.height(ivec2(-1, 0)) != 0.0f || !World::map.oob(ivec2(nx, 0))) ++bad;
  // node grid: ceil(10/4) x ceil(7/4), ragged last nodes, every cell visited once
  size_t visited = 0;
  for (auto& n : World::map.nodes) visited += n.s.size();
  if (World::map.nodes.size() != 6 || visited != (size_t)(nx * ny)) ++bad;
  if (World::map.nodes[5].s.res.x != 2 || World::map.nodes[5].s.res.y != 3) ++bad;
  std::printf("%d\n", bad);
  return bad != 0;
}
'''


def test_facade_host_view_follows_reference_call_sites(tmp_path):
    src = tmp_path / "facade.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "facade"
    lib_dir = os.path.join(ROOT, "simplehydrology_amd")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I",
                        os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        "-L" + lib_dir, "-lshy", "-Wl,-rpath," + lib_dir],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "0", (r.stdout, r.stderr)
