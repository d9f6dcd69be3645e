"""The serving arena's offset allocator (csrc/arena.h), compiled on the host with g++ and
driven through random admit/release sequences: blocks never overlap, stay inside the
capacity and are 256-byte aligned; freeing everything coalesces back to one block; a
request fails only when no free block is large enough."""
import subprocess

from conftest import PKG

DRIVER = r'''
#include "arena.h"
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
using phpc::OffsetArena;
struct Blk { size_t off, bytes; };
int main() {
  std::mt19937_64 rng(7);
  for (int trial = 0; trial < 200; ++trial) {
    OffsetArena a;
    const size_t cap = a.reset(1000000 + rng() % 5000000);
    std::vector<Blk> live;
    for (int step = 0; step < 2000; ++step) {
      if (live.empty() || rng() % 3) {
        const size_t want = 1 + rng() % 200000;
        const size_t off = a.alloc(want);
        if (off == OffsetArena::kNone) {
          // a failure is allowed only if no gap between live blocks fits the request
          std::vector<Blk> s = live;
          std::sort(s.begin(), s.end(), [](const Blk &x, const Blk &y) { return x.off < y.off; });
          size_t pos = 0, need = OffsetArena::round_up(want);
          for (const Blk &b : s) { if (b.off - pos >= need) { printf("spurious failure\n"); return 1; } pos = b.off + OffsetArena::round_up(b.bytes); }
          if (cap - pos >= need) { printf("spurious failure at end\n"); return 1; }
          continue;
        }
        if (off % OffsetArena::kAlign || off + want > cap) { printf("bad block\n"); return 1; }
        for (const Blk &b : live)
          if (off < b.off + b.bytes && b.off < off + want) { printf("overlap\n"); return 1; }
        live.push_back({off, want});
      } else {
        const size_t i = rng() % live.size();
        a.release(live[i].off, live[i].bytes);
        live.erase(live.begin() + i);
      }
    }
    for (const Blk &b : live) a.release(b.off, b.bytes);
    if (a.free_blocks() != 1 || a.alloc(cap) != 0) { printf("no full coalescing\n"); return 1; }
  }
  printf("ok\n");
  return 0;
}
'''


def test_offset_arena(tmp_path):
    src = tmp_path / "arena_check.cpp"
    src.write_text(DRIVER.replace('#include <vector>', '#include <vector>\n#include <algorithm>'))
    exe = tmp_path / "arena_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", str(PKG / "csrc"), str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
