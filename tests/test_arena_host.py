"""Host logic of the serving arena (csrc/hdg_arena.h): first-fit pieces of
one device allocation, 256-byte granules, neighbours merged on return.  A
randomized sequence of takes and gives is checked against a granule bitmap
(no overlap, nothing outside the arena, alignment, every free granule
reachable again once all pieces are back).  The struct is host C++ only, so
the check compiles it with g++ against a fake base address; no device call."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>
#include "hdg_arena.h"

int main() {
  const size_t G = 256, NG = 4096;  // 1 MiB arena
  hdg_arena a;
  a.base = reinterpret_cast<char*>(uintptr_t(1) << 40);
  a.size = NG * G;
  a.free_[0] = a.size;
  std::vector<int> owner(NG, -1);
  struct Piece { char* p; size_t bytes; int id; };
  std::vector<Piece> live;
  std::mt19937_64 rng(12345);
  int next_id = 0, fails = 0;
  for (int step = 0; step < 200000; ++step) {
    const bool take = live.empty() || (rng() % 100) < 55;
    if (take) {
      const size_t bytes = 1 + rng() % (rng() % 4 == 0 ? 64 * G : 3 * G);
      char* p = static_cast<char*>(a.take(bytes));
      // first fit: the lowest run of free granules that is long enough
      const size_t need = (bytes + G - 1) / G;
      size_t expect = NG, run = 0;
      for (size_t g = 0; g < NG; ++g) {
        run = owner[g] < 0 ? run + 1 : 0;
        if (run == need) { expect = g + 1 - need; break; }
      }
      if (!p) {
        if (expect != NG) { std::printf("take failed with room at granule %zu\n", expect); return 1; }
        ++fails;
        continue;
      }
      const size_t off = size_t(p - a.base);
      if (off % G || off + need * G > a.size) { std::printf("bad piece %zu\n", off); return 1; }
      if (off / G != expect) { std::printf("not first fit: %zu vs %zu\n", off / G, expect); return 1; }
      for (size_t g = off / G; g < off / G + need; ++g) {
        if (owner[g] >= 0) { std::printf("overlap at granule %zu\n", g); return 1; }
        owner[g] = next_id;
      }
      live.push_back({p, bytes, next_id++});
    } else {
      const size_t i = rng() % live.size();
      Piece q = live[i];
      live[i] = live.back();
      live.pop_back();
      if (!a.owns(q.p)) { std::printf("owns() false for a piece\n"); return 1; }
      a.give(q.p, q.bytes);
      const size_t off = size_t(q.p - a.base), need = (q.bytes + G - 1) / G;
      for (size_t g = off / G; g < off / G + need; ++g) owner[g] = -1;
    }
    // the free list covers exactly the free granules, in maximal runs
    size_t covered = 0, prev_end = SIZE_MAX;
    for (auto& [off, bytes] : a.free_) {
      if (off == prev_end) { std::printf("unmerged neighbours at %zu\n", off); return 1; }
      for (size_t g = off / G; g < (off + bytes) / G; ++g)
        if (owner[g] >= 0) { std::printf("free list holds a live granule\n"); return 1; }
      covered += bytes;
      prev_end = off + bytes;
    }
    size_t nfree = 0;
    for (int o : owner) nfree += o < 0;
    if (covered != nfree * G) { std::printf("free list misses granules\n"); return 1; }
  }
  for (auto& q : live) a.give(q.p, q.bytes);
  if (a.free_.size() != 1 || a.free_.begin()->first != 0 || a.free_.begin()->second != a.size) {
    std::printf("arena not whole again\n");
    return 1;
  }
  if (a.owns(a.base + a.size) || a.owns(a.base - 1)) { std::printf("owns() past the ends\n"); return 1; }
  std::printf("ok %d full\n", fails);
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_arena_first_fit_and_merge(tmp_path):
    src = tmp_path / "arena_check.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "arena_check"
    csrc = os.path.join(REPO, "hdiscontinuousgalerkin.jl_amd", "csrc")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", csrc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok")
