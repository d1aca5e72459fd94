// Device memory arena for the contexts and batches of a serving loop
// (hdg_server.cpp).
// Not part of the public ABI.
//
// The serving slots build a new set-up for almost every request (new mesh
// size, new order), so their contexts free and allocate device buffers all
// the time.  Through the device's stream-ordered pool those buffers go back
// to the driver at synchronisations and come back at the next allocation,
// and while the runtime trims and regrows the pool every HIP call of every
// slot waits (serving stream of 835 requests: stalls of 1-3 s).  An arena
// takes the budget from the device once and hands out pieces of it with a
// host-side first-fit free list, so the slots never call the allocator while
// requests run.
#pragma once
#include <cstddef>
#include <iterator>
#include <map>
#include <mutex>

#include "../../include/hdg.h"

// First-fit sub-allocator over one device allocation; free pieces are kept
// by offset and merged with their neighbours when returned.
struct hdg_arena {
  std::mutex mu;
  int device = 0;
  char* base = nullptr;
  size_t size = 0;
  std::map<size_t, size_t> free_;  // offset -> bytes
  static size_t round(size_t b) { return (b + 255) & ~size_t(255); }
  bool owns(const void* p) const { return p >= base && p < base + size; }
  void* take(size_t bytes) {
    bytes = round(bytes);
    std::lock_guard<std::mutex> lk(mu);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, rest = it->second - bytes;
      free_.erase(it);
      if (rest) free_[off + bytes] = rest;
      return base + off;
    }
    return nullptr;
  }
  void give(void* p, size_t bytes) {
    bytes = round(bytes);
    size_t off = (size_t)((char*)p - base);
    std::lock_guard<std::mutex> lk(mu);
    auto next = free_.lower_bound(off);
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        off = prev->first;
        bytes += prev->second;
        free_.erase(prev);
      }
    }
    if (next != free_.end() && off + bytes == next->first) {
      bytes += next->second;
      free_.erase(next);
    }
    free_[off] = bytes;
  }
};

// One hipMalloc of `bytes` on `device`; HDG_OK or HDG_ERR_HIP.
int hdg_arena_create(hdg_arena** out, int device, size_t bytes);
// Frees the arena (every context that used it must be destroyed first).
void hdg_arena_destroy(hdg_arena* a);
// Later allocations of the context come from the arena while it has room
// (and from the device pool otherwise); buffers return to the arena when the
// context frees them, after the context's stream has drained.
int hdg_attach_arena(hdg_ctx* c, hdg_arena* a);
// hdg_batch_create whose device buffers come from the arena while it has
// room (the serving loop's continuous batching).
int hdg_batch_create_in_arena(hdg_batch** out, int device, size_t hbm_budget, int n, const hdg_request* reqs,
                              hdg_arena* arena);
