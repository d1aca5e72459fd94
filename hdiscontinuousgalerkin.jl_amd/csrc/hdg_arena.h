// Device memory arena for the contexts of a serving loop (hdg_server.cpp).
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

struct hdg_ctx;
struct hdg_arena;

// One hipMalloc of `bytes` on `device`; HDG_OK or HDG_ERR_HIP.
int hdg_arena_create(hdg_arena** out, int device, size_t bytes);
// Frees the arena (every context that used it must be destroyed first).
void hdg_arena_destroy(hdg_arena* a);
// Later allocations of the context come from the arena while it has room
// (and from the device pool otherwise); buffers return to the arena when the
// context frees them, after the context's stream has drained.
int hdg_attach_arena(hdg_ctx* c, hdg_arena* a);
