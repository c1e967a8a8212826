// Per-launch device workspace of the persistent relaxation kernels: the
// control block and the tile queue (ring, per-tile state and visited flags).
// Allocated stream-ordered (hipMallocAsync), or laid out in memory the
// caller provides, and initialised by a kernel, so a launch sequence
// contains no memset node.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"
#include "tile_engine.h"

namespace tt {

__global__ void zero_words(unsigned* p, size_t n);

// Bytes of the queue workspace of ntiles tiles.
inline size_t queue_workspace_bytes(int ntiles) {
  return (64 + (size_t)ntiles * 4) * sizeof(unsigned);
}

struct Workspace {
  void* base = nullptr;
  bool owned = false;  // base was allocated here (released by finish)
  TileQueue q{};
  hipStream_t stream = nullptr;

  // All ntiles tiles start queued, in index order.
  hipError_t alloc(int ntiles, hipStream_t s);
  // All tiles of an ntx x nty grid start queued, boundary rings first.
  hipError_t alloc(int ntiles, int ntx, int nty, hipStream_t s);
  // The same layout in caller memory `mem` (device, 256-byte aligned, at
  // least queue_workspace_bytes(ntiles)); nothing is allocated or freed.
  hipError_t attach(void* mem, int ntiles, int ntx, int nty, hipStream_t s);
  // Copies the control block into `st` (synchronises the stream) when st
  // is non-null, then releases allocated memory stream-ordered.
  hipError_t finish(RunStats* st);
};

// Device address of this GPU's sticky error word (looked up once per
// device) and the wait bound new launches carry; both fill a TileQueue.
hipError_t sticky_error_word(unsigned** p);
unsigned long long wait_limit_ticks();
// Sets the wait bound of later launches (0: the default, 10 s).
void set_wait_limit_ticks(unsigned long long ticks);
// Reads the current GPU's sticky error word after synchronising the
// device, and clears it when `clear` is set.
hipError_t take_sticky_error(bool clear, unsigned* bits);

// Workgroups of the persistent grid: co-resident blocks, capped by tiles.
hipError_t persistent_grid(const void* kernel, int threads, int ntiles, int* grid);

// Allocates the workspace (or lays it out in `scratch`, device memory of
// queue_workspace_bytes(ntiles), when non-null), launches `kern` as a
// persistent grid over `ntiles` tiles (Args must carry a TileQueue `q`)
// and releases an allocated workspace stream-ordered; fills `stats`
// (synchronising) when non-null. Tiles of an ntx x nty grid start in
// boundary-first order.
// With max_visits > 0 the launch stops after about that many tile visits
// (kStopVisitCap in the control block; every tile is visited at least once).
template <class Args>
hipError_t launch_persistent(void (*kern)(Args), int threads, int ntiles, Args a,
                             hipStream_t stream, RunStats* stats, int tile_x, int tile_y,
                             int ntx = 0, int nty = 0, void* scratch = nullptr,
                             unsigned max_visits = 0) {
  int grid = 0;
  hipError_t e = persistent_grid((const void*)kern, threads, ntiles, &grid);
  if (e != hipSuccess) return e;
  Workspace ws;
  e = scratch ? ws.attach(scratch, ntiles, ntx, nty, stream)
              : ws.alloc(ntiles, ntx, nty, stream);
  if (e != hipSuccess) return e;
  a.q = ws.q;
  a.q.stats = stats ? 1 : 0;
  a.q.max_visits = max_visits;
  kern<<<grid, threads, 0, stream>>>(a);
  e = hipGetLastError();
  if (stats) {
    stats->grid = grid;
    stats->tile_x = tile_x;
    stats->tile_y = tile_y;
  }
  hipError_t f = ws.finish(stats);
  return e != hipSuccess ? e : f;
}

}  // namespace tt
