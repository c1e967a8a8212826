// Per-launch device workspace of the persistent relaxation kernels:
// the control block, the per-tile queue marks and the two round lists.
// Allocated stream-ordered (hipMallocAsync) and cleared by a kernel, so a
// launch sequence contains no memset node.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"
#include "tile_engine.h"

namespace tt {

__global__ void zero_words(unsigned* p, size_t n);

struct Workspace {
  void* base = nullptr;
  Sync* sync = nullptr;
  unsigned* mark = nullptr;
  unsigned* list = nullptr;
  size_t words = 0;
  hipStream_t stream = nullptr;

  hipError_t alloc(int ntiles, hipStream_t s);
  // Copies the control block into `st` (synchronises the stream) when st
  // is non-null, then releases the memory stream-ordered.
  hipError_t finish(RunStats* st);
};

// Workgroups of the persistent grid: co-resident blocks, capped by tiles.
hipError_t persistent_grid(const void* kernel, int threads, int ntiles, int* grid);

}  // namespace tt
