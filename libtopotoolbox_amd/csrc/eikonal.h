// Godunov upwind update of the constrained eikonal solver and its local
// linearisation (shared by the forward sweep and the adjoint).
#pragma once

#include <hip/hip_runtime.h>

namespace tt {

// Proposal for a cell from its x-neighbour minimum u1, y-neighbour minimum
// u2 and fi = cellsize * slope; the float operations of
// src/excesstopography.c:36-40 in the same order (built without FMA
// contraction).
__device__ __forceinline__ float godunov(float u1, float u2, float fi) {
  if (fabsf(u1 - u2) < fi) {
    return (u1 + u2) / 2 + sqrtf(-(u1 - u2) * (u1 - u2) + 2 * (fi * fi)) / 2;
  }
  return fminf(u1 + fi, u2 + fi);
}

}  // namespace tt
