// Godunov upwind update of the constrained eikonal solver and its local
// linearisation (shared by the forward sweep and the adjoint).
#pragma once

#include <hip/hip_runtime.h>

namespace tt {

// Proposal for a cell from its x-neighbour minimum u1, y-neighbour minimum
// u2 and fi = cellsize * slope: the expression of
// src/excesstopography.c:36-40 with its operations in the same order (the
// library is built without FMA contraction), evaluated branch-free (both
// sides, then a select) and with the hardware square root (v_sqrt_f32,
// within 1 ulp) in place of a correctly rounded one. NaN results (the
// square root of a negative number when fi is tiny) are never accepted,
// since proposal < u is false for NaN, as in the reference.
__device__ __forceinline__ float godunov(float u1, float u2, float fi) {
  const float d = u1 - u2;
  const float two = (u1 + u2) / 2 + __builtin_amdgcn_sqrtf(-d * d + 2 * (fi * fi)) / 2;
  const float one = fminf(u1 + fi, u2 + fi);
  return fabsf(d) < fi ? two : one;
}

}  // namespace tt
