// Tile visit of the fast-sweeping excess-topography relaxation (constrained
// eikonal), shared by the single-grid kernel (fsm.hip) and the batched
// kernel (batched.hip).
//
// Reference semantics (src/excesstopography.c:16-41, 47-111): u starts at
// the DEM; a cell accepts the Godunov upwind proposal
//     u1 = fminf(u[i+1], u[i-1])    (x neighbours, +INF outside the grid)
//     u2 = fminf(u[j+1], u[j-1])    (y neighbours)
//     |u1-u2| < fi ? (u1+u2)/2 + sqrtf(2 fi^2 - (u1-u2)^2)/2
//                  : fminf(u1+fi, u2+fi),      fi = cellsize * slope
// only when it is strictly lower than its current value, until no cell
// changes. The proposal uses the reference's float operations in the same
// order (no FMA contraction) except that the square root is the hardware
// one (within 1 ulp), and the visiting order differs (tiles relaxed in
// parallel, lanes sweeping rows or columns side by side): the converged
// state matches the reference up to rounding, which the parity contract
// bounds at 1e-4 relative to an FP64 solution.
//
// NaN cells never change (proposal < NaN is false) and fminf skips NaN
// neighbours, exactly as in the reference. Cells of a partial tile beyond
// the grid hold +INF with fi = NaN, so they never accept a proposal.
#pragma once

#include "eikonal.h"
#include "kernels.h"
#include "tile_engine.h"

namespace tt {

// One grid as seen by a tile visit.
struct FsmView {
  float* u;
  const float* dem;
  const float* slopes;
  float cellsize;
  long long nx, ny;
};

// One Gauss-Seidel sweep over a 64x64 quadrant. VERT: lanes are columns
// and the sweep walks rows (+y when FWD); the carried value is the updated
// neighbour behind the sweep, the one ahead ("far") is the old value.
// Otherwise lanes are rows and the sweep walks columns. LDS operands are
// fetched two steps ahead (none of them is written in between) and the
// lane's cell is stored every step.
template <class G, bool VERT, bool FWD>
__device__ __forceinline__ bool fsm_sweep(float* su, const float* sfi, int R0, int C0, int lane) {
  constexpr int step = FWD ? 1 : -1;
  bool changed = false;
  const int own = VERT ? C0 + lane : R0 + lane;
  const int side = (lane == 0) ? own - 1 : own + 1;
  const int k0 = VERT ? (FWD ? R0 : R0 + 63) : (FWD ? C0 : C0 + 63);
  auto cell = [&](int k) { return VERT ? G::at(k, own) : G::at(own, k); };
  auto edge_cell = [&](int k) { return VERT ? G::at(k, side) : G::at(side, k); };
  float pc = su[cell(k0 - step)];
  float c0 = su[cell(k0)], r0 = su[cell(k0 + step)], e0 = su[edge_cell(k0)], i0 = sfi[cell(k0)];
  float c1 = r0, r1 = su[cell(k0 + 2 * step)], e1 = su[edge_cell(k0 + step)],
        i1 = sfi[cell(k0 + step)];
  int k = k0;
#pragma unroll 2
  for (int s = 0; s < 64; s++, k += step) {
    const float cur = c0, far = r0, edge = e0, fi = i0;
    c0 = c1;
    r0 = r1;
    e0 = e1;
    i0 = i1;
    if (s < 62) {
      c1 = r1;  // the cell two steps ahead is read once, as "far" then "cur"
      r1 = su[cell(k + 3 * step)];
      e1 = su[edge_cell(k + 2 * step)];
      i1 = sfi[cell(k + 2 * step)];
    }
    // across-lane neighbours (columns C-1 / C+1, or rows R-1 / R+1)
    const float lo = lane_from_below(cur, edge);
    const float hi = lane_from_above(cur, edge);
    // u1 = fminf(x+1, x-1), u2 = fminf(y+1, y-1) as in the reference
    const float along = FWD ? fminf(far, pc) : fminf(pc, far);
    const float across = fminf(hi, lo);
    const float prop = VERT ? godunov(across, along, fi) : godunov(along, across, fi);
    const bool upd = prop < cur;
    changed |= upd;
    pc = upd ? prop : cur;
    su[cell(k)] = pc;
  }
  return changed;
}

// Fixed-point test of a wavefront's quadrant: true (in some lane) if a cell
// would accept its Godunov proposal. Lanes are columns walking the rows
// (rows above / below in registers, side columns by DPP); the steps are
// independent.
template <class G>
__device__ __forceinline__ bool fsm_violation(const float* su, const float* sfi, int R0, int C0,
                                              int lane) {
  const int C = C0 + lane;
  const int S = (lane == 0) ? C - 1 : C + 1;
  float up = su[G::at(R0 - 1, C)], mid = su[G::at(R0, C)];
  bool v = false;
#pragma unroll 4
  for (int r = 0; r < 64; r++) {
    const float dn = su[G::at(R0 + r + 1, C)];
    const float edge = su[G::at(R0 + r, S)];
    const float fi = sfi[G::at(R0 + r, C)];
    const float u1 = fminf(lane_from_above(mid, edge), lane_from_below(mid, edge));
    const float u2 = fminf(dn, up);
    v |= godunov(u1, u2, fi) < mid;
    up = mid;
    mid = dn;
  }
  return v;
}

// LDS of one excess-topography visit.
template <class G>
struct FsmLds {
  float su[G::CELLS];
  float sfi[G::CELLS];
  unsigned sper[2 * (G::TX + G::TY)];
  unsigned bits;
};

// Relaxes the tile at (x0, y0) of grid v (at most max_local sweep
// iterations; kSelf is returned if it stopped there) and writes it back.
// Returns the neighbours to queue; called by all threads of the workgroup.
template <class G>
__device__ inline unsigned fsm_visit(const FsmView& v, long long x0, long long y0, bool first,
                                     unsigned nvis, unsigned max_local, Sync* sync,
                                     FsmLds<G>& lds) {
  float* su = lds.su;
  float* sfi = lds.sfi;
  const unsigned* su_bits = reinterpret_cast<const unsigned*>(su);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int R0 = (wave / G::QX) * 64;
  const int C0 = (wave % G::QX) * 64;
  const unsigned long long c0 = wall_clock64();
  // u from the DEM on the first visit, from the stored state afterwards;
  // fi = cellsize * slope (NaN outside the grid: never accepts a proposal)
  stage_tile<G>(StageF32{first ? v.dem : v.u, su, kInf},
                StageF32{v.slopes, sfi, __builtin_nanf("")}, true, x0, y0, v.nx, v.ny);
  __syncthreads();
  for (int k = tid; k < G::CELLS; k += G::THREADS) sfi[k] = v.cellsize * sfi[k];
  for_stale_halo<G>(first, nvis, [&](int r, int c, bool from_state) {
    const long long gx = x0 + c, gy = y0 + r;
    if (gx < 0 || gx >= v.nx || gy < 0 || gy >= v.ny) return;
    const size_t g = (size_t)gy * v.nx + gx;
    su[G::at(r, c)] = from_state ? v.u[g] : v.dem[g];
  });
  __syncthreads();
  save_perimeter<G>(lds.sper, su_bits);
  const unsigned long long c1 = wall_clock64();
  unsigned it = 0;
  int any = __syncthreads_or(fsm_violation<G>(su, sfi, R0, C0, lane) ? 1 : 0);
  int any_change = 0;
  while (any && it < max_local) {
    bool ch = fsm_sweep<G, true, true>(su, sfi, R0, C0, lane);
    ch |= fsm_sweep<G, true, false>(su, sfi, R0, C0, lane);
    ch |= fsm_sweep<G, false, true>(su, sfi, R0, C0, lane);
    ch |= fsm_sweep<G, false, false>(su, sfi, R0, C0, lane);
    any = __syncthreads_or(ch ? 1 : 0);
    any_change |= any;
    it++;
    if (any) any = __syncthreads_or(fsm_violation<G>(su, sfi, R0, C0, lane) ? 1 : 0);
  }
  const unsigned long long c2 = wall_clock64();
  // the first visit initialises the output; later visits write changed
  // tiles only
  if (first || any_change) {
    for (int idx = tid; idx < G::TY * G::TX; idx += G::THREADS) {
      const int r = idx / G::TX, c = idx % G::TX;
      const long long gx = x0 + c, gy = y0 + r;
      if (gx < v.nx && gy < v.ny) v.u[(size_t)gy * v.nx + gx] = su[G::at(r, c)];
    }
  }
  // NaN cells compare as words, so they never count as changed. Every
  // Godunov proposal through p is at least u_p, so a neighbour cell q can
  // only be lowered by p if u_p < u_q.
  unsigned bits = block_or(
      pushes_from_perimeter<G, false>(lds.sper, su_bits, [&](int p, int q) { return su[p] < su[q]; }),
      &lds.bits);
  if (any) bits |= kSelf;
  if (tid == 0) visit_stats(sync, it, c0, c1, c2);
  return bits;
}

}  // namespace tt
