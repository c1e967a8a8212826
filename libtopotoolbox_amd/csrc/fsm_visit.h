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

// Proposal of tile cell (r, c) from its x-neighbour minimum u1 and
// y-neighbour minimum u2: the Godunov update with the cell's fi from LDS.
template <class G>
struct FsmProp {
  const float* sfi;
  __device__ __forceinline__ float operator()(int r, int c, float u1, float u2) const {
    return godunov(u1, u2, sfi[G::at(r, c)]);
  }
};

// Fixed-point test of a wavefront's quadrant: true (in some lane) if a cell
// would accept its Godunov proposal. Lanes are columns walking the rows
// (rows above / below in registers, side columns by DPP); the steps are
// independent.
template <class G, class Prop>
__device__ __forceinline__ bool fsm_violation(const float* su, const Prop& prop, int R0, int C0,
                                              int lane) {
  const int C = C0 + lane;
  const int S = (lane == 0) ? C - 1 : C + 1;
  float up = su[G::at(R0 - 1, C)], mid = su[G::at(R0, C)];
  bool v = false;
#pragma unroll 4
  for (int r = 0; r < 64; r++) {
    const float dn = su[G::at(R0 + r + 1, C)];
    const float edge = su[G::at(R0 + r, S)];
    const float u1 = fminf(lane_from_above(mid, edge), lane_from_below(mid, edge));
    const float u2 = fminf(dn, up);
    v |= prop(R0 + r, C, u1, u2) < mid;
    up = mid;
    mid = dn;
  }
  return v;
}

// ---------------------------------------------------------------- register sweeps
// The wavefront's 64x64 quadrant of u in registers: u[i] is cell
// (R0 + i, C0 + lane) in the column layout (vertical sweeps) and cell
// (R0 + lane, C0 + i) in the row layout (horizontal sweeps), switched
// through the tile in LDS; fi comes from LDS (independent loads). Along
// the sweep the carried value is the updated cell behind and the register
// ahead the old one (the outside row / column at the ends); across it the
// lanes' old values of the same row / column by DPP. Lanes 0 and 63 take
// the quadrant's outside line across the sweep only at a vertical sweep's
// first and last step (the quadrant's corner cells); every other edge cell
// has its full 4-neighbour proposal evaluated by the perpendicular sweeps,
// and the corner cells by these steps, so every cell's full Godunov
// proposal is applied by some sweep. Two interleaved chains of 32 steps
// (the second starts from the old value at the split) overlap each
// other's dependent update.
template <class G, bool VERT, bool FWD, class Prop>
__device__ __forceinline__ bool fsm_sweep_regs(float (&u)[64], const float* su, const Prop& prop,
                                               int R0, int C0, int lane) {
  constexpr int step = FWD ? 1 : -1;
  const int out_k = FWD ? -1 : 64, end_k = FWD ? 64 : -1;
  float pc[2];
  pc[0] = VERT ? su[G::at(R0 + out_k, C0 + lane)] : su[G::at(R0 + lane, C0 + out_k)];
  pc[1] = u[FWD ? 31 : 32];
  const float beyond = VERT ? su[G::at(R0 + end_k, C0 + lane)] : su[G::at(R0 + lane, C0 + end_k)];
  float e_first = kInf, e_last = kInf;
  if (VERT) {
    const int rf = FWD ? 0 : 63, rl = FWD ? 63 : 0;
    e_first = su[G::at(R0 + rf, lane == 0 ? C0 - 1 : C0 + 64)];
    e_last = su[G::at(R0 + rl, lane == 0 ? C0 - 1 : C0 + 64)];
  }
  bool changed = false;
#pragma unroll
  for (int s = 0; s < 32; s++) {
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const int i = c == 0 ? (FWD ? s : 63 - s) : (FWD ? 32 + s : 31 - s);
      const int gs = c * 32 + s;  // step of the whole sweep
      const float cur = u[i];
      const int j = i + step;
      const float far = (j >= 0 && j < 64) ? u[j] : beyond;
      const float old = (VERT && gs == 0) ? e_first : ((VERT && gs == 63) ? e_last : kInf);
      const float lo = __int_as_float(__builtin_amdgcn_update_dpp(
          __float_as_int(old), __float_as_int(cur), 0x138, 0xf, 0xf, false));
      const float hi = __int_as_float(__builtin_amdgcn_update_dpp(
          __float_as_int(old), __float_as_int(cur), 0x130, 0xf, 0xf, false));
      const float across = fminf(hi, lo);
      const float along = FWD ? fminf(far, pc[c]) : fminf(pc[c], far);
      const float p = VERT ? prop(R0 + i, C0 + lane, across, along)
                           : prop(R0 + lane, C0 + i, along, across);
      const bool upd = p < cur;
      changed |= upd;
      const float nv = upd ? p : cur;
      u[i] = nv;
      pc[c] = nv;
    }
  }
  return changed;
}

// Relaxes the tile to its local fixed point with the register sweeps (state
// in su before and after); returns the sweep iterations.
template <class G, class Prop>
__device__ inline unsigned fsm_relax_regs(float* su, const Prop& prop, int R0, int C0, int lane,
                                          unsigned max_local, int* any_out, int* changed_out) {
  float u[64];
  unsigned it = 0;
  int any = __syncthreads_or(fsm_violation<G>(su, prop, R0, C0, lane) ? 1 : 0);
  int any_change = 0;
  while (any && it < max_local) {
#pragma unroll
    for (int r = 0; r < 64; r++) u[r] = su[G::at(R0 + r, C0 + lane)];
    bool ch = fsm_sweep_regs<G, true, true>(u, su, prop, R0, C0, lane);
    ch |= fsm_sweep_regs<G, true, false>(u, su, prop, R0, C0, lane);
#pragma unroll
    for (int r = 0; r < 64; r++) su[G::at(R0 + r, C0 + lane)] = u[r];
#pragma unroll
    for (int c = 0; c < 64; c++) u[c] = su[G::at(R0 + lane, C0 + c)];
    ch |= fsm_sweep_regs<G, false, true>(u, su, prop, R0, C0, lane);
    ch |= fsm_sweep_regs<G, false, false>(u, su, prop, R0, C0, lane);
#pragma unroll
    for (int c = 0; c < 64; c++) su[G::at(R0 + lane, C0 + c)] = u[c];
    any = __syncthreads_or(ch ? 1 : 0);
    any_change |= any;
    it++;
    if (any) any = __syncthreads_or(fsm_violation<G>(su, prop, R0, C0, lane) ? 1 : 0);
  }
  *any_out = any;
  *changed_out = any_change;
  return it;
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
                StageF32Scaled{v.slopes, sfi, __builtin_nanf(""), v.cellsize}, true, x0, y0, v.nx,
                v.ny);
  __syncthreads();
  for_stale_halo<G>(first, nvis, [&](int r, int c, bool from_state) {
    const long long gx = x0 + c, gy = y0 + r;
    if (gx < 0 || gx >= v.nx || gy < 0 || gy >= v.ny) return;
    const size_t g = (size_t)gy * v.nx + gx;
    su[G::at(r, c)] = from_state ? v.u[g] : v.dem[g];
  });
  __syncthreads();
  save_perimeter<G>(lds.sper, su_bits);
  const unsigned long long c1 = wall_clock64();
  int any = 0, any_change = 0;
  const unsigned it =
      fsm_relax_regs<G>(su, FsmProp<G>{sfi}, R0, C0, lane, max_local, &any, &any_change);
  const unsigned long long c2 = wall_clock64();
  // the first visit initialises the output; later visits write changed
  // tiles only
  if (first || any_change) write_tile_f32<G>(v.u, su, x0, y0, v.nx, v.ny);
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
