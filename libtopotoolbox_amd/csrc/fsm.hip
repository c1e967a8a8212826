// Excess topography by fast sweeping (constrained eikonal) on gfx950.
//
// Reference semantics (src/excesstopography.c:16-41, 47-111): u starts at
// the DEM; a cell accepts the Godunov upwind proposal
//     u1 = fminf(u[i+1], u[i-1])    (x neighbours, +INF outside the grid)
//     u2 = fminf(u[j+1], u[j-1])    (y neighbours)
//     |u1-u2| < fi ? (u1+u2)/2 + sqrtf(2 fi^2 - (u1-u2)^2)/2
//                  : fminf(u1+fi, u2+fi),      fi = cellsize * slope
// only when it is strictly lower than its current value, until no cell
// changes. The proposal is evaluated with the same float operations in the
// same order (the library is built with -ffp-contract=off), so every
// accepted value is what the reference would compute from the same
// neighbours; only the visiting order differs (tiles relaxed in parallel,
// lanes sweeping rows or columns side by side), which moves the converged
// state by rounding only.
//
// NaN cells never change (proposal < NaN is false) and fminf skips NaN
// neighbours, exactly as in the reference. Cells of a partial tile beyond
// the grid hold +INF with fi = NaN, so they never accept a proposal.
#include "eikonal.h"
#include "kernels.h"
#include "tile_engine.h"
#include "workspace.h"

namespace tt {

struct FsmArgs {
  float* u;
  const float* dem;
  const float* slopes;
  float cellsize;
  long long nx, ny;
  int ntx, nty;
  TileQueue q;
  unsigned max_local;
};

// One Gauss-Seidel sweep over a 64x64 quadrant. VERT: lanes are columns
// and the sweep walks rows (+y when FWD); the carried value is the updated
// y neighbour. Otherwise lanes are rows and the sweep walks columns.
template <class G, bool VERT, bool FWD>
__device__ __forceinline__ bool fsm_sweep(float* su, const float* sfi, int R0, int C0, int lane) {
  constexpr int step = FWD ? 1 : -1;
  bool changed = false;
  if (VERT) {
    const int C = C0 + lane;
    const int ec = (lane == 0) ? C - 1 : C + 1;
    int R = FWD ? R0 : R0 + 63;
    float pc = su[G::at(R - step, C)];
    float cur_n = su[G::at(R, C)], far_n = su[G::at(R + step, C)];
    float edge_n = su[G::at(R, ec)], fi_n = sfi[G::at(R, C)];
#pragma unroll 2
    for (int k = 0; k < 64; k++, R += step) {
      const float cur = cur_n, far = far_n, edge = edge_n, fi = fi_n;
      if (k < 63) {  // prefetch the next row before this row's store
        cur_n = su[G::at(R + step, C)];
        far_n = su[G::at(R + 2 * step, C)];
        edge_n = su[G::at(R + step, ec)];
        fi_n = sfi[G::at(R + step, C)];
      }
      const float xm = lane_from_below(cur, edge);  // (R, C-1)
      const float xp = lane_from_above(cur, edge);  // (R, C+1)
      const float u1 = fminf(xp, xm);
      const float u2 = FWD ? fminf(far, pc) : fminf(pc, far);
      const float prop = godunov(u1, u2, fi);
      const bool upd = prop < cur;
      if (upd) su[G::at(R, C)] = prop;
      changed |= upd;
      pc = upd ? prop : cur;
    }
  } else {
    const int R = R0 + lane;
    const int er = (lane == 0) ? R - 1 : R + 1;
    int C = FWD ? C0 : C0 + 63;
    float pc = su[G::at(R, C - step)];
    float cur_n = su[G::at(R, C)], far_n = su[G::at(R, C + step)];
    float edge_n = su[G::at(er, C)], fi_n = sfi[G::at(R, C)];
#pragma unroll 2
    for (int k = 0; k < 64; k++, C += step) {
      const float cur = cur_n, far = far_n, edge = edge_n, fi = fi_n;
      if (k < 63) {
        cur_n = su[G::at(R, C + step)];
        far_n = su[G::at(R, C + 2 * step)];
        edge_n = su[G::at(er, C + step)];
        fi_n = sfi[G::at(R, C + step)];
      }
      const float ym = lane_from_below(cur, edge);  // (R-1, C)
      const float yp = lane_from_above(cur, edge);  // (R+1, C)
      const float u1 = FWD ? fminf(far, pc) : fminf(pc, far);
      const float u2 = fminf(yp, ym);
      const float prop = godunov(u1, u2, fi);
      const bool upd = prop < cur;
      if (upd) su[G::at(R, C)] = prop;
      changed |= upd;
      pc = upd ? prop : cur;
    }
  }
  return changed;
}

template <class G>
__global__ __launch_bounds__(G::THREADS) void fsm_persistent(FsmArgs a) {
  __shared__ float su[G::CELLS];
  __shared__ float sfi[G::CELLS];
  __shared__ unsigned sper[2 * (G::TX + G::TY)];
  __shared__ unsigned s_bits;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int R0 = (wave / G::QX) * 64;
  const int C0 = (wave % G::QX) * 64;
  const unsigned* su_bits = reinterpret_cast<const unsigned*>(su);

  auto visit = [&](int t, bool first, unsigned nvis) -> unsigned {
    const unsigned long long c0 = wall_clock64();
    const int tx = t % a.ntx, ty = t / a.ntx;
    const long long x0 = (long long)tx * G::TX, y0 = (long long)ty * G::TY;
    // u from the DEM on the first visit, from the stored state afterwards;
    // fi = cellsize * slope (NaN outside the grid: never accepts a proposal)
    stage_tile<G>(StageF32{first ? a.dem : a.u, su, kInf},
                  StageF32{a.slopes, sfi, __builtin_nanf("")}, true, x0, y0, a.nx, a.ny);
    __syncthreads();
    for (int k = tid; k < G::CELLS; k += G::THREADS) sfi[k] = a.cellsize * sfi[k];
    for_stale_halo<G>(first, nvis, [&](int r, int c, bool from_state) {
      const long long gx = x0 + c, gy = y0 + r;
      if (gx < 0 || gx >= a.nx || gy < 0 || gy >= a.ny) return;
      const size_t g = (size_t)gy * a.nx + gx;
      su[G::at(r, c)] = from_state ? a.u[g] : a.dem[g];
    });
    __syncthreads();
    save_perimeter<G>(sper, su_bits);
    const unsigned long long c1 = wall_clock64();
    unsigned it = 0;
    int any, any_change = 0;
    do {
      bool ch = fsm_sweep<G, true, true>(su, sfi, R0, C0, lane);
      ch |= fsm_sweep<G, true, false>(su, sfi, R0, C0, lane);
      ch |= fsm_sweep<G, false, true>(su, sfi, R0, C0, lane);
      ch |= fsm_sweep<G, false, false>(su, sfi, R0, C0, lane);
      any = __syncthreads_or(ch ? 1 : 0);
      any_change |= any;
      it++;
    } while (any && it < a.max_local);
    const unsigned long long c2 = wall_clock64();
    // the first visit initialises the output; later visits write changed
    // tiles only
    if (first || any_change) {
      for (int idx = tid; idx < G::TY * G::TX; idx += G::THREADS) {
        const int r = idx / G::TX, c = idx % G::TX;
        const long long gx = x0 + c, gy = y0 + r;
        if (gx < a.nx && gy < a.ny) a.u[(size_t)gy * a.nx + gx] = su[G::at(r, c)];
      }
    }
    // NaN cells compare as words, so they never count as changed
    const unsigned bits = block_or(changed_perimeter<G, false>(sper, su_bits), &s_bits);
    if (tid == 0) visit_stats(a.q.sync, it, any != 0, c0, c1, c2);
    return bits;
  };

  run_queue(a.q, a.ntx, a.nty, visit);
}

template <class G>
static hipError_t launch_fsm(FsmArgs a, hipStream_t stream, RunStats* stats) {
  a.ntx = (int)((a.nx + G::TX - 1) / G::TX);
  a.nty = (int)((a.ny + G::TY - 1) / G::TY);
  a.max_local = 1u << 16;
  return launch_persistent(fsm_persistent<G>, G::THREADS, a.ntx * a.nty, a, stream, stats, G::TX,
                           G::TY);
}

hipError_t run_fsm(float* u, const float* dem, const float* slopes, float cellsize, int64_t nx,
                   int64_t ny, TileShape ts, hipStream_t stream, RunStats* stats) {
  if (nx <= 0 || ny <= 0) return hipSuccess;
  FsmArgs a{};
  a.u = u;
  a.dem = dem;
  a.slopes = slopes;
  a.cellsize = cellsize;
  a.nx = nx;
  a.ny = ny;
  if (ts.qx == 2 && ts.qy == 2) return launch_fsm<TileGeom<2, 2>>(a, stream, stats);
  if (ts.qx == 2 && ts.qy == 1) return launch_fsm<TileGeom<2, 1>>(a, stream, stats);
  if (ts.qx == 1 && ts.qy == 2) return launch_fsm<TileGeom<1, 2>>(a, stream, stats);
  return launch_fsm<TileGeom<1, 1>>(a, stream, stats);
}

}  // namespace tt
