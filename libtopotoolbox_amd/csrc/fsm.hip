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
  unsigned* mark;
  unsigned* list;
  Sync* sync;
  unsigned max_rounds;
  unsigned max_local;
};

__device__ __forceinline__ float godunov(float u1, float u2, float fi) {
  if (fabsf(u1 - u2) < fi) {
    return (u1 + u2) / 2 + sqrtf(-(u1 - u2) * (u1 - u2) + 2 * (fi * fi)) / 2;
  }
  return fminf(u1 + fi, u2 + fi);
}

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
#pragma unroll 2
    for (int k = 0; k < 64; k++, R += step) {
      const float cur = su[G::at(R, C)];
      const float far = su[G::at(R + step, C)];
      const float edge = su[G::at(R, ec)];
      const float fi = sfi[G::at(R, C)];
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
#pragma unroll 2
    for (int k = 0; k < 64; k++, C += step) {
      const float cur = su[G::at(R, C)];
      const float far = su[G::at(R, C + step)];
      const float edge = su[G::at(er, C)];
      const float fi = sfi[G::at(R, C)];
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
  const int ntiles = a.ntx * a.nty;
  const unsigned nb = gridDim.x;

  auto visit = [&](int t, unsigned round) {
    const int tx = t % a.ntx, ty = t / a.ntx;
    const long long x0 = (long long)tx * G::TX, y0 = (long long)ty * G::TY;
    for (int idx = tid; idx < G::ROWS * (G::TX + 2); idx += G::THREADS) {
      const int r = idx / (G::TX + 2) - 1;
      const int c = idx % (G::TX + 2) - 1;
      const long long gx = x0 + c, gy = y0 + r;
      float u = kInf, fi = __builtin_nanf("");
      if (gx >= 0 && gx < a.nx && gy >= 0 && gy < a.ny) {
        const size_t g = (size_t)gy * a.nx + gx;
        u = (round == 0) ? a.dem[g] : a.u[g];
        fi = a.cellsize * a.slopes[g];
      }
      su[G::at(r, c)] = u;
      sfi[G::at(r, c)] = fi;
    }
    if (tid == 0) s_bits = 0u;
    __syncthreads();
    for (int k = tid; k < G::TX; k += G::THREADS) {
      sper[k] = __float_as_uint(su[G::at(0, k)]);
      sper[G::TX + k] = __float_as_uint(su[G::at(G::TY - 1, k)]);
    }
    for (int k = tid; k < G::TY; k += G::THREADS) {
      sper[2 * G::TX + k] = __float_as_uint(su[G::at(k, 0)]);
      sper[2 * G::TX + G::TY + k] = __float_as_uint(su[G::at(k, G::TX - 1)]);
    }
    unsigned it = 0;
    int any;
    do {
      bool ch = fsm_sweep<G, true, true>(su, sfi, R0, C0, lane);
      ch |= fsm_sweep<G, true, false>(su, sfi, R0, C0, lane);
      ch |= fsm_sweep<G, false, true>(su, sfi, R0, C0, lane);
      ch |= fsm_sweep<G, false, false>(su, sfi, R0, C0, lane);
      any = __syncthreads_or(ch ? 1 : 0);
      it++;
    } while (any && it < a.max_local);
    for (int idx = tid; idx < G::TY * G::TX; idx += G::THREADS) {
      const int r = idx / G::TX, c = idx % G::TX;
      const long long gx = x0 + c, gy = y0 + r;
      if (gx < a.nx && gy < a.ny) a.u[(size_t)gy * a.nx + gx] = su[G::at(r, c)];
    }
    unsigned bits = 0;
    for (int k = tid; k < G::TX; k += G::THREADS) {
      if (sper[k] != __float_as_uint(su[G::at(0, k)])) bits |= kEdgeN;
      if (sper[G::TX + k] != __float_as_uint(su[G::at(G::TY - 1, k)])) bits |= kEdgeS;
    }
    for (int k = tid; k < G::TY; k += G::THREADS) {
      if (sper[2 * G::TX + k] != __float_as_uint(su[G::at(k, 0)])) bits |= kEdgeW;
      if (sper[2 * G::TX + G::TY + k] != __float_as_uint(su[G::at(k, G::TX - 1)]))
        bits |= kEdgeE;
    }
    if (bits) atomicOr(&s_bits, bits);
    __syncthreads();
    if (tid == 0) {
      queue_neighbours(s_bits, tx, ty, a.ntx, a.nty, a.mark, a.list, a.sync, round + 1);
      atomicAdd(&a.sync->visits, 1ull);
      atomicAdd(&a.sync->local_iters, (unsigned long long)it);
      atomicMax(&a.sync->max_local, it);
      if (any) atomicOr(&a.sync->error, kErrLocalCap);
    }
    __syncthreads();
  };

  for (unsigned round = 0;; round++) {
    if (round == 0) {
      for (int t = blockIdx.x; t < ntiles; t += nb) visit(t, 0u);
    } else {
      const unsigned n = ld_relaxed(&a.sync->count[round & 1u]);
      const unsigned* L = a.list + (size_t)(round & 1u) * ntiles;
      for (unsigned k = blockIdx.x; k < n; k += nb) visit((int)L[k], round);
    }
    const unsigned consumed = round & 1u;
    bool ok = grid_barrier(a.sync, nb, [&] {
      __hip_atomic_store(&a.sync->count[consumed], 0u, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    });
    if (!ok) return;
    const unsigned next = ld_relaxed(&a.sync->count[(round + 1) & 1u]);
    if (next == 0u || round + 1 >= a.max_rounds) {
      if (blockIdx.x == 0 && tid == 0) {
        a.sync->rounds = round + 1;
        if (next != 0u) atomicOr(&a.sync->error, kErrRoundCap);
      }
      return;
    }
  }
}

template <class G>
static hipError_t launch_fsm(float* u, const float* dem, const float* slopes, float cellsize,
                             int64_t nx, int64_t ny, hipStream_t stream, RunStats* stats) {
  FsmArgs a{};
  a.u = u;
  a.dem = dem;
  a.slopes = slopes;
  a.cellsize = cellsize;
  a.nx = nx;
  a.ny = ny;
  a.ntx = (int)((nx + G::TX - 1) / G::TX);
  a.nty = (int)((ny + G::TY - 1) / G::TY);
  a.max_rounds = 1u << 20;
  a.max_local = 1u << 16;
  const int ntiles = a.ntx * a.nty;
  int grid = 0;
  hipError_t e = persistent_grid((const void*)fsm_persistent<G>, G::THREADS, ntiles, &grid);
  if (e != hipSuccess) return e;
  Workspace ws;
  e = ws.alloc(ntiles, stream);
  if (e != hipSuccess) return e;
  a.mark = ws.mark;
  a.list = ws.list;
  a.sync = ws.sync;
  fsm_persistent<G><<<grid, G::THREADS, 0, stream>>>(a);
  e = hipGetLastError();
  if (stats) {
    stats->grid = grid;
    stats->tile_x = G::TX;
    stats->tile_y = G::TY;
  }
  hipError_t f = ws.finish(stats);
  return e != hipSuccess ? e : f;
}

hipError_t run_fsm(float* u, const float* dem, const float* slopes, float cellsize, int64_t nx,
                   int64_t ny, TileShape ts, hipStream_t stream, RunStats* stats) {
  if (nx <= 0 || ny <= 0) return hipSuccess;
  if (ts.qx == 2 && ts.qy == 2)
    return launch_fsm<TileGeom<2, 2>>(u, dem, slopes, cellsize, nx, ny, stream, stats);
  if (ts.qx == 2 && ts.qy == 1)
    return launch_fsm<TileGeom<2, 1>>(u, dem, slopes, cellsize, nx, ny, stream, stats);
  if (ts.qx == 1 && ts.qy == 2)
    return launch_fsm<TileGeom<1, 2>>(u, dem, slopes, cellsize, nx, ny, stream, stats);
  return launch_fsm<TileGeom<1, 1>>(u, dem, slopes, cellsize, nx, ny, stream, stats);
}

}  // namespace tt
