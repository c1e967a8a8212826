// Sink filling and grayscale reconstruction on gfx950.
//
// Reference semantics (src/fillsinks.c:22-48, src/reconstruct.c:68-112,
// 134-202, 270-279): fillsinks negates the DEM, seeds the marker with the
// negated DEM on bc==1 cells and -INF elsewhere, reconstructs by dilation
// (marker = min(mask, max(marker, 8 neighbours)) until nothing changes) and
// negates back. That fixed point is unique and every value is a copy of an
// input value, so any update order reaches the same bits.
//
// This kernel works directly in the "fill domain" f = -marker, d = -mask:
//     f_p <- max(d_p, min(f_p, f_q for the 8 neighbours q))
// iterated from f0 = d on seeds (+INF elsewhere) down to the greatest fixed
// point. NaN cells of d are inert: they hold +INF (so fminf ignores them as
// fmaxf ignores NaN in the reference) and are written back as the input NaN.
//
// fillsinks: d = dem, f0 = (bc == 1 && dem is not NaN) ? dem : +INF; the
//            output array holds f between rounds.
// reconstruct: d = -mask, f0 = max(d, -marker) (a NaN marker acts as -INF);
//            the marker array holds -f between rounds and at the end.
#include "kernels.h"
#include "tile_engine.h"
#include "workspace.h"

namespace tt {

struct FillArgs {
  float* state;
  const float* src;
  const uint8_t* bc;
  long long nx, ny;
  int ntx, nty;
  unsigned* mark;
  unsigned* list;
  Sync* sync;
  unsigned max_rounds;
  unsigned max_local;
};

// One directional sweep over a 64x64 quadrant with lanes along columns
// (vertical sweep: down = +y) or along rows (horizontal: right = +x).
// The value carried between steps is the lane's previous cell; its two
// side neighbours come from the adjacent lanes by DPP, with the quadrant's
// outside column/row read from LDS for the edge lanes.
template <class G, bool VERT, bool FWD>
__device__ __forceinline__ bool fill_sweep(float* sf, const float* sd, int R0, int C0, int lane) {
  constexpr int step = FWD ? 1 : -1;
  bool changed = false;
  if (VERT) {
    const int C = C0 + lane;
    const int ec = (lane == 0) ? C - 1 : C + 1;
    int R = FWD ? R0 : R0 + 63;
    float pc = sf[G::at(R - step, C)];
#pragma unroll 4
    for (int k = 0; k < 64; k++, R += step) {
      float edge = sf[G::at(R - step, ec)];
      float fc = sf[G::at(R, C)];
      float dv = sd[G::at(R, C)];
      float pl = lane_from_below(pc, edge);
      float pr = lane_from_above(pc, edge);
      float m = fminf(fminf(pl, pr), fminf(pc, fc));
      float nv = fmaxf(dv, m);
      bool upd = (nv < fc) && (dv == dv);
      if (upd) sf[G::at(R, C)] = nv;
      changed |= upd;
      pc = upd ? nv : fc;
    }
  } else {
    const int R = R0 + lane;
    const int er = (lane == 0) ? R - 1 : R + 1;
    int C = FWD ? C0 : C0 + 63;
    float pc = sf[G::at(R, C - step)];
#pragma unroll 4
    for (int k = 0; k < 64; k++, C += step) {
      float edge = sf[G::at(er, C - step)];
      float fc = sf[G::at(R, C)];
      float dv = sd[G::at(R, C)];
      float pu = lane_from_below(pc, edge);
      float pd = lane_from_above(pc, edge);
      float m = fminf(fminf(pu, pd), fminf(pc, fc));
      float nv = fmaxf(dv, m);
      bool upd = (nv < fc) && (dv == dv);
      if (upd) sf[G::at(R, C)] = nv;
      changed |= upd;
      pc = upd ? nv : fc;
    }
  }
  return changed;
}

template <class G, int MODE>
__global__ __launch_bounds__(G::THREADS) void fill_persistent(FillArgs a) {
  __shared__ float sf[G::CELLS];
  __shared__ float sd[G::CELLS];
  __shared__ float sper[2 * (G::TX + G::TY)];
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
    // ---- stage tile + halo
    for (int idx = tid; idx < G::ROWS * (G::TX + 2); idx += G::THREADS) {
      const int r = idx / (G::TX + 2) - 1;
      const int c = idx % (G::TX + 2) - 1;
      const long long gx = x0 + c, gy = y0 + r;
      float f = kInf, d = __builtin_nanf("");
      if (gx >= 0 && gx < a.nx && gy >= 0 && gy < a.ny) {
        const size_t g = (size_t)gy * a.nx + gx;
        const float raw = a.src[g];
        if (MODE == kModeFill) {
          d = raw;
          if (round == 0)
            f = (a.bc[g] == 1) ? raw : kInf;
          else
            f = a.state[g];
        } else {
          d = -raw;
          const float m = a.state[g];
          f = fmaxf(d, (m == m) ? -m : kInf);
        }
        if (d != d) f = kInf;
      }
      sf[G::at(r, c)] = f;
      sd[G::at(r, c)] = d;
    }
    if (tid == 0) s_bits = 0u;
    __syncthreads();
    for (int k = tid; k < G::TX; k += G::THREADS) {
      sper[k] = sf[G::at(0, k)];
      sper[G::TX + k] = sf[G::at(G::TY - 1, k)];
    }
    for (int k = tid; k < G::TY; k += G::THREADS) {
      sper[2 * G::TX + k] = sf[G::at(k, 0)];
      sper[2 * G::TX + G::TY + k] = sf[G::at(k, G::TX - 1)];
    }
    // ---- relax to the local fixed point
    unsigned it = 0;
    int any;
    do {
      bool ch = fill_sweep<G, true, true>(sf, sd, R0, C0, lane);
      ch |= fill_sweep<G, true, false>(sf, sd, R0, C0, lane);
      ch |= fill_sweep<G, false, true>(sf, sd, R0, C0, lane);
      ch |= fill_sweep<G, false, false>(sf, sd, R0, C0, lane);
      any = __syncthreads_or(ch ? 1 : 0);
      it++;
    } while (any && it < a.max_local);
    // ---- write back
    for (int idx = tid; idx < G::TY * G::TX; idx += G::THREADS) {
      const int r = idx / G::TX, c = idx % G::TX;
      const long long gx = x0 + c, gy = y0 + r;
      if (gx < a.nx && gy < a.ny) {
        const float d = sd[G::at(r, c)];
        float v = (d != d) ? d : sf[G::at(r, c)];
        if (MODE == kModeRecon) v = -v;
        a.state[(size_t)gy * a.nx + gx] = v;
      }
    }
    // ---- changed perimeter -> neighbour tiles for the next round
    unsigned bits = 0;
    for (int k = tid; k < G::TX; k += G::THREADS) {
      if (sper[k] != sf[G::at(0, k)])
        bits |= kEdgeN | (k == 0 ? kEdgeNW : 0u) | (k == G::TX - 1 ? kEdgeNE : 0u);
      if (sper[G::TX + k] != sf[G::at(G::TY - 1, k)])
        bits |= kEdgeS | (k == 0 ? kEdgeSW : 0u) | (k == G::TX - 1 ? kEdgeSE : 0u);
    }
    for (int k = tid; k < G::TY; k += G::THREADS) {
      if (sper[2 * G::TX + k] != sf[G::at(k, 0)])
        bits |= kEdgeW | (k == 0 ? kEdgeNW : 0u) | (k == G::TY - 1 ? kEdgeSW : 0u);
      if (sper[2 * G::TX + G::TY + k] != sf[G::at(k, G::TX - 1)])
        bits |= kEdgeE | (k == 0 ? kEdgeNE : 0u) | (k == G::TY - 1 ? kEdgeSE : 0u);
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
static hipError_t launch_fill(float* state, const float* src, const uint8_t* bc, int64_t nx,
                              int64_t ny, int mode, hipStream_t stream, RunStats* stats) {
  FillArgs a{};
  a.state = state;
  a.src = src;
  a.bc = bc;
  a.nx = nx;
  a.ny = ny;
  a.ntx = (int)((nx + G::TX - 1) / G::TX);
  a.nty = (int)((ny + G::TY - 1) / G::TY);
  a.max_rounds = 1u << 20;
  a.max_local = 1u << 16;
  const int ntiles = a.ntx * a.nty;
  const void* kern = mode == kModeFill ? (const void*)fill_persistent<G, kModeFill>
                                       : (const void*)fill_persistent<G, kModeRecon>;
  int grid = 0;
  hipError_t e = persistent_grid(kern, G::THREADS, ntiles, &grid);
  if (e != hipSuccess) return e;
  Workspace ws;
  e = ws.alloc(ntiles, stream);
  if (e != hipSuccess) return e;
  a.mark = ws.mark;
  a.list = ws.list;
  a.sync = ws.sync;
  if (mode == kModeFill)
    fill_persistent<G, kModeFill><<<grid, G::THREADS, 0, stream>>>(a);
  else
    fill_persistent<G, kModeRecon><<<grid, G::THREADS, 0, stream>>>(a);
  e = hipGetLastError();
  if (stats) {
    stats->grid = grid;
    stats->tile_x = G::TX;
    stats->tile_y = G::TY;
  }
  hipError_t f = ws.finish(stats);
  return e != hipSuccess ? e : f;
}

hipError_t run_fill(float* state, const float* src, const uint8_t* bc, int64_t nx, int64_t ny,
                    int mode, TileShape ts, hipStream_t stream, RunStats* stats) {
  if (nx <= 0 || ny <= 0) return hipSuccess;
  if (ts.qx == 2 && ts.qy == 2)
    return launch_fill<TileGeom<2, 2>>(state, src, bc, nx, ny, mode, stream, stats);
  if (ts.qx == 2 && ts.qy == 1)
    return launch_fill<TileGeom<2, 1>>(state, src, bc, nx, ny, mode, stream, stats);
  if (ts.qx == 1 && ts.qy == 2)
    return launch_fill<TileGeom<1, 2>>(state, src, bc, nx, ny, mode, stream, stats);
  return launch_fill<TileGeom<1, 1>>(state, src, bc, nx, ny, mode, stream, stats);
}

}  // namespace tt
