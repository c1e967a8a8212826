// Persistent tiled relaxation engine shared by the sink-filling and the
// excess-topography kernels (gfx950, wave64).
//
// A grid of nx x ny cells (x = dims[0] fastest) is cut into tiles of
// TX x TY cells. One workgroup owns one tile at a time: it stages the tile
// plus a one-cell halo in LDS, relaxes the tile to its local fixed point
// with directional sweeps (one 64x64 quadrant per wavefront), writes the
// tile back and queues the neighbour tiles whose halo it changed.
//
// The grid is persistent: every workgroup loops over rounds. Round 0 visits
// every tile; round r > 0 visits the tiles queued during round r-1. Rounds
// are separated by a software grid barrier. The kernel ends when a round
// queues nothing. All workgroups must be co-resident; the host sizes the
// grid from the occupancy query, and every wait is bounded by a wall-clock
// limit that turns a stuck barrier into an error flag instead of a hang.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tt {

constexpr float kInf = __builtin_inff();

// Control block in device memory, zeroed before each launch.
struct Sync {
  unsigned int bar_count;   // arrivals at the current barrier
  unsigned int bar_gen;     // barrier generation
  unsigned int error;       // kErr* bits
  unsigned int rounds;      // rounds executed (written by workgroup 0)
  unsigned int count[2];    // queued tiles for the next round, by parity
  unsigned int max_local;   // largest local iteration count of any visit
  unsigned int pad;
  unsigned long long visits;       // tile visits
  unsigned long long local_iters;  // sum of local iterations over visits
  // wall-clock ticks (100 MHz) summed over workgroups, by phase
  unsigned long long t_stage;      // tile + halo staging into LDS
  unsigned long long t_relax;      // local sweeps
  unsigned long long t_write;      // write-back and neighbour queueing
  unsigned long long t_wait;       // grid barrier waits
};

enum : unsigned int {
  kErrBarrierTimeout = 1u,
  kErrRoundCap = 2u,
  kErrLocalCap = 4u,
};

// Upper bound on a single wait at a grid barrier (wall clock ticks of the
// 100 MHz constant clock): 10 s.
constexpr unsigned long long kBarrierTimeoutTicks = 1000000000ull;

__device__ inline unsigned ld_acquire(const unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned ld_relaxed(const unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Software grid barrier. Returns false (for every thread of the workgroup)
// when the wait timed out or another workgroup reported an error. The last
// workgroup to arrive runs `on_last` before releasing the others.
template <class F>
__device__ inline bool grid_barrier(Sync* s, unsigned nblocks, F on_last) {
  __shared__ int ok;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) {
    ok = 1;
    unsigned gen = ld_acquire(&s->bar_gen);
    unsigned arrived = __hip_atomic_fetch_add(&s->bar_count, 1u, __ATOMIC_ACQ_REL,
                                              __HIP_MEMORY_SCOPE_AGENT);
    if (arrived == nblocks - 1) {
      on_last();
      __hip_atomic_store(&s->bar_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&s->bar_gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      unsigned long long t0 = wall_clock64();
      while (ld_acquire(&s->bar_gen) == gen) {
        __builtin_amdgcn_s_sleep(1);
        if (ld_relaxed(&s->error) != 0u) {
          ok = 0;
          break;
        }
        if (wall_clock64() - t0 > kBarrierTimeoutTicks) {
          __hip_atomic_fetch_or(&s->error, kErrBarrierTimeout, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
      }
    }
    if (ld_relaxed(&s->error) != 0u) ok = 0;
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return ok != 0;
}

// Queue tile `t` for round `next` (deduplicated through mark[t]).
__device__ inline void queue_tile(unsigned* mark, unsigned* list, Sync* s, int ntiles,
                                  int t, unsigned next) {
  unsigned old = atomicMax(&mark[t], next);
  if (old < next) {
    unsigned pos = atomicAdd(&s->count[next & 1u], 1u);
    list[(size_t)(next & 1u) * ntiles + pos] = (unsigned)t;
  }
}

// Cross-lane shifts within a wavefront (DPP wave_shr:1 / wave_shl:1).
// Lane 0 (shr) or lane 63 (shl) receives `edge` instead.
__device__ inline float lane_from_below(float v, float edge) {  // lane i <- lane i-1
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v),
                                                    0x138, 0xf, 0xf, false));
}
__device__ inline float lane_from_above(float v, float edge) {  // lane i <- lane i+1
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v),
                                                    0x130, 0xf, 0xf, false));
}

// Perimeter bits of a tile: which neighbour tiles see a changed cell.
enum : unsigned {
  kEdgeN = 1u, kEdgeS = 2u, kEdgeW = 4u, kEdgeE = 8u,
  kEdgeNW = 16u, kEdgeNE = 32u, kEdgeSW = 64u, kEdgeSE = 128u,
};

// Tile geometry: QX x QY quadrants of 64x64 cells, one wavefront each.
// "row" = y = dims[1] index, "col" = x = dims[0] index.
template <int QX_, int QY_>
struct TileGeom {
  static constexpr int QX = QX_, QY = QY_;
  static constexpr int TX = 64 * QX, TY = 64 * QY;
  static constexpr int P = TX + 3;  // LDS pitch: odd, so columns hit distinct banks
  static constexpr int ROWS = TY + 2;
  static constexpr int CELLS = ROWS * P;
  static constexpr int THREADS = 64 * QX * QY;
  // LDS index of tile-local cell (r, c), r in [-1, TY], c in [-1, TX]
  __device__ static constexpr int at(int r, int c) { return (r + 1) * P + (c + 1); }
};

// Bits of neighbour tiles affected when tile-local cell (r, c) changes.
// `diag` adds the corner tiles (8-connected stencils).
template <class G, bool diag>
__device__ inline unsigned edge_bits(int r, int c) {
  unsigned b = 0;
  b |= (r == 0) ? kEdgeN : 0u;
  b |= (r == G::TY - 1) ? kEdgeS : 0u;
  b |= (c == 0) ? kEdgeW : 0u;
  b |= (c == G::TX - 1) ? kEdgeE : 0u;
  if (diag) {
    b |= (r == 0 && c == 0) ? kEdgeNW : 0u;
    b |= (r == 0 && c == G::TX - 1) ? kEdgeNE : 0u;
    b |= (r == G::TY - 1 && c == 0) ? kEdgeSW : 0u;
    b |= (r == G::TY - 1 && c == G::TX - 1) ? kEdgeSE : 0u;
  }
  return b;
}

// Queue the neighbours of tile (tx, ty) named by `bits`.
__device__ inline void queue_neighbours(unsigned bits, int tx, int ty, int ntx, int nty,
                                        unsigned* mark, unsigned* list, Sync* s,
                                        unsigned next) {
  const int ntiles = ntx * nty;
  const int dx[8] = {0, 0, -1, 1, -1, 1, -1, 1};
  const int dy[8] = {-1, 1, 0, 0, -1, -1, 1, 1};
  for (int k = 0; k < 8; k++) {
    if (!(bits & (1u << k))) continue;
    int x = tx + dx[k], y = ty + dy[k];
    if (x < 0 || x >= ntx || y < 0 || y >= nty) continue;
    queue_tile(mark, list, s, ntiles, y * ntx + x, next);
  }
}

// The persistent round loop. `visit(t, round)` processes tile t (all
// threads of the workgroup call it together) and queues the neighbours it
// changed for round + 1. Round 0 visits every tile; round r > 0 the tiles
// queued during round r - 1. Returns when a round queues nothing.
template <class Visit>
__device__ inline void run_rounds(Sync* s, const unsigned* list, int ntiles,
                                  unsigned max_rounds, Visit visit) {
  const unsigned nb = gridDim.x;
  for (unsigned round = 0;; round++) {
    if (round == 0) {
      for (int t = blockIdx.x; t < ntiles; t += nb) visit(t, 0u);
    } else {
      const unsigned n = ld_relaxed(&s->count[round & 1u]);
      const unsigned* L = list + (size_t)(round & 1u) * ntiles;
      for (unsigned k = blockIdx.x; k < n; k += nb) visit((int)L[k], round);
    }
    const unsigned consumed = round & 1u;
    const unsigned long long w0 = wall_clock64();
    bool ok = grid_barrier(s, nb, [&] {
      __hip_atomic_store(&s->count[consumed], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    });
    if (!ok) return;
    if (threadIdx.x == 0) atomicAdd(&s->t_wait, wall_clock64() - w0);
    const unsigned next = ld_relaxed(&s->count[(round + 1) & 1u]);
    if (next == 0u || round + 1 >= max_rounds) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        s->rounds = round + 1;
        if (next != 0u) atomicOr(&s->error, kErrRoundCap);
      }
      return;
    }
  }
}

// Per-visit bookkeeping by thread 0: queue neighbours, counters, phase
// timers (c0..c3: start, staged, relaxed, written).
__device__ inline void finish_visit(Sync* s, unsigned* mark, unsigned* list, unsigned bits,
                                    int tx, int ty, int ntx, int nty, unsigned round,
                                    unsigned it, bool capped, unsigned long long c0,
                                    unsigned long long c1, unsigned long long c2) {
  queue_neighbours(bits, tx, ty, ntx, nty, mark, list, s, round + 1);
  atomicAdd(&s->visits, 1ull);
  atomicAdd(&s->local_iters, (unsigned long long)it);
  atomicMax(&s->max_local, it);
  if (capped) atomicOr(&s->error, kErrLocalCap);
  const unsigned long long c3 = wall_clock64();
  atomicAdd(&s->t_stage, c1 - c0);
  atomicAdd(&s->t_relax, c2 - c1);
  atomicAdd(&s->t_write, c3 - c2);
}

}  // namespace tt
