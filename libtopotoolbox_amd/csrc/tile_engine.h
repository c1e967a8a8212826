// Persistent tiled relaxation engine shared by the sink-filling, the
// excess-topography and the adjoint kernels (gfx950, wave64).
//
// A grid of nx x ny cells (x = dims[0] fastest) is cut into tiles of
// TX x TY cells. One workgroup owns one tile at a time: it stages the tile
// plus a one-cell halo in LDS, relaxes the tile to its local fixed point
// with directional sweeps (one 64x64 quadrant per wavefront), writes the
// tile back and queues the neighbour tiles whose halo it changed.
//
// Scheduling is a persistent grid over a lock-free FIFO of tile ids in
// device memory. Each tile has a state word (QUEUED | RUNNING bits): every
// tile starts queued; a workgroup pops a tile and marks it RUNNING (it is
// then the tile's only owner), relaxes it, and pushes each neighbour whose
// halo it changed. A push onto an idle tile enqueues it; a push onto a
// running tile only sets QUEUED, and the owner, seeing it when it finishes,
// appends the tile to the queue. A counter of queued-or-running tiles
// reaches zero exactly when the grid has converged, and every workgroup
// then exits. There are no rounds and no grid barriers.
//
// All workgroups must be co-resident (the host sizes the grid from the
// occupancy query). Every wait is bounded by a wall-clock limit that turns
// a stuck wait into an error flag instead of a hang.
//
// Races: a tile's halo is read while neighbours may be writing their
// tiles. The relaxations are monotone and every value a read can return is
// a valid bound, and a neighbour that changes cells we read queues us
// again, so the fixed point is reached regardless (state words must be read
// and written whole: one aligned 32- or 64-bit word per cell).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tt {

constexpr float kInf = __builtin_inff();

// Control block in device memory, initialised before each launch.
struct Sync {
  unsigned int head;        // pop tickets handed out
  unsigned int tail;        // push positions handed out
  unsigned int pending;     // tiles queued or being relaxed
  unsigned int error;       // kErr* bits
  unsigned int max_local;   // largest local iteration count of a visit
  unsigned int pops;        // tiles popped (counted only under a visit cap)
  unsigned int pad[2];
  unsigned long long visits;       // tile visits
  unsigned long long local_iters;  // sum of local iterations over visits
  // wall-clock ticks (100 MHz) summed over workgroups, by phase
  unsigned long long t_stage;      // tile + halo staging into LDS
  unsigned long long t_relax;      // local sweeps
  unsigned long long t_write;      // write-back and neighbour queueing
  unsigned long long t_wait;       // waiting for a tile to pop
};

enum : unsigned int {
  kErrTimeout = 1u,
  // not an error: the launch reached its visit cap and stopped early (set
  // in the control block only, never in the sticky word)
  kStopVisitCap = 0x100u,
};

// Default upper bound on one wait (wall-clock ticks of the 100 MHz
// constant clock): 10 s. A launch carries its own bound (TileQueue).
constexpr unsigned long long kWaitTimeoutTicks = 1000000000ull;

__device__ inline unsigned ld_acquire(const unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned ld_relaxed(const unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Work-queue state of one launch (device memory).
struct TileQueue {
  Sync* sync;
  unsigned long long* ring;  // ntiles slots: position << 32 | (tile id + 1); 0 = empty
  unsigned* state;    // per tile: kQueued | kRunning bits
  unsigned* visited;  // per tile: 1 once its first visit has been written
  int ntiles;
  // 1: count visits and time their phases in *sync (atomics each visit;
  // only when the caller reads them back)
  int stats;
  // the Sync block visits report their statistics to, or null
  __device__ Sync* stats_sink() const { return stats ? sync : nullptr; }
  // per-GPU word that keeps the error bits of every launch until the host
  // reads and clears it (may be null)
  unsigned* sticky;
  // upper bound on one wait, in wall-clock ticks
  unsigned long long wait_limit;
  // 0, or the number of tile visits after which the launch stops taking
  // new work (a partial result); see run_queue
  unsigned max_visits;
};

// Records error bits in the launch's control block (which stops the grid)
// and in the per-GPU sticky word.
__device__ inline void raise_error(const TileQueue& q, unsigned bits) {
  atomicOr(&q.sync->error, bits);
  if (q.sticky) atomicOr(q.sticky, bits);
}

enum : unsigned { kQueued = 1u, kRunning = 2u };

// Neighbour tiles (bit k of a queue mask); kSelf queues the visited tile
// itself again (a visit that stopped at its sweep cap before converging).
enum : unsigned {
  kEdgeN = 1u, kEdgeS = 2u, kEdgeW = 4u, kEdgeE = 8u,
  kEdgeNW = 16u, kEdgeNE = 32u, kEdgeSW = 64u, kEdgeSE = 128u,
  kSelf = 256u,
};

// Spinning uses relaxed loads only: an acquire load at agent scope
// invalidates the XCD's L2, which hundreds of spinning workgroups would do
// continuously. Ordering comes from one fence per visit in run_queue.

// Push tile t (called by one thread, after a release fence that published
// the caller's writes). The fetch_or always writes, so the tile's owner or
// next popper, whose own read-modify-write of the state word follows it,
// synchronises with it through the fences.
__device__ inline void enqueue_tile(const TileQueue& q, int t);

// `on_new(t)` runs when the push turns an idle tile into a queued one.
template <class OnNew>
__device__ inline void push_tile(const TileQueue& q, int t, OnNew on_new) {
  const unsigned old =
      __hip_atomic_fetch_or(&q.state[t], kQueued, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (old != 0u) return;  // already queued, or running (its owner requeues it)
  __hip_atomic_fetch_add(&q.sync->pending, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  on_new(t);
  enqueue_tile(q, t);
}
__device__ inline void push_tile(const TileQueue& q, int t) {
  push_tile(q, t, [](int) {});
}

// Write tile t into the ring at push position pos (reserved from tail).
__device__ inline void fill_slot(const TileQueue& q, unsigned pos, int t) {
  unsigned long long* slot = &q.ring[pos % (unsigned)q.ntiles];
  // the slot's previous entry (pos - ntiles) has a ticket holder, since at
  // most ntiles tiles are in the ring; wait until it has emptied the slot
  // once an error is set the poppers stop taking tickets, so the slot may
  // never empty: the push is dropped (the grid is ending anyway)
  const unsigned long long t0 = wall_clock64();
  while (__hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0ull) {
    __builtin_amdgcn_s_sleep(1);
    if (ld_relaxed(&q.sync->error) != 0u) return;
    if (wall_clock64() - t0 > q.wait_limit) {
      raise_error(q, kErrTimeout);
      return;
    }
  }
  __hip_atomic_store(slot, ((unsigned long long)pos << 32) | ((unsigned)t + 1u),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Append tile t to the ring (its state already says queued).
__device__ inline void enqueue_tile(const TileQueue& q, int t) {
  fill_slot(q, __hip_atomic_fetch_add(&q.sync->tail, 1u, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT), t);
}

// Pop a tile (called by one thread): take a ticket, then wait on that
// ticket's own ring slot until the push with the same position fills it.
// Returns -1 when the grid has converged (nothing queued, nothing
// running, so no push can come) or after an error. The caller issues an
// acquire fence before reading the tile's inputs.
__device__ inline int pop_tile(const TileQueue& q) {
  Sync* s = q.sync;
  const unsigned h =
      __hip_atomic_fetch_add(&s->head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  unsigned long long* slot = &q.ring[h % (unsigned)q.ntiles];
  const unsigned long long t0 = wall_clock64();
  for (unsigned spin = 0;; spin++) {
    const unsigned long long v =
        __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v != 0ull && (unsigned)(v >> 32) == h) {
      __hip_atomic_store(slot, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return (int)(unsigned)v - 1;
    }
    if ((spin & 7u) == 7u) {
      if (ld_relaxed(&s->pending) == 0u || ld_relaxed(&s->error) != 0u) return -1;
      if (wall_clock64() - t0 > q.wait_limit) {
        raise_error(q, kErrTimeout);
        return -1;
      }
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

// on_release result meaning "nothing to expand".
constexpr unsigned long long kNoExpand = ~0ull;

// Neighbour offsets, indexed like the kEdge* bits below.
__device__ constexpr int kNbrDx[8] = {0, 0, -1, 1, -1, 1, -1, 1};
__device__ constexpr int kNbrDy[8] = {-1, 1, 0, 0, -1, -1, 1, 1};

// Tile topology of one grid of ntx x nty tiles: neighbour k of tile t
// (kEdge* order), -1 outside; no per-tile bookkeeping.
struct GridTopo {
  int ntx, nty;
  __device__ int neighbour(int t, int k) const {
    const int x = t % ntx + kNbrDx[k], y = t / ntx + kNbrDy[k];
    return (x < 0 || x >= ntx || y < 0 || y >= nty) ? -1 : y * ntx + x;
  }
  __device__ void on_new(int) const {}
  __device__ unsigned long long on_release(const TileQueue&, int) const { return kNoExpand; }
  __device__ void expand(const TileQueue&, unsigned long long) const {}
};

// The persistent loop. visit(t, first, nvis) relaxes tile t (all threads
// of the workgroup call it together). `first` is true on the tile's first
// visit, when its stored state is not initialised yet; bit k of `nvis` is
// set when neighbour k (kEdge* order) has been visited, i.e. its stored
// state may be read, and is meaningful only for neighbours that exist.
// Returns the bits of the neighbours to queue (same value in every thread).
//
// The flags are read after the tile's state word is taken (a
// read-modify-write): a neighbour that published its first visit before
// that point has pushed us before it too, and we synchronise with that
// push; one that publishes later pushes us while we run, and the tile is
// queued again when we finish.
//
// Topo supplies neighbour(t, k), on_new(t) (a push made tile t queued),
// on_release(q, t) (thread 0: tile t went idle; runs before the global
// pending count drops, so work it counts in keeps the grid alive) and
// expand(q, token): when on_release returns a token other than kNoExpand,
// every thread of the workgroup calls expand with it (bulk queueing of
// tiles that on_release has counted and reserved ring positions for).
template <class Topo, class Visit>
__device__ inline void run_queue(const TileQueue& q, const Topo& topo, Visit visit) {
  __shared__ int s_tile;
  __shared__ int s_first;
  __shared__ unsigned s_nvis;
  __shared__ unsigned long long s_expand;
  auto on_new = [&](int n) { topo.on_new(n); };
  for (;;) {
    if (threadIdx.x == 0) {
      const unsigned long long w0 = wall_clock64();
      const int t = pop_tile(q);
      if (t >= 0)
        __hip_atomic_exchange(&q.state[t], kRunning, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      // Visit cap: the pop that reaches it sets kStopVisitCap, which, like
      // an error, makes waiting poppers return -1 and drops pushes that
      // would wait for a ring slot. A popper whose slot is already filled
      // still takes and visits its tile, so every tile of the initial queue
      // is visited once and the grid drains after a bounded number of
      // further visits. Each visit writes a complete relaxation step, so
      // the stored state is a valid partial result.
      if (t >= 0 && q.max_visits &&
          __hip_atomic_fetch_add(&q.sync->pops, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) +
                  1u >= q.max_visits)
        atomicOr(&q.sync->error, kStopVisitCap);
      s_tile = t;
      if (t >= 0) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_first = ld_relaxed(&q.visited[t]) == 0u;
        unsigned nvis = 0;
        for (int k = 0; k < 8; k++) {
          const int n = topo.neighbour(t, k);
          if (n >= 0 && ld_relaxed(&q.visited[n])) nvis |= 1u << k;
        }
        s_nvis = nvis;
      }
      if (q.stats) atomicAdd(&q.sync->t_wait, wall_clock64() - w0);
    }
    __syncthreads();
    const int tt = s_tile;
    const bool first = s_first != 0;
    const unsigned nvis = s_nvis;
    if (tt < 0) return;
    // thread 0's agent-scope acquire after the pop, ordered before every
    // wave's loads by the barrier, covers the workgroup: the caches it
    // invalidates (the CU's L1, the XCD's L2) are the ones all its waves
    // read through
    const unsigned bits = visit(tt, first, nvis);
    // every wave's stores complete (in L2) at the barrier; one agent-scope
    // release by thread 0 then writes back the L2 that holds them all,
    // before anyone can pop a tile we queue
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    if (threadIdx.x == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      s_expand = kNoExpand;
      if (first) {
        __hip_atomic_store(&q.visited[tt], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // the visited flag is ordered before the state words we touch next
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      }
      for (int k = 0; k < 8; k++) {
        if (!(bits & (1u << k))) continue;
        const int n = topo.neighbour(tt, k);
        if (n >= 0) push_tile(q, n, on_new);
      }
      if (bits & kSelf) push_tile(q, tt, on_new);  // running: requeued when released
      // release the tile; if a neighbour queued it while it ran, append it
      // to the queue (it stays outstanding: pending is unchanged), so that
      // further neighbour updates coalesce into its next visit
      for (;;) {
        unsigned cur = ld_relaxed(&q.state[tt]);
        if (cur & kQueued) {
          __hip_atomic_exchange(&q.state[tt], kQueued, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
          enqueue_tile(q, tt);
          break;
        }
        if (__hip_atomic_compare_exchange_strong(&q.state[tt], &cur, 0u, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          s_expand = topo.on_release(q, tt);
          __hip_atomic_fetch_sub(&q.sync->pending, 1u, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
      }
    }
    __syncthreads();
    const unsigned long long ex = s_expand;
    if (ex != kNoExpand) topo.expand(q, ex);
  }
}

// Per-visit counters and phase timers (thread 0; c0..c2: start, staged,
// relaxed).
__device__ inline void visit_stats(Sync* s, unsigned it, unsigned long long c0,
                                   unsigned long long c1, unsigned long long c2) {
  if (!s) return;  // statistics not requested
  atomicAdd(&s->visits, 1ull);
  atomicAdd(&s->local_iters, (unsigned long long)it);
  atomicMax(&s->max_local, it);
  const unsigned long long c3 = wall_clock64();
  atomicAdd(&s->t_stage, c1 - c0);
  atomicAdd(&s->t_relax, c2 - c1);
  atomicAdd(&s->t_write, c3 - c2);
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

// ---------------------------------------------------------------- staging
// Copies the tile + halo of one or two global arrays into LDS (float
// layout of TileGeom). Cells outside the grid get `oob`. Full-width tiles
// of grids with nx % 4 == 0 and aligned base pointers use 128-bit loads
// (uint8: 32-bit), in two batches per array. Values are loaded raw (V, S)
// and converted (to4, cv) only at the LDS stores: a conversion next to its
// load, inside the load's bounds branch, would make every load wait there
// and serialise the batch.
struct StageF32 {
  using V = float4;
  using S = float;
  const float* p;
  float* lds;
  float oob;
  __device__ V load4(size_t g) const { return *reinterpret_cast<const float4*>(p + g); }
  __device__ S load1(size_t g) const { return p[g]; }
  __device__ float4 to4(V v) const { return v; }
  __device__ float cv(S v) const { return v; }
  __device__ bool aligned() const { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }
};
// float cells staged multiplied by `scale` (one rounding, as a separate
// multiply in LDS would give)
struct StageF32Scaled {
  using V = float4;
  using S = float;
  const float* p;
  float* lds;
  float oob;
  float scale;
  __device__ V load4(size_t g) const { return *reinterpret_cast<const float4*>(p + g); }
  __device__ S load1(size_t g) const { return p[g]; }
  __device__ float4 to4(V v) const {
    return make_float4(scale * v.x, scale * v.y, scale * v.z, scale * v.w);
  }
  __device__ float cv(S v) const { return scale * v; }
  __device__ bool aligned() const { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }
};
// uint8 cells staged as floats (bc flags)
struct StageU8 {
  using V = uchar4;
  using S = uint8_t;
  const uint8_t* p;
  float* lds;
  float oob;
  __device__ V load4(size_t g) const { return *reinterpret_cast<const uchar4*>(p + g); }
  __device__ S load1(size_t g) const { return p[g]; }
  __device__ float4 to4(V v) const { return make_float4(v.x, v.y, v.z, v.w); }
  __device__ float cv(S v) const { return v; }
  __device__ bool aligned() const { return (reinterpret_cast<uintptr_t>(p) & 3) == 0; }
};

template <class G, class A, class B>
__device__ inline void stage_tile(const A& a, const B& b, bool two, long long x0, long long y0,
                                  long long nx, long long ny) {
  // opaque per call: keeps the compiler from computing every batch's
  // addresses once, outside a persistent loop, and holding them
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const bool vec = ((nx & 3) == 0) && (x0 + G::TX <= nx) && a.aligned() && (!two || b.aligned());
  auto put = [&](int at, float fa, float fb) {
    a.lds[at] = fa;
    if (two) b.lds[at] = fb;
  };
  if (vec) {
    // two batches of loads
    constexpr int Q = G::TX / 4, NV = G::ROWS * Q;
    constexpr int UV = (NV + 2 * G::THREADS - 1) / (2 * G::THREADS);
    for (int base = 0; base < NV; base += G::THREADS * UV) {
      typename A::V va[UV];
      typename B::V vb[UV];
      bool ok[UV];
#pragma unroll
      for (int u = 0; u < UV; u++) {
        const int idx = base + u * G::THREADS + tid;
        const long long gy = y0 + idx / Q - 1;
        ok[u] = idx < NV && gy >= 0 && gy < ny;
        va[u] = typename A::V{};
        vb[u] = typename B::V{};
        if (ok[u]) {
          const size_t g = (size_t)gy * nx + x0 + 4 * (idx % Q);
          va[u] = a.load4(g);
          if (two) vb[u] = b.load4(g);
        }
      }
#pragma unroll
      for (int u = 0; u < UV; u++) {
        const int idx = base + u * G::THREADS + tid;
        if (idx < NV) {
          const int at = G::at(idx / Q - 1, 4 * (idx % Q));
          const float4 fa = ok[u] ? a.to4(va[u]) : make_float4(a.oob, a.oob, a.oob, a.oob);
          const float4 fb = ok[u] ? b.to4(vb[u]) : make_float4(b.oob, b.oob, b.oob, b.oob);
          put(at, fa.x, fb.x);
          put(at + 1, fa.y, fb.y);
          put(at + 2, fa.z, fb.z);
          put(at + 3, fa.w, fb.w);
        }
      }
    }
    // halo columns -1 and TX
    for (int idx = tid; idx < 2 * G::ROWS; idx += G::THREADS) {
      const int r = idx % G::ROWS - 1, c = (idx < G::ROWS) ? -1 : G::TX;
      const long long gx = x0 + c, gy = y0 + r;
      const bool ok = gx >= 0 && gx < nx && gy >= 0 && gy < ny;
      typename A::S sa{};
      typename B::S sb{};
      if (ok) {
        const size_t g = (size_t)gy * nx + gx;
        sa = a.load1(g);
        if (two) sb = b.load1(g);
      }
      put(G::at(r, c), ok ? a.cv(sa) : a.oob, ok ? b.cv(sb) : b.oob);
    }
  } else {
    constexpr int E = G::ROWS * (G::TX + 2), U = 8;
    for (int base = 0; base < E; base += G::THREADS * U) {
      typename A::S sa[U];
      typename B::S sb[U];
      bool ok[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int idx = base + u * G::THREADS + tid;
        const long long gx = x0 + idx % (G::TX + 2) - 1, gy = y0 + idx / (G::TX + 2) - 1;
        ok[u] = idx < E && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
        sa[u] = typename A::S{};
        sb[u] = typename B::S{};
        if (ok[u]) {
          const size_t g = (size_t)gy * nx + gx;
          sa[u] = a.load1(g);
          if (two) sb[u] = b.load1(g);
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int idx = base + u * G::THREADS + tid;
        if (idx < E)
          put(G::at(idx / (G::TX + 2) - 1, idx % (G::TX + 2) - 1), ok[u] ? a.cv(sa[u]) : a.oob,
              ok[u] ? b.cv(sb[u]) : b.oob);
      }
    }
  }
}

// Writes the tile's cells (float layout of TileGeom, no halo) that lie
// inside the grid to p. Full-width tiles of grids with nx % 4 == 0 and an
// aligned p store four cells per lane with 16-byte stores.
template <class G>
__device__ inline void write_tile_f32(float* p, const float* lds, long long x0, long long y0,
                                      long long nx, long long ny) {
  int tid = threadIdx.x;  // opaque per call, as in stage_tile
  asm volatile("" : "+v"(tid));
  if (((nx & 3) == 0) && (x0 + G::TX <= nx) && (reinterpret_cast<uintptr_t>(p) & 15) == 0) {
    constexpr int Q = G::TX / 4, NQ = G::TY * Q;
#pragma unroll 4
    for (int q = tid; q < NQ; q += G::THREADS) {
      const int r = q / Q, c = 4 * (q % Q);
      if (y0 + r >= ny) continue;
      const int at = G::at(r, c);
      *reinterpret_cast<float4*>(p + (size_t)(y0 + r) * nx + x0 + c) =
          make_float4(lds[at], lds[at + 1], lds[at + 2], lds[at + 3]);
    }
  } else {
    for (int idx = tid; idx < G::TY * G::TX; idx += G::THREADS) {
      const int r = idx / G::TX, c = idx % G::TX;
      const long long gx = x0 + c, gy = y0 + r;
      if (gx < nx && gy < ny) p[(size_t)gy * nx + gx] = lds[G::at(r, c)];
    }
  }
}

// Neighbour (kEdge* bit index) owning halo cell (r, c) of a tile.
template <class G>
__device__ inline int halo_owner(int r, int c) {
  if (r < 0) return c < 0 ? 4 : (c >= G::TX ? 5 : 0);
  if (r >= G::TY) return c < 0 ? 6 : (c >= G::TX ? 7 : 1);
  return c < 0 ? 2 : 3;
}

// Calls fn(r, c, from_state) for every halo cell whose owner's state
// availability differs from what the tile was staged with: on a first
// visit (staged from the initial values) the halo of visited neighbours
// (from_state = true); on later visits (staged from the stored state) the
// halo of neighbours not visited yet (from_state = false). fn must skip
// cells outside the grid.
template <class G, class F>
__device__ inline void for_stale_halo(bool first, unsigned nvis, F fn) {
  constexpr int H = 2 * (G::TX + 2) + 2 * G::TY;
  for (int k = threadIdx.x; k < H; k += G::THREADS) {
    int r, c;
    if (k < G::TX + 2) {
      r = -1;
      c = k - 1;
    } else if (k < 2 * (G::TX + 2)) {
      r = G::TY;
      c = k - (G::TX + 2) - 1;
    } else {
      const int m = k - 2 * (G::TX + 2);
      r = m < G::TY ? m : m - G::TY;
      c = m < G::TY ? -1 : G::TX;
    }
    const bool vis = (nvis >> halo_owner<G>(r, c)) & 1u;
    if (vis == first) fn(r, c, vis);
  }
}

// Neighbour bits for a changed perimeter: compares the saved perimeter
// (sper: top row, bottom row, left column, right column) with the relaxed
// tile, word by word. `diag` adds the corner tiles (8-connected).
template <class G, bool diag, class W>
__device__ inline unsigned changed_perimeter(const W* sper, const W* lds) {
  unsigned bits = 0;
  for (int k = threadIdx.x; k < G::TX; k += G::THREADS) {
    if (sper[k] != lds[G::at(0, k)])
      bits |= kEdgeN | (diag && k == 0 ? kEdgeNW : 0u) | (diag && k == G::TX - 1 ? kEdgeNE : 0u);
    if (sper[G::TX + k] != lds[G::at(G::TY - 1, k)])
      bits |= kEdgeS | (diag && k == 0 ? kEdgeSW : 0u) | (diag && k == G::TX - 1 ? kEdgeSE : 0u);
  }
  for (int k = threadIdx.x; k < G::TY; k += G::THREADS) {
    if (sper[2 * G::TX + k] != lds[G::at(k, 0)])
      bits |= kEdgeW | (diag && k == 0 ? kEdgeNW : 0u) | (diag && k == G::TY - 1 ? kEdgeSW : 0u);
    if (sper[2 * G::TX + G::TY + k] != lds[G::at(k, G::TX - 1)])
      bits |= kEdgeE | (diag && k == 0 ? kEdgeNE : 0u) | (diag && k == G::TY - 1 ? kEdgeSE : 0u);
  }
  return bits;
}

// Neighbour bits for a relaxed tile, selective version: for every
// perimeter cell p whose value changed (lds vs the saved sper), each halo
// cell q adjacent to p (8- or 4-connected) is checked with lowers(p, q)
// (LDS indices); only the owners of cells that the new value of p may
// still lower are queued. The halo holds the neighbours' values as
// staged, which can only be higher than their current ones, so "cannot
// lower the staged value" implies "cannot lower the current value".
template <class G, bool diag, class W, class Lowers>
__device__ inline unsigned pushes_from_perimeter(const W* sper, const W* lds, Lowers lowers) {
  unsigned bits = 0;
  auto check = [&](int pr, int pc, int qr, int qc) {
    if (lowers(G::at(pr, pc), G::at(qr, qc))) bits |= 1u << halo_owner<G>(qr, qc);
  };
  for (int k = threadIdx.x; k < G::TX; k += G::THREADS) {
    if (sper[k] != lds[G::at(0, k)]) {
      check(0, k, -1, k);
      if (diag) {
        check(0, k, -1, k - 1);
        check(0, k, -1, k + 1);
      }
    }
    if (sper[G::TX + k] != lds[G::at(G::TY - 1, k)]) {
      check(G::TY - 1, k, G::TY, k);
      if (diag) {
        check(G::TY - 1, k, G::TY, k - 1);
        check(G::TY - 1, k, G::TY, k + 1);
      }
    }
  }
  for (int k = threadIdx.x; k < G::TY; k += G::THREADS) {
    if (sper[2 * G::TX + k] != lds[G::at(k, 0)]) {
      check(k, 0, k, -1);
      if (diag) {
        check(k, 0, k - 1, -1);
        check(k, 0, k + 1, -1);
      }
    }
    if (sper[2 * G::TX + G::TY + k] != lds[G::at(k, G::TX - 1)]) {
      check(k, G::TX - 1, k, G::TX);
      if (diag) {
        check(k, G::TX - 1, k - 1, G::TX);
        check(k, G::TX - 1, k + 1, G::TX);
      }
    }
  }
  return bits;
}

template <class G, class W>
__device__ inline void save_perimeter(W* sper, const W* lds) {
  for (int k = threadIdx.x; k < G::TX; k += G::THREADS) {
    sper[k] = lds[G::at(0, k)];
    sper[G::TX + k] = lds[G::at(G::TY - 1, k)];
  }
  for (int k = threadIdx.x; k < G::TY; k += G::THREADS) {
    sper[2 * G::TX + k] = lds[G::at(k, 0)];
    sper[2 * G::TX + G::TY + k] = lds[G::at(k, G::TX - 1)];
  }
}

// OR-reduce a per-thread bit mask over the workgroup (result in all threads).
__device__ inline unsigned block_or(unsigned bits, unsigned* s_bits) {
  if (threadIdx.x == 0) *s_bits = 0u;
  __syncthreads();
  if (bits) atomicOr(s_bits, bits);
  __syncthreads();
  return *s_bits;
}

}  // namespace tt
