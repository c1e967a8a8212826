// Lowest-level-first scheduling of the persistent tile relaxation of one
// grid: sink filling, and reconstruction on request (levels in its
// negated domain, seeds at the marker cells).
//
// The FIFO queue of tile_engine.h visits tiles in the order their halos
// changed. Sink filling lowers levels from the seeds upwards, like a
// priority flood: a tile visited while a lower level is still on its way
// to it is visited again when that level arrives. Visiting the queued
// tiles whose pending levels are lowest first saves many of those revisits
// (about 30% fewer visits on 8192 x 8192 grids of 128 x 128 tiles).
//
// A queued tile carries a bucket: the level of the lowest value a
// neighbour pushed onto its halo since its last visit, quantised linearly
// over [lo, hi] of the finite DEM into kPrioBuckets - 1 buckets. The last
// bucket holds the tiles no seed has reached yet (visited only when
// nothing else is queued). The state word of tile t is kQueued | kRunning
// | bucket << 8; a push with a lower bucket than the queued one moves the
// tile down (decrease-key).
//
// The queue itself is one bitmap per bucket (bit t: tile t may be queued
// in that bucket) and a per-bucket count of set bits. A push sets the bit
// of its bucket; a moved tile leaves its old bit behind, which a popper
// discards when the tile's state no longer names that bucket. A pop ranks
// the set bits by bucket (from the counts), picks one of the lowest
// `window` of them at random, lower ranks more likely (window = twice the
// number of workgroups: with strictly lowest-first pops, hundreds of idle
// workgroups race for the same few tiles), clears it and claims the tile
// if its state still says "queued in this bucket". The counts may lag the
// bitmaps for a moment; they agree again whenever no push or pop is in
// flight.
//
// Termination and ordering follow tile_engine.h: a pending counter of
// queued-or-running tiles (zero exactly at convergence), an acquire fence
// after each pop and a release fence after each visit, bounded waits.
#pragma once

#include "fill_visit.h"

namespace tt {

constexpr int kPrioBuckets = 1024;
constexpr int kPrioUnreached = kPrioBuckets - 1;

// Queue state of one launch (device memory). The finite range of the
// bounds (the DEM; the negated mask for reconstruction) is kept in the
// control block's spare words as order keys (ord_bits), each raised with
// atomicMax from zero: pad[0] = ~key(min), pad[1] = key(max).
struct PrioQueue {
  Sync* sync;
  int* count;                // per bucket: set bits (may lag for a moment)
  unsigned long long* bits;  // kPrioBuckets x nwords
  unsigned* state;           // per tile: kQueued | kRunning | bucket << 8
  unsigned* visited;         // per tile: 1 once its first visit has been written
  unsigned* seed;            // per tile: key of its lowest seed level, ~0u none
  int ntiles, nwords;
  // a pop picks at random among this many of the lowest queued tiles
  int window;
  int stats;  // 1: count visits and time their phases in *sync
  unsigned* sticky;
  unsigned long long wait_limit;
  __device__ Sync* stats_sink() const { return stats ? sync : nullptr; }
};

// Words of the queue workspace: control block, counts, bitmaps (two
// words per 64-bit word), state, visited flags, seeds.
inline size_t prio_workspace_words(int ntiles) {
  const size_t nwords = ((size_t)ntiles + 63) / 64;
  return 64 + (size_t)kPrioBuckets + 2 * (size_t)kPrioBuckets * nwords + 3 * (size_t)ntiles;
}

inline PrioQueue prio_layout(void* base, int ntiles) {
  PrioQueue q{};
  const size_t nwords = ((size_t)ntiles + 63) / 64;
  unsigned* w = static_cast<unsigned*>(base);
  q.sync = reinterpret_cast<Sync*>(w);
  q.count = reinterpret_cast<int*>(w + 64);
  q.bits = reinterpret_cast<unsigned long long*>(w + 64 + kPrioBuckets);
  q.state = w + 64 + kPrioBuckets + 2 * kPrioBuckets * nwords;
  q.visited = q.state + ntiles;
  q.seed = q.visited + ntiles;
  q.ntiles = ntiles;
  q.nwords = (int)nwords;
  return q;
}

__device__ inline void pq_raise_error(const PrioQueue& q, unsigned bits) {
  atomicOr(&q.sync->error, bits);
  if (q.sticky) atomicOr(q.sticky, bits);
}

// Level -> bucket map of one launch.
struct PrioScale {
  float lo, scale;
  __device__ int bucket(unsigned key) const {
    // NaN (no finite range) and -INF go to bucket 0
    const float x = (unord_bits(key) - lo) * scale;
    return (int)fminf(fmaxf(x, 0.f), (float)(kPrioBuckets - 2));
  }
};

__device__ inline PrioScale prio_scale(const Sync* s) {
  const unsigned lk = ~ld_relaxed(&s->pad[0]), hk = ld_relaxed(&s->pad[1]);
  if (lk >= hk) return PrioScale{0.f, 0.f};  // no finite cell, or one level
  const float lo = unord_bits(lk), hi = unord_bits(hk);
  return PrioScale{lo, (float)(kPrioBuckets - 1) / (hi - lo)};
}

__device__ inline bool pq_cas(unsigned* p, unsigned& expected, unsigned desired) {
  return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sets tile t's bit in bucket b (counted if it was clear).
__device__ inline void pq_mark(const PrioQueue& q, int t, int b) {
  unsigned long long* w = &q.bits[(size_t)b * q.nwords + (t >> 6)];
  const unsigned long long m = 1ull << (t & 63);
  const unsigned long long old =
      __hip_atomic_fetch_or(w, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (!(old & m))
    __hip_atomic_fetch_add(&q.count[b], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Queues tile t at bucket b, or moves it down to b (called after a release
// fence that published the caller's writes). A running tile only records
// the bucket; its owner queues it when it releases the tile.
__device__ inline void pq_push(const PrioQueue& q, int t, int b) {
  unsigned cur = ld_relaxed(&q.state[t]);
  for (;;) {
    if ((cur & kQueued) && (int)(cur >> 8) <= b) return;  // queued at or below b already
    if (pq_cas(&q.state[t], cur, (cur & kRunning) | kQueued | ((unsigned)b << 8))) break;
  }
  if (cur & kRunning) return;
  if (!(cur & kQueued)) {
    __hip_atomic_fetch_add(&q.sync->pending, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (q.stats) atomicAdd(&q.sync->tail, 1u);
  }
  pq_mark(q, t, b);
}

// Releases tile t after its visit: queued again (at the bucket pushes
// recorded while it ran) or idle.
__device__ inline void pq_release(const PrioQueue& q, int t) {
  unsigned cur = ld_relaxed(&q.state[t]);
  for (;;) {
    if (pq_cas(&q.state[t], cur, (cur & kQueued) ? (cur & ~kRunning) : 0u)) break;
  }
  if (cur & kQueued) {
    if (q.stats) atomicAdd(&q.sync->tail, 1u);
    pq_mark(q, t, (int)(cur >> 8));
    return;
  }
  // the pushes made before are performed before the count can reach zero
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __hip_atomic_fetch_sub(&q.sync->pending, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Pops one of the `window` lowest queued tiles, chosen at random by `rnd`
// (called by all 64 lanes of one wavefront; the result is the same in
// every lane): concurrent poppers then rarely race for the same tile,
// while the order stays close to lowest-first. The unreached bucket is
// taken only when every other bucket is empty. Returns -1 when the grid
// has converged or after an error.
__device__ inline int pq_pop(const PrioQueue& q, unsigned rnd, int* bucket) {
  const int lane = threadIdx.x & 63;
  Sync* s = q.sync;
  const unsigned long long t0 = wall_clock64();
  static_assert(kPrioBuckets == 64 * 16, "16 buckets per lane");
  for (unsigned spin = 0;; spin++, rnd = rnd * 1664525u + 1013904223u) {
    // lane l holds the counts of buckets 16 l .. 16 l + 15
    int c[16];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
      c[k] = __hip_atomic_load(&q.count[lane * 16 + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      c[k] = c[k] > 0 ? c[k] : 0;
    }
    const int unreached = __shfl(c[15], 63);
    if (lane == 63) c[15] = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) sum += c[k];
    int incl = sum;  // inclusive prefix over lanes
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o);
      if (lane >= o) incl += y;
    }
    const int total = __shfl(incl, 63);
    int b = -1, rank = 0;
    if (total > 0) {
      // rank r among the lowest `window` counted bits -> its bucket
      const int window = total < q.window ? total : q.window;
      // the lower of two uniform ranks: lower ranks are more likely
      const unsigned r2 = (rnd * 747796405u + 2891336453u) >> 8;
      const int r = min((int)((rnd >> 8) % (unsigned)window), (int)(r2 % (unsigned)window));
      const int src = __ffsll((long long)__ballot(incl > r)) - 1;
      if (lane == src) {
        int acc = incl - sum;
#pragma unroll
        for (int k = 0; k < 16; k++) {
          if (b < 0 && acc + c[k] > r) {
            b = lane * 16 + k;
            rank = r - acc;
          }
          acc += c[k];
        }
      }
      b = __shfl(b, src);
      rank = __shfl(rank, src);
    } else if (unreached > 0) {
      b = kPrioUnreached;
      rank = (int)((rnd >> 8) % (unsigned)unreached);
    }
    bool found = false;
    if (b >= 0) {
      const unsigned long long* row = q.bits + (size_t)b * q.nwords;
      for (int w0 = 0; w0 < q.nwords; w0 += 64) {
        const int wi = w0 + lane;
        const unsigned long long word =
            wi < q.nwords ? __hip_atomic_load(&row[wi], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                          : 0ull;
        const int pc = __popcll(word);
        int pincl = pc;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const int y = __shfl_up(pincl, o);
          if (lane >= o) pincl += y;
        }
        const int nbits = __shfl(pincl, 63);
        if (nbits == 0) continue;
        found = true;
        // the (rank mod nbits)-th set bit of this stretch of the bucket
        const int j = rank % nbits;
        const int pick = __ffsll((long long)__ballot(pincl > j)) - 1;
        int got = -1;
        if (lane == pick) {
          unsigned long long w = word;
          for (int k = j - (pincl - pc); k > 0; k--) w &= w - 1;
          const int bit = __ffsll((long long)w) - 1;
          const unsigned long long m = 1ull << bit;
          const unsigned long long old =
              __hip_atomic_fetch_and(const_cast<unsigned long long*>(&row[wi]), ~m,
                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (old & m) {
            __hip_atomic_fetch_sub(&q.count[b], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int t = wi * 64 + bit;
            // claim it if it is still queued in this bucket (a stale bit
            // of a tile moved down or already taken is dropped)
            unsigned cur = kQueued | ((unsigned)b << 8);
            if (pq_cas(&q.state[t], cur, kRunning)) got = t;
          }
        }
        got = __shfl(got, pick);
        if (got >= 0) {
          *bucket = b;
          return got;
        }
        break;
      }
    }
    if (found && (spin & 7u) != 7u) continue;  // lost a race: look again at once
    // one lane's reading decides for the whole wavefront
    const unsigned pending = __builtin_amdgcn_readfirstlane(ld_relaxed(&s->pending));
    const unsigned error = __builtin_amdgcn_readfirstlane(ld_relaxed(&s->error));
    if (pending == 0u || error != 0u) return -1;
    if (wall_clock64() - t0 > q.wait_limit) {
      if (lane == 0) pq_raise_error(q, kErrTimeout);
      return -1;
    }
    if (!found) __builtin_amdgcn_s_sleep(2);
  }
}

// The persistent loop (priority version of run_queue): visit(t, first,
// nvis, lev) relaxes tile t and returns the kEdge* bits of the neighbours
// to queue; for each such neighbour k it has lowered lev[k] (LDS, preset
// to ~0u) to the key of the lowest level it pushes onto that neighbour.
// kSelf requeues the tile at the bucket it was popped from.
template <class Topo, class Visit>
__device__ inline void run_prio_queue(const PrioQueue& q, const Topo& topo, Visit visit) {
  __shared__ int s_tile, s_first, s_bucket;
  __shared__ unsigned s_nvis;
  __shared__ unsigned s_lev[8];
  const PrioScale scale = prio_scale(q.sync);
  unsigned rot = blockIdx.x * 0x9e3779b9u;
  for (;;) {
    if (threadIdx.x < 64) {
      const unsigned long long w0 = wall_clock64();
      int b = 0;
      const int t = pq_pop(q, rot, &b);
      rot = rot * 1664525u + 1013904223u;
      if (threadIdx.x < 8) s_lev[threadIdx.x] = ~0u;
      if (threadIdx.x == 0) {
        s_tile = t;
        s_bucket = b;
        if (t >= 0) {
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
          s_first = ld_relaxed(&q.visited[t]) == 0u;
          unsigned nvis = 0;
          for (int k = 0; k < 8; k++) {
            const int n = topo.neighbour(t, k);
            if (n >= 0 && ld_relaxed(&q.visited[n])) nvis |= 1u << k;
          }
          s_nvis = nvis;
          if (q.stats) atomicAdd(&q.sync->t_wait, wall_clock64() - w0);
        }
      }
    }
    __syncthreads();
    const int tt = s_tile;
    const bool first = s_first != 0;
    const unsigned nvis = s_nvis;
    if (tt < 0) return;
    const unsigned bits = visit(tt, first, nvis, s_lev);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    if (threadIdx.x < 64) {
      const int lane = threadIdx.x;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      if (first) {
        if (lane == 0)
          __hip_atomic_store(&q.visited[tt], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      }
      if (lane < 8 && ((bits >> lane) & 1u)) {
        const int n = topo.neighbour(tt, lane);
        if (n >= 0) pq_push(q, n, scale.bucket(s_lev[lane]));
      }
      if (lane == 0) {
        if (bits & kSelf) pq_push(q, tt, s_bucket);
        pq_release(q, tt);
      }
    }
    __syncthreads();
  }
}

}  // namespace tt
