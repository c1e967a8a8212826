// Sink filling by a tile graph: three passes, each tile staged twice,
// instead of revisiting tiles until drainage information has crossed the
// grid.
//
// Fill domain as in fill_visit.h: F_p = max(d_p, min(F_p, F_q over the 8
// neighbours)) from f0 = dem on bc cells, +INF elsewhere; F_p is the
// minimax level of the cheapest path from p to a bc cell ("source" s).
// Every source has f0_s = d_s, the lowest level any path through s can
// have, so a source is always its own cheapest exit and can be a fixed
// seed below. (Reconstruction with an arbitrary marker has sources with
// f0_s > d_s, which a path through s can undercut; it keeps the work-queue
// algorithm.)
//
// A (label flood, per tile, no halo): every ring cell b of the tile (its
// outermost rows and columns) is a seed with key (d_b, b); every interior
// source is a seed (d_s, OUT). Keys relax in the lexicographic order
//     key_p = min over neighbours q of (max(d_p, level_q), label_q)
// so level_p is the in-tile minimax level from p to its cheapest seed and
// label_p that seed. Two adjacent cells with different labels give an
// edge between their labels of weight max(level_p, level_q). For seeds
// x, y the minimax in-tile path cost equals the minimax over this label
// graph (any cell path from x to y maps to a label walk whose edges weigh
// at most the path's maximum, and every edge is realised by a cell path of
// exactly its weight), so the label graph keeps all the in-tile
// information the ring cells need.
//
// B (ring graph, persistent work queue over tiles): nodes are the ring
// cells of all tiles, F_b starts at f0_b and relaxes by
//     F_a <- min(F_a, max(w, F_b))   for a label edge (a, b, w); OUT = -INF
//     F_b <- min(F_b, max(d_b, F_q)) for q a ring cell of an adjacent tile
// to the greatest fixed point, which is the exact F of every ring cell.
//
// C (per tile, no halo needed): relax the tile with its ring fixed at the
// exact F; every path from an interior cell leaves through the ring, so
// the interior converges to the exact F. Same bits as any other schedule
// (unique fixed point), checked against the oracle bit for bit.
//
// A tile whose label graph does not fit its edge table sets an overflow
// bit; passes B and C then do nothing and the work-queue kernel (fill.hip),
// launched behind them and gated on that bit, computes the result instead,
// with no host synchronisation in between.
#include "fill_visit.h"
#include "workspace.h"

namespace tt {

constexpr unsigned kErrOverflow = 2u;  // label graph larger than the edge table
constexpr unsigned kLabelOut = 0xfffeu;
constexpr unsigned kNoLabel = 0xffffffffu;
constexpr unsigned kPairEmpty = 0xffffffffu;
constexpr int kEdgeCap = 2048;   // label edges per tile
constexpr int kHashSlots = 2048;
constexpr unsigned kOrdNegInf = 0x007fffffu;  // ord_bits(-INF)

// ---------------------------------------------------------------- ring geometry
// Ring of a w x h tile: top row, bottom row, left column, right column
// (corners in the rows); a single row or column is all ring.
__device__ inline int ring_size(int w, int h) {
  if (h == 1) return w;
  if (w == 1) return h;
  return 2 * w + 2 * (h - 2);
}
__device__ inline int ring_pos(int w, int h, int r, int c) {
  if (h == 1) return c;
  if (w == 1) return r;
  if (r == 0) return c;
  if (r == h - 1) return w + c;
  if (c == 0) return 2 * w + (r - 1);
  if (c == w - 1) return 2 * w + (h - 2) + (r - 1);
  return -1;
}
__device__ inline void ring_rc(int w, int h, int k, int* r, int* c) {
  if (h == 1) {
    *r = 0;
    *c = k;
  } else if (w == 1) {
    *r = k;
    *c = 0;
  } else if (k < w) {
    *r = 0;
    *c = k;
  } else if (k < 2 * w) {
    *r = h - 1;
    *c = k - w;
  } else if (k < 2 * w + (h - 2)) {
    *r = k - 2 * w + 1;
    *c = 0;
  } else {
    *r = k - 2 * w - (h - 2) + 1;
    *c = w - 1;
  }
}

// ---------------------------------------------------------------- label keys
// (ord(level) << 32) | label. The "bound" word of a cell is ord(d) << 32,
// or all ones for cells that never change (seeds, NaN cells, outside).
//
// Only levels are compared: on equal levels the second operand (the
// cell's current key) is kept, so labels change only when a level drops.
// Levels still reach the exact minimax; each label is a seed reachable at
// a cost no higher than the cell's level, i.e. exactly its level, which is
// all the edge construction needs (ties between seeds are immaterial).
struct LabelKey {
  using T = unsigned long long;
  static __device__ __forceinline__ T lo(T a, T b) { return (a >> 32) < (b >> 32) ? a : b; }
  static __device__ __forceinline__ bool less(T a, T b) { return (a >> 32) < (b >> 32); }
  // the cell's key given the smallest neighbour key m: level max(d, level_m),
  // label of m
  static __device__ __forceinline__ T hi(T dv, T m) {
    if (dv == ~0ull) return ~0ull;
    return (m >> 32) < (dv >> 32) ? (dv | (m & 0xffffffffull)) : m;
  }
  static __device__ __forceinline__ T from_below(T v, T e) { return KeyLevel::shift(v, e, 0x138); }
  static __device__ __forceinline__ T from_above(T v, T e) { return KeyLevel::shift(v, e, 0x130); }
  static constexpr T kUnreached = (0xff800000ull << 32) | 0xffffffffull;
  static constexpr T kFixed = ~0ull;
};

struct GraphArgs {
  const float* src;     // dem
  const uint8_t* bc;
  float* state;         // output
  long long nx, ny;
  int ntx, nty, rmax;
  unsigned* fring;      // per tile, rmax: ord(F) of ring cells
  float* dring;         // per tile, rmax: ring cell levels d (NaN: no-data)
  unsigned* ecount;     // per tile: label edges
  uint2* edges;         // per tile, kEdgeCap: (a << 16 | b, ord(weight))
  unsigned cap;         // edges a tile may use (<= kEdgeCap)
  TileQueue q;
};

// Edge capacity per tile used by the next launches (lowered only by tests,
// to exercise the overflow path).
static int g_edge_cap = kEdgeCap;

bool set_graph_edge_cap(int cap) {
  if (cap < 1 || cap > kEdgeCap) return false;
  g_edge_cap = cap;
  return true;
}

// Level d and initial f0 of cell g; d is NaN for no-data.
template <int MODE>
__device__ inline void cell_levels(const GraphArgs& a, size_t g, float* d, float* f0) {
  static_assert(MODE == kModeFill, "sources must sit at their own level");
  const float s = a.src[g];
  *d = s;
  *f0 = (a.bc[g] == 1) ? s : kInf;
}

template <class G>
__device__ inline void tile_extent(const GraphArgs& a, int t, long long* x0, long long* y0,
                                   int* w, int* h) {
  *x0 = (long long)(t % a.ntx) * G::TX;
  *y0 = (long long)(t / a.ntx) * G::TY;
  const long long rw = a.nx - *x0, rh = a.ny - *y0;
  *w = (int)(rw < G::TX ? rw : G::TX);
  *h = (int)(rh < G::TY ? rh : G::TY);
}

// ---------------------------------------------------------------- A: label flood
template <class G>
struct LabelLds {
  unsigned long long sf[G::CELLS];
  unsigned long long sd[G::CELLS];  // after the flood: the edge hash table
  unsigned count, overflow;
};

template <class G, int MODE>
__device__ void label_tile(const GraphArgs& a, int t, LabelLds<G>& lds) {
  using L = LabelKey;
  using T = L::T;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int R0 = (wave / G::QX) * 64, C0 = (wave % G::QX) * 64;
  long long x0, y0;
  int w, h;
  tile_extent<G>(a, t, &x0, &y0, &w, &h);
  unsigned* fr = a.fring + (size_t)t * a.rmax;
  float* dr = a.dring + (size_t)t * a.rmax;
  const unsigned long long c0 = wall_clock64();
  // 8 cells per thread in flight: loads first, then keys
  constexpr int E = G::ROWS * (G::TX + 2), U = 8;
  for (int base = 0; base < E; base += G::THREADS * U) {
    float dv[U], fv[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int idx = base + u * G::THREADS + tid;
      const int r = idx / (G::TX + 2) - 1, c = idx % (G::TX + 2) - 1;
      dv[u] = __builtin_nanf("");
      fv[u] = kInf;
      if (idx < E && r >= 0 && r < h && c >= 0 && c < w)
        cell_levels<MODE>(a, (size_t)(y0 + r) * a.nx + (size_t)(x0 + c), &dv[u], &fv[u]);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int idx = base + u * G::THREADS + tid;
      if (idx >= E) continue;
      const int r = idx / (G::TX + 2) - 1, c = idx % (G::TX + 2) - 1;
      const float d = dv[u], f0 = fv[u];
      T f = L::kFixed, dk = L::kFixed;
      if (r >= 0 && r < h && c >= 0 && c < w) {
        const int k = ring_pos(w, h, r, c);
        if (d == d) {
          if (k >= 0)
            f = ((T)ord_bits(d) << 32) | (unsigned)k;
          else if (f0 < kInf)
            f = ((T)ord_bits(f0) << 32) | kLabelOut;
          else {
            f = L::kUnreached;
            dk = (T)ord_bits(d) << 32;
          }
        }
        if (k >= 0) {
          dr[k] = d;
          fr[k] = (d == d) ? ord_bits(f0) : 0xffffffffu;
        }
      }
      lds.sf[G::at(r, c)] = f;
      lds.sd[G::at(r, c)] = dk;
    }
  }
  __syncthreads();
  const unsigned long long c1 = wall_clock64();
  unsigned it = 0;
  int any = __syncthreads_or(fill_violation<G, L>(lds.sf, lds.sd, R0, C0, lane) ? 1 : 0);
  while (any) {
    bool ch = fill_sweep<G, L, true, true>(lds.sf, lds.sd, R0, C0, lane);
    ch |= fill_sweep<G, L, true, false>(lds.sf, lds.sd, R0, C0, lane);
    ch |= fill_sweep<G, L, false, true>(lds.sf, lds.sd, R0, C0, lane);
    ch |= fill_sweep<G, L, false, false>(lds.sf, lds.sd, R0, C0, lane);
    any = __syncthreads_or(ch ? 1 : 0);
    it++;
    if (any) any = __syncthreads_or(fill_violation<G, L>(lds.sf, lds.sd, R0, C0, lane) ? 1 : 0);
  }
  const unsigned long long c2 = wall_clock64();
  // label edges: pairs of adjacent cells (E, S, SE, SW: each pair once)
  unsigned* hk = reinterpret_cast<unsigned*>(lds.sd);
  unsigned* hw = hk + kHashSlots;
  for (int k = tid; k < kHashSlots; k += G::THREADS) {
    hk[k] = kPairEmpty;
    hw[k] = 0xffffffffu;
  }
  if (tid == 0) {
    lds.count = 0u;
    lds.overflow = 0u;
  }
  __syncthreads();
  const int cells = w * h;
  for (int idx = tid; idx < cells; idx += G::THREADS) {
    const int r = idx / w, c = idx % w;
    const T kp = lds.sf[G::at(r, c)];
    const unsigned lp = (unsigned)kp;
    if (lp == kNoLabel) continue;
#pragma unroll
    for (int dir = 0; dir < 4; dir++) {
      const int rr = r + (dir == 0 ? 0 : 1);
      const int cc = c + (dir == 0 ? 1 : (dir == 1 ? 0 : (dir == 2 ? 1 : -1)));
      if (rr >= h || cc < 0 || cc >= w) continue;
      const T kq = lds.sf[G::at(rr, cc)];
      const unsigned lq = (unsigned)kq;
      if (lq == kNoLabel || lq == lp) continue;
      const unsigned hp = (unsigned)(kp >> 32), hq = (unsigned)(kq >> 32);
      const unsigned wv = hp > hq ? hp : hq;
      const unsigned pair = lp < lq ? (lp << 16) | lq : (lq << 16) | lp;
      unsigned s = (pair * 2654435761u) >> 21;  // 11 bits
      bool done = false;
      for (int probe = 0; probe < kHashSlots; probe++) {
        const unsigned prev = atomicCAS(&hk[s], kPairEmpty, pair);
        if (prev == kPairEmpty || prev == pair) {
          atomicMin(&hw[s], wv);
          done = true;
          break;
        }
        s = (s + 1) & (kHashSlots - 1);
      }
      if (!done) lds.overflow = 1u;
    }
  }
  __syncthreads();
  uint2* ed = a.edges + (size_t)t * kEdgeCap;
  for (int k = tid; k < kHashSlots; k += G::THREADS) {
    if (hk[k] == kPairEmpty) continue;
    const unsigned i = atomicAdd(&lds.count, 1u);
    if (i < a.cap) ed[i] = make_uint2(hk[k], hw[k]);
  }
  __syncthreads();
  if (tid == 0) {
    const unsigned n = lds.count;
    a.ecount[t] = n < a.cap ? n : a.cap;
    if (n > a.cap || lds.overflow) atomicOr(&a.q.sync->error, kErrOverflow);
    atomicAdd(&a.q.sync->t_stage, wall_clock64() - c0);
    (void)c1;
    (void)c2;
    (void)it;
  }
  __syncthreads();
}

template <class G, int MODE>
__global__ __launch_bounds__(G::THREADS) void graph_label(GraphArgs a) {
  __shared__ LabelLds<G> lds;
  const int ntiles = a.ntx * a.nty;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) label_tile<G, MODE>(a, t, lds);
}

// ---------------------------------------------------------------- B: ring graph
constexpr int kGraphThreads = 256;

template <class G>
struct RingLds {
  unsigned F[2 * G::TX + 2 * G::TY];
  unsigned F0[2 * G::TX + 2 * G::TY];
  float d[2 * G::TX + 2 * G::TY];
  uint2 e[kEdgeCap];
  unsigned ne, bits;
};

// Bit (kEdge* order) of the neighbour tile holding tile-local cell (rr, cc)
// just outside a w x h tile.
__device__ inline unsigned nbr_bit(int rr, int cc, int w, int h) {
  const int ty = rr < 0 ? -1 : (rr >= h ? 1 : 0);
  const int tx = cc < 0 ? -1 : (cc >= w ? 1 : 0);
  for (int k = 0; k < 8; k++)
    if (kNbrDx[k] == tx && kNbrDy[k] == ty) return 1u << k;
  return 0u;
}

template <class G>
__device__ unsigned ring_visit(const GraphArgs& a, int t, RingLds<G>& lds) {
  const int tid = threadIdx.x;
  const unsigned long long c0 = wall_clock64();
  long long x0, y0;
  int w, h;
  tile_extent<G>(a, t, &x0, &y0, &w, &h);
  const int R = ring_size(w, h);
  const unsigned* frt = a.fring + (size_t)t * a.rmax;
  const float* drt = a.dring + (size_t)t * a.rmax;
  for (int k = tid; k < R; k += kGraphThreads) {
    const unsigned v = frt[k];
    lds.F[k] = v;
    lds.F0[k] = v;
    lds.d[k] = drt[k];
  }
  const unsigned ne = a.ecount[t];
  const uint2* ed = a.edges + (size_t)t * kEdgeCap;
  for (int k = tid; k < (int)ne; k += kGraphThreads) lds.e[k] = ed[k];
  __syncthreads();
  // adjacent tiles' ring cells (read once per visit; they only decrease)
  for (int k = tid; k < R; k += kGraphThreads) {
    const float dk = lds.d[k];
    if (dk != dk) continue;
    int r, c;
    ring_rc(w, h, k, &r, &c);
    unsigned best = 0xffffffffu;
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        const int rr = r + dy, cc = c + dx;
        if (rr >= 0 && rr < h && cc >= 0 && cc < w) continue;
        const long long gx = x0 + cc, gy = y0 + rr;
        if (gx < 0 || gx >= a.nx || gy < 0 || gy >= a.ny) continue;
        const int tx2 = (int)(gx / G::TX), ty2 = (int)(gy / G::TY);
        const int t2 = ty2 * a.ntx + tx2;
        long long xx0, yy0;
        int w2, h2;
        tile_extent<G>(a, t2, &xx0, &yy0, &w2, &h2);
        const int k2 = ring_pos(w2, h2, (int)(gy - yy0), (int)(gx - xx0));
        if (k2 < 0) continue;  // cannot happen: adjacent cells are ring cells
        const float d2 = a.dring[(size_t)t2 * a.rmax + k2];
        if (d2 != d2) continue;
        const unsigned v = __hip_atomic_load(&a.fring[(size_t)t2 * a.rmax + k2],
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        best = v < best ? v : best;
      }
    if (best != 0xffffffffu) {
      const unsigned od = ord_bits(dk);
      const unsigned cand = od > best ? od : best;
      if (cand < lds.F[k]) lds.F[k] = cand;
    }
  }
  __syncthreads();
  // label edges to their fixed point (atomics in LDS: chaotic relaxation)
  unsigned rounds = 0;
  for (;;) {
    int ch = 0;
    for (int k = tid; k < (int)ne; k += kGraphThreads) {
      const uint2 e = lds.e[k];
      const unsigned la = e.x >> 16, lb = e.x & 0xffffu;
      const bool va = la < (unsigned)R, vb = lb < (unsigned)R;
      if ((!va && la != kLabelOut) || (!vb && lb != kLabelOut)) continue;
      const unsigned fa = va ? lds.F[la] : kOrdNegInf;
      const unsigned fb = vb ? lds.F[lb] : kOrdNegInf;
      if (va) {
        const unsigned cand = e.y > fb ? e.y : fb;
        if (cand < fa) {
          atomicMin(&lds.F[la], cand);
          ch = 1;
        }
      }
      if (vb) {
        const unsigned cand = e.y > fa ? e.y : fa;
        if (cand < fb) {
          atomicMin(&lds.F[lb], cand);
          ch = 1;
        }
      }
    }
    rounds++;
    if (!__syncthreads_or(ch)) break;
  }
  // write back; neighbours reading a changed ring cell are queued
  unsigned bits = 0;
  unsigned* frw = a.fring + (size_t)t * a.rmax;
  for (int k = tid; k < R; k += kGraphThreads) {
    const unsigned v = lds.F[k];
    if (v == lds.F0[k]) continue;
    __hip_atomic_store(&frw[k], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int r, c;
    ring_rc(w, h, k, &r, &c);
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        const int rr = r + dy, cc = c + dx;
        if (rr >= 0 && rr < h && cc >= 0 && cc < w) continue;
        bits |= nbr_bit(rr, cc, w, h);
      }
  }
  bits = block_or(bits, &lds.bits);
  if (tid == 0) {
    atomicAdd(&a.q.sync->visits, 1ull);
    atomicAdd(&a.q.sync->local_iters, (unsigned long long)rounds);
    atomicMax(&a.q.sync->max_local, rounds);
    atomicAdd(&a.q.sync->t_relax, wall_clock64() - c0);
  }
  return bits;
}

template <class G>
__global__ __launch_bounds__(kGraphThreads) void graph_ring(GraphArgs a) {
  __shared__ RingLds<G> lds;
  run_queue(a.q, GridTopo{a.ntx, a.nty},
            [&](int t, bool, unsigned) { return ring_visit<G>(a, t, lds); });
}

// ---------------------------------------------------------------- C: final fill
template <class G>
struct FinalLds {
  float sf[G::CELLS];
  float sd[G::CELLS];
};

template <class G, int MODE>
__device__ void final_tile(const GraphArgs& a, int t, FinalLds<G>& lds) {
  using L = FloatLevel;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int R0 = (wave / G::QX) * 64, C0 = (wave % G::QX) * 64;
  long long x0, y0;
  int w, h;
  tile_extent<G>(a, t, &x0, &y0, &w, &h);
  const unsigned long long c0 = wall_clock64();
  float* sf = lds.sf;
  float* sd = lds.sd;
  // raw inputs (halo included: its initial values never undercut the
  // exact ring levels, which already account for the neighbours)
  stage_tile<G>(StageF32{a.src, sd, __builtin_nanf("")}, StageU8{a.bc, sf, 0.f}, true, x0, y0,
                a.nx, a.ny);
  __syncthreads();
  for (int k = tid; k < G::CELLS; k += G::THREADS) {
    float d = sd[k], f = sf[k];
    f = (f == 1.f) ? d : kInf;
    if (d != d) {
      d = kInf;
      f = kInf;
    }
    sd[k] = d;
    sf[k] = f;
  }
  __syncthreads();
  const int R = ring_size(w, h);
  const unsigned* frt = a.fring + (size_t)t * a.rmax;
  const float* drt = a.dring + (size_t)t * a.rmax;
  for (int k = tid; k < R; k += G::THREADS) {
    const float dk = drt[k];
    if (dk != dk) continue;
    int r, c;
    ring_rc(w, h, k, &r, &c);
    sf[G::at(r, c)] = unord_bits(frt[k]);
  }
  __syncthreads();
  const unsigned long long c1 = wall_clock64();
  unsigned it = 0;
  int any = __syncthreads_or(fill_violation<G, L>(sf, sd, R0, C0, lane) ? 1 : 0);
  while (any) {
    bool ch = fill_sweep<G, L, true, true>(sf, sd, R0, C0, lane);
    ch |= fill_sweep<G, L, true, false>(sf, sd, R0, C0, lane);
    ch |= fill_sweep<G, L, false, true>(sf, sd, R0, C0, lane);
    ch |= fill_sweep<G, L, false, false>(sf, sd, R0, C0, lane);
    any = __syncthreads_or(ch ? 1 : 0);
    it++;
    if (any) any = __syncthreads_or(fill_violation<G, L>(sf, sd, R0, C0, lane) ? 1 : 0);
  }
  const unsigned long long c2 = wall_clock64();
  for (int idx = tid; idx < G::TY * G::TX; idx += G::THREADS) {
    const int r = idx / G::TX, c = idx % G::TX;
    if (r >= h || c >= w) continue;
    const size_t g = (size_t)(y0 + r) * a.nx + (size_t)(x0 + c);
    float out = sf[G::at(r, c)];
    if (sd[G::at(r, c)] == kInf) {  // no-data (or +INF) cell: input bits
      const float raw = a.src[g];
      if (raw != raw) out = raw;
    }
    a.state[g] = out;
  }
  if (tid == 0) {
    atomicAdd(&a.q.sync->t_write, wall_clock64() - c0);
    (void)c1;
    (void)c2;
    (void)it;
  }
  __syncthreads();
}

template <class G, int MODE>
__global__ __launch_bounds__(G::THREADS) void graph_final(GraphArgs a) {
  __shared__ FinalLds<G> lds;
  if (__hip_atomic_load(&a.q.sync->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
    return;
  const int ntiles = a.ntx * a.nty;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) final_tile<G, MODE>(a, t, lds);
}

// ---------------------------------------------------------------- host
template <class G, int MODE>
static hipError_t graph_fill_launch(float* state, const float* src, const uint8_t* bc,
                                    int64_t nx, int64_t ny, hipStream_t stream,
                                    RunStats* stats) {
  GraphArgs a{};
  a.src = src;
  a.bc = bc;
  a.state = state;
  a.nx = nx;
  a.ny = ny;
  a.ntx = (int)((nx + G::TX - 1) / G::TX);
  a.nty = (int)((ny + G::TY - 1) / G::TY);
  a.rmax = 2 * G::TX + 2 * G::TY;
  a.cap = (unsigned)g_edge_cap;
  const int ntiles = a.ntx * a.nty;
  Workspace ws;
  hipError_t e = ws.alloc(ntiles, stream);
  if (e != hipSuccess) return e;
  a.q = ws.q;
  void* gbuf = nullptr;
  const size_t ring_bytes = (size_t)ntiles * a.rmax * 4;
  const size_t bytes = 2 * ring_bytes + (size_t)ntiles * 4 + 256 + (size_t)ntiles * kEdgeCap * 8;
  e = hipMallocAsync(&gbuf, bytes, stream);
  if (e != hipSuccess) {
    (void)ws.finish(nullptr);
    return e;
  }
  char* p = static_cast<char*>(gbuf);
  a.fring = reinterpret_cast<unsigned*>(p);
  a.dring = reinterpret_cast<float*>(p + ring_bytes);
  a.ecount = reinterpret_cast<unsigned*>(p + 2 * ring_bytes);
  a.edges = reinterpret_cast<uint2*>(
      p + ((2 * ring_bytes + (size_t)ntiles * 4 + 255) & ~(size_t)255));
  int ga = 0, gb = 0, gc = 0;
  e = persistent_grid((const void*)graph_label<G, MODE>, G::THREADS, ntiles, &ga);
  if (e == hipSuccess)
    e = persistent_grid((const void*)graph_ring<G>, kGraphThreads, ntiles, &gb);
  if (e == hipSuccess)
    e = persistent_grid((const void*)graph_final<G, MODE>, G::THREADS, ntiles, &gc);
  if (e == hipSuccess) {
    graph_label<G, MODE><<<ga, G::THREADS, 0, stream>>>(a);
    graph_ring<G><<<gb, kGraphThreads, 0, stream>>>(a);
    graph_final<G, MODE><<<gc, G::THREADS, 0, stream>>>(a);
    e = hipGetLastError();
  }
  // the work-queue kernel runs only if a tile overflowed its edge table
  if (e == hipSuccess)
    e = run_fill_gated(state, src, bc, nx, ny, MODE, TileShape{2, 2}, stream, stats,
                       &ws.q.sync->error, kErrOverflow);
  if (stats && e == hipSuccess) {
    RunStats g;
    g.grid = gb;
    g.tile_x = G::TX;
    g.tile_y = G::TY;
    hipError_t f = ws.finish(&g);
    if (f == hipSuccess && !(g.error & kErrOverflow)) *stats = g;
    e = f;
  } else {
    hipError_t f = ws.finish(nullptr);
    if (e == hipSuccess) e = f;
  }
  hipError_t f = hipFreeAsync(gbuf, stream);
  return e != hipSuccess ? e : f;
}

hipError_t run_fill_graph(float* state, const float* src, const uint8_t* bc, int64_t nx,
                          int64_t ny, TileShape ts, hipStream_t stream, RunStats* stats) {
  if (nx <= 0 || ny <= 0) return hipSuccess;
  // 64-bit label keys: at most two quadrants per tile
  if (ts.qx == 1 && ts.qy == 1)
    return graph_fill_launch<TileGeom<1, 1>, kModeFill>(state, src, bc, nx, ny, stream, stats);
  if (ts.qx == 1 && ts.qy == 2)
    return graph_fill_launch<TileGeom<1, 2>, kModeFill>(state, src, bc, nx, ny, stream, stats);
  return graph_fill_launch<TileGeom<2, 1>, kModeFill>(state, src, bc, nx, ny, stream, stats);
}

}  // namespace tt
