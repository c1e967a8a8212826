// Sink filling by a tile graph: every tile is staged once, a small graph of
// the tiles' ring cells carries the drainage between tiles, and a per-cell
// map finishes, instead of revisiting whole tiles until drainage
// information has crossed the grid.
//
// Fill domain as in fill_visit.h: F_p = max(d_p, min over the 8 neighbours
// q of F_q), F = d on bc cells ("sources"): F_p is the minimax level of the
// cheapest path from p to a source (reference src/fillsinks.c:22-48 and
// src/reconstruct.c:270-279 reach the same unique fixed point).
//
// A (per tile, no halo): the tile's ring cells (its outermost rows and
// columns) and its interior sources are seeds at their own level. The tile
// relaxes to L_p = min over seeds s of m(s, p), the in-tile minimax level,
// and every cell gets a label: a seed s with m(s, p) = L_p (ring cell k:
// label k; interior sources: one label OUT). Labels are the smallest seed
// reachable through "allowed" neighbours (q with max(d_p, L_q) == L_p),
// which makes them unique and independent of the update order. Two
// adjacent cells with different labels give an edge between their labels
// of weight max(L_p, L_q); for two seeds x, y, m(x, y) equals the minimax
// over these edges, so the label graph keeps everything the ring cells
// need about the tile's interior. Edges to OUT bound the ring cell's level
// directly (OUT drains at -INF). L goes to the output array, the labels to
// a 16-bit array.
//
// B (ring graph, persistent work queue over tiles): the ring cells of all
// tiles, with the label edges inside tiles and the grid adjacency between
// ring cells of neighbouring tiles (F_p <- min(F_p, max(d_p, F_q))), relax
// to the greatest fixed point: the exact F of every ring cell.
//
// C (per cell): F_p = max(L_p, F[label_p]) (= L_p for OUT, unlabelled and
// NaN cells). Exact for any valid labels: for the label a of p and any seed
// b, F_a <= max(F_b, m(a, b)) <= max(F_b, m(b, p), L_p), so max(F_a, L_p)
// <= min over b of max(F_b, m(b, p)) = F_p <= max(F_a, L_p).
//
// A tile whose label graph does not fit its edge table sets an overflow
// bit; passes B and C then do nothing and the work-queue kernel (fill.hip),
// launched behind them and gated on that bit, computes the result instead,
// with no host synchronisation in between.
#include "fill_visit.h"
#include "workspace.h"

namespace tt {

namespace {

constexpr unsigned kErrOverflow = 2u;  // a label graph larger than its edge table
using GG = TileGeom<2, 2>;             // 128 x 128 tiles, four wavefronts
constexpr int kRingStride = 512;       // ring entries per tile (2 * 128 + 2 * 126 = 508 used)
constexpr int kEdgeCap = 2048;         // label edges per tile
constexpr int kHashSlots = 2048;
constexpr unsigned kPairEmpty = 0xffffffffu;
// kind bits of a cell during pass A (two per cell)
constexpr unsigned kKindSource = 1u;  // bc == 1
constexpr unsigned kKindNan = 2u;     // no-data cell of the tile
constexpr unsigned kOrdInf = 0xff800000u;  // ord_bits(+INF)
// Labels during pass A: ring positions below kLabOut.
constexpr unsigned kLabOut = 0xfffffffdu;   // an interior source
constexpr unsigned kLabNan = 0xfffffffeu;   // no-data cell
constexpr unsigned kLabNone = 0xffffffffu;  // unlabelled (no seed reached yet)
// 16-bit labels stored for pass C
constexpr unsigned short kLab16Out = 0xfffe, kLab16None = 0xffff;

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

}  // namespace

struct GraphArgs {
  const float* src;  // dem
  const uint8_t* bc;
  float* out;              // pass A: L; pass C: F
  unsigned short* lab;     // per cell: ring position, kLab16Out or kLab16None
  long long nx, ny;
  int ntx, nty;
  unsigned* ringF;  // per tile, kRingStride: ord(F) of ring cells (pass B state)
  unsigned* ringD;  // per tile, kRingStride: ord(d) of ring cells (kOrdInf: no-data)
  unsigned* ecount;  // per tile: label edges
  uint2* edges;      // per tile, kEdgeCap: (a << 16 | b, ord(weight))
  unsigned cap;      // edges a tile may use (<= kEdgeCap)
  TileQueue q;       // pass B's queue; its Sync carries the overflow bit
};

// Edge capacity per tile used by the next launches (lowered only by tests,
// to exercise the overflow path).
static int g_edge_cap = kEdgeCap;

bool set_graph_edge_cap(int cap) {
  if (cap < 1 || cap > kEdgeCap) return false;
  g_edge_cap = cap;
  return true;
}

namespace {

__device__ inline void tile_extent(const GraphArgs& a, int t, long long* x0, long long* y0,
                                   int* w, int* h) {
  *x0 = (long long)(t % a.ntx) * GG::TX;
  *y0 = (long long)(t / a.ntx) * GG::TY;
  const long long rw = a.nx - *x0, rh = a.ny - *y0;
  *w = (int)(rw < GG::TX ? rw : GG::TX);
  *h = (int)(rh < GG::TY ? rh : GG::TY);
}

// ---------------------------------------------------------------- A: joint flood
// Levels and labels relax together. A cell's state is (level, label) and
// every state is realisable: its label is a seed from which a path reaches
// the cell with no d above the level (seeds start at (d, own label), other
// cells at (+INF, none)). A step lowers a cell's level only from a strictly
// lower candidate max(d, level of a neighbour) and then takes that
// neighbour's label, which keeps the state realisable; levels reach the
// in-tile minimax L, so at the fixed point every label is a seed s with
// m(s, p) = L_p. Which such seed a cell ends with depends on the schedule;
// the filled output does not (pass C is exact for any valid labels). A
// state is one 64-bit LDS word (level << 32 | label), so a cell read by
// another wavefront is never seen half updated.
using Key = unsigned long long;
__device__ __forceinline__ unsigned key_level(Key k) { return (unsigned)(k >> 32); }
__device__ __forceinline__ unsigned key_label(Key k) { return (unsigned)k; }
__device__ __forceinline__ Key make_key(unsigned level, unsigned label) {
  return ((Key)level << 32) | label;
}

// One directional sweep of the joint relaxation over the lane's 64 cells
// (the structure of ord_sweep_regs in fill_visit.h: two interleaved chains,
// the carried state the updated cell behind, its side neighbours by DPP,
// the outside line and the corner cells from LDS). f: levels, lb: labels,
// dl: the cells' bounds d.
template <class G, bool VERT, bool FWD>
__device__ __forceinline__ bool joint_sweep_regs(unsigned (&f)[64], unsigned (&lb)[64],
                                                 const Key* K, const unsigned (&dl)[64], int R0,
                                                 int C0, int lane) {
  const int out_k = FWD ? -1 : 64;
  const Key s0 = VERT ? K[G::at(R0 + out_k, C0 + lane)] : K[G::at(R0 + lane, C0 + out_k)];
  unsigned pc0 = key_level(s0), pl0 = key_label(s0);
  unsigned corner = kNoNbr, cornerl = kLabNone;
  if (VERT) {
    const Key lo = K[G::at(R0 + out_k, C0 - 1)], hi = K[G::at(R0 + out_k, C0 + 64)];
    const Key c = (lane == 0) ? lo : hi;
    corner = key_level(c);
    cornerl = key_label(c);
  }
  unsigned pc1 = f[FWD ? 31 : 32], pl1 = lb[FWD ? 31 : 32];
  unsigned changed = 0u;
  auto step = [&](int i, unsigned& pc, unsigned& pl, unsigned a, unsigned la, unsigned b,
                  unsigned lbb) {
    unsigned m = a, ml = la;
    if (b < m) {
      m = b;
      ml = lbb;
    }
    if (pc < m) {
      m = pc;
      ml = pl;
    }
    const unsigned nv = max(dl[i], m);
    const bool lower = nv < f[i];
    changed |= (unsigned)lower;
    f[i] = lower ? nv : f[i];
    lb[i] = lower ? ml : lb[i];
    pc = f[i];
    pl = lb[i];
  };
#pragma unroll
  for (int s = 0; s < 32; s++) {
    const int i0 = FWD ? s : 63 - s;
    const int i1 = FWD ? 32 + s : 31 - s;
    const unsigned old0 = (s == 0) ? corner : kNoNbr;
    const unsigned oldl0 = (s == 0) ? cornerl : kLabNone;
    const unsigned a0 = dpp_u32(old0, pc0, 0x138), b0 = dpp_u32(old0, pc0, 0x130);
    const unsigned la0 = dpp_u32(oldl0, pl0, 0x138), lb0 = dpp_u32(oldl0, pl0, 0x130);
    const unsigned a1 = dpp_u32(kNoNbr, pc1, 0x138), b1 = dpp_u32(kNoNbr, pc1, 0x130);
    const unsigned la1 = dpp_u32(kLabNone, pl1, 0x138), lb1 = dpp_u32(kLabNone, pl1, 0x130);
    step(i0, pc0, pl0, a0, la0, b0, lb0);
    step(i1, pc1, pl1, a1, la1, b1, lb1);
  }
  return changed != 0u;
}

// Fixed-point test of the levels from the row layout in registers (as
// fill_violation_rows in fill_visit.h, reading the levels of the 64-bit
// states). Labels move only with levels, so the levels' fixed point is the
// joint one.
template <class G>
__device__ __forceinline__ bool joint_violation_rows(const unsigned (&f)[64],
                                                     const unsigned (&d)[64], const Key* K,
                                                     int R0, int C0, int lane) {
  auto lev = [&](int r, int c) { return key_level(K[G::at(r, c)]); };
  const int R = R0 + lane;
  const int er = (lane == 0) ? R0 - 1 : R0 + 64;
  const unsigned left = lev(R, C0 - 1), right = lev(R, C0 + 64);
  unsigned ep = lev(er, C0 - 1), ec = lev(er, C0);
  bool v = false;
#pragma unroll
  for (int i = 0; i < 64; i++) {
    const unsigned op = (i == 0) ? left : f[i - 1];
    const unsigned on = (i == 63) ? right : f[i + 1];
    const unsigned en = lev(er, C0 + i + 1);
    const unsigned row = min(min(op, on), f[i]);
    const unsigned erow = min(min(ep, en), ec);
    const unsigned m = min(row, min(dpp_u32(erow, row, 0x138), dpp_u32(erow, row, 0x130)));
    v |= max(d[i], m) < f[i];
    ep = ec;
    ec = en;
  }
  return v;
}

template <class G, bool VERT>
__device__ __forceinline__ void load_states(unsigned (&f)[64], unsigned (&lb)[64], const Key* K,
                                            int R0, int C0, int lane) {
#pragma unroll
  for (int i = 0; i < 64; i++) {
    const Key k = VERT ? K[G::at(R0 + i, C0 + lane)] : K[G::at(R0 + lane, C0 + i)];
    f[i] = key_level(k);
    lb[i] = key_label(k);
  }
}
template <class G, bool VERT>
__device__ __forceinline__ void store_states(const unsigned (&f)[64], const unsigned (&lb)[64],
                                             Key* K, int R0, int C0, int lane) {
#pragma unroll
  for (int i = 0; i < 64; i++) {
    const Key k = make_key(f[i], lb[i]);
    if (VERT)
      K[G::at(R0 + i, C0 + lane)] = k;
    else
      K[G::at(R0 + lane, C0 + i)] = k;
  }
}

// ---------------------------------------------------------------- A: one tile
constexpr int kKindWords = (GG::CELLS + 15) / 16;
static_assert(kKindWords <= kHashSlots, "kind bits live in the hash table's space");

struct LabelLds {
  Key K[GG::CELLS];  // (level, label) states; during staging (start level, d)
  unsigned hkey[kHashSlots];
  unsigned hw[kHashSlots];
  unsigned F0[kRingStride];  // ring cells: starting F (bc: d; OUT edges)
  unsigned count, overflow;
};

// Quadrant coordinates of the calling lane, from a thread index the
// compiler cannot see through: every phase of a tile recomputes its LDS
// addresses instead of having all of them hoisted out of the tile loop.
struct LaneQuad {
  int lane, R0, C0;
  __device__ LaneQuad() {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    lane = t & 63;
    R0 = ((t >> 6) / GG::QX) * 64;
    C0 = ((t >> 6) % GG::QX) * 64;
  }
};

// Records a label edge (a, b) of weight w, or, for an edge to OUT, bounds
// the ring cell's starting level.
__device__ inline void add_edge(LabelLds& lds, unsigned la, unsigned lb, unsigned wv) {
  if (la == kLabOut || lb == kLabOut) {
    atomicMin(&lds.F0[la == kLabOut ? lb : la], wv);
    return;
  }
  const unsigned pair = la < lb ? (la << 16) | lb : (lb << 16) | la;
  unsigned s = (pair * 2654435761u) >> 21;  // 11 bits
  for (int probe = 0; probe < 64; probe++) {
    const unsigned prev = atomicCAS(&lds.hkey[s], kPairEmpty, pair);
    if (prev == kPairEmpty || prev == pair) {
      atomicMin(&lds.hw[s], wv);
      return;
    }
    s = (s + 1) & (kHashSlots - 1);
  }
  lds.overflow = 1u;
}

__device__ void label_tile(const GraphArgs& a, int t, LabelLds& lds) {
  const int tid = threadIdx.x;
  long long x0, y0;
  int w, h;
  tile_extent(a, t, &x0, &y0, &w, &h);
  Key* K = lds.K;
  // ---- stage: (start level, d) of the tile; the halo and cells outside
  // the grid are inert (+INF, +INF). Two kind bits per cell (bc == 1, no
  // data) go to a bit array in the hash table's space, for the labels.
  unsigned* kind = lds.hkey;  // 2 bits per LDS cell index, until the hash table is built
  for (int k = tid; k < kKindWords; k += GG::THREADS) kind[k] = 0u;
  for (int k = tid; k < 2 * (GG::TX + 2) + 2 * GG::TY; k += GG::THREADS) {
    int r, c;  // the halo ring of the LDS tile
    if (k < GG::TX + 2) {
      r = -1;
      c = k - 1;
    } else if (k < 2 * (GG::TX + 2)) {
      r = GG::TY;
      c = k - (GG::TX + 2) - 1;
    } else {
      const int m = k - 2 * (GG::TX + 2);
      r = m % GG::TY;
      c = m < GG::TY ? -1 : GG::TX;
    }
    K[GG::at(r, c)] = make_key(kOrdInf, kLabNone);
  }
  __syncthreads();
  {
    int it_ = tid;
    asm volatile("" : "+v"(it_));  // no addresses hoisted out of the tile loop
    auto put = [&](int r, int c, float d, unsigned char b) {
      unsigned od = kOrdInf, of = kOrdInf, kb = 0u;
      if (r < h && c < w) {
        if (d == d) {
          const bool ring = r == 0 || r == h - 1 || c == 0 || c == w - 1;
          od = ord_bits(d);
          of = (ring || b == 1) ? od : kOrdInf;
          kb = (b == 1) ? kKindSource : 0u;
        } else {
          kb = kKindNan;
        }
      }
      const int at = GG::at(r, c);
      K[at] = make_key(of, od);
      if (kb) atomicOr(&kind[at >> 4], kb << (2 * (at & 15)));
    };
    const bool vec = ((a.nx & 3) == 0) && w == GG::TX &&
                     (reinterpret_cast<uintptr_t>(a.src) & 15) == 0 &&
                     (reinterpret_cast<uintptr_t>(a.bc) & 3) == 0;
    if (vec) {
      constexpr int Q = GG::TX / 4, NQ = GG::TY * Q, U = 8;
      for (int base = 0; base < NQ; base += GG::THREADS * U) {
        float4 dv[U];
        uchar4 bv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int q = base + u * GG::THREADS + it_;
          const int r = q / Q, c = 4 * (q % Q);
          const float nan = __builtin_nanf("");
          dv[u] = make_float4(nan, nan, nan, nan);
          bv[u] = make_uchar4(0, 0, 0, 0);
          if (q < NQ && r < h) {
            const size_t g = (size_t)(y0 + r) * a.nx + (size_t)(x0 + c);
            dv[u] = *reinterpret_cast<const float4*>(a.src + g);
            bv[u] = *reinterpret_cast<const uchar4*>(a.bc + g);
          }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int q = base + u * GG::THREADS + it_;
          if (q >= NQ) continue;
          const int r = q / Q, c = 4 * (q % Q);
          put(r, c, dv[u].x, bv[u].x);
          put(r, c + 1, dv[u].y, bv[u].y);
          put(r, c + 2, dv[u].z, bv[u].z);
          put(r, c + 3, dv[u].w, bv[u].w);
        }
      }
    } else {
      constexpr int E = GG::TY * GG::TX, U = 8;
      for (int base = 0; base < E; base += GG::THREADS * U) {
        float dv[U];
        unsigned char bv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int idx = base + u * GG::THREADS + it_;
          const int r = idx / GG::TX, c = idx % GG::TX;
          dv[u] = __builtin_nanf("");
          bv[u] = 0;
          if (idx < E && r < h && c < w) {
            const size_t g = (size_t)(y0 + r) * a.nx + (size_t)(x0 + c);
            dv[u] = a.src[g];
            bv[u] = a.bc[g];
          }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int idx = base + u * GG::THREADS + it_;
          if (idx < E) put(idx / GG::TX, idx % GG::TX, dv[u], bv[u]);
        }
      }
    }
  }
  __syncthreads();
  // ---- the lane's bounds in both layouts, from the states' low halves
  const LaneQuad lq;
  const int lane = lq.lane, R0 = lq.R0, C0 = lq.C0;
  unsigned dcol[64], drow[64];
#pragma unroll
  for (int i = 0; i < 64; i++) {
    dcol[i] = key_label(K[GG::at(R0 + i, C0 + lane)]);
    drow[i] = key_label(K[GG::at(R0 + lane, C0 + i)]);
  }
  __syncthreads();
  // ---- starting labels (ring position, OUT, no data or none) in the low
  // halves; the ring cells' bounds and starting levels for pass B
  const int R = ring_size(w, h);
  for (int idx = tid; idx < GG::CELLS; idx += GG::THREADS) {
    const int r = idx / GG::P - 1, c = idx % GG::P - 1;
    if (r < 0 || r >= h || c < 0 || c >= w) {  // halo and cells outside the grid
      K[idx] = make_key(kOrdInf, kLabNone);
      continue;
    }
    const unsigned kb = (kind[idx >> 4] >> (2 * (idx & 15))) & 3u;
    const int k = ring_pos(w, h, r, c);
    const Key st = K[idx];
    unsigned v = kLabNone;
    if (kb & kKindNan)
      v = kLabNan;
    else if (k >= 0)
      v = (unsigned)k;
    else if (kb & kKindSource)
      v = kLabOut;
    if (k >= 0) {
      const unsigned dp = key_label(st);
      a.ringD[(size_t)t * kRingStride + k] = (kb & kKindNan) ? kOrdInf : dp;
      lds.F0[k] = (kb == kKindSource) ? dp : kOrdInf;
    }
    K[idx] = make_key(key_level(st), v);
  }
  __syncthreads();  // the kind bits are read; their space becomes the hash table
  for (int k = tid; k < kHashSlots; k += GG::THREADS) {
    lds.hkey[k] = kPairEmpty;
    lds.hw[k] = 0xffffffffu;
  }
  if (tid == 0) {
    lds.count = 0u;
    lds.overflow = 0u;
  }
  // ---- joint flood to the tile's fixed point
  {
    unsigned f[64], lb[64];
    load_states<GG, false>(f, lb, K, R0, C0, lane);
    int any = __syncthreads_or(joint_violation_rows<GG>(f, drow, K, R0, C0, lane) ? 1 : 0);
    while (any) {
      load_states<GG, true>(f, lb, K, R0, C0, lane);
      bool ch = joint_sweep_regs<GG, true, true>(f, lb, K, dcol, R0, C0, lane);
      ch |= joint_sweep_regs<GG, true, false>(f, lb, K, dcol, R0, C0, lane);
      store_states<GG, true>(f, lb, K, R0, C0, lane);
      load_states<GG, false>(f, lb, K, R0, C0, lane);
      ch |= joint_sweep_regs<GG, false, true>(f, lb, K, drow, R0, C0, lane);
      ch |= joint_sweep_regs<GG, false, false>(f, lb, K, drow, R0, C0, lane);
      store_states<GG, false>(f, lb, K, R0, C0, lane);
      any = __syncthreads_or(ch ? 1 : 0);
      if (any) any = __syncthreads_or(joint_violation_rows<GG>(f, drow, K, R0, C0, lane) ? 1 : 0);
    }
  }
  __syncthreads();
  // ---- label edges: adjacent cells (E, S, SE, SW: each pair once) with
  // different labels. Each thread walks half a row and records an edge
  // only when its label pair changes along the walk (runs of one pair are
  // merged first), so the LDS atomics see a few entries per boundary run.
  {
    int it_ = tid;
    asm volatile("" : "+v"(it_));
    const int r = it_ >> 1, cb = (it_ & 1) * (GG::TX / 2);
    if (r < h) {
      unsigned pk[4], pw[4];  // pending pair per direction (kPairEmpty: none)
#pragma unroll
      for (int dir = 0; dir < 4; dir++) pk[dir] = kPairEmpty;
      auto flush = [&](int dir) {
        if (pk[dir] != kPairEmpty) add_edge(lds, pk[dir] >> 16 == 0xffffu ? kLabOut : pk[dir] >> 16,
                                            (pk[dir] & 0xffffu) == 0xffffu ? kLabOut : pk[dir] & 0xffffu,
                                            pw[dir]);
        pk[dir] = kPairEmpty;
      };
      for (int c = cb; c < cb + GG::TX / 2 && c < w; c++) {
        const Key kp = K[GG::at(r, c)];
        const unsigned lp = key_label(kp), vp = key_level(kp);
#pragma unroll
        for (int dir = 0; dir < 4; dir++) {
          const int rr = r + (dir == 0 ? 0 : 1);
          const int cc = c + (dir == 0 ? 1 : (dir == 1 ? 0 : (dir == 2 ? 1 : -1)));
          unsigned key = kPairEmpty, wv = 0;
          if (lp <= kLabOut && rr < h && cc >= 0 && cc < w) {
            const Key kq = K[GG::at(rr, cc)];
            const unsigned lq = key_label(kq);
            if (lq <= kLabOut && lq != lp) {
              const unsigned vq = key_level(kq);
              wv = vp > vq ? vp : vq;
              const unsigned a16 = lp == kLabOut ? 0xffffu : lp, b16 = lq == kLabOut ? 0xffffu : lq;
              key = a16 < b16 ? (a16 << 16) | b16 : (b16 << 16) | a16;
            }
          }
          if (key != pk[dir]) {
            flush(dir);
            pk[dir] = key;
            pw[dir] = wv;
          } else if (key != kPairEmpty && wv < pw[dir]) {
            pw[dir] = wv;
          }
        }
      }
#pragma unroll
      for (int dir = 0; dir < 4; dir++) flush(dir);
    }
  }
  __syncthreads();
  // ---- outputs: L (NaN cells keep their input bits), 16-bit labels, the
  // ring's starting levels and the edge list
  const int cells = w * h;
  {
    auto cell_out = [&](int r, int c, size_t g, float* o, unsigned short* l) {
      const Key k = K[GG::at(r, c)];
      const unsigned lv = key_label(k);
      *o = lv == kLabNan ? a.src[g] : unord_bits(key_level(k));
      *l = lv < kLabOut ? (unsigned short)lv : (lv == kLabOut ? kLab16Out : kLab16None);
    };
    const bool vec = ((a.nx & 3) == 0) && w == GG::TX &&
                     (reinterpret_cast<uintptr_t>(a.out) & 15) == 0;
    if (vec) {
      constexpr int Q = GG::TX / 4;
      for (int q = tid; q < h * Q; q += GG::THREADS) {
        const int r = q / Q, c = 4 * (q % Q);
        const size_t g = (size_t)(y0 + r) * a.nx + (size_t)(x0 + c);
        float4 o;
        ushort4 l;
        cell_out(r, c, g, &o.x, &l.x);
        cell_out(r, c + 1, g + 1, &o.y, &l.y);
        cell_out(r, c + 2, g + 2, &o.z, &l.z);
        cell_out(r, c + 3, g + 3, &o.w, &l.w);
        *reinterpret_cast<float4*>(a.out + g) = o;
        *reinterpret_cast<ushort4*>(a.lab + g) = l;
      }
    } else {
      for (int idx = tid; idx < cells; idx += GG::THREADS) {
        const int r = idx / w, c = idx % w;
        const size_t g = (size_t)(y0 + r) * a.nx + (size_t)(x0 + c);
        cell_out(r, c, g, &a.out[g], &a.lab[g]);
      }
    }
  }
  for (int k = tid; k < R; k += GG::THREADS) a.ringF[(size_t)t * kRingStride + k] = lds.F0[k];
  uint2* ed = a.edges + (size_t)t * kEdgeCap;
  for (int k = tid; k < kHashSlots; k += GG::THREADS) {
    if (lds.hkey[k] == kPairEmpty) continue;
    const unsigned i = atomicAdd(&lds.count, 1u);
    if (i < a.cap) ed[i] = make_uint2(lds.hkey[k], lds.hw[k]);
  }
  __syncthreads();
  if (tid == 0) {
    const unsigned n = lds.count;
    a.ecount[t] = n < a.cap ? n : a.cap;
    if (n > a.cap || lds.overflow) atomicOr(&a.q.sync->error, kErrOverflow);
  }
  __syncthreads();
}

}  // namespace

__global__ __launch_bounds__(GG::THREADS) void graph_label(GraphArgs a) {
  __shared__ LabelLds lds;
  const int ntiles = a.ntx * a.nty;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) label_tile(a, t, lds);
}

// ---------------------------------------------------------------- B: ring graph
namespace {

constexpr int kRingThreads = 256;

struct RingLds {
  unsigned F[kRingStride];
  unsigned F0[kRingStride];
  unsigned D[kRingStride];
  uint2 e[kEdgeCap];
  unsigned bits;
};

// The tile being visited and its neighbour tiles' extents (no divisions
// per cell).
struct RingTile {
  int t, tx, ty, w, h;
  long long x0, y0;
  int wl, hl;  // width of the last tile column, height of the last tile row
  __device__ RingTile(const GraphArgs& a, int t_) : t(t_) {
    tx = t % a.ntx;
    ty = t / a.ntx;
    x0 = (long long)tx * GG::TX;
    y0 = (long long)ty * GG::TY;
    wl = (int)(a.nx - (long long)(a.ntx - 1) * GG::TX);
    hl = (int)(a.ny - (long long)(a.nty - 1) * GG::TY);
    w = tx == a.ntx - 1 ? wl : GG::TX;
    h = ty == a.nty - 1 ? hl : GG::TY;
  }
};

// Calls fn(index into ringF / ringD, kEdge* bit of the neighbour tile) for
// every grid neighbour of ring cell (r, c) that lies in another tile.
template <class Fn>
__device__ inline void for_outside(const GraphArgs& a, const RingTile& T, int r, int c, Fn fn) {
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int rr = r + kNbrDy[k], cc = c + kNbrDx[k];
    const int dx = cc < 0 ? -1 : (cc >= T.w ? 1 : 0), dy = rr < 0 ? -1 : (rr >= T.h ? 1 : 0);
    if (dx == 0 && dy == 0) continue;
    const int tx2 = T.tx + dx, ty2 = T.ty + dy;
    if (tx2 < 0 || tx2 >= a.ntx || ty2 < 0 || ty2 >= a.nty) continue;
    const int w2 = tx2 == a.ntx - 1 ? T.wl : GG::TX, h2 = ty2 == a.nty - 1 ? T.hl : GG::TY;
    const int r2 = dy < 0 ? rr + h2 : (dy > 0 ? rr - T.h : rr);
    const int c2 = dx < 0 ? cc + w2 : (dx > 0 ? cc - T.w : cc);
    const int k2 = ring_pos(w2, h2, r2, c2);
    unsigned bit = 0;
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (kNbrDx[j] == dx && kNbrDy[j] == dy) bit = 1u << j;
    fn((size_t)(ty2 * a.ntx + tx2) * kRingStride + k2, bit);
  }
}

// Neighbours' ring levels are read with plain loads: the pop's acquire
// fence makes their published values visible, a value written meanwhile
// can only be higher than the current one (a valid bound), and its writer
// queues this tile again.
__device__ unsigned ring_visit(const GraphArgs& a, int t, RingLds& lds, Sync* sync) {
  const int tid = threadIdx.x;
  const unsigned long long c0 = wall_clock64();
  const RingTile T(a, t);
  const int R = ring_size(T.w, T.h);
  const unsigned ne = a.ecount[t];
  const unsigned* frt = a.ringF + (size_t)t * kRingStride;
  const unsigned* drt = a.ringD + (size_t)t * kRingStride;
  for (int k = tid; k < R; k += kRingThreads) {
    const unsigned v = frt[k], dk = drt[k];
    lds.F0[k] = v;
    lds.D[k] = dk;
    // the cross-tile relation F_k <- max(d_k, F_q) for the outside
    // neighbours q (their current values: upper bounds of their final ones)
    unsigned best = kOrdInf;
    int r, c;
    ring_rc(T.w, T.h, k, &r, &c);
    for_outside(a, T, r, c, [&](size_t q, unsigned) {
      const unsigned fq = a.ringF[q];
      best = fq < best ? fq : best;
    });
    const unsigned cand = dk > best ? dk : best;
    lds.F[k] = (dk != kOrdInf && cand < v) ? cand : v;
  }
  const uint2* ed = a.edges + (size_t)t * kEdgeCap;
  for (int k = tid; k < (int)ne; k += kRingThreads) lds.e[k] = ed[k];
  __syncthreads();
  const unsigned long long c1 = wall_clock64();
  // label edges to the fixed point (chaotic relaxation with LDS atomics)
  unsigned rounds = 0;
  for (;;) {
    int ch = 0;
    for (int k = tid; k < (int)ne; k += kRingThreads) {
      const uint2 e = lds.e[k];
      const unsigned la = e.x >> 16, lb = e.x & 0xffffu;
      const unsigned fa = lds.F[la], fb = lds.F[lb];
      const unsigned ca = e.y > fb ? e.y : fb;
      const unsigned cb = e.y > fa ? e.y : fa;
      if (ca < fa) {
        atomicMin(&lds.F[la], ca);
        ch = 1;
      }
      if (cb < fb) {
        atomicMin(&lds.F[lb], cb);
        ch = 1;
      }
    }
    rounds++;
    if (!__syncthreads_or(ch)) break;
  }
  const unsigned long long c2 = wall_clock64();
  // write back; queue the neighbour tiles whose cells a changed ring cell
  // can still lower
  unsigned bits = 0;
  unsigned* frw = a.ringF + (size_t)t * kRingStride;
  for (int k = tid; k < R; k += kRingThreads) {
    const unsigned v = lds.F[k];
    if (v == lds.F0[k]) continue;
    frw[k] = v;
    int r, c;
    ring_rc(T.w, T.h, k, &r, &c);
    for_outside(a, T, r, c, [&](size_t q, unsigned bit) {
      const unsigned dq = a.ringD[q], fq = a.ringF[q];
      const unsigned cand = dq > v ? dq : v;
      if (dq != kOrdInf && cand < fq) bits |= bit;
    });
  }
  bits = block_or(bits, &lds.bits);
  if (tid == 0) visit_stats(sync, rounds, c0, c1, c2);
  return bits;
}

}  // namespace

__global__ __launch_bounds__(kRingThreads) void graph_ring(GraphArgs a) {
  __shared__ RingLds lds;
  run_queue(a.q, GridTopo{a.ntx, a.nty},
            [&](int t, bool, unsigned) { return ring_visit(a, t, lds, a.q.stats_sink()); });
}

// ---------------------------------------------------------------- C: per-cell map
__global__ void graph_final(GraphArgs a) {
  if (__hip_atomic_load(&a.q.sync->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
    return;
  const long long n = a.nx * a.ny;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (long long)gridDim.x * blockDim.x) {
    const unsigned short l = a.lab[p];
    if (l >= kLab16Out) continue;
    const long long y = p / a.nx, x = p - y * a.nx;
    const size_t t = (size_t)(y / GG::TY) * a.ntx + (size_t)(x / GG::TX);
    const unsigned f = a.ringF[t * kRingStride + l];
    const unsigned o = ord_bits(a.out[p]);
    if (f > o) a.out[p] = unord_bits(f);
  }
}

// ---------------------------------------------------------------- host
hipError_t run_fill_graph(float* state, const float* src, const uint8_t* bc, int64_t nx,
                          int64_t ny, TileShape ts, hipStream_t stream, RunStats* stats) {
  (void)ts;  // the tile graph always uses 128 x 128 tiles
  if (nx <= 0 || ny <= 0) return hipSuccess;
  GraphArgs a{};
  a.src = src;
  a.bc = bc;
  a.out = state;
  a.nx = nx;
  a.ny = ny;
  a.ntx = (int)((nx + GG::TX - 1) / GG::TX);
  a.nty = (int)((ny + GG::TY - 1) / GG::TY);
  a.cap = (unsigned)g_edge_cap;
  const int ntiles = a.ntx * a.nty;
  const size_t n = (size_t)nx * (size_t)ny;
  Workspace ws;
  hipError_t e = ws.alloc(ntiles, a.ntx, a.nty, stream);
  if (e != hipSuccess) return e;
  a.q = ws.q;
  a.q.stats = stats ? 1 : 0;
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t lab_bytes = up(n * 2), ring_bytes = up((size_t)ntiles * kRingStride * 4);
  const size_t cnt_bytes = up((size_t)ntiles * 4), edge_bytes = (size_t)ntiles * kEdgeCap * 8;
  void* gbuf = nullptr;
  e = hipMallocAsync(&gbuf, lab_bytes + 2 * ring_bytes + cnt_bytes + edge_bytes, stream);
  if (e != hipSuccess) {
    (void)ws.finish(nullptr);
    return e;
  }
  char* p = static_cast<char*>(gbuf);
  a.lab = reinterpret_cast<unsigned short*>(p);
  a.ringF = reinterpret_cast<unsigned*>(p + lab_bytes);
  a.ringD = reinterpret_cast<unsigned*>(p + lab_bytes + ring_bytes);
  a.ecount = reinterpret_cast<unsigned*>(p + lab_bytes + 2 * ring_bytes);
  a.edges = reinterpret_cast<uint2*>(p + lab_bytes + 2 * ring_bytes + cnt_bytes);
  int ga = 0, gb = 0;
  e = persistent_grid((const void*)graph_label, GG::THREADS, ntiles, &ga);
  if (e == hipSuccess) e = persistent_grid((const void*)graph_ring, kRingThreads, ntiles, &gb);
  // one ring workgroup per CU: the chain of dependent visits, not the
  // number of workgroups, bounds pass B, and every idle workgroup spins on
  // the queue (more of them only slow the memory the visits read)
  if (e == hipSuccess) {
    int dev = 0, cus = 0;
    e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess && cus > 0 && gb > cus) gb = cus;
  }
  if (e == hipSuccess) {
    graph_label<<<ga, GG::THREADS, 0, stream>>>(a);
    graph_ring<<<gb, kRingThreads, 0, stream>>>(a);
    long long blocks = ((long long)n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    graph_final<<<(int)blocks, 256, 0, stream>>>(a);
    e = hipGetLastError();
  }
  // the work-queue kernel runs only if a tile overflowed its edge table
  if (e == hipSuccess)
    e = run_fill_gated(state, src, bc, nx, ny, kModeFill, TileShape{2, 2}, stream, stats,
                       &ws.q.sync->error, kErrOverflow);
  if (stats && e == hipSuccess) {
    RunStats g;
    hipError_t f = ws.finish(&g);
    if (f == hipSuccess && !(g.error & kErrOverflow)) {
      g.grid = gb;
      g.tile_x = GG::TX;
      g.tile_y = GG::TY;
      *stats = g;
    }
    e = f;
  } else {
    hipError_t f = ws.finish(nullptr);
    if (e == hipSuccess) e = f;
  }
  hipError_t f = hipFreeAsync(gbuf, stream);
  return e != hipSuccess ? e : f;
}

}  // namespace tt
