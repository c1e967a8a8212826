// Spill indices of fillsinks from its final levels by connected components
// (the forward half of the fillsinks adjoint; no reference counterpart).
//
// The spill key of a cell is its lexicographic (level, index) minimax over
// paths from the seeds (fill_visit.h, kModeSpill): the level half is the
// fill level F, the index half the cell whose DEM value F copies. With F
// known, most index halves follow from F without any relaxation:
//
//  * a surface cell (F == dem, bitwise) that is a bc cell or has a
//    neighbour at a lower level is a "source": its own spill cell;
//  * a submerged cell (F > dem) takes the least key of its neighbours, all
//    of which lie at its level or above, so 8-neighbouring submerged cells
//    share their level, and every cell of an 8-connected component of
//    submerged cells has the same spill index: the least index among the
//    component's equal-level neighbours' keys;
//  * a surface cell with no lower neighbour ("flat", a plateau at its own
//    level) is settled only by the relaxation.
//
// The labelling below writes the exact index of every source, and for each
// component the least index of its equal-level source neighbours. That is
// exact unless a flat cell offers the component a smaller key; every value
// is an upper bound of the key (an index some path reaches, or none), so
// the spill relaxation (fill_visit, kModeSpill) reaches the unique fixed
// point from it, and only tiles holding flat cells can violate it at the
// start: those are marked dirty and relaxed first.
//
// Steps, each a pass over tiles or cells that follows the previous one:
//   spill_label_tile  classes, union-find of the tile's submerged cells in
//                     LDS (over row runs), each local root's least source
//                     candidate; the spill array holds the union-find
//                     parents (grid indices) of submerged cells, cmin the
//                     candidates at local roots and a class marker at every
//                     other cell (kNotRoot, kNotSubmerged)
//   spill_border_tile unions across the tile's top and left borders, in
//                     the spill array (lock-free, linking the larger root
//                     under the smaller)
//   spill_root_cell   local roots: the path to the root compressed, the
//                     candidate folded into the root's
//   spill_final_cell  other submerged cells: spill := the root's candidate
//                     (their parent is a local root, which now points at
//                     the root)
//   spill_final_root  local roots: spill := the root's candidate
// A component spanning many tiles leaves a chain of local roots after the
// border unions; compressing it once per local root, rather than walking
// it from every cell, keeps the passes at a few loads per cell.
#pragma once

#include "fill_visit.h"

namespace tt {

constexpr unsigned kNoIndex = 0xffffffffu;    // no spill index (int -1)
constexpr unsigned kNotRoot = 0xfffffffeu;    // cmin of a submerged cell that is no local root
constexpr unsigned kNotSubmerged = 0xfffffffdu;  // cmin of every other cell
constexpr unsigned kInertLevel = 0xffffffffu;  // order image above every real level

// Cell classes.
enum : unsigned char {
  kClsInert = 0,       // NaN DEM or outside the grid
  kClsSubmerged = 1,   // F > dem
  kClsSource = 2,      // surface, bc or a lower neighbour
  kClsFlat = 3         // surface, neither
};

struct SpillView {
  const float* F;    // final fill levels
  const float* dem;
  const uint8_t* bc;
  int* spill;        // output; union-find parents of submerged cells before the last step
  unsigned* cmin;    // per cell: least candidate index (meaningful at roots)
  unsigned* dirty;   // per relaxation tile: holds a flat cell
  long long nx, ny;
  int rtx, rty;      // relaxation tile size in cells
  int rntx;          // relaxation tiles along x
};

// ---------------------------------------------------------------- local union-find
// 16-bit parents, two per 32-bit LDS word; every access is a 32-bit atomic
// on the word, so concurrent updates of the two halves never lose one.
__device__ inline unsigned uf_get(unsigned* par, unsigned i) {
  const unsigned w = __hip_atomic_load(&par[i >> 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  return (i & 1u) ? (w >> 16) : (w & 0xffffu);
}

// Sets parent(i) := to if parent(i) == expect; returns whether it did.
__device__ inline bool uf_cas(unsigned* par, unsigned i, unsigned expect, unsigned to) {
  const int sh = (i & 1u) ? 16 : 0;
  unsigned old = __hip_atomic_load(&par[i >> 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  for (;;) {
    if (((old >> sh) & 0xffffu) != expect) return false;
    const unsigned nw = (old & ~(0xffffu << sh)) | (to << sh);
    if (__hip_atomic_compare_exchange_strong(&par[i >> 1], &old, nw, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
      return true;
  }
}

// Root of i with path halving: a node whose parent is not a root gets its
// grandparent as parent (only if its parent is unchanged meanwhile), so
// the chains that unions in row order would build stay short.
__device__ inline unsigned uf_find(unsigned* par, unsigned i) {
  for (;;) {
    const unsigned p = uf_get(par, i);
    if (p == i) return i;
    const unsigned g = uf_get(par, p);
    if (g == p) return p;
    uf_cas(par, i, p, g);
    i = g;
  }
}

__device__ inline void uf_unite(unsigned* par, unsigned a, unsigned b) {
  for (;;) {
    a = uf_find(par, a);
    b = uf_find(par, b);
    if (a == b) return;
    if (a < b) {
      const unsigned t = a;
      a = b;
      b = t;
    }
    if (uf_cas(par, a, a, b)) return;  // a still a root: linked under b
  }
}

// ---------------------------------------------------------------- global union-find
__device__ inline int gp_load(const int* P, long long x) {
  return __hip_atomic_load(&P[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving (a non-root only ever gets an ancestor
// stored; roots change only by the CAS in gp_unite).
__device__ inline int gp_find(int* P, int x) {
  for (;;) {
    const int p = gp_load(P, x);
    if (p == x) return x;
    const int g = gp_load(P, p);
    if (g == p) return p;
    __hip_atomic_store(&P[x], g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = g;
  }
}

// Root of x without writes (once no union runs).
__device__ inline int gp_root(const int* P, int x) {
  for (;;) {
    const int p = gp_load(P, x);
    if (p == x) return x;
    x = p;
  }
}

__device__ inline void gp_unite(int* P, int a, int b) {
  for (;;) {
    a = gp_find(P, a);
    b = gp_find(P, b);
    if (a == b) return;
    if (a < b) {
      const int t = a;
      a = b;
      b = t;
    }
    int expect = a;
    if (__hip_atomic_compare_exchange_strong(&P[a], &expect, b, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      return;
  }
}

// ---------------------------------------------------------------- tile labelling
// The tile is processed by rows: a wavefront takes one 64-cell segment of a
// row at a time (lane = column), so the submerged cells of a segment are a
// 64-bit ballot and a cell's run start is a bit scan of it.
template <int TX, int TY>
struct SpillLabelLds {
  static constexpr int QX = TX / 64;      // 64-cell segments per row
  static constexpr int P2 = TX + 4;       // pitch of the tile + 2-cell halo (levels)
  static constexpr int C2 = (TY + 4) * P2;
  static constexpr int P1 = TX + 2;       // pitch of the tile + 1-cell halo (classes)
  static constexpr int C1 = (TY + 2) * P1;
  static_assert(TX % 64 == 0 && TX * TY <= 65536, "whole segments, 16-bit local parents");
  union {
    unsigned lev[C2];        // levels, tile + 2-cell halo (classification)
    // unions to make, two 16-bit run starts each: at most one per cell (a
    // cell makes two only as a one-cell run, whose right neighbour makes none)
    unsigned task[TX * TY];
    unsigned cand[TX * TY];  // least candidate of each local root (afterwards)
  };
  unsigned char cls[C1];
  unsigned char eq[TX * TY];  // submerged tile cells: bit k = neighbour k is an equal-level source
  unsigned par[TX * TY / 2];  // 16-bit parents, two per word
  unsigned short starts[TX * TY / 2];  // the run starts (runs are separated by a cell)
  unsigned long long sub[TY * QX];  // submerged cells of each segment
  unsigned flat, ntask, nstart;
  __device__ static constexpr int l2(int r, int c) { return (r + 2) * P2 + (c + 2); }
  __device__ static constexpr int l1(int r, int c) { return (r + 1) * P1 + (c + 1); }
};

// Neighbour k = 0..7 of the eq bits: (dr, dc) in row-major order, centre skipped.
__device__ constexpr int kEqDr[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
__device__ constexpr int kEqDc[8] = {-1, 0, 1, -1, 1, -1, 0, 1};

// Run starts (a run: consecutive submerged cells of a row) of segment k of
// row r, and the column of the start of the run through submerged cell
// (r, c).
template <int TX, int TY>
__device__ inline unsigned long long run_starts(const SpillLabelLds<TX, TY>& s, int r, int k) {
  const unsigned long long m = s.sub[r * (TX / 64) + k];
  const unsigned long long carry = k > 0 ? (s.sub[r * (TX / 64) + k - 1] >> 63) : 0ull;
  return m & ~((m << 1) | carry);
}
// (r, c) must be submerged, so a start exists at or before c; the scan
// stops at the row's first segment regardless.
template <int TX, int TY>
__device__ inline int run_start(const SpillLabelLds<TX, TY>& s, int r, int c) {
  int k = c >> 6;
  unsigned long long st = run_starts(s, r, k) & (~0ull >> (63 - (c & 63)));
  while (st == 0ull && k > 0) st = run_starts(s, r, --k);  // began in an earlier segment
  return st ? 64 * k + 63 - __clzll(st) : 0;
}
template <int TX, int TY>
__device__ inline bool sub_at(const SpillLabelLds<TX, TY>& s, int r, int c) {
  return (s.sub[r * (TX / 64) + (c >> 6)] >> (c & 63)) & 1ull;
}

// Labels the tile at (x0, y0) (TX x TY cells, NT threads, all threads of the
// workgroup call it together).
// GLOBAL: the first of the whole-grid steps (a label tile lies inside a
// relaxation tile and only sets that tile's dirty flag, cleared
// beforehand, when it holds a flat cell).
// Otherwise the tile is a relaxation tile of its own and labelled alone:
// each component's cells get the component's least candidate as their
// spill index at once, exact for a component inside the tile, an upper
// bound for one that continues across the tile's edge; the dirty flag is
// written either way, set when the tile holds a flat cell or a submerged
// cell with a submerged neighbour outside the tile. Grid indices fit in
// 32 bits (nx * ny < 2^31).
template <int TX, int TY, int NT, bool GLOBAL>
__device__ void spill_label_tile(const SpillView& v, long long x0l, long long y0l,
                                 SpillLabelLds<TX, TY>& s) {
  using S = SpillLabelLds<TX, TY>;
  constexpr int QX = S::QX, NW = NT / 64;
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));  // opaque: no addresses hoisted out of a persistent loop
  const int lane = tid & 63, wave = tid >> 6;
  const int x0 = (int)x0l, y0 = (int)y0l, nx = (int)v.nx, ny = (int)v.ny;
  if (tid == 0) {
    s.flat = 0u;
    s.ntask = 0u;
    s.nstart = 0u;
  }
  // levels of the tile + 2-cell halo; first classes of the tile + 1-cell
  // halo: inert, submerged, or surface (| 8 for a bc cell). Every load of a
  // batch is issued before any value is used.
  auto put = [&](int r, int c, float d, float f, unsigned char b) {
    unsigned lev = kInertLevel;
    unsigned char cls = kClsInert;
    if (d == d) {  // false outside the grid too
      lev = ord_bits(f);
      cls = lev != ord_bits(d) ? kClsSubmerged : (unsigned char)(kClsSource | (b == 1 ? 8 : 0));
    }
    s.lev[S::l2(r, c)] = lev;
    if (r >= -1 && r <= TY && c >= -1 && c <= TX) s.cls[S::l1(r, c)] = cls;
  };
  auto al = [](const void* p, uintptr_t m) { return (reinterpret_cast<uintptr_t>(p) & m) == 0; };
  if ((nx & 3) == 0 && x0 + TX <= nx && al(v.dem, 15) && al(v.F, 15) && al(v.bc, 3)) {
    // the tile's columns four cells per lane (16-byte loads), rows -2..TY+1
    constexpr int Q = TX / 4, NV = (TY + 4) * Q, UV = (NV + 2 * NT - 1) / (2 * NT);
    for (int base = 0; base < NV; base += NT * UV) {
      float4 dv[UV], fv[UV];
      uchar4 bv[UV];
#pragma unroll
      for (int u = 0; u < UV; u++) {
        const int k = base + u * NT + tid, gy = y0 + k / Q - 2;
        const float nan = __builtin_nanf("");
        dv[u] = make_float4(nan, nan, nan, nan);
        fv[u] = make_float4(0.f, 0.f, 0.f, 0.f);
        bv[u] = make_uchar4(0, 0, 0, 0);
        if (k < NV && gy >= 0 && gy < ny) {
          const unsigned g = (unsigned)gy * (unsigned)nx + (unsigned)(x0 + 4 * (k % Q));
          dv[u] = *reinterpret_cast<const float4*>(v.dem + g);
          fv[u] = *reinterpret_cast<const float4*>(v.F + g);
          bv[u] = *reinterpret_cast<const uchar4*>(v.bc + g);
        }
      }
#pragma unroll
      for (int u = 0; u < UV; u++) {
        const int k = base + u * NT + tid;
        if (k >= NV) continue;
        const int r = k / Q - 2, c = 4 * (k % Q);
        put(r, c, dv[u].x, fv[u].x, bv[u].x);
        put(r, c + 1, dv[u].y, fv[u].y, bv[u].y);
        put(r, c + 2, dv[u].z, fv[u].z, bv[u].z);
        put(r, c + 3, dv[u].w, fv[u].w, bv[u].w);
      }
    }
    // the halo columns -2, -1, TX, TX + 1
    for (int k = tid; k < 4 * (TY + 4); k += NT) {
      const int r = k / 4 - 2, j = k % 4, c = j < 2 ? j - 2 : TX + j - 2;
      const int gx = x0 + c, gy = y0 + r;
      float d = __builtin_nanf(""), f = 0.f;
      unsigned char b = 0;
      if (gx >= 0 && gy >= 0 && gx < nx && gy < ny) {
        const unsigned g = (unsigned)gy * (unsigned)nx + (unsigned)gx;
        d = v.dem[g];
        f = v.F[g];
        b = v.bc[g];
      }
      put(r, c, d, f, b);
    }
  } else {
    // a row is QX segments plus the four halo columns (lanes 0..3 of one
    // more pass)
    constexpr int ROWS = TY + 4, PASSES = QX + 1, U = (ROWS + 2 * NW - 1) / (2 * NW);
    for (int r0 = wave * U; r0 < ROWS; r0 += NW * U) {
      float d[U][PASSES], f[U][PASSES];
      unsigned char b[U][PASSES];
#pragma unroll
      for (int u = 0; u < U; u++)
#pragma unroll
        for (int j = 0; j < PASSES; j++) {
          const int r = r0 + u - 2;
          const int c = j < QX ? 64 * j + lane : (lane < 2 ? lane - 2 : TX + lane - 2);
          const int gx = x0 + c, gy = y0 + r;
          d[u][j] = __builtin_nanf("");
          f[u][j] = 0.f;
          b[u][j] = 0;
          if (r0 + u < ROWS && (j < QX || lane < 4) && gx >= 0 && gy >= 0 && gx < nx &&
              gy < ny) {
            const unsigned g = (unsigned)gy * (unsigned)nx + (unsigned)gx;
            d[u][j] = v.dem[g];
            f[u][j] = v.F[g];
            b[u][j] = v.bc[g];
          }
        }
#pragma unroll
      for (int u = 0; u < U; u++)
#pragma unroll
        for (int j = 0; j < PASSES; j++) {
          const int r = r0 + u - 2;
          const int c = j < QX ? 64 * j + lane : (lane < 2 ? lane - 2 : TX + lane - 2);
          if (r0 + u >= ROWS || (j == QX && lane >= 4)) continue;
          put(r, c, d[u][j], f[u][j], b[u][j]);
        }
    }
  }
  __syncthreads();
  // The row loops below take U rows (or segments) of a wavefront at a time
  // and issue all their LDS loads before using any: the rows are
  // independent, and a wavefront that walked them one by one would wait
  // out each row's dependent loads.
  constexpr int U = (TY * QX) / NW >= 32 ? 8 : 4;
  // surface cells of the tile + 1-cell halo: source (bc, or a lower
  // neighbour) or flat
  {
    constexpr int NB = (TY + 2) * (QX + 1);
    bool any_flat = false;
    for (int w0 = wave; w0 < NB; w0 += NW * U) {
      unsigned char c0[U];
      unsigned lv[U], nb[U][8];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int w = w0 + u * NW, r = w / (QX + 1) - 1, j = w % (QX + 1);
        const int c = j < QX ? 64 * j + lane : (lane == 0 ? -1 : TX);
        c0[u] = kClsInert;
        lv[u] = 0u;
#pragma unroll
        for (int n = 0; n < 8; n++) nb[u][n] = kInertLevel;
        if (w < NB && (j < QX || lane < 2)) {
          c0[u] = s.cls[S::l1(r, c)];
          lv[u] = s.lev[S::l2(r, c)];
#pragma unroll
          for (int n = 0; n < 8; n++) nb[u][n] = s.lev[S::l2(r + kEqDr[n], c + kEqDc[n])];
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int w = w0 + u * NW, r = w / (QX + 1) - 1, j = w % (QX + 1);
        const int c = j < QX ? 64 * j + lane : (lane == 0 ? -1 : TX);
        if ((c0[u] & 7) != kClsSource) continue;  // inert, submerged, or outside this pass
        bool src = (c0[u] & 8) != 0;
#pragma unroll
        for (int n = 0; n < 8; n++) src |= nb[u][n] < lv[u];
        s.cls[S::l1(r, c)] = src ? kClsSource : kClsFlat;
        any_flat |= !src && r >= 0 && r < TY && c >= 0 && c < TX && w < NB;
      }
    }
    if (any_flat) s.flat = 1u;
  }
  __syncthreads();
  // equal-level source neighbours of the submerged tile cells, and the
  // submerged mask of each segment
  constexpr int NS = TY * QX;  // segments
  {
    bool crosses = false;
    for (int w0 = wave; w0 < NS; w0 += NW * U) {
      unsigned char c0[U], nc[U][8];
      unsigned lv[U], nl[U][8];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int w = w0 + u * NW, r = w / QX, c = 64 * (w % QX) + lane;
        c0[u] = kClsInert;
        lv[u] = 0u;
#pragma unroll
        for (int n = 0; n < 8; n++) {
          nc[u][n] = kClsInert;
          nl[u][n] = kInertLevel;
        }
        if (w < NS) {
          c0[u] = s.cls[S::l1(r, c)];
          lv[u] = s.lev[S::l2(r, c)];
#pragma unroll
          for (int n = 0; n < 8; n++) {
            nc[u][n] = s.cls[S::l1(r + kEqDr[n], c + kEqDc[n])];
            nl[u][n] = s.lev[S::l2(r + kEqDr[n], c + kEqDc[n])];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int w = w0 + u * NW, r = w / QX, c = 64 * (w % QX) + lane;
        if (w >= NS) continue;  // uniform over the wavefront
        const bool sub = c0[u] == kClsSubmerged;
        unsigned char m = 0;
#pragma unroll
        for (int n = 0; n < 8; n++) {
          const int rr = r + kEqDr[n], cc = c + kEqDc[n];
          if (sub && nc[u][n] == kClsSource && nl[u][n] == lv[u]) m |= (unsigned char)(1u << n);
          crosses |= sub && nc[u][n] == kClsSubmerged && (rr < 0 || cc < 0 || rr >= TY || cc >= TX);
        }
        s.eq[r * TX + c] = m;
        const unsigned long long mask = __ballot(sub);
        if (lane == 0) s.sub[w] = mask;
      }
    }
    if (!GLOBAL && crosses) s.flat = 1u;
  }
  __syncthreads();
  // levels no longer needed. Parents: every submerged cell points at the
  // start of its run (a root). The run starts and the unions to make are
  // collected in lists, so that the dependent LDS round trips of the unions
  // and finds run across the whole workgroup at once rather than one row
  // after another: unions of the runs of rows r - 1 and r that touch
  // (8-connected), once per overlap, at its first column, and at a run's
  // end where the touching run above does not reach over it (else the
  // overlap covered it). A segment's masks are wavefront-uniform words;
  // bit `lane` is the cell. One LDS atomic per list reserves the slots of
  // U segments.
  for (int w0 = wave; w0 < NS; w0 += NW * U) {
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned a[U];
    unsigned long long sm[U], b1[U], b2[U];
    bool vl[U], ll[U], rl[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int w = w0 + u * NW, r = w / QX, k = w % QX, c = 64 * k + lane, i = r * TX + c;
      sm[u] = b1[u] = b2[u] = 0ull;
      vl[u] = ll[u] = rl[u] = false;
      a[u] = (unsigned)i;
      if (w >= NS) continue;  // uniform over the wavefront
      const unsigned long long m = s.sub[w];
      const bool sub = (m >> lane) & 1ull;
      const unsigned long long ml = (m << 1) | (k > 0 ? s.sub[w - 1] >> 63 : 0ull);
      const unsigned long long mr = (m >> 1) | (k + 1 < QX ? s.sub[w + 1] << 63 : 0ull);
      const bool start = sub && !((ml >> lane) & 1ull), end = sub && !((mr >> lane) & 1ull);
      if (sub) a[u] = (unsigned)(r * TX + run_start(s, r, c));
      const unsigned hi = (unsigned)__shfl_down((int)a[u], 1);
      if (!(lane & 1)) s.par[i >> 1] = a[u] | (hi << 16);
      sm[u] = __ballot(start);
      if (r > 0 && sub) {
        const unsigned long long up = s.sub[w - QX];
        const unsigned long long ul = (up << 1) | (k > 0 ? s.sub[w - QX - 1] >> 63 : 0ull);
        const unsigned long long ur = (up >> 1) | (k + 1 < QX ? s.sub[w - QX + 1] << 63 : 0ull);
        const bool u0 = (up >> lane) & 1ull, u_l = (ul >> lane) & 1ull, u_r = (ur >> lane) & 1ull;
        vl[u] = u0 && (start || !u_l);
        ll[u] = start && u_l && !u0;
        rl[u] = end && u_r && !u0;
      }
      const int nl = (int)vl[u] + (int)ll[u] + (int)rl[u];
      b1[u] = __ballot(nl >= 1);
      b2[u] = __ballot(nl >= 2);
    }
    unsigned nst = 0u, ntk = 0u;
#pragma unroll
    for (int u = 0; u < U; u++) {
      nst += (unsigned)__popcll(sm[u]);
      ntk += (unsigned)(__popcll(b1[u]) + __popcll(b2[u]));
    }
    unsigned bs = 0u, bt = 0u;
    if (lane == 0) {
      if (nst) bs = atomicAdd(&s.nstart, nst);
      if (ntk) bt = atomicAdd(&s.ntask, ntk);
    }
    bs = (unsigned)__shfl((int)bs, 0);
    bt = (unsigned)__shfl((int)bt, 0);
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int w = w0 + u * NW, r = w / QX, c = 64 * (w % QX) + lane;
      if ((sm[u] >> lane) & 1ull)
        s.starts[bs + (unsigned)__popcll(sm[u] & below)] = (unsigned short)a[u];
      unsigned t = bt + (unsigned)(__popcll(b1[u] & below) + __popcll(b2[u] & below));
      if (vl[u]) s.task[t++] = a[u] | ((unsigned)((r - 1) * TX + run_start(s, r - 1, c)) << 16);
      if (ll[u]) s.task[t++] = a[u] | ((unsigned)((r - 1) * TX + run_start(s, r - 1, c - 1)) << 16);
      if (rl[u]) s.task[t++] = a[u] | ((unsigned)((r - 1) * TX + run_start(s, r - 1, c + 1)) << 16);
      bs += (unsigned)__popcll(sm[u]);
      bt += (unsigned)(__popcll(b1[u]) + __popcll(b2[u]));
    }
  }
  __syncthreads();
  const int ntask = (int)s.ntask, nstart = (int)s.nstart;
  for (int t = tid; t < ntask; t += NT) {
    const unsigned tk = s.task[t];
    uf_unite(s.par, tk & 0xffffu, tk >> 16);
  }
  __syncthreads();
  // run starts point straight at their root (no union runs any more, so
  // concurrent updates all store the root); the task words become the
  // candidates
  for (int t = tid; t < nstart; t += NT) {
    const unsigned i = s.starts[t];
    const unsigned root = uf_find(s.par, i);
    for (;;) {
      const unsigned p = uf_get(s.par, i);
      if (p == root || uf_cas(s.par, i, p, root)) break;
    }
  }
  for (int i = tid; i < TX * TY; i += NT) s.cand[i] = kNoIndex;
  __syncthreads();
  // root of submerged cell i: its run start's parent (the parents no
  // longer change: plain loads)
  auto par_of = [&](unsigned i) {
    const unsigned wd = s.par[i >> 1];
    return (i & 1u) ? (wd >> 16) : (wd & 0xffffu);
  };
  // each submerged cell's least candidate onto its local root
  for (int i = tid; i < TX * TY; i += NT) {
    const unsigned m = s.eq[i];
    if (!m) continue;
    const int r = i / TX, c = i % TX;
    unsigned best = kNoIndex;
#pragma unroll
    for (int n = 0; n < 8; n++)
      if (m & (1u << n)) {
        const unsigned g =
            (unsigned)(y0 + r + kEqDr[n]) * (unsigned)nx + (unsigned)(x0 + c + kEqDc[n]);
        best = g < best ? g : best;
      }
    atomicMin(&s.cand[par_of(par_of((unsigned)i))], best);
  }
  __syncthreads();
  // write back: spill = own index (sources), none (flat, inert), and for a
  // submerged cell the local root's grid index with cmin = the candidate at
  // local roots, a class marker elsewhere (GLOBAL), or the component's
  // candidate
  for (int w0 = wave; w0 < NS; w0 += NW * U) {
    unsigned char cl[U];
    unsigned root[U], cd[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int w = w0 + u * NW, r = w / QX, c = 64 * (w % QX) + lane;
      cl[u] = w < NS ? s.cls[S::l1(r, c)] : (unsigned char)kClsInert;
      root[u] = cl[u] == kClsSubmerged ? par_of((unsigned)(r * TX + c)) : 0u;
    }
#pragma unroll
    for (int u = 0; u < U; u++) root[u] = cl[u] == kClsSubmerged ? par_of(root[u]) : 0u;
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int w = w0 + u * NW, r = w / QX, c = 64 * (w % QX) + lane;
      cd[u] = cl[u] == kClsSubmerged ? s.cand[GLOBAL ? (unsigned)(r * TX + c) : root[u]] : 0u;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int w = w0 + u * NW, r = w / QX, c = 64 * (w % QX) + lane;
      const int gx = x0 + c, gy = y0 + r;
      if (w >= NS || gx >= nx || gy >= ny) continue;
      const unsigned g = (unsigned)gy * (unsigned)nx + (unsigned)gx;
      const unsigned i = (unsigned)(r * TX + c);
      int sp;
      unsigned cm = kNotSubmerged;
      if (cl[u] == kClsSubmerged) {
        if (GLOBAL) {
          sp = (int)((unsigned)(y0 + (int)root[u] / TX) * (unsigned)nx +
                     (unsigned)(x0 + (int)root[u] % TX));
          cm = root[u] == i ? cd[u] : kNotRoot;
        } else {
          sp = (int)cd[u];
        }
      } else {
        sp = cl[u] == kClsSource ? (int)g : -1;
      }
      v.spill[g] = sp;
      if (GLOBAL) v.cmin[g] = cm;
    }
  }
  if (tid == 0 && (!GLOBAL || s.flat)) {
    const long long rt = (y0l / v.rty) * v.rntx + x0l / v.rtx;
    __hip_atomic_store(&v.dirty[rt], s.flat, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Unions across the tile's top border (its row 0 against the row above)
// and left border (its column 0 against the column to the left): every
// cross-tile pair of 8-neighbours is taken by the lower or the right tile.
// Along a border, runs of submerged cells are united once per overlap (at
// its first cell) and at the runs' diagonal ends, as within a tile (runs
// cut at the tile's corners still cover every pair).
template <int TX, int TY, int NT>
__device__ void spill_border_tile(const SpillView& v, long long x0, long long y0) {
  auto sub = [&](long long x, long long y) {
    return x >= 0 && y >= 0 && x < v.nx && y < v.ny && v.cmin[(size_t)y * v.nx + x] != kNotSubmerged;
  };
  for (int k = threadIdx.x; k < TX + TY; k += NT) {
    const bool top = k < TX;
    // own line: the tile's row 0 (along x) or column 0 (along y); the other
    // line is the row above or the column to the left
    const int i = top ? k : k - TX, len = top ? TX : TY;
    const long long gx = top ? x0 + i : x0, gy = top ? y0 : y0 + i;
    if (gx >= v.nx || gy >= v.ny || (top ? gy == 0 : gx == 0)) continue;
    if (!sub(gx, gy)) continue;
    // cells of the own and the other line at offset j along the border
    auto own = [&](int j) { return top ? sub(gx + j, gy) : sub(gx, gy + j); };
    auto oth = [&](int j) { return top ? sub(gx + j, gy - 1) : sub(gx - 1, gy + j); };
    auto link = [&](int j) {
      const long long qx = top ? gx + j : gx - 1, qy = top ? gy - 1 : gy + j;
      gp_unite(v.spill, (int)((size_t)gy * v.nx + gx), (int)((size_t)qy * v.nx + qx));
    };
    const bool start = i == 0 || !own(-1);
    const bool end = i == len - 1 || !own(1);
    if (oth(0) && (start || !oth(-1))) link(0);
    if (start && oth(-1)) link(-1);
    if (end && oth(1)) link(1);
  }
}

// Local root g: every node on its path gets the root as parent (once no
// union runs, every store to a node writes the same value, its root, so
// concurrent compressions agree), and its candidate folds into the root's.
// A local root that is not a global root receives no candidate, so its own
// cmin is read unchanged.
__device__ inline void spill_root_cell(const SpillView& v, size_t g, unsigned c) {
  if (c == kNotRoot || c == kNotSubmerged) return;
  const int root = gp_root(v.spill, (int)g);
  if (root == (int)g) return;
  for (int x = (int)g; x != root;) {
    const int p = gp_load(v.spill, x);
    __hip_atomic_store(&v.spill[x], root, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = p;
  }
  if (c != kNoIndex) atomicMin(&v.cmin[root], c);
}

// Submerged cell g that is no local root: spill := the candidate of its
// root, its local root's parent (local roots are only read in this step,
// and a cell's own word only by its own thread).
__device__ inline void spill_final_cell(const SpillView& v, size_t g, unsigned c) {
  if (c != kNotRoot) return;
  v.spill[g] = (int)v.cmin[v.spill[v.spill[g]]];
}

// Local root g (the last step): spill := its root's candidate (its parent
// is the root, or itself).
__device__ inline void spill_final_root(const SpillView& v, size_t g, unsigned c) {
  if (c == kNotRoot || c == kNotSubmerged) return;
  v.spill[g] = (int)v.cmin[v.spill[g]];
}

}  // namespace tt
