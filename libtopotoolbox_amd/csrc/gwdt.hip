// Costs of the gray-weighted distance transform (gwdt_computecosts):
// connected components of the flats (8-connected) with the largest depth
// (filled - original) of each, every flat's cost
// (max depth - own depth) + 0.1 and the component's label.
//
// Reference semantics, src/gwdt.c:83-274: one raster scan builds a
// union-find over the flats, unifying each flat with its flat neighbours
// among (i-1,j-1), (i,j-1), (i+1,j-1), (i-1,j) in that order; a merge keeps
// the neighbour's root unless the current pixel's root is strictly deeper.
// The label is that final root: a pixel of largest depth, and when several
// pixels share it, the one the scan's merge order picks.
//
// Components come from a tile-parallel union-find: each 64 x 16 tile is
// labelled in LDS, the tiles are joined across their borders in global
// memory, and every flat then learns its component, whose representative
// is its first pixel in scan order. The largest depth of each component,
// the number of pixels that reach it and the first of them are reduced per
// tile root in LDS, then over the tile roots of each component. A
// component with one pixel of largest depth is labelled with it, and the
// last pass writes its costs and labels. The other components ("tie"
// components: two or more deepest pixels, or depths mixing NaN and
// numbers) have all their pixels collected as keys component << cb |
// pixel, and their label is replayed from the few scan steps that can
// change which deepest pixel is the root, sorted into scan order, one
// component per thread:
//   birth     a flat with no flat scanned neighbour starts a set;
//   merge     a flat whose scanned neighbours lie in two different sets
//             (only possible when (i,j-1) is not flat, (i+1,j-1) is, and
//             (i-1,j-1) or (i-1,j) is); the set met second in the
//             neighbour order keeps its deepest root, else the first,
//             else the pixel itself if it is one of the deepest;
//   first     the first deepest pixel that joins a set grown from one
//             birth makes it a set with a deepest root, if it has none.
// All other scan steps leave every deepest root where it is, since only a
// strictly deeper pixel can displace a root. Sets are tracked by their
// birth pixel: each flat points at one flat scanned neighbour (its chain
// parent), and the chains end at births; the chains are resolved over the
// keys of the tie components only.
//
// NaN depths (e.g. flats of -INF cells, where filled - original is NaN):
// the scan's "strictly deeper" test is false whenever a NaN is involved, so
// a component whose depths are all NaN behaves as one whose depths all tie,
// and is replayed the same way with every pixel a deepest one. A component
// mixing NaN and numbers has no such order; its scan is replayed in full,
// pixel by pixel, with the reference's merge rule.
#include <hipcub/hipcub.hpp>

#include "kernels.h"

namespace tt {
namespace {

constexpr unsigned kNone = 0xffffffffu;
// component kinds (bits): depths that are numbers, depths that are NaN
constexpr unsigned kHasNumber = 1u, kHasNaN = 2u, kMixed = 3u;
// tile of the component labelling: one wavefront lane per column
constexpr unsigned kTileW = 64, kTileH = 16;

__device__ __forceinline__ unsigned gw_ord(float f) {
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float gw_unord(unsigned u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__device__ __forceinline__ unsigned ld(const unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st(unsigned* p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving; every pointer points at a smaller or equal
// index, so concurrent halving and linking only ever store ancestors.
__device__ unsigned find_halving(unsigned* P, unsigned x) {
  for (;;) {
    const unsigned px = ld(P + x);
    if (px == x) return x;
    const unsigned ppx = ld(P + px);
    if (ppx != px) st(P + x, ppx);
    x = px;
  }
}

// Cached load, for guards that tolerate a stale word: a word that only
// grows (or only shrinks, or only gains bits) read stale only lets an
// atomic through that was not needed.
__device__ __forceinline__ unsigned ldc(const unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Links the sets of a and b, the larger root under the smaller.
__device__ void unite(unsigned* P, unsigned a, unsigned b) {
  for (;;) {
    a = find_halving(P, a);
    b = find_halving(P, b);
    if (a == b) return;
    if (a < b) {
      const unsigned t = a;
      a = b;
      b = t;
    }
    if (atomicCAS(P + a, a, b) == a) return;
  }
}

// The same in LDS, for the tile labelling: pointers only decrease, and a
// link (atomicMin on the larger root) that finds the root already linked
// goes on to join the root it was linked to.
__device__ __forceinline__ unsigned lds_find(unsigned* lab, unsigned x) {
  for (;;) {
    const unsigned px = __hip_atomic_load(lab + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (px == x) return x;
    x = px;
  }
}

__device__ void lds_union(unsigned* lab, unsigned a, unsigned b) {
  for (;;) {
    a = lds_find(lab, a);
    b = lds_find(lab, b);
    if (a == b) return;
    if (a > b) {
      const unsigned t = a;
      a = b;
      b = t;
    }
    const unsigned old = atomicMin(lab + b, a);
    if (old == b) return;
    b = old;
  }
}

// Flat test on the labels (kNone marks a pixel that is not a flat).
struct Grid {
  const unsigned* L;
  unsigned nx, ny;
  __device__ bool flat(long long i, long long j) const {
    return i >= 0 && i < nx && j >= 0 && j < ny &&
           L[(unsigned)j * nx + (unsigned)i] != kNone;
  }
};

// Merge candidate at (i, j): the neighbour of the first set (A: (i-1,j-1)
// if flat, else (i-1,j)) and whether (i-1,j-1) is flat, which decides the
// order in which the scan meets the two sets. Returns false if the pixel
// cannot join two sets.
__device__ __forceinline__ bool merge_site(const Grid& g, long long i, long long j,
                                           unsigned* a_rep, bool* nw) {
  if (g.flat(i, j - 1) || !g.flat(i + 1, j - 1)) return false;
  *nw = g.flat(i - 1, j - 1);
  if (!*nw && !g.flat(i - 1, j)) return false;
  *a_rep = *nw ? (unsigned)((j - 1) * g.nx + (i - 1)) : (unsigned)(j * g.nx + (i - 1));
  return true;
}

// One flat scanned neighbour of (i, j), or the pixel itself if it has none
// (a birth). The trees these links form are the chain groups; their
// smallest pixel, the root, is their birth.
__device__ __forceinline__ unsigned scanned_parent(const Grid& g, long long i, long long j) {
  const unsigned p = (unsigned)(j * g.nx + i);
  if (g.flat(i - 1, j)) return p - 1;
  if (g.flat(i, j - 1)) return p - g.nx;
  if (g.flat(i - 1, j - 1)) return p - g.nx - 1;
  if (g.flat(i + 1, j - 1)) return p - g.nx + 1;
  return p;
}

__device__ __forceinline__ float depth_at(const float* original, const float* filled,
                                          unsigned p) {
  return filled[p] - original[p];
}

// Whether p is one of the deepest pixels of its component of kind k.
__device__ __forceinline__ bool deepest_pixel(unsigned k, float d, unsigned wbits) {
  return k == kHasNaN ? d != d : d == gw_unord(wbits);  // float ==: -0 and +0 tie
}

// Arrays of the labelling. Component arrays are indexed by the
// component's representative pixel and initialised for every tile root (a
// superset of the representatives). Each tile root also has an entry in
// its tile's list (tile t's entries from t * cap, TC[t] of them) with its
// tile's share of the reductions.
struct Ws {
  unsigned* L;   // per pixel: tile root (global index); kNone off flats
  unsigned* G;   // per tile root: union-find across tiles; then the tie keys' positions
  unsigned* Kc;  // per component: kind bits
  unsigned* Wc;  // per component: largest depth (order-preserving bits)
  unsigned* Cc;  // per component: number of deepest pixels (at least up to 2)
  unsigned* Fc;  // per component: first deepest pixel, then the label
  unsigned* RL;  // per entry: the tile root
  unsigned* RH;  // per entry: its component
  unsigned* RK;  // per entry: kind bits of the root's pixels in its tile
  unsigned* RW;  // per entry: their largest depth
  unsigned* RC;  // per entry: how many of them reach it (NaN pixels if all are NaN)
  unsigned* RF;  // per entry: the first of those
  unsigned* TC;  // per tile: its number of entries
  unsigned cap;  // entries per tile: 32 * ceil(tile rows / 2) bounds its components
  unsigned ntx, nty;
  unsigned* cnt; // [0] tie keys, [1] consistency errors
};

// Tile labelling and the tile's share of every reduction: 64 x kTileH
// pixels per workgroup, one lane per column, wavefront w on rows w, w + 4,
// ... Each row's runs are labelled with their first pixel from a ballot;
// runs of neighbouring rows are joined by the decision tree below (a run
// start joins the pixel above it, or both upper diagonals; a later pixel
// of a run joins its upper right neighbour only where a run above starts
// there). Edges to pixels of other tiles are left to gw_border. Per tile
// root, LDS atomics then gather the kind bits and largest depth of its
// pixels, then how many reach that depth and the first of them. Every flat
// gets its tile root as a global index; tile roots initialise their
// union-find and component entries and write their list entry.
__global__ __launch_bounds__(256) void gw_local(const int32_t* flats, const float* original,
                                                const float* filled, unsigned nx, unsigned ny,
                                                Ws w) {
  __shared__ unsigned lab[kTileH * kTileW];
  __shared__ unsigned sK[kTileH * kTileW], sW[kTileH * kTileW], sC[kTileH * kTileW],
      sF[kTileH * kTileW];
  __shared__ unsigned long long fbits[kTileH];
  __shared__ unsigned nroots;
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const unsigned x0 = blockIdx.x * kTileW, x = x0 + lane;
  constexpr unsigned R = kTileH / 4;  // rows per wavefront
  for (unsigned ty = blockIdx.y; ty < w.nty; ty += gridDim.y) {
    const unsigned y0 = ty * kTileH;
    int fv[R];
    float dv[R];
#pragma unroll
    for (unsigned k = 0; k < R; k++) {
      const unsigned y = y0 + wave + 4 * k;
      fv[k] = 0;
      dv[k] = 0.0f;
      if (x < nx && y < ny) {
        const unsigned p = y * nx + x;
        fv[k] = flats[p];
        dv[k] = depth_at(original, filled, p);
      }
    }
#pragma unroll
    for (unsigned k = 0; k < R; k++) {
      const unsigned r = wave + 4 * k, me = r * kTileW + lane;
      const bool f = fv[k] & 1;
      const unsigned long long fb = __ballot(f);
      if (lane == 0) fbits[r] = fb;
      const unsigned long long left = ~fb & ((1ull << lane) - 1ull);
      const unsigned st = left ? 64u - (unsigned)__clzll((long long)left) : 0u;  // run start
      lab[me] = f ? r * kTileW + st : kNone;
      if (f && st == lane) {  // only run starts can become roots
        sK[me] = 0u;
        sW[me] = 0u;  // below every ordered float
        sC[me] = 0u;
        sF[me] = kNone;
      }
    }
    if (threadIdx.x == 0) nroots = 0u;
    __syncthreads();
#pragma unroll
    for (unsigned k = 0; k < R; k++) {
      const unsigned r = wave + 4 * k;
      if (r == 0) continue;
      const unsigned long long fb = fbits[r], up = fbits[r - 1];
      if (!((fb >> lane) & 1ull)) continue;
      const unsigned me = r * kTileW + lane;
      const bool u = (up >> lane) & 1ull;
      const bool ul = lane > 0 && ((up >> (lane - 1)) & 1ull);
      const bool ur = lane < 63 && ((up >> (lane + 1)) & 1ull);
      const bool lf = lane > 0 && ((fb >> (lane - 1)) & 1ull);
      if (lf) {
        if (ur && !u) lds_union(lab, me, me - kTileW + 1);
      } else if (u) {
        lds_union(lab, me, me - kTileW);
      } else {
        if (ul) lds_union(lab, me, me - kTileW - 1);
        if (ur) lds_union(lab, me, me - kTileW + 1);
      }
    }
    __syncthreads();
    // Per tile root, in LDS: kind bits and largest depth of its pixels,
    // then how many reach that depth (or are NaN, where all are) and the
    // first of them.
    unsigned rt[R];
#pragma unroll
    for (unsigned k = 0; k < R; k++) {
      const unsigned me = (wave + 4 * k) * kTileW + lane;
      rt[k] = (fv[k] & 1) ? lds_find(lab, me) : kNone;
    }
#pragma unroll
    for (unsigned k = 0; k < R; k++) {
      const unsigned c = rt[k];
      if (c == kNone) continue;
      const float d = dv[k];
      if (d != d) {
        atomicOr(sK + c, kHasNaN);
      } else {
        atomicOr(sK + c, kHasNumber);
        atomicMax(sW + c, gw_ord(d));
      }
    }
    __syncthreads();
#pragma unroll
    for (unsigned k = 0; k < R; k++) {
      const unsigned c = rt[k];
      if (c == kNone) continue;
      const unsigned kind = sK[c];
      if (kind == kMixed || !deepest_pixel(kind, dv[k], sW[c])) continue;
      atomicAdd(sC + c, 1u);
      atomicMin(sF + c, (y0 + wave + 4 * k) * nx + x);
    }
    __syncthreads();
    const unsigned tile = ty * w.ntx + blockIdx.x;
#pragma unroll
    for (unsigned k = 0; k < R; k++) {
      const unsigned r = wave + 4 * k, y = y0 + r;
      if (x >= nx || y >= ny) continue;
      const unsigned p = y * nx + x;
      if (rt[k] == kNone) {
        w.L[p] = kNone;
        continue;
      }
      const unsigned c = rt[k];
      w.L[p] = (y0 + c / kTileW) * nx + x0 + c % kTileW;
      if (c != r * kTileW + lane) continue;  // not a tile root
      w.G[p] = p;
      w.Kc[p] = 0u;
      w.Wc[p] = 0u;
      w.Cc[p] = 0u;
      w.Fc[p] = kNone;
      const unsigned slot = atomicAdd(&nroots, 1u);
      if (slot < w.cap) {
        const unsigned e = tile * w.cap + slot;
        w.RL[e] = p;
        w.RK[e] = sK[c];
        w.RW[e] = sW[c];
        w.RC[e] = sC[c];
        w.RF[e] = sF[c];
      } else {
        atomicOr(w.cnt + 1, 1u);  // more components than a tile can hold: cannot happen
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) w.TC[tile] = nroots < w.cap ? nroots : w.cap;
    __syncthreads();  // LDS is reused by the next tile
  }
}

// Edges between tiles, each joined through the tile roots of its ends.
// Runs of a row are one set once a tile's left edge is joined, so the
// rules skip edges another join already covers:
//   left edge (first column): joins the left neighbour, unless the pixels
//     above both ends are flats too (the row above joined the same sets);
//   upper left edge (first column): only where neither the left neighbour
//     nor the pixel above is a flat;
//   upper right edge (last column): only where neither the right neighbour
//     nor the pixel above is a flat;
//   first row: the run decision tree of the tile labelling against the
//     row above, with the neighbours of the adjacent tiles.
__global__ __launch_bounds__(128) void gw_border(unsigned nx, unsigned ny, Ws w) {
  const unsigned t = threadIdx.x;
  unsigned r, c;
  if (t < kTileW) {
    r = 0;
    c = t;
  } else if (t < kTileW + kTileH - 1) {
    r = t - kTileW + 1;
    c = 0;
  } else if (t < kTileW + 2 * (kTileH - 1)) {
    r = t - (kTileW + kTileH - 1) + 1;
    c = kTileW - 1;
  } else {
    return;
  }
  const unsigned x = blockIdx.x * kTileW + c;
  if (x >= nx) return;
  for (unsigned ty = blockIdx.y; ty < w.nty; ty += gridDim.y) {
    const unsigned y = ty * kTileH + r;
    if (y >= ny) continue;
    const unsigned p = y * nx + x;
    const unsigned lp = w.L[p];
    if (lp == kNone) continue;
    auto fl = [&](unsigned q) { return w.L[q] != kNone; };
    auto join = [&](unsigned q) { unite(w.G, lp, w.L[q]); };
    if (c == 0 && x > 0) {
      if (fl(p - 1)) {
        if (!(r >= 1 && fl(p - nx) && fl(p - nx - 1))) join(p - 1);
      } else if (r >= 1 && fl(p - nx - 1) && !fl(p - nx)) {
        join(p - nx - 1);
      }
    }
    if (c == kTileW - 1 && r >= 1 && x + 1 < nx && fl(p - nx + 1) && !fl(p + 1) && !fl(p - nx))
      join(p - nx + 1);
    if (r == 0 && y > 0) {
      const bool lf = x > 0 && fl(p - 1), u = fl(p - nx);
      const bool ul = x > 0 && fl(p - nx - 1), ur = x + 1 < nx && fl(p - nx + 1);
      if (lf) {
        if (ur && !u) join(p - nx + 1);
      } else if (u) {
        join(p - nx);
      } else {
        if (ul) join(p - nx - 1);
        if (ur) join(p - nx + 1);
      }
    }
  }
}

// Loops over the list entries: one wavefront per tile.
#define GW_FOR_ENTRIES                                                                    \
  for (unsigned tile = blockIdx.x * 4u + (threadIdx.x >> 6); tile < w.ntx * w.nty;         \
       tile += gridDim.x * 4u)                                                            \
    for (unsigned e = tile * w.cap + (threadIdx.x & 63u), e_end = tile * w.cap + w.TC[tile]; \
         e < e_end; e += 64u)

// Component of every tile root (into RH: a root stored into G itself could
// be overwritten by another thread's path halving that read its old parent
// just before), and the component's kind bits and largest depth from its
// tile roots' shares. Each atomic is issued only where it changes the word
// (a large lake has thousands of tile roots, and atomics on one word queue
// up).
__global__ __launch_bounds__(256) void gw_roots(Ws w) {
  GW_FOR_ENTRIES {
    const unsigned r = w.RL[e], c = find_halving(w.G, r);
    w.RH[e] = c;
    const unsigned k = w.RK[e], wb = w.RW[e];
    if ((ldc(w.Kc + c) & k) != k) atomicOr(w.Kc + c, k);
    if (wb > ldc(w.Wc + c)) atomicMax(w.Wc + c, wb);
  }
}

// Number of deepest pixels (counted up to 2: only whether there are two or
// more matters) and the first of them per component, from the tile roots
// whose pixels reach the component's largest depth; guarded atomics as
// above.
__global__ __launch_bounds__(256) void gw_merge_deepest(Ws w) {
  GW_FOR_ENTRIES {
    const unsigned c = w.RH[e], kind = w.Kc[c];
    if (kind == kMixed) continue;
    if (kind == kHasNumber && w.RW[e] != w.Wc[c]) continue;
    if (ldc(w.Cc + c) < 2u) atomicAdd(w.Cc + c, w.RC[e]);
    if (w.RF[e] < ldc(w.Fc + c)) atomicMin(w.Fc + c, w.RF[e]);
  }
}

__device__ __forceinline__ float gw_cost(float label_depth, float depth) {
  const float gap = label_depth - depth;
  return (float)((double)gap + 0.1);
}

// Every pixel's outputs, one workgroup per tile: 0 and 0 for non-flats;
// for the flats of a component with one deepest pixel, that label and
// (root depth - own depth, in float) + 0.1 rounded from double, as the
// reference. Each tile root's component record is fetched once into LDS.
// The pixels of tie components are appended as keys component << cb |
// pixel (one atomic per wavefront), for the replay.
__global__ __launch_bounds__(256) void gw_final(const float* original, const float* filled,
                                                unsigned nx, unsigned ny, Ws w, float* costs,
                                                int64_t* conncomps, unsigned long long* keys,
                                                unsigned* parent, unsigned cb) {
  __shared__ unsigned sComp[kTileH * kTileW], sLab[kTileH * kTileW];
  __shared__ float sDep[kTileH * kTileW];
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const unsigned x0 = blockIdx.x * kTileW, x = x0 + lane;
  const unsigned n = nx * ny;
  constexpr unsigned R = kTileH / 4;
  for (unsigned ty = blockIdx.y; ty < w.nty; ty += gridDim.y) {
    const unsigned y0 = ty * kTileH, tile = ty * w.ntx + blockIdx.x;
    unsigned lv[R];
    float dv[R];
#pragma unroll
    for (unsigned k = 0; k < R; k++) {
      const unsigned y = y0 + wave + 4 * k;
      lv[k] = kNone;
      dv[k] = 0.0f;
      if (x < nx && y < ny) {
        const unsigned p = y * nx + x;
        lv[k] = w.L[p];
        dv[k] = depth_at(original, filled, p);
      }
    }
    for (unsigned e = tile * w.cap + threadIdx.x, e_end = tile * w.cap + w.TC[tile]; e < e_end;
         e += blockDim.x) {
      const unsigned r = w.RL[e], c = w.RH[e];
      const unsigned slot = (r / nx - y0) * kTileW + (r % nx - x0);
      const bool tie = w.Kc[c] == kMixed || w.Cc[c] >= 2u;
      const unsigned label = tie ? kNone : w.Fc[c];
      sComp[slot] = c;
      sLab[slot] = label;
      sDep[slot] = label < n ? depth_at(original, filled, label) : 0.0f;
    }
    __syncthreads();
    unsigned long long tm[R];
    unsigned par[R];
    unsigned ties = 0u;
#pragma unroll
    for (unsigned k = 0; k < R; k++) {
      const unsigned y = y0 + wave + 4 * k;
      bool tie = false;
      par[k] = kNone;
      if (x < nx && y < ny) {
        const unsigned p = y * nx + x, l = lv[k];
        if (l == kNone) {
          costs[p] = 0.0f;
          conncomps[p] = 0;
        } else {
          const unsigned slot = (l / nx - y0) * kTileW + (l % nx - x0);
          const unsigned label = sLab[slot];
          if (label < n) {
            costs[p] = gw_cost(sDep[slot], dv[k]);
            conncomps[p] = (int64_t)label;
          } else if (w.Kc[sComp[slot]] == kMixed || w.Cc[sComp[slot]] >= 2u) {
            tie = true;
          } else {
            atomicOr(w.cnt + 1, 1u);  // no deepest pixel: cannot happen
          }
        }
      }
      tm[k] = __ballot(tie);
      ties += (unsigned)__popcll(tm[k]);
      // chain parent of a tie pixel: the start of its run in this tile row,
      // or at the start, its first flat scanned neighbour (itself if none)
      const unsigned long long fb = __ballot(lv[k] != kNone);
      if (tie) {
        const unsigned long long left = ~fb & ((1ull << lane) - 1ull);
        const unsigned st = left ? 64u - (unsigned)__clzll((long long)left) : 0u;
        const unsigned p = y * nx + x;
        par[k] = st < lane ? p - (lane - st) : scanned_parent(Grid{w.L, nx, ny}, x, y);
      }
    }
    if (ties) {
      unsigned base = 0u;
      if (lane == 0) base = atomicAdd(w.cnt, ties);
      base = __shfl(base, 0);
#pragma unroll
      for (unsigned k = 0; k < R; k++) {
        if ((tm[k] >> lane) & 1ull) {
          const unsigned p = (y0 + wave + 4 * k) * nx + x;
          const unsigned slot = (lv[k] / nx - y0) * kTileW + (lv[k] % nx - x0);
          const unsigned q = base + (unsigned)__popcll(tm[k] & ((1ull << lane) - 1ull));
          keys[q] = ((unsigned long long)sComp[slot] << cb) | (unsigned long long)p;
          parent[q] = par[k];
        }
        base += (unsigned)__popcll(tm[k]);
      }
    }
    __syncthreads();  // LDS is reused by the next tile
  }
}

// ---- tie components, over their keys (nk of them, in append order) ----

#define GW_FOR_KEYS(n) \
  for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < (n); q += gridDim.x * blockDim.x)

struct Tie {
  const unsigned long long* keys;  // component << cb | pixel, in no particular order
  unsigned nk;
  unsigned n;      // pixels
  unsigned cb;     // bits of a pixel index in a key
  const unsigned* parent;  // per key: its chain parent pixel (itself at a birth)
  unsigned* P;    // per pixel: position of its key (the G array)
  unsigned* B;    // per key: chain parent, as a position
  unsigned* R;    // per key: its birth, as a position
  unsigned* D;    // per birth: first deepest pixel of its chain group
  unsigned* F;    // per birth: union-find of the replayed scan
  unsigned* Hs;   // per replayed set: its deepest root pixel
  unsigned char* flag;  // per key: a step to replay
};

// keys: component << cb | pixel, with 2^cb >= the number of pixels
__device__ __forceinline__ unsigned key_pixel(unsigned long long k, unsigned cb) {
  return (unsigned)(k & ((1ull << cb) - 1ull));
}
__device__ __forceinline__ unsigned key_comp(unsigned long long k, unsigned cb) {
  return (unsigned)(k >> cb);
}

__global__ void gw_tie_pos(Tie t) {
  GW_FOR_KEYS(t.nk) t.P[key_pixel(t.keys[q], t.cb)] = q;
}

// Position of pixel p's key, or kNone (and the error bit) if it has none.
__device__ __forceinline__ unsigned key_pos(const Tie& t, unsigned p, unsigned* err) {
  const unsigned q = t.P[p];
  if (q < t.nk && key_pixel(t.keys[q], t.cb) == p) return q;
  atomicOr(err, 1u);
  return kNone;
}

// Chain parent of every key, as a position: the parent pixel is a flat of
// the same component with a smaller index (an earlier pixel of its run, or
// a flat scanned neighbour), so it has a key and the chains end at births
// whatever the order of the keys; a pixel without a key marks an
// inconsistency and is not followed.
__global__ void gw_tie_parent(Tie t, const unsigned* K, unsigned* err) {
  GW_FOR_KEYS(t.nk) {
    const unsigned long long key = t.keys[q];
    const unsigned p = key_pixel(key, t.cb);
    t.D[q] = kNone;
    unsigned b = q;
    if (K[key_comp(key, t.cb)] != kMixed) {
      const unsigned par = t.parent[q];
      if (par < p) {
        b = key_pos(t, par, err);
        if (b == kNone) b = q;
      }
    }
    t.B[q] = b;
  }
}

__global__ void gw_tie_births(Tie t) { GW_FOR_KEYS(t.nk) t.R[q] = find_halving(t.B, q); }

// First deepest pixel of every chain group: per wavefront, one guarded
// atomic per group present (a large all-NaN flat is one group of millions
// of deepest pixels).
__global__ void gw_tie_deepest(Tie t, const unsigned* K, const unsigned* W, const float* original,
                               const float* filled) {
  const unsigned lane = threadIdx.x & 63u;
  for (unsigned q0 = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u; q0 < t.nk;
       q0 += gridDim.x * blockDim.x) {
    const unsigned q = q0 + lane;
    unsigned b = kNone, p = kNone;
    if (q < t.nk) {
      const unsigned long long key = t.keys[q];
      const unsigned c = key_comp(key, t.cb), k = K[c];
      p = key_pixel(key, t.cb);
      if (k != kMixed && deepest_pixel(k, depth_at(original, filled, p), W[c])) b = t.R[q];
    }
    unsigned long long pending = __ballot(b != kNone);
    while (pending) {
      const int leader = __ffsll((long long)pending) - 1;
      const unsigned b0 = __shfl(b, leader);
      const bool mine = b == b0;
      unsigned m = mine ? p : kNone;
      for (int dd = 32; dd > 0; dd >>= 1) m = min(m, __shfl_xor(m, dd));
      if ((int)lane == leader && m < ldc(t.D + b0)) atomicMin(t.D + b0, m);
      pending &= ~__ballot(mine);
    }
  }
}

// Scan steps to replay: births, first deepest pixels of chain groups and
// merges of two chain groups; every pixel of a mixed component.
__global__ void gw_tie_flags(Tie t, Grid g, const unsigned* K, unsigned* err) {
  GW_FOR_KEYS(t.nk) {
    const unsigned long long key = t.keys[q];
    const unsigned p = key_pixel(key, t.cb);
    bool ev = K[key_comp(key, t.cb)] == kMixed;
    if (!ev) {
      const unsigned b = t.R[q];
      ev = b == q || t.D[b] == p;
      unsigned a_rep;
      bool nw;
      if (!ev && merge_site(g, p % g.nx, p / g.nx, &a_rep, &nw)) {
        const unsigned pa = key_pos(t, a_rep, err), pb = key_pos(t, p - g.nx + 1, err);
        if (pa != kNone && pb != kNone) ev = t.R[pa] != t.R[pb];
      }
    }
    t.flag[q] = ev ? 1 : 0;
  }
}

// Root with path compression, for one thread's sequential replay; bounded,
// so that an inconsistent forest ends the walk (with the error bit) rather
// than the kernel.
__device__ unsigned find_seq(unsigned* P, unsigned x, unsigned n, unsigned* err) {
  for (unsigned steps = 0; steps <= n; steps++) {
    if (x >= n) break;
    const unsigned px = P[x];
    if (px == x) return x;
    if (px >= n) break;
    P[x] = P[px] < n ? P[px] : px;
    x = px;
  }
  atomicOr(err, 1u);
  return 0u;
}

// The reference's scan over every pixel of a mixed component, in scan
// order: each pixel starts a set and is merged with its flat scanned
// neighbours in order; a merge keeps the neighbour's root unless the
// current root is strictly deeper. F is the union-find, Hs the root (label)
// of each set, both by key position.
__device__ void replay_full(const Tie& t, const Grid& g, const unsigned long long* ev,
                            unsigned k, unsigned e, const float* original, const float* filled,
                            unsigned* Fc, unsigned* err) {
  const unsigned c = key_comp(ev[k], t.cb);
  for (unsigned q = k; q < e; q++) {
    const unsigned p = key_pixel(ev[q], t.cb);
    const unsigned pq = key_pos(t, p, err);
    if (pq == kNone) return;
    t.F[pq] = pq;
    t.Hs[pq] = p;
    const long long i = p % g.nx, j = p / g.nx;
    const long long nb[4][2] = {{i - 1, j - 1}, {i, j - 1}, {i + 1, j - 1}, {i - 1, j}};
    for (int s = 0; s < 4; s++) {
      if (!g.flat(nb[s][0], nb[s][1])) continue;
      const unsigned qn = key_pos(t, (unsigned)(nb[s][1] * g.nx + nb[s][0]), err);
      if (qn == kNone) return;
      const unsigned a = find_seq(t.F, qn, t.nk, err);
      const unsigned r = find_seq(t.F, pq, t.nk, err);
      if (a == r) continue;
      const unsigned ra = t.Hs[a], rs = t.Hs[r];
      if (ra >= t.n || rs >= t.n) {
        atomicOr(err, 1u);
        return;
      }
      const bool current_deeper = depth_at(original, filled, ra) < depth_at(original, filled, rs);
      t.F[r] = a;
      t.Hs[a] = current_deeper ? rs : ra;
    }
  }
  const unsigned qc = key_pos(t, c, err);
  if (qc != kNone) Fc[c] = t.Hs[find_seq(t.F, qc, t.nk, err)];
}

// One thread per tie component: the replay of its scan steps (the nev
// flagged keys, in scan order).
__global__ void gw_tie_replay(Tie t, Grid g, const unsigned long long* ev, const unsigned* nev,
                              const unsigned* K, const unsigned* W, unsigned* Fc,
                              const float* original, const float* filled, unsigned* err) {
  const unsigned n = *nev;
  GW_FOR_KEYS(n) {
    const unsigned c = key_comp(ev[q], t.cb);
    if (q > 0 && key_comp(ev[q - 1], t.cb) == c) continue;  // not the first of its component
    unsigned e = q;
    while (e < n && key_comp(ev[e], t.cb) == c) e++;
    const unsigned kind = K[c];
    if (kind == kMixed) {
      replay_full(t, g, ev, q, e, original, filled, Fc, err);
      continue;
    }
    const unsigned wbits = W[c];
    bool ok = true;
    for (unsigned s = q; s < e && ok; s++) {
      const unsigned p = key_pixel(ev[s], t.cb);
      const unsigned pq = key_pos(t, p, err);
      if (pq == kNone) {
        ok = false;
        break;
      }
      const bool pm = deepest_pixel(kind, depth_at(original, filled, p), wbits);
      const unsigned b = t.R[pq];
      if (b == pq) {
        t.F[pq] = pq;
        t.Hs[pq] = pm ? p : kNone;
        continue;  // a birth is the first deepest pixel of its own group if it is one
      }
      unsigned a_rep;
      bool nw;
      if (merge_site(g, p % g.nx, p / g.nx, &a_rep, &nw)) {
        const unsigned pa = key_pos(t, a_rep, err), pb = key_pos(t, p - g.nx + 1, err);
        if (pa == kNone || pb == kNone) {
          ok = false;
          break;
        }
        unsigned a = t.R[pa], bb = t.R[pb];
        if (a != bb) {
          a = find_seq(t.F, a, t.nk, err);
          bb = find_seq(t.F, bb, t.nk, err);
          if (a != bb) {
            const unsigned first = nw ? a : bb, second = nw ? bb : a;
            const unsigned r = t.Hs[second] != kNone ? t.Hs[second]
                               : t.Hs[first] != kNone ? t.Hs[first]
                               : pm                   ? p
                                                      : kNone;
            t.F[first] = second;
            t.Hs[second] = r;
          }
        }
      }
      if (t.D[b] == p) {
        const unsigned r = find_seq(t.F, b, t.nk, err);
        if (t.Hs[r] == kNone) t.Hs[r] = p;
      }
    }
    if (!ok) continue;
    const unsigned qc = key_pos(t, c, err);
    if (qc != kNone) Fc[c] = t.Hs[find_seq(t.F, qc, t.nk, err)];
  }
}

// Outputs of the tie components' pixels, from their replayed labels.
__global__ void gw_tie_outputs(Tie t, const unsigned* Fc, unsigned n, const float* original,
                               const float* filled, float* costs, int64_t* conncomps,
                               unsigned* err) {
  GW_FOR_KEYS(t.nk) {
    const unsigned long long key = t.keys[q];
    const unsigned p = key_pixel(key, t.cb), label = Fc[key_comp(key, t.cb)];
    if (label >= n) {
      atomicOr(err, 1u);
      continue;
    }
    costs[p] = gw_cost(depth_at(original, filled, label), depth_at(original, filled, p));
    conncomps[p] = (int64_t)label;
  }
}

__global__ void gw_zero(unsigned* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

unsigned capped(unsigned long long v, unsigned cap) { return (unsigned)(v < cap ? v : cap); }

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

hipError_t run_gwdt_computecosts(float* costs, int64_t* conncomps, const int32_t* flats,
                                 const float* original, const float* filled, int64_t nx_,
                                 int64_t ny_, hipStream_t s) {
  if (nx_ <= 0 || ny_ <= 0) return hipSuccess;
  const unsigned nx = (unsigned)nx_, ny = (unsigned)ny_;
  const size_t n = (size_t)nx * (size_t)ny;
  unsigned cb = 1;  // bits of a pixel index (and of a component) in a tie key
  while (cb < 32 && ((size_t)1 << cb) < n) cb++;
  const unsigned ntx = (nx + kTileW - 1) / kTileW, nty = (ny + kTileH - 1) / kTileH;
  const size_t ntiles = (size_t)ntx * nty;
  const unsigned tile_rows = ny < kTileH ? ny : kTileH;
  // a tile of w x h pixels has at most ceil(w/2) * ceil(h/2) components
  const unsigned cap = (kTileW / 2) * ((tile_rows + 1) / 2);
  const size_t ne = ntiles * cap;
  // per pixel six words, per list entry six, per tile one, counters, and up
  // to one tie key (and its chain parent) per pixel, in one workspace
  const size_t words = 7 * n + 6 * ne + ntiles + 16;
  const size_t keys_off = up256(words * sizeof(unsigned));
  unsigned* ws = nullptr;
  hipError_t e = hipMallocAsync((void**)&ws, keys_off + n * sizeof(unsigned long long), s);
  if (e != hipSuccess) return e;
  Ws w;
  unsigned* m0 = ws;
  auto take = [&](size_t k) {
    unsigned* r = m0;
    m0 += k;
    return r;
  };
  w.L = take(n);
  w.G = take(n);
  w.Kc = take(n);
  w.Wc = take(n);
  w.Cc = take(n);
  w.Fc = take(n);
  w.RL = take(ne);
  w.RH = take(ne);
  w.RK = take(ne);
  w.RW = take(ne);
  w.RC = take(ne);
  w.RF = take(ne);
  w.TC = take(ntiles);
  w.cnt = take(16);  // [0] tie keys, [1] consistency errors
  unsigned* parent = take(n);
  w.cap = cap;
  w.ntx = ntx;
  w.nty = nty;
  unsigned long long* keys = (unsigned long long*)((char*)ws + keys_off);
  void* tie_ws = nullptr;
  auto done = [&](hipError_t r) {
    if (tie_ws) (void)hipFreeAsync(tie_ws, s);
    hipError_t f = hipFreeAsync(ws, s);
    return r != hipSuccess ? r : f;
  };
  const dim3 tgrid(ntx, capped(nty, 65535));
  const unsigned egrid = capped((ntiles + 3) / 4, 1u << 20);
  gw_zero<<<1, 64, 0, s>>>(w.cnt, 16);
  gw_local<<<tgrid, 256, 0, s>>>(flats, original, filled, nx, ny, w);
  gw_border<<<tgrid, 128, 0, s>>>(nx, ny, w);
  gw_roots<<<egrid, 256, 0, s>>>(w);
  gw_merge_deepest<<<egrid, 256, 0, s>>>(w);
  gw_final<<<tgrid, 256, 0, s>>>(original, filled, nx, ny, w, costs, conncomps, keys, parent,
                                 cb);
  if ((e = hipGetLastError()) != hipSuccess) return done(e);
  unsigned host[2] = {0, 0};
  e = hipMemcpyAsync(host, w.cnt, sizeof(host), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return done(e);
  if (host[1]) return done(hipErrorIllegalState);
  if (host[0] == 0) return done(hipSuccess);
  if (host[0] > 0x7fffffffu) return done(hipErrorInvalidValue);  // hipcub counts in int
  const unsigned nk = host[0];

  // Tie components: chain groups over their keys (in append order), then
  // the replay over the scan steps it needs, sorted (component, then
  // pixel), and their outputs.
  size_t sort_bytes = 0, flag_bytes = 0;
  e = hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, (unsigned long long*)nullptr,
                                        (unsigned long long*)nullptr, (int)nk, 0, (int)(2 * cb),
                                        s);
  if (e == hipSuccess)
    e = hipcub::DeviceSelect::Flagged(nullptr, flag_bytes, (unsigned long long*)nullptr,
                                      (unsigned char*)nullptr, (unsigned long long*)nullptr,
                                      (unsigned*)nullptr, (int)nk, s);
  if (e != hipSuccess) return done(e);
  const size_t key_bytes = up256((size_t)nk * sizeof(unsigned long long));
  const size_t word_bytes = up256((size_t)nk * sizeof(unsigned));
  const size_t total = 2 * key_bytes + 5 * word_bytes + up256(nk) + 256 +
                       up256(sort_bytes > flag_bytes ? sort_bytes : flag_bytes);
  e = hipMallocAsync(&tie_ws, total, s);
  if (e != hipSuccess) return done(e);
  char* m = (char*)tie_ws;
  Tie t;
  unsigned long long* ev = (unsigned long long*)m;  // the flagged keys
  m += key_bytes;
  unsigned long long* evs = (unsigned long long*)m;  // the same, sorted
  m += key_bytes;
  t.keys = keys;
  t.parent = parent;
  t.nk = nk;
  t.n = (unsigned)n;
  t.cb = cb;
  t.P = w.G;  // the cross-tile union-find is no longer needed
  t.B = (unsigned*)m;
  m += word_bytes;
  t.R = (unsigned*)m;
  m += word_bytes;
  t.D = (unsigned*)m;
  m += word_bytes;
  t.F = (unsigned*)m;
  m += word_bytes;
  t.Hs = (unsigned*)m;
  m += word_bytes;
  t.flag = (unsigned char*)m;
  m += up256(nk);
  unsigned* nev = (unsigned*)m;
  m += 256;
  void* temp = m;
  const Grid g{w.L, nx, ny};
  unsigned* err = w.cnt + 1;
  const unsigned kb = capped((nk + 255) / 256, 16384);
  gw_tie_pos<<<kb, 256, 0, s>>>(t);
  gw_tie_parent<<<kb, 256, 0, s>>>(t, w.Kc, err);
  gw_tie_births<<<kb, 256, 0, s>>>(t);
  gw_tie_deepest<<<kb, 256, 0, s>>>(t, w.Kc, w.Wc, original, filled);
  gw_tie_flags<<<kb, 256, 0, s>>>(t, g, w.Kc, err);
  size_t tb = flag_bytes;
  e = hipcub::DeviceSelect::Flagged(temp, tb, keys, t.flag, ev, nev, (int)nk, s);
  unsigned nev_host = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&nev_host, nev, sizeof(unsigned), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return done(e);
  if (nev_host) {
    tb = sort_bytes;
    e = hipcub::DeviceRadixSort::SortKeys(temp, tb, ev, evs, (int)nev_host, 0, (int)(2 * cb), s);
    if (e != hipSuccess) return done(e);
    gw_tie_replay<<<capped((nev_host + 255) / 256, 16384), 256, 0, s>>>(
        t, g, evs, nev, w.Kc, w.Wc, w.Fc, original, filled, err);
  }
  gw_tie_outputs<<<kb, 256, 0, s>>>(t, w.Fc, (unsigned)n, original, filled, costs, conncomps,
                                    err);
  if ((e = hipGetLastError()) != hipSuccess) return done(e);
  e = hipMemcpyAsync(host, w.cnt, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return done(e);
  // a key without its position, or a chain that does not lead back: an
  // internal inconsistency, reported rather than returned as labels
  return done(host[1] ? hipErrorIllegalState : hipSuccess);
}

}  // namespace tt
