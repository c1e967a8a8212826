// Tile visit of the sink-filling / reconstruction relaxation, shared by the
// single-grid kernels (fill.hip) and the batched kernel (batched.hip).
//
// Reference semantics (src/fillsinks.c:22-48, src/reconstruct.c:68-112,
// 134-202, 270-279): fillsinks negates the DEM, seeds the marker with the
// negated DEM on bc==1 cells and -INF elsewhere, reconstructs by dilation
// (marker = min(mask, max(marker, 8 neighbours)) until nothing changes) and
// negates back. That fixed point is unique and every value is a copy of an
// input value, so any update order reaches the same bits.
//
// The visit works directly in the "fill domain" f = -marker, d = -mask:
//     f_p <- max(d_p, min(f_p, f_q for the 8 neighbours q))
// iterated from f0 = d on seeds (+INF elsewhere) down to the greatest fixed
// point. NaN cells of d are inert: level +INF, bound +INF (fminf skips them
// as fmaxf skips NaN in the reference), written back as the input NaN.
//
// kModeFill:  d = dem, f0 = (bc == 1) ? dem : +INF; the output array holds
//             f between visits.
// kModeRecon: d = -mask, f0 = max(d, -marker) (a NaN marker acts as -INF);
//             the marker array holds -f between visits and at the end.
// kModeSpill: fillsinks over lexicographic (level, cell index) keys packed
//             in 64 bits; the index half of the fixed point is the cell
//             whose DEM value the output copies (the spill cell), which the
//             adjoint scatters to. Unique fixed point, so deterministic.
//             The levels are final (state) and never change, so a key is
//             staged from the level and the cell's spill index (one 32-bit
//             word): a halo read racing with a neighbour's write-back sees
//             an old or a new index, both valid upper bounds. The relaxation
//             starts from the indices of spill_ccl.h's labelling (upper
//             bounds, exact except near flat cells).
#pragma once

#include <type_traits>

#include "kernels.h"
#include "tile_engine.h"

namespace tt {

enum : int { kModeSpill = 2 };

// One grid as seen by a tile visit (spill indices are cell positions in
// this grid).
struct FillView {
  float* state;       // kModeSpill: the final fill levels (read only)
  int* spill;         // kModeSpill only: the iteration state's index halves
  const float* src;
  const uint8_t* bc;
  long long nx, ny;
};

// ---------------------------------------------------------------- level types
struct FloatLevel {
  using T = float;
  static __device__ __forceinline__ T lo(T a, T b) { return fminf(a, b); }
  static __device__ __forceinline__ T hi(T a, T b) { return fmaxf(a, b); }
  static __device__ __forceinline__ bool less(T a, T b) { return a < b; }
  static __device__ __forceinline__ T from_below(T v, T e) { return lane_from_below(v, e); }
  static __device__ __forceinline__ T from_above(T v, T e) { return lane_from_above(v, e); }
  static constexpr T kUnreached = __builtin_inff();
  static constexpr T kInert = __builtin_inff();
};

// Order-preserving map of a float's bits to uint32 (x < y <=> ord(x) < ord(y)
// for non-NaN x, y, with -0 ordered below +0).
__device__ __forceinline__ unsigned ord_bits(float f) {
  const unsigned u = __float_as_uint(f);
  return u ^ ((u >> 31) ? 0xffffffffu : 0x80000000u);
}
__device__ __forceinline__ float unord_bits(unsigned u) {
  return __uint_as_float((u >> 31) ? (u ^ 0x80000000u) : ~u);
}

__device__ __forceinline__ unsigned dpp_u32(unsigned old, unsigned v, int ctrl) {
  return ctrl == 0x138
             ? (unsigned)__builtin_amdgcn_update_dpp((int)old, (int)v, 0x138, 0xf, 0xf, false)
             : (unsigned)__builtin_amdgcn_update_dpp((int)old, (int)v, 0x130, 0xf, 0xf, false);
}

struct KeyLevel {
  using T = unsigned long long;  // ord(level) << 32 | cell index
  static __device__ __forceinline__ T lo(T a, T b) { return a < b ? a : b; }
  static __device__ __forceinline__ T hi(T a, T b) { return a < b ? b : a; }
  static __device__ __forceinline__ bool less(T a, T b) { return a < b; }
  static __device__ __forceinline__ T shift(T v, T e, int ctrl) {
    const unsigned l = dpp_u32((unsigned)e, (unsigned)v, ctrl);
    const unsigned h = dpp_u32((unsigned)(e >> 32), (unsigned)(v >> 32), ctrl);
    return ((T)h << 32) | l;
  }
  static __device__ __forceinline__ T from_below(T v, T e) { return shift(v, e, 0x138); }
  static __device__ __forceinline__ T from_above(T v, T e) { return shift(v, e, 0x130); }
  // (+INF, no index): a cell no seed has reached yet; all ones: inert bound
  static constexpr T kUnreached = (0xff800000ull << 32) | 0xffffffffull;
  static constexpr T kInert = ~0ull;
  static __device__ __forceinline__ T make(float f, unsigned idx) {
    return ((T)ord_bits(f) << 32) | idx;
  }
};

// Float levels as their order-preserving uint32 images (ord_bits): the
// relaxation only compares, and integer min / max / min3 need none of the
// NaN quieting that IEEE-mode float min / max cost. Fill levels are never
// NaN (no-data cells are +INF), so the order is the float order.
struct OrdLevel {
  using T = unsigned;
  static __device__ __forceinline__ T lo(T a, T b) { return a < b ? a : b; }
  static __device__ __forceinline__ T hi(T a, T b) { return a < b ? b : a; }
  static __device__ __forceinline__ bool less(T a, T b) { return a < b; }
  static __device__ __forceinline__ T from_below(T v, T e) { return dpp_u32(e, v, 0x138); }
  static __device__ __forceinline__ T from_above(T v, T e) { return dpp_u32(e, v, 0x130); }
  static constexpr T kUnreached = 0xff800000u;  // ord_bits(+INF)
  static constexpr T kInert = 0xff800000u;
};

template <int MODE>
using FillLevel = typename std::conditional<MODE == kModeSpill, KeyLevel, OrdLevel>::type;

// One directional sweep over a 64x64 quadrant with lanes along columns
// (vertical sweep: down = +y) or along rows (horizontal: right = +x).
// The sweep runs as NCH independent chains of 64 / NCH steps, interleaved
// step by step so the dependent min / max / DPP sequence of one chain
// overlaps the others' (a lone wavefront per SIMD cannot hide that
// latency otherwise). Each chain starts from the value its predecessor's
// last cell held before the sweep; the relaxation reaches the same fixed
// point in any order, information only crosses a chain boundary one
// sweep later. The value carried between steps is the lane's previous
// cell; its two side neighbours come from the adjacent lanes by DPP, with
// the quadrant's outside column/row read from LDS for the edge lanes. LDS
// operands are fetched kPrefetch steps ahead (the loop is fully unrolled,
// so the windows live in registers); none of them is written in between
// (a lane's cell is written only at its own step, the outside line belongs
// to another wavefront or the halo), and the carried value is stored every
// step.
constexpr int kPrefetch = 4;
constexpr int kChains = 2;

template <class G, class L, bool VERT, bool FWD>
__device__ __forceinline__ bool fill_sweep(typename L::T* sf, const typename L::T* sd, int R0,
                                           int C0, int lane) {
  using T = typename L::T;
  constexpr int NCH = kChains, LEN = 64 / kChains;
  constexpr int step = FWD ? 1 : -1;
  bool changed = false;
  const int own = VERT ? C0 + lane : R0 + lane;      // lane's column / row
  const int side = (lane == 0) ? own - 1 : own + 1;  // outside neighbour line
  const int k0 = VERT ? (FWD ? R0 : R0 + 63) : (FWD ? C0 : C0 + 63);
  auto cell = [&](int k) { return VERT ? G::at(k, own) : G::at(own, k); };
  auto edge_cell = [&](int k) { return VERT ? G::at(k, side) : G::at(side, k); };
  T pc[NCH];
  T fw[NCH][kPrefetch], dw[NCH][kPrefetch], ew[NCH][kPrefetch];
#pragma unroll
  for (int c = 0; c < NCH; c++) {
    const int kc = k0 + c * LEN * step;
    pc[c] = sf[cell(kc - step)];
#pragma unroll
    for (int i = 0; i < kPrefetch; i++) {
      fw[c][i] = sf[cell(kc + i * step)];
      dw[c][i] = sd[cell(kc + i * step)];
      ew[c][i] = sf[edge_cell(kc + (i - 1) * step)];
    }
  }
#pragma unroll
  for (int s = 0; s < LEN; s++) {
    const int slot = s % kPrefetch;
#pragma unroll
    for (int c = 0; c < NCH; c++) {
      const int k = k0 + (c * LEN + s) * step;
      const T fc = fw[c][slot], dv = dw[c][slot], edge = ew[c][slot];
      if (s + kPrefetch < LEN) {
        fw[c][slot] = sf[cell(k + kPrefetch * step)];
        dw[c][slot] = sd[cell(k + kPrefetch * step)];
        ew[c][slot] = sf[edge_cell(k + (kPrefetch - 1) * step)];
      }
      const T a = L::from_below(pc[c], edge);
      const T b = L::from_above(pc[c], edge);
      const T nv = L::hi(dv, L::lo(L::lo(a, b), L::lo(pc[c], fc)));
      changed |= L::less(nv, fc);
      pc[c] = L::lo(nv, fc);
      sf[cell(k)] = pc[c];
    }
  }
  return changed;
}

// kModeSpill bound of a cell from its kind byte. With the levels known, a
// cell's bound key (ord(d), own index) matters only where d equals its
// level (kSurface: the bound is (own level, own index)); a submerged cell's
// bound lies below every key its neighbours can offer (their levels are
// at least its own), and NaN cells are inert.
enum : unsigned char { kSubmerged = 0, kSurface = 1, kInertCell = 2 };

struct SpillBound {
  const unsigned char* kind;  // LDS, one byte per cell (TileGeom layout)
  unsigned base;              // grid index of tile cell (0, 0)
  unsigned nx;
  // bound of the cell at LDS index `at`, tile row r, column c, whose key is `own`
  __device__ __forceinline__ unsigned long long operator()(int at, int r, int c,
                                                           unsigned long long own) const {
    const unsigned k = kind[at];
    if (k == kSubmerged) return 0ull;
    if (k == kInertCell) return KeyLevel::kInert;
    return (own & 0xffffffff00000000ull) | (unsigned)(base + (unsigned)r * nx + (unsigned)c);
  }
};

// Surface bits of a lane's 64 cells in the column layout (bit i: cell
// (R0 + i, C0 + lane)) or the row layout (bit i: cell (R0 + lane, C0 + i)),
// taken once per visit from the kind bytes, which do not change during it.
template <class G, bool VERT>
__device__ __forceinline__ unsigned long long surface_bits(const unsigned char* kind, int R0,
                                                           int C0, int lane) {
  unsigned long long m = 0ull;
#pragma unroll 8
  for (int i = 0; i < 64; i++) {
    const int at = VERT ? G::at(R0 + i, C0 + lane) : G::at(R0 + lane, C0 + i);
    m |= (unsigned long long)(kind[at] == kSurface) << i;
  }
  return m;
}

// Fixed-point test of a wavefront's quadrant: true (in some lane) if a
// cell would still decrease. Lanes are columns walking the rows with a
// three-row window in registers; the side columns come by DPP (the
// quadrant's outside columns from LDS for the edge lanes). The steps are
// independent, so this costs a fraction of a sweep.
template <class G>
__device__ __forceinline__ bool spill_violation(const unsigned long long* sf,
                                                unsigned long long vsurf, const SpillBound& bound,
                                                int R0, int C0, int lane) {
  using L = KeyLevel;
  using T = typename L::T;
  const int C = C0 + lane;
  const int S = (lane == 0) ? C - 1 : C + 1;  // outside column for the edge lanes
  const unsigned line0 = bound.base + (unsigned)R0 * bound.nx + (unsigned)C;
  T up = sf[G::at(R0 - 1, C)], mid = sf[G::at(R0, C)];
  T eup = sf[G::at(R0 - 1, S)], emid = sf[G::at(R0, S)];
  bool v = false;
#pragma unroll 8
  for (int r = 0; r < 64; r++) {
    const T dn = sf[G::at(R0 + r + 1, C)];
    const T edn = sf[G::at(R0 + r + 1, S)];
    const T dv = mid == L::kInert             ? L::kInert
                 : ((vsurf >> r) & 1ull) != 0 ? ((mid & 0xffffffff00000000ull) |
                                                (line0 + (unsigned)r * bound.nx))
                                              : 0ull;
    const T col = L::lo(L::lo(up, dn), mid);       // own column, rows r-1..r+1
    const T ecol = L::lo(L::lo(eup, edn), emid);   // outside column (edge lanes)
    const T m = L::lo(col, L::lo(L::from_below(col, ecol), L::from_above(col, ecol)));
    v |= L::less(L::hi(dv, m), mid);
    up = mid;
    mid = dn;
    eup = emid;
    emid = edn;
  }
  return v;
}

constexpr unsigned kNoNbr = 0xffffffffu;  // above every level / index: min() ignores it

// ---------------------------------------------------------------- spill index sweeps
// With the levels fixed, a key's relaxation only moves its index half, and
// only between neighbours of equal level: a lower neighbour gives a key
// below every bound (such surface cells sit at their own index from the
// labelling on), a higher one a key above the cell's own. The register
// sweeps below therefore relax 32-bit indices as the order-domain sweeps
// relax levels, a neighbour counting only where a per-visit mask says its
// level equals the cell's. Inert cells hold all ones and differ in level
// (all ones) from every real cell, so they neither move nor spread.
//
// Masks of a lane's 64 cells in the column (VERT: cell i = (R0 + i,
// C0 + lane)) or row layout (cell i = (R0 + lane, C0 + i)): byte i of
// m[i / 4] (from bit 8 * (i % 4)) has a bit SET where the neighbour does
// NOT count: bit 0 (i - 1, lane), 1 (i - 1, lane - 1), 2 (i - 1, lane + 1),
// 3 (i + 1, lane), 4 (i + 1, lane - 1), 5 (i + 1, lane + 1), 6 (i, lane - 1),
// 7 (i, lane + 1), lane +-1 meaning the neighbouring line (the edge lanes'
// outside line from LDS).
// Levels are the high words of the LDS keys, which a visit never changes.
template <class G, bool VERT>
__device__ __forceinline__ void level_masks(unsigned (&m)[16], const unsigned long long* sf,
                                            int R0, int C0, int lane) {
  const unsigned* w = reinterpret_cast<const unsigned*>(sf);  // w[2 at + 1]: level
  const int own = VERT ? C0 + lane : R0 + lane;
  const int side = (lane == 0) ? own - 1 : own + 1;
  auto lvl = [&](int i, int line) {
    return VERT ? w[2 * G::at(R0 + i, line) + 1] : w[2 * G::at(line, C0 + i) + 1];
  };
  // line position k: own level v, lane - 1's lo, lane + 1's hi
  unsigned pv = lvl(-1, own);
  unsigned pe = lvl(-1, side);
  unsigned plo = dpp_u32(pe, pv, 0x138), phi = dpp_u32(pe, pv, 0x130);
  unsigned cv = lvl(0, own);
  unsigned ce = lvl(0, side);
  unsigned clo = dpp_u32(ce, cv, 0x138), chi = dpp_u32(ce, cv, 0x130);
#pragma unroll
  for (int i = 0; i < 16; i++) m[i] = 0u;
#pragma unroll
  for (int i = 0; i < 64; i++) {
    const unsigned nv = lvl(i + 1, own), ne = lvl(i + 1, side);
    const unsigned nlo = dpp_u32(ne, nv, 0x138), nhi = dpp_u32(ne, nv, 0x130);
    const unsigned bits = (unsigned)(pv != cv) | ((unsigned)(plo != cv) << 1) |
                          ((unsigned)(phi != cv) << 2) | ((unsigned)(nv != cv) << 3) |
                          ((unsigned)(nlo != cv) << 4) | ((unsigned)(nhi != cv) << 5) |
                          ((unsigned)(clo != cv) << 6) | ((unsigned)(chi != cv) << 7);
    m[i >> 2] |= bits << (8 * (i & 3));
    pv = cv;
    plo = clo;
    phi = chi;
    cv = nv;
    clo = nlo;
    chi = nhi;
  }
}

// One directional index sweep over the lane's 64 cells x (as
// ord_sweep_regs: two interleaved chains, the carried value the updated
// cell behind, its side neighbours by DPP, the outside line and corner
// cells from LDS). Cell i's bound is its own grid index line0 + i * stride
// on surface cells (bit i of surf), 0 elsewhere. Returns whether a cell
// decreased.
template <class G, bool VERT, bool FWD>
__device__ __forceinline__ bool spill_sweep_regs(unsigned (&x)[64], const unsigned long long* sf,
                                                 const unsigned (&m)[16], unsigned long long surf,
                                                 unsigned line0, unsigned stride, int R0, int C0,
                                                 int lane) {
  const unsigned* w = reinterpret_cast<const unsigned*>(sf);  // w[2 at]: index
  const int out_k = FWD ? -1 : 64;
  unsigned pc0 = VERT ? w[2 * G::at(R0 + out_k, C0 + lane)] : w[2 * G::at(R0 + lane, C0 + out_k)];
  unsigned corner = kNoNbr;
  if (VERT) {
    const unsigned lo = w[2 * G::at(R0 + out_k, C0 - 1)], hi = w[2 * G::at(R0 + out_k, C0 + 64)];
    corner = (lane == 0) ? lo : hi;
  }
  constexpr int sh = FWD ? 0 : 3;  // mask bits of the cells behind
  unsigned pc1 = x[FWD ? 31 : 32];
  // grid indices of the two chains' current cells
  unsigned g0 = line0 + (unsigned)(FWD ? 0 : 63) * stride;
  unsigned g1 = line0 + (unsigned)(FWD ? 32 : 31) * stride;
  const unsigned dg = FWD ? stride : 0u - stride;
  auto relax = [&](int i, unsigned g, unsigned pc, unsigned a, unsigned b) {
    const unsigned mb = m[i >> 2] >> (8 * (i & 3) + sh);
    const unsigned cn = pc | (0u - (mb & 1u));
    const unsigned ca = a | (0u - ((mb >> 1) & 1u));
    const unsigned cb = b | (0u - ((mb >> 2) & 1u));
    const unsigned mm = min(min(ca, cb), min(cn, x[i]));
    const unsigned bnd = g & (0u - (unsigned)((surf >> i) & 1ull));
    return max(bnd, mm);
  };
  unsigned changed = 0u;
#pragma unroll
  for (int s = 0; s < 32; s++) {
    const int i0 = FWD ? s : 63 - s;
    const int i1 = FWD ? 32 + s : 31 - s;
    const unsigned old0 = (s == 0) ? corner : kNoNbr;
    const unsigned a0 = dpp_u32(old0, pc0, 0x138), b0 = dpp_u32(old0, pc0, 0x130);
    const unsigned a1 = dpp_u32(kNoNbr, pc1, 0x138), b1 = dpp_u32(kNoNbr, pc1, 0x130);
    const unsigned n0 = relax(i0, g0, pc0, a0, b0), n1 = relax(i1, g1, pc1, a1, b1);
    changed |= (unsigned)(n0 < x[i0]) | (unsigned)(n1 < x[i1]);
    x[i0] = n0;
    x[i1] = n1;
    pc0 = n0;
    pc1 = n1;
    g0 += dg;
    g1 += dg;
  }
  return changed != 0u;
}

// Fixed-point test of the index relaxation from the row layout in
// registers (x[i]: cell (R0 + lane, C0 + i), m its masks, surf its surface
// bits, line0 + i its grid index): own row by register, the rows above /
// below by DPP (the edge lanes' outside row and every lane's outside
// columns from LDS). True (in some lane) if a cell would still decrease.
template <class G>
__device__ __forceinline__ bool spill_violation_rows(const unsigned (&x)[64],
                                                     const unsigned long long* sf,
                                                     const unsigned (&m)[16],
                                                     unsigned long long surf, unsigned line0,
                                                     int R0, int C0, int lane) {
  const unsigned* w = reinterpret_cast<const unsigned*>(sf);  // w[2 at]: index
  const int R = R0 + lane;
  const int er = (lane == 0) ? R0 - 1 : R0 + 64;  // outside row for the edge lanes
  auto lds = [&](int r, int c) { return w[2 * G::at(r, c)]; };
  // position k: own value, lane - 1's (lo) and lane + 1's (hi)
  unsigned pv = lds(R, C0 - 1);
  unsigned pe = lds(er, C0 - 1);
  unsigned plo = dpp_u32(pe, pv, 0x138), phi = dpp_u32(pe, pv, 0x130);
  unsigned cv = x[0];
  unsigned ce = lds(er, C0);
  unsigned clo = dpp_u32(ce, cv, 0x138), chi = dpp_u32(ce, cv, 0x130);
  bool v = false;
#pragma unroll
  for (int i = 0; i < 64; i++) {
    const unsigned nv = (i == 63) ? lds(R, C0 + 64) : x[i + 1];
    const unsigned ne = lds(er, C0 + i + 1);
    const unsigned nlo = dpp_u32(ne, nv, 0x138), nhi = dpp_u32(ne, nv, 0x130);
    const unsigned mb = m[i >> 2] >> (8 * (i & 3));
    auto c = [&](unsigned val, int bit) { return val | (0u - ((mb >> bit) & 1u)); };
    const unsigned mm = min(min(min(c(pv, 0), c(plo, 1)), min(c(phi, 2), c(nv, 3))),
                            min(min(c(nlo, 4), c(nhi, 5)), min(c(clo, 6), c(chi, 7))));
    const unsigned bnd = (line0 + (unsigned)i) & (0u - (unsigned)((surf >> i) & 1ull));
    v |= max(bnd, min(mm, cv)) < cv;
    pv = cv;
    plo = clo;
    phi = chi;
    cv = nv;
    clo = nlo;
    chi = nhi;
  }
  return v;
}

// The same test for the fill / reconstruction order domain (bounds in sd).
template <class G, class L>
__device__ __forceinline__ bool fill_violation(const typename L::T* sf, const typename L::T* sd,
                                               int R0, int C0, int lane) {
  using T = typename L::T;
  const int C = C0 + lane;
  const int S = (lane == 0) ? C - 1 : C + 1;  // outside column for the edge lanes
  T up = sf[G::at(R0 - 1, C)], mid = sf[G::at(R0, C)];
  T eup = sf[G::at(R0 - 1, S)], emid = sf[G::at(R0, S)];
  bool v = false;
#pragma unroll 8
  for (int r = 0; r < 64; r++) {
    const T dn = sf[G::at(R0 + r + 1, C)];
    const T edn = sf[G::at(R0 + r + 1, S)];
    const T dv = sd[G::at(R0 + r, C)];
    const T col = L::lo(L::lo(up, dn), mid);
    const T ecol = L::lo(L::lo(eup, edn), emid);
    const T m = L::lo(col, L::lo(L::from_below(col, ecol), L::from_above(col, ecol)));
    v |= L::less(L::hi(dv, m), mid);
    up = mid;
    mid = dn;
    eup = emid;
    emid = edn;
  }
  return v;
}

// ---------------------------------------------------------------- register sweeps
// Order-domain (uint32) sweeps with the wavefront's 64x64 quadrant held in
// registers: f[i] is cell (R0 + i, C0 + lane) in the column layout
// (vertical sweeps) and cell (R0 + lane, C0 + i) in the row layout
// (horizontal sweeps); the layouts switch through the tile in LDS. Only
// the bound d comes from LDS during a sweep (independent loads, hoisted
// freely), so no step waits on LDS. Each sweep applies the neighbours
// behind it (down: N, NW, NE; up: S, SW, SE; right: W, NW, SW; left: E,
// NE, SE) with lanes' side values by DPP. Lanes 0 and 63 take the
// quadrant's outside line only where the perpendicular sweeps cannot
// supply it: the carried start value of each sweep is the outside row /
// column, and the vertical sweeps' first step reads the four corner cells;
// every other relation to outside cells is applied by a perpendicular
// sweep, so together the four sweeps apply all 8 neighbour relations of
// every cell (the fixed-point test then finds nothing a sweep cannot fix).
// Each sweep runs as two interleaved chains of 32 steps (the second starts
// from the old value at the split) so one chain's dependent DPP / min /
// med3 sequence overlaps the other's.

template <class G, bool VERT, bool FWD>
__device__ __forceinline__ bool ord_sweep_regs(unsigned (&f)[64], const unsigned* sf,
                                               const unsigned (&dl)[64], int R0, int C0,
                                               int lane) {
  // start value: the outside row (vertical) or column (horizontal)
  const int out_k = FWD ? -1 : 64;
  unsigned pc0 = VERT ? sf[G::at(R0 + out_k, C0 + lane)] : sf[G::at(R0 + lane, C0 + out_k)];
  // first-step DPP fill-in for lanes 0 / 63: the corner cells (vertical only)
  unsigned corner = kNoNbr;
  if (VERT) {
    const unsigned lo = sf[G::at(R0 + out_k, C0 - 1)], hi = sf[G::at(R0 + out_k, C0 + 64)];
    corner = (lane == 0) ? lo : hi;
  }
  unsigned changed = 0u;
  unsigned pc1 = f[FWD ? 31 : 32];  // second chain starts from the old split value
#pragma unroll
  for (int s = 0; s < 32; s++) {
    const int i0 = FWD ? s : 63 - s;         // chain 0
    const int i1 = FWD ? 32 + s : 31 - s;    // chain 1
    const unsigned d0 = dl[i0], d1 = dl[i1];
    const unsigned old0 = (s == 0) ? corner : kNoNbr;
    const unsigned a0 = dpp_u32(old0, pc0, 0x138), b0 = dpp_u32(old0, pc0, 0x130);
    const unsigned a1 = dpp_u32(kNoNbr, pc1, 0x138), b1 = dpp_u32(kNoNbr, pc1, 0x130);
    const unsigned m0 = min(min(a0, b0), pc0), m1 = min(min(a1, b1), pc1);
    const unsigned f0 = f[i0], f1 = f[i1];
    const unsigned n0 = max(d0, min(m0, f0)), n1 = max(d1, min(m1, f1));
    changed |= (unsigned)(n0 < f0) | (unsigned)(n1 < f1);
    f[i0] = n0;
    f[i1] = n1;
    pc0 = n0;
    pc1 = n1;
  }
  return changed != 0u;
}

// Fixed-point test of a wavefront's quadrant from the row layout held in
// registers (f[i]: cell (R0 + lane, C0 + i), d the bounds of the same
// cells): lanes are rows walking the columns; the rows above / below come
// by DPP, the edge lanes' outside row and every lane's outside columns
// from LDS. True (in some lane) if a cell would still decrease.
template <class G>
__device__ __forceinline__ bool fill_violation_rows(const unsigned (&f)[64], const unsigned (&d)[64],
                                                    const unsigned* sf, int R0, int C0, int lane) {
  const int R = R0 + lane;
  const int er = (lane == 0) ? R0 - 1 : R0 + 64;  // outside row for the edge lanes
  const unsigned left = sf[G::at(R, C0 - 1)], right = sf[G::at(R, C0 + 64)];
  unsigned ep = sf[G::at(er, C0 - 1)], ec = sf[G::at(er, C0)];
  bool v = false;
#pragma unroll
  for (int i = 0; i < 64; i++) {
    const unsigned op = (i == 0) ? left : f[i - 1];
    const unsigned on = (i == 63) ? right : f[i + 1];
    const unsigned en = sf[G::at(er, C0 + i + 1)];
    const unsigned row = min(min(op, on), f[i]);  // own row, columns i-1..i+1
    const unsigned erow = min(min(ep, en), ec);   // outside row (edge lanes)
    const unsigned m = min(row, min(dpp_u32(erow, row, 0x138), dpp_u32(erow, row, 0x130)));
    v |= max(d[i], m) < f[i];
    ep = ec;
    ec = en;
  }
  return v;
}

// Relaxes a tile to its local fixed point with the register sweeps; the
// tile's state is in sf before and after. Returns the sweep iterations.
template <class G>
__device__ inline unsigned ord_relax_regs(unsigned* sf, const unsigned* sd, int R0, int C0,
                                          int lane, unsigned max_local, int* any_out,
                                          int* changed_out, int* wave_changed_out) {
  unsigned f[64];
  unsigned it = 0;
  // the bounds of the lane's cells in both layouts, fixed during the visit;
  // the first fixed-point test runs on the row layout
  unsigned dcol[64], drow[64];
#pragma unroll
  for (int i = 0; i < 64; i++) {
    drow[i] = sd[G::at(R0 + lane, C0 + i)];
    f[i] = sf[G::at(R0 + lane, C0 + i)];
  }
  int any = __syncthreads_or(fill_violation_rows<G>(f, drow, sf, R0, C0, lane) ? 1 : 0);
  int any_change = 0, wave_change = 0;
  if (any) {
#pragma unroll
    for (int i = 0; i < 64; i++) dcol[i] = sd[G::at(R0 + i, C0 + lane)];
  }
  while (any && it < max_local) {
#pragma unroll
    for (int r = 0; r < 64; r++) f[r] = sf[G::at(R0 + r, C0 + lane)];
    bool ch = ord_sweep_regs<G, true, true>(f, sf, dcol, R0, C0, lane);
    ch |= ord_sweep_regs<G, true, false>(f, sf, dcol, R0, C0, lane);
#pragma unroll
    for (int r = 0; r < 64; r++) sf[G::at(R0 + r, C0 + lane)] = f[r];
#pragma unroll
    for (int c = 0; c < 64; c++) f[c] = sf[G::at(R0 + lane, C0 + c)];
    ch |= ord_sweep_regs<G, false, true>(f, sf, drow, R0, C0, lane);
    ch |= ord_sweep_regs<G, false, false>(f, sf, drow, R0, C0, lane);
#pragma unroll
    for (int c = 0; c < 64; c++) sf[G::at(R0 + lane, C0 + c)] = f[c];
    wave_change |= __any(ch ? 1 : 0);
    any = __syncthreads_or(ch ? 1 : 0);
    any_change |= any;
    it++;
    if (any) any = __syncthreads_or(fill_violation_rows<G>(f, drow, sf, R0, C0, lane) ? 1 : 0);
  }
  *any_out = any;
  *changed_out = any_change;
  *wave_changed_out = wave_change;
  return it;
}

// LDS of one fill visit.
// kModeSpill keeps one kind byte per cell instead of a bound key (see
// SpillBound), so that a 128x64 tile fits two workgroups per CU.
template <class G, int MODE>
struct FillLds {
  using T = typename FillLevel<MODE>::T;
  using B = typename std::conditional<MODE == kModeSpill, unsigned char, T>::type;
  T sf[G::CELLS];
  B sd[G::CELLS];
  T sper[2 * (G::TX + G::TY)];
  unsigned bits;
};

// Stages the keys and kinds of a spill tile + halo (kModeSpill): key =
// (level, spill index) from the levels in state and the current indices;
// NaN and outside cells are inert. Full-width tiles of grids with
// nx % 4 == 0 load four cells per lane (16-byte loads of the DEM, the
// levels and the indices); every load of a batch is issued before any
// value is used.
template <class G>
__device__ inline void stage_spill(const FillView& v, unsigned long long* sf, unsigned char* sd,
                                   long long x0, long long y0) {
  using T = unsigned long long;
  // opaque per call: keeps the compiler from computing every batch's cell
  // addresses once, outside the persistent loop, and spilling them
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  auto store = [&](int r, int c, float raw, float lev, int idx) {
    T f = KeyLevel::kInert;  // NaN and outside cells: never lowered, never lower
    unsigned char kind = kInertCell;
    if (raw == raw) {  // false for outside cells too
      f = KeyLevel::make(lev, (unsigned)idx);
      kind = ord_bits(lev) == ord_bits(raw) ? kSurface : kSubmerged;
    }
    sf[G::at(r, c)] = f;
    sd[G::at(r, c)] = kind;
  };
  auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  const bool vec = ((v.nx & 3) == 0) && (x0 + G::TX <= v.nx) && al(v.src) && al(v.state) &&
                   al(v.spill);
  const float nan = __builtin_nanf("");
  if (vec) {
    // quads of the tile's columns 0..TX-1 over rows -1..TY, in two batches
    constexpr int Q = G::TX / 4, NV = G::ROWS * Q;
    constexpr int UV = (NV + 2 * G::THREADS - 1) / (2 * G::THREADS);
    for (int base = 0; base < NV; base += G::THREADS * UV) {
      float4 raw[UV], lev[UV];
      int4 idx[UV];
#pragma unroll
      for (int u = 0; u < UV; u++) {
        const int k = base + u * G::THREADS + tid;
        const long long gy = y0 + k / Q - 1;
        raw[u] = make_float4(nan, nan, nan, nan);
        lev[u] = make_float4(0.f, 0.f, 0.f, 0.f);
        idx[u] = make_int4(-1, -1, -1, -1);
        if (k < NV && gy >= 0 && gy < v.ny) {
          const size_t g = (size_t)gy * v.nx + x0 + 4 * (k % Q);
          raw[u] = *reinterpret_cast<const float4*>(v.src + g);
          lev[u] = *reinterpret_cast<const float4*>(v.state + g);
          idx[u] = *reinterpret_cast<const int4*>(v.spill + g);
        }
      }
#pragma unroll
      for (int u = 0; u < UV; u++) {
        const int k = base + u * G::THREADS + tid;
        if (k < NV) {
          const int r = k / Q - 1, c = 4 * (k % Q);
          store(r, c, raw[u].x, lev[u].x, idx[u].x);
          store(r, c + 1, raw[u].y, lev[u].y, idx[u].y);
          store(r, c + 2, raw[u].z, lev[u].z, idx[u].z);
          store(r, c + 3, raw[u].w, lev[u].w, idx[u].w);
        }
      }
    }
    // halo columns -1 and TX
    for (int k = tid; k < 2 * G::ROWS; k += G::THREADS) {
      const int r = k % G::ROWS - 1, c = (k < G::ROWS) ? -1 : G::TX;
      const long long gx = x0 + c, gy = y0 + r;
      float raw = nan, lev = 0.f;
      int idx = -1;
      if (gx >= 0 && gx < v.nx && gy >= 0 && gy < v.ny) {
        const size_t g = (size_t)gy * v.nx + gx;
        raw = v.src[g];
        lev = v.state[g];
        idx = v.spill[g];
      }
      store(r, c, raw, lev, idx);
    }
  } else {
    constexpr int E = G::ROWS * (G::TX + 2), U = 8;
    for (int base = 0; base < E; base += G::THREADS * U) {
      float raw[U], lev[U];
      int idx[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int k = base + u * G::THREADS + tid;
        const long long gx = x0 + k % (G::TX + 2) - 1, gy = y0 + k / (G::TX + 2) - 1;
        raw[u] = nan;
        lev[u] = 0.f;
        idx[u] = -1;
        if (k < E && gx >= 0 && gx < v.nx && gy >= 0 && gy < v.ny) {
          const size_t g = (size_t)gy * v.nx + gx;
          raw[u] = v.src[g];
          lev[u] = v.state[g];
          idx[u] = v.spill[g];
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int k = base + u * G::THREADS + tid;
        if (k < E) store(k / (G::TX + 2) - 1, k % (G::TX + 2) - 1, raw[u], lev[u], idx[u]);
      }
    }
  }
}

// Stages the (level, bound) order words of a fill / reconstruction tile +
// halo, converted in registers:
//   kModeFill:  d = dem, f = (bc == 1) ? d : +INF on the first visit, else
//               the stored state;
//   kModeRecon: d = -mask, f = max(d, -marker) (a NaN marker acts as -INF);
// NaN and outside cells are inert (+INF, +INF). Full-width tiles of grids
// with nx % 4 == 0 load four cells per lane and array (16-byte loads; bc
// 4 bytes) in two batches; every load of a batch is issued before any
// value is used.
template <class G, int MODE>
__device__ inline void stage_levels(const FillView& v, unsigned* sf, unsigned* sd, long long x0,
                                    long long y0, bool first) {
  // opaque per call, as in stage_spill
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const bool use_bc = MODE == kModeFill && first;
  const float nan = __builtin_nanf("");
  // s: the bc flag (first fillsinks visit) or the stored state
  auto put = [&](int r, int c, float d, float s) {
    float f;
    if (MODE == kModeFill) {
      f = use_bc ? ((s == 1.f) ? d : kInf) : s;
    } else {
      d = -d;
      f = fmaxf(d, (s == s) ? -s : kInf);
    }
    if (d != d) {
      d = kInf;
      f = kInf;
    }
    sd[G::at(r, c)] = ord_bits(d);
    sf[G::at(r, c)] = ord_bits(f);
  };
  const float s_oob = use_bc ? 0.f : kInf;
  auto al = [](const void* p, uintptr_t m) { return (reinterpret_cast<uintptr_t>(p) & m) == 0; };
  const bool vec = ((v.nx & 3) == 0) && (x0 + G::TX <= v.nx) && al(v.src, 15) &&
                   (use_bc ? al(v.bc, 3) : al(v.state, 15));
  if (vec) {
    constexpr int Q = G::TX / 4, NV = G::ROWS * Q;
    constexpr int UV = (NV + 2 * G::THREADS - 1) / (2 * G::THREADS);
    for (int base = 0; base < NV; base += G::THREADS * UV) {
      // raw loads only here; conversions at the LDS stores (a conversion
      // inside the bounds branch would make each load wait there)
      // defaults (outside cells) are the values to stage; a bc flag of 0
      // stages as s_oob does on first visits
      float4 dv[UV], sv[UV];
      uchar4 bv[UV];
#pragma unroll
      for (int u = 0; u < UV; u++) {
        const int idx = base + u * G::THREADS + tid;
        const long long gy = y0 + idx / Q - 1;
        dv[u] = make_float4(nan, nan, nan, nan);
        sv[u] = make_float4(s_oob, s_oob, s_oob, s_oob);
        bv[u] = make_uchar4(0, 0, 0, 0);
        if (idx < NV && gy >= 0 && gy < v.ny) {
          const size_t g = (size_t)gy * v.nx + x0 + 4 * (idx % Q);
          dv[u] = *reinterpret_cast<const float4*>(v.src + g);
          if (use_bc)
            bv[u] = *reinterpret_cast<const uchar4*>(v.bc + g);
          else
            sv[u] = *reinterpret_cast<const float4*>(v.state + g);
        }
      }
#pragma unroll
      for (int u = 0; u < UV; u++) {
        const int idx = base + u * G::THREADS + tid;
        if (idx < NV) {
          const int r = idx / Q - 1, c = 4 * (idx % Q);
          const float4 d4 = dv[u];
          const float4 s4 = use_bc ? make_float4(bv[u].x, bv[u].y, bv[u].z, bv[u].w) : sv[u];
          put(r, c, d4.x, s4.x);
          put(r, c + 1, d4.y, s4.y);
          put(r, c + 2, d4.z, s4.z);
          put(r, c + 3, d4.w, s4.w);
        }
      }
    }
    // halo columns -1 and TX
    for (int idx = tid; idx < 2 * G::ROWS; idx += G::THREADS) {
      const int r = idx % G::ROWS - 1, c = (idx < G::ROWS) ? -1 : G::TX;
      const long long gx = x0 + c, gy = y0 + r;
      float d = nan, sv = s_oob;
      uint8_t bv = 0;
      if (gx >= 0 && gx < v.nx && gy >= 0 && gy < v.ny) {
        const size_t g = (size_t)gy * v.nx + gx;
        d = v.src[g];
        if (use_bc)
          bv = v.bc[g];
        else
          sv = v.state[g];
      }
      put(r, c, d, use_bc ? (float)bv : sv);
    }
  } else {
    constexpr int E = G::ROWS * (G::TX + 2), U = 8;
    for (int base = 0; base < E; base += G::THREADS * U) {
      float dv[U], sv[U];
      uint8_t bv[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int idx = base + u * G::THREADS + tid;
        const long long gx = x0 + idx % (G::TX + 2) - 1, gy = y0 + idx / (G::TX + 2) - 1;
        dv[u] = nan;
        sv[u] = s_oob;
        bv[u] = 0;
        if (idx < E && gx >= 0 && gx < v.nx && gy >= 0 && gy < v.ny) {
          const size_t g = (size_t)gy * v.nx + gx;
          dv[u] = v.src[g];
          if (use_bc)
            bv[u] = v.bc[g];
          else
            sv[u] = v.state[g];
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int idx = base + u * G::THREADS + tid;
        if (idx < E)
          put(idx / (G::TX + 2) - 1, idx % (G::TX + 2) - 1, dv[u], use_bc ? (float)bv[u] : sv[u]);
      }
    }
  }
}

// Writes the wavefront's own 64x64 quadrant (tile rows R0.., columns C0..)
// back, for fill / reconstruction: a quadrant the visit did not change is
// left as it is in memory. Same values as write_tile.
template <class G, int MODE>
__device__ inline void write_quadrant(const FillView& v, const unsigned* sf, const unsigned* sd,
                                      long long x0, long long y0, int R0, int C0, int lane) {
  auto level_out = [&](int at, size_t g) {
    float out = unord_bits(sf[at]);
    if (sd[at] == OrdLevel::kInert) {
      const float raw = v.src[g];
      if (raw != raw) out = (MODE == kModeFill) ? raw : -raw;
    }
    return (MODE == kModeRecon) ? -out : out;
  };
  const long long qx = x0 + C0, qy = y0 + R0;
  if (((v.nx & 3) == 0) && (qx + 64 <= v.nx) &&
      (reinterpret_cast<uintptr_t>(v.state) & 15) == 0) {
#pragma unroll 4
    for (int k = 0; k < 16; k++) {
      const int q = lane + 64 * k;
      const int r = q >> 4, c = 4 * (q & 15);
      if (qy + r >= v.ny) continue;
      const size_t g = (size_t)(qy + r) * v.nx + qx + c;
      const int at = G::at(R0 + r, C0 + c);
      *reinterpret_cast<float4*>(v.state + g) =
          make_float4(level_out(at, g), level_out(at + 1, g + 1), level_out(at + 2, g + 2),
                      level_out(at + 3, g + 3));
    }
  } else {
    for (int r = 0; r < 64; r++) {
      const long long gx = qx + lane, gy = qy + r;
      if (gx >= v.nx || gy >= v.ny) continue;
      const size_t g = (size_t)gy * v.nx + gx;
      v.state[g] = level_out(G::at(R0 + r, C0 + lane), g);
    }
  }
}

// The spill pass's version: the wavefront's quadrant of spill indices
// (the key's index halves; the levels never change), -1 for inert cells.
template <class G>
__device__ inline void write_spill_quadrant(const FillView& v, const unsigned long long* sf,
                                            const unsigned char* sd, long long x0, long long y0,
                                            int R0, int C0, int lane) {
  const long long qx = x0 + C0, qy = y0 + R0;
  auto spill_of = [&](int at) { return sd[at] == kInertCell ? -1 : (int)(unsigned)sf[at]; };
  if (((v.nx & 3) == 0) && (qx + 64 <= v.nx) && (reinterpret_cast<uintptr_t>(v.spill) & 15) == 0) {
#pragma unroll 4
    for (int k = 0; k < 16; k++) {
      const int q = lane + 64 * k;
      const int r = q >> 4, c = 4 * (q & 15);
      if (qy + r >= v.ny) continue;
      const size_t g = (size_t)(qy + r) * v.nx + qx + c;
      const int at = G::at(R0 + r, C0 + c);
      *reinterpret_cast<int4*>(v.spill + g) =
          make_int4(spill_of(at), spill_of(at + 1), spill_of(at + 2), spill_of(at + 3));
    }
  } else {
    for (int r = 0; r < 64; r++) {
      const long long gx = qx + lane, gy = qy + r;
      if (gx >= v.nx || gy >= v.ny) continue;
      v.spill[(size_t)gy * v.nx + gx] = spill_of(G::at(R0 + r, C0 + lane));
    }
  }
}

// Relaxes the tile at (x0, y0) of grid v to its local fixed point (at most
// max_local sweep iterations; kSelf is returned if it stopped there) and
// writes it back. Returns the neighbours to queue. Called by all threads
// of the workgroup; `sync` receives the visit statistics. With `lev`
// (kModeFill, kModeRecon; LDS, 8 words), lev[k] is lowered to the order
// key of the lowest level (in the visit's domain) the tile pushes onto
// neighbour k.
template <class G, int MODE>
__device__ inline unsigned fill_visit(const FillView& v, long long x0, long long y0, bool first,
                                      unsigned nvis, unsigned max_local, Sync* sync,
                                      FillLds<G, MODE>& lds, unsigned* lev = nullptr) {
  using L = FillLevel<MODE>;
  using T = typename L::T;
  T* sf = lds.sf;
  auto* sd = lds.sd;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int R0 = (wave / G::QX) * 64;
  const int C0 = (wave % G::QX) * 64;
  const unsigned long long c0 = wall_clock64();
  // ---- stage tile + halo
  if constexpr (MODE == kModeSpill) {
    stage_spill<G>(v, sf, sd, x0, y0);
  } else {
    stage_levels<G, MODE>(v, sf, sd, x0, y0, first);
  }
  __syncthreads();
  // halo cells whose owner tile's stored state availability differs from
  // what was staged (the recon marker and the spill indices are always
  // initialised)
  if constexpr (MODE == kModeFill) {
    for_stale_halo<G>(first, nvis, [&](int r, int c, bool from_state) {
      const long long gx = x0 + c, gy = y0 + r;
      if (gx < 0 || gx >= v.nx || gy < 0 || gy >= v.ny) return;
      const size_t g = (size_t)gy * v.nx + gx;
      const float raw = v.src[g];
      if (raw != raw) return;  // inert either way
      sf[G::at(r, c)] = ord_bits(from_state ? v.state[g] : ((v.bc[g] == 1) ? raw : kInf));
    });
    __syncthreads();
  }
  // the perimeter as the neighbours last saw it (staged state), before
  // anything below lowers it: every lowering from here on is compared with
  // it to decide which neighbours to queue
  save_perimeter<G>(lds.sper, sf);
  // the bound of each cell (kModeSpill); tile cell (0, 0) is grid cell
  // y0 * nx + x0, below 2^31 like every spill index
  const SpillBound bound{reinterpret_cast<const unsigned char*>(lds.sd),
                         (unsigned)(y0 * v.nx + x0), (unsigned)v.nx};
  const unsigned long long c1 = wall_clock64();
  // ---- relax: sweep iterations until the parallel fixed-point test
  // passes (also tried before the first one)
  unsigned it = 0;
  int any = 0, any_change = 0, wave_change = 0;
  if constexpr (MODE == kModeSpill) {
    const unsigned long long vsurf = surface_bits<G, true>(bound.kind, R0, C0, lane);
    const unsigned long long hsurf = surface_bits<G, false>(bound.kind, R0, C0, lane);
    any = __syncthreads_or(spill_violation<G>(sf, vsurf, bound, R0, C0, lane) ? 1 : 0);
    // indices in registers (spill_sweep_regs); the keys' index halves in
    // LDS are updated after each layout's sweeps
    unsigned* w = reinterpret_cast<unsigned*>(sf);
    unsigned x[64], mcol[16], mrow[16];
    if (any) {
      level_masks<G, true>(mcol, sf, R0, C0, lane);
      level_masks<G, false>(mrow, sf, R0, C0, lane);
    }
    const unsigned vline0 = bound.base + (unsigned)R0 * bound.nx + (unsigned)(C0 + lane);
    const unsigned hline0 = bound.base + (unsigned)(R0 + lane) * bound.nx + (unsigned)C0;
    while (any && it < max_local) {
      // opaque per iteration: the per-step mask bits and bounds are
      // recomputed in the sweeps instead of being hoisted out of this loop
      // (about a thousand registers' worth)
      unsigned long long vs = vsurf, hs = hsurf;
      unsigned vl = vline0, hl = hline0;
      asm volatile("" : "+v"(vs), "+v"(hs), "+v"(vl), "+v"(hl));
#pragma unroll
      for (int k = 0; k < 16; k++) asm volatile("" : "+v"(mcol[k]), "+v"(mrow[k]));
#pragma unroll
      for (int r = 0; r < 64; r++) x[r] = w[2 * G::at(R0 + r, C0 + lane)];
      bool ch = spill_sweep_regs<G, true, true>(x, sf, mcol, vs, vl, bound.nx, R0, C0, lane);
      ch |= spill_sweep_regs<G, true, false>(x, sf, mcol, vs, vl, bound.nx, R0, C0, lane);
#pragma unroll
      for (int r = 0; r < 64; r++) w[2 * G::at(R0 + r, C0 + lane)] = x[r];
#pragma unroll
      for (int c = 0; c < 64; c++) x[c] = w[2 * G::at(R0 + lane, C0 + c)];
      ch |= spill_sweep_regs<G, false, true>(x, sf, mrow, hs, hl, 1u, R0, C0, lane);
      ch |= spill_sweep_regs<G, false, false>(x, sf, mrow, hs, hl, 1u, R0, C0, lane);
#pragma unroll
      for (int c = 0; c < 64; c++) w[2 * G::at(R0 + lane, C0 + c)] = x[c];
      wave_change |= __any(ch ? 1 : 0);
      any = __syncthreads_or(ch ? 1 : 0);
      any_change |= any;
      it++;
      if (any) {
        // opaque again: nothing derived from the masks stays live from the
        // sweeps into the test
        asm volatile("" : "+v"(hs), "+v"(hl));
#pragma unroll
        for (int k = 0; k < 16; k++) asm volatile("" : "+v"(mrow[k]));
        any = __syncthreads_or(
            spill_violation_rows<G>(x, sf, mrow, hs, hl, R0, C0, lane) ? 1 : 0);
      }
    }
  } else {
    it = ord_relax_regs<G>(sf, sd, R0, C0, lane, max_local, &any, &any_change, &wave_change);
  }
  const unsigned long long c2 = wall_clock64();
  // ---- write back: the first visit initialises the tile's state; later
  // visits skip tiles whose relaxation changed nothing. Inert cells (NaN
  // input) take the input bits.
  if constexpr (MODE == kModeSpill) {
    if (wave_change) write_spill_quadrant<G>(v, sf, sd, x0, y0, R0, C0, lane);
  } else {
    // each wavefront writes its own quadrant if its sweeps changed it
    if (first || wave_change) write_quadrant<G, MODE>(v, sf, sd, x0, y0, R0, C0, lane);
  }
  // ---- changed perimeter -> neighbour tiles whose cells it may lower
  unsigned bits = block_or(pushes_from_perimeter<G, true>(lds.sper, sf,
                                                          [&](int p, int q) {
                                                            T dq;
                                                            if constexpr (MODE == kModeSpill)
                                                              dq = bound(q, q / G::P - 1,
                                                                         q % G::P - 1, sf[q]);
                                                            else
                                                              dq = sd[q];
                                                            const T to = L::hi(dq, sf[p]);
                                                            const bool lowers = L::less(to, sf[q]);
                                                            if constexpr (MODE != kModeSpill)
                                                              if (lowers && lev)
                                                                atomicMin(&lev[halo_owner<G>(
                                                                              q / G::P - 1,
                                                                              q % G::P - 1)],
                                                                          to);
                                                            return lowers;
                                                          }),
                           &lds.bits);
  if (any) bits |= kSelf;
  if (tid == 0) visit_stats(sync, it, c0, c1, c2);
  return bits;
}

// ---------------------------------------------------------------- spill relaxation schedule
// The spill relaxation starts from the labelling's indices (spill_ccl.h),
// at which only tiles holding a flat cell can break the fixed point: their
// dirty flag is set. A first visit that finds its tile's flag clear has
// nothing to do and is skipped; a visit that queues neighbours sets their
// flags first (spill_mark, before the release fence that precedes the
// pushes), so a tile whose neighbour changed before its first visit sees
// the flag set, and one changed later is queued again (a later visit is
// never skipped). Returns true in every thread when the visit is skipped.
__device__ inline bool spill_skip(const unsigned* dirty, int t, bool first, unsigned* s_flag) {
  if (!first) return false;
  if (threadIdx.x == 0) *s_flag = ld_relaxed(&dirty[t]);
  __syncthreads();
  return *s_flag == 0u;
}

// Thread 0: flags the neighbours a visit queues (bits as run_queue takes them).
template <class Topo>
__device__ inline void spill_mark(unsigned* dirty, const Topo& topo, int t, unsigned bits) {
  if (threadIdx.x != 0) return;
  for (int k = 0; k < 8; k++) {
    if (!(bits & (1u << k))) continue;
    const int n = topo.neighbour(t, k);
    if (n >= 0) __hip_atomic_store(&dirty[n], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

}  // namespace tt
