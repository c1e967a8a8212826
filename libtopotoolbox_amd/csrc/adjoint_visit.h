// Tile visit and per-cell helpers of the excess-topography adjoint, shared
// by the single-grid kernels (adjoint.hip) and the batched kernel
// (batched.hip).
//
// At the converged state u, every cell is either constrained (u == dem:
// du/ddem = 1) or eikonal (u = G(u1, u2, fi) with fi = cellsize * slope),
// and its upwind parents hold strictly lower values, so L = sum(g_out * u)
// has adjoint lambda = g_out + J^T lambda, a triangular system solved by
// the persistent tiled relaxation (pulling lambda from the cells that use a
// cell as parent). Then g_dem = lambda on constrained cells, g_slopes =
// lambda * du/dfi * cellsize on eikonal cells; NaN (no-data) cells get zero
// gradients. The local linearisation follows the selection rules of the
// FP64 oracle (x parent: i+1 unless i-1 is strictly lower; y parent: j+1
// unless j-1 is strictly lower; one-sided ties to x).
#pragma once

#include "eikonal.h"
#include "kernels.h"
#include "tile_engine.h"

namespace tt {

// Per cell: sx = +w / -w for an x parent at i+1 / i-1 (0: none), sy likewise
// for j+1 / j-1, dfic = du/dfi * cellsize on eikonal cells, NaN on
// constrained cells, -INF on NaN (no-data) cells, whose gradients are 0.
// Local linearisation of cell q of the converged state u (see above).
__device__ __forceinline__ void fsm_linearize_cell(size_t q, float* sx, float* sy, float* dfic,
                                                 const float* u, const float* dem,
                                                 const float* slopes, float cellsize,
                                                 long long nx, long long ny) {
  const long long i = (long long)(q % nx), j = (long long)(q / nx);
  const float uq = u[q];
  float wx = 0.f, wy = 0.f, dc = (uq == uq) ? __builtin_nanf("") : -kInf;
  if (uq < dem[q]) {
    const float S = i < nx - 1 ? u[q + 1] : kInf;
    const float N = i > 0 ? u[q - 1] : kInf;
    const float E = j < ny - 1 ? u[q + nx] : kInf;
    const float W = j > 0 ? u[q - nx] : kInf;
    float u1 = kInf, u2 = kInf;
    float ax = 0.f, ay = 0.f;
    if (S == S && !(N < S)) {
      u1 = S;
      ax = S < kInf ? 1.f : 0.f;
    } else if (N == N) {
      u1 = N;
      ax = N < kInf ? -1.f : 0.f;
    }
    if (E == E && !(W < E)) {
      u2 = E;
      ay = E < kInf ? 1.f : 0.f;
    } else if (W == W) {
      u2 = W;
      ay = W < kInf ? -1.f : 0.f;
    }
    const float fi = cellsize * slopes[q];
    const float d = u1 - u2;
    float dfi;
    if (fabsf(d) < fi) {
      const float D = sqrtf(2.f * fi * fi - d * d);
      wx = ax * ((D - d) / (2.f * D));
      wy = ay * ((D + d) / (2.f * D));
      dfi = fi / D;
    } else {
      if (u1 <= u2)
        wx = ax;
      else
        wy = ay;
      dfi = 1.f;
    }
    dc = dfi * cellsize;
  }
  sx[q] = wx;
  sy[q] = wy;
  dfic[q] = dc;
}

// Gradients of cell q from its converged lambda.
__device__ __forceinline__ void fsm_adjoint_finish_cell(size_t q, float* g_dem, float* g_slopes,
                                                      const float* lam, const float* dfic) {
  const float l = lam[q], dc = dfic[q];
  const bool constrained = dc != dc;
  const bool nodata = dc == -kInf;
  g_dem[q] = constrained ? l : 0.f;
  g_slopes[q] = (constrained || nodata) ? 0.f : l * dc;
}

// One grid as seen by an adjoint tile visit.
struct AdjView {
  float* lam;
  const float* g;
  const float* sx;
  const float* sy;
  long long nx, ny;
};

// Contribution of neighbour lambda l through weight w (0 when w <= 0).
__device__ __forceinline__ float pull(float w, float l) { return w > 0.f ? w * l : 0.f; }

// One directional lambda sweep: lambda = g + up + down + left + right,
// always summed in that order, each term pulled from a neighbour whose
// weight points at the cell. The wavefront's 64 lambdas of its line are in
// registers (l[i]: cell (R0 + i, C0 + lane) in the column layout,
// (R0 + lane, C0 + i) in the row layout); g, the weights and the
// quadrant's outside line come from LDS (independent loads); the side
// neighbours by DPP. Two interleaved chains of 32 steps, the second
// starting from the old value at the split, as in the other register
// sweeps: the triangular system's fixed point is reached from any update
// order.
template <class G, bool VERT, bool FWD>
__device__ __forceinline__ bool adj_sweep_regs(float (&l)[64], const float* sl, const float* sg,
                                               const float* ssx, const float* ssy, int R0,
                                               int C0, int lane) {
  constexpr int step = FWD ? 1 : -1;
  const int own = VERT ? C0 + lane : R0 + lane;
  const int side = (lane == 0) ? own - 1 : own + 1;
  const int base = VERT ? R0 : C0;
  // line position i in [-1, 64]
  auto cell = [&](int i) { return VERT ? G::at(base + i, own) : G::at(own, base + i); };
  auto edge_cell = [&](int i) { return VERT ? G::at(base + i, side) : G::at(side, base + i); };
  const float* s_al = VERT ? ssy : ssx;
  const float* s_ac = VERT ? ssx : ssy;
  const int out_i = FWD ? -1 : 64, end_i = FWD ? 64 : -1;
  const float beyond = sl[cell(end_i)];
  // per chain: lambda and along-axis weight of the cell behind, and the
  // weight of the current cell (the previous step's cell ahead)
  float pl[2], ps[2], sc[2];
  pl[0] = sl[cell(out_i)];
  ps[0] = s_al[cell(out_i)];
  sc[0] = s_al[cell(FWD ? 0 : 63)];
  pl[1] = l[FWD ? 31 : 32];
  ps[1] = s_al[cell(FWD ? 31 : 32)];
  sc[1] = s_al[cell(FWD ? 32 : 31)];
  bool changed = false;
#pragma unroll
  for (int s = 0; s < 32; s++) {
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const int i = FWD ? c * 32 + s : 63 - (c * 32 + s);
      const int j = i + step;
      const float cur = l[i];
      const float g = sg[cell(i)];
      const float lf = (j >= 0 && j < 64) ? l[j] : beyond;
      const float sf = s_al[cell(j)];
      const float sa = s_ac[cell(i)];
      const float le = sl[edge_cell(i)];
      const float se = s_ac[edge_cell(i)];
      const float l_lo = lane_from_below(cur, le), s_lo = lane_from_below(sa, se);
      const float l_hi = lane_from_above(cur, le), s_hi = lane_from_above(sa, se);
      const float l_bf = FWD ? pl[c] : lf, s_bf = FWD ? ps[c] : sf;
      const float l_af = FWD ? lf : pl[c], s_af = FWD ? sf : ps[c];
      float up, dn, lt, rt;
      if (VERT) {
        up = pull(s_bf, l_bf);
        dn = pull(-s_af, l_af);
        lt = pull(s_lo, l_lo);
        rt = pull(-s_hi, l_hi);
      } else {
        lt = pull(s_bf, l_bf);
        rt = pull(-s_af, l_af);
        up = pull(s_lo, l_lo);
        dn = pull(-s_hi, l_hi);
      }
      const float nv = (((g + up) + dn) + lt) + rt;
      changed |= __float_as_uint(nv) != __float_as_uint(cur);
      l[i] = nv;
      pl[c] = nv;
      ps[c] = sc[c];
      sc[c] = sf;
    }
  }
  return changed;
}

// Fixed-point test of a wavefront's quadrant: true (in some lane) if some
// lambda would change. Lanes are columns walking the rows; the steps are
// independent.
template <class G>
__device__ __forceinline__ bool adj_violation(const float* sl, const float* sg, const float* ssx,
                                              const float* ssy, int R0, int C0, int lane) {
  const int C = C0 + lane;
  const int S = (lane == 0) ? C - 1 : C + 1;
  float lup = sl[G::at(R0 - 1, C)], lmid = sl[G::at(R0, C)];
  float yup = ssy[G::at(R0 - 1, C)], ymid = ssy[G::at(R0, C)];
  bool v = false;
#pragma unroll 4
  for (int r = 0; r < 64; r++) {
    const float ldn = sl[G::at(R0 + r + 1, C)];
    const float ydn = ssy[G::at(R0 + r + 1, C)];
    const float g = sg[G::at(R0 + r, C)];
    const float x = ssx[G::at(R0 + r, C)];
    const float le = sl[G::at(R0 + r, S)], xe = ssx[G::at(R0 + r, S)];
    const float up = pull(yup, lup);
    const float dn = pull(-ydn, ldn);
    const float lt = pull(lane_from_below(x, xe), lane_from_below(lmid, le));
    const float rt = pull(-lane_from_above(x, xe), lane_from_above(lmid, le));
    const float nv = (((g + up) + dn) + lt) + rt;
    v |= __float_as_uint(nv) != __float_as_uint(lmid);
    lup = lmid;
    lmid = ldn;
    yup = ymid;
    ymid = ydn;
  }
  return v;
}

// LDS of one adjoint visit.
template <class G>
struct AdjLds {
  float sl[G::CELLS];
  float sg[G::CELLS];
  float ssx[G::CELLS];
  float ssy[G::CELLS];
  unsigned sper[2 * (G::TX + G::TY)];
  unsigned bits;
};

// Relaxes lambda on the tile at (x0, y0) of grid v (at most max_local
// sweep iterations; kSelf is returned if it stopped there) and writes it
// back. Returns the neighbours to queue; called by all threads.
template <class G>
__device__ inline unsigned adj_visit(const AdjView& v, long long x0, long long y0, bool first,
                                     unsigned nvis, unsigned max_local, Sync* sync,
                                     AdjLds<G>& lds) {
  float* sl = lds.sl;
  float* sg = lds.sg;
  float* ssx = lds.ssx;
  float* ssy = lds.ssy;
  const unsigned* sl_bits = reinterpret_cast<const unsigned*>(sl);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int R0 = (wave / G::QX) * 64;
  const int C0 = (wave % G::QX) * 64;
  const unsigned long long c0 = wall_clock64();
  // lambda starts at g; outside the grid everything is 0 (no pulls)
  stage_tile<G>(StageF32{first ? v.g : v.lam, sl, 0.f}, StageF32{v.g, sg, 0.f}, true, x0, y0,
                v.nx, v.ny);
  stage_tile<G>(StageF32{v.sx, ssx, 0.f}, StageF32{v.sy, ssy, 0.f}, true, x0, y0, v.nx, v.ny);
  __syncthreads();
  for_stale_halo<G>(first, nvis, [&](int r, int c, bool from_state) {
    const long long gx = x0 + c, gy = y0 + r;
    if (gx < 0 || gx >= v.nx || gy < 0 || gy >= v.ny) return;
    const size_t q = (size_t)gy * v.nx + gx;
    sl[G::at(r, c)] = from_state ? v.lam[q] : v.g[q];
  });
  __syncthreads();
  save_perimeter<G>(lds.sper, sl_bits);
  const unsigned long long c1 = wall_clock64();
  unsigned it = 0;
  int any = __syncthreads_or(adj_violation<G>(sl, sg, ssx, ssy, R0, C0, lane) ? 1 : 0);
  int any_change = 0;
  float l[64];
  while (any && it < max_local) {
    // opaque per iteration: the LDS reads of g and the weights (which the
    // sweeps never write) stay in the sweeps instead of being hoisted out
    // of this loop into hundreds of registers
    int ln = lane;
    asm volatile("" : "+v"(ln));
#pragma unroll
    for (int r = 0; r < 64; r++) l[r] = sl[G::at(R0 + r, C0 + ln)];
    bool ch = adj_sweep_regs<G, true, true>(l, sl, sg, ssx, ssy, R0, C0, ln);
    ch |= adj_sweep_regs<G, true, false>(l, sl, sg, ssx, ssy, R0, C0, ln);
#pragma unroll
    for (int r = 0; r < 64; r++) sl[G::at(R0 + r, C0 + ln)] = l[r];
#pragma unroll
    for (int c = 0; c < 64; c++) l[c] = sl[G::at(R0 + ln, C0 + c)];
    ch |= adj_sweep_regs<G, false, true>(l, sl, sg, ssx, ssy, R0, C0, ln);
    ch |= adj_sweep_regs<G, false, false>(l, sl, sg, ssx, ssy, R0, C0, ln);
#pragma unroll
    for (int c = 0; c < 64; c++) sl[G::at(R0 + ln, C0 + c)] = l[c];
    any = __syncthreads_or(ch ? 1 : 0);
    any_change |= any;
    it++;
    if (any) any = __syncthreads_or(adj_violation<G>(sl, sg, ssx, ssy, R0, C0, lane) ? 1 : 0);
  }
  const unsigned long long c2 = wall_clock64();
  if (first || any_change) write_tile_f32<G>(v.lam, sl, x0, y0, v.nx, v.ny);
  unsigned bits = block_or(changed_perimeter<G, false>(lds.sper, sl_bits), &lds.bits);
  if (any) bits |= kSelf;
  if (tid == 0) visit_stats(sync, it, c0, c1, c2);
  return bits;
}

// ---------------------------------------------------------------- fill backward
// out[p] = dem[spill[p]], so g_dem[q] = sum of g_out[p] over p with
// spill[p] == q. The finite contributions are summed in 64-bit fixed point,
// where addition is associative: atomics in any order give the same bits.
// The scale 2^e is chosen per call (per batch item) from max |g_out| over
// its n cells: with max < 2^E and n < 2^B, e = 62 - E - B, so every
// contribution is below 2^(62-B) and any sum of at most n of them stays
// below 2^62: no overflow whatever the lakes. A contribution is rounded to
// 2^-e = 2^(E+B-62), i.e. relative resolution 2^(B-62) of max |g_out|
// (2^-35 on 2^26 cells). Non-finite contributions are left out of the sum
// and combined afterwards with combine_nonfinite.
__device__ __forceinline__ unsigned abs_bits_if_finite(float g) {
  return isfinite(g) ? (__float_as_uint(g) & 0x7fffffffu) : 0u;
}

// e for max |g| given by its bits (|g| >= 0 compares as unsigned) over n
// cells; 0 when every g is zero.
__device__ __forceinline__ int fix_exponent(unsigned max_abs_bits, unsigned long long n) {
  if (max_abs_bits == 0u) return 0;
  int E;
  (void)frexpf(__uint_as_float(max_abs_bits), &E);  // max < 2^E
  const int B = 64 - __clzll(n);                     // n < 2^B
  return 62 - E - B;
}

__device__ __forceinline__ unsigned long long to_fixed(float g, int e) {
  return isfinite(g) ? (unsigned long long)llrint(ldexp((double)g, e)) : 0ull;
}
__device__ __forceinline__ float from_fixed(unsigned long long a, int e) {
  return (float)ldexp((double)(long long)a, -e);
}

// IEEE addition of a non-finite g to a sum: NaN when either is NaN or the
// two are opposite infinities, else the infinity. Commutative and
// associative over the classes, so CAS loops in any order give one result.
__device__ __forceinline__ float add_nonfinite(float acc, float g) {
  const float nan = __uint_as_float(0x7fc00000u);
  if (isnan(acc) || isnan(g)) return nan;
  if (isinf(acc) && acc != g) return nan;
  return g;
}

// g_dem[q] := g_dem[q] + g for a non-finite g, after the finite sum is in.
__device__ __forceinline__ void combine_nonfinite(float* g_dem, size_t q, float g) {
  unsigned* w = reinterpret_cast<unsigned*>(g_dem + q);
  unsigned old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (;;) {  // at most two changes: finite -> infinity -> NaN
    const unsigned nb = __float_as_uint(add_nonfinite(__uint_as_float(old), g));
    if (nb == old) return;
    const unsigned prev = atomicCAS(w, old, nb);
    if (prev == old) return;
    old = prev;
  }
}

}  // namespace tt
