// Sink filling and grayscale reconstruction on gfx950.
//
// Reference semantics (src/fillsinks.c:22-48, src/reconstruct.c:68-112,
// 134-202, 270-279): fillsinks negates the DEM, seeds the marker with the
// negated DEM on bc==1 cells and -INF elsewhere, reconstructs by dilation
// (marker = min(mask, max(marker, 8 neighbours)) until nothing changes) and
// negates back. That fixed point is unique and every value is a copy of an
// input value, so any update order reaches the same bits.
//
// The kernels work directly in the "fill domain" f = -marker, d = -mask:
//     f_p <- max(d_p, min(f_p, f_q for the 8 neighbours q))
// iterated from f0 = d on seeds (+INF elsewhere) down to the greatest fixed
// point. NaN cells of d are inert: level +INF, bound +INF (fminf skips them
// as fmaxf skips NaN in the reference), written back as the input NaN.
//
// fillsinks:   d = dem, f0 = (bc == 1) ? dem : +INF; the output array
//              holds f between visits.
// reconstruct: d = -mask, f0 = max(d, -marker) (a NaN marker acts as -INF);
//              the marker array holds -f between visits and at the end.
// spill:       fillsinks over lexicographic (level, cell index) keys packed
//              in 64 bits; the index half of the fixed point is the cell
//              whose DEM value the output copies (the spill cell), which
//              the adjoint scatters to. Unique fixed point, so deterministic.
//              Halo reads race with neighbour write-backs,
//              which is harmless only because every read value is a valid
//              upper bound: hence one 64-bit word per key, never two arrays.
#include <type_traits>

#include "kernels.h"
#include "tile_engine.h"
#include "workspace.h"

namespace tt {

enum : int { kModeSpill = 2 };

struct FillArgs {
  float* state;
  int* spill;                // kModeSpill only
  unsigned long long* keys;  // kModeSpill only: the iteration state
  const float* src;
  const uint8_t* bc;
  long long nx, ny;
  int ntx, nty;
  TileQueue q;
  unsigned max_local;
};

// ---------------------------------------------------------------- level types
struct FloatLevel {
  using T = float;
  static __device__ __forceinline__ T lo(T a, T b) { return fminf(a, b); }
  static __device__ __forceinline__ T hi(T a, T b) { return fmaxf(a, b); }
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

// One directional sweep over a 64x64 quadrant with lanes along columns
// (vertical sweep: down = +y) or along rows (horizontal: right = +x).
// The value carried between steps is the lane's previous cell; its two
// side neighbours come from the adjacent lanes by DPP, with the quadrant's
// outside column/row read from LDS for the edge lanes. The next step's LDS
// operands are fetched before this step's store.
template <class G, class L, bool VERT, bool FWD>
__device__ __forceinline__ bool fill_sweep(typename L::T* sf, const typename L::T* sd, int R0,
                                           int C0, int lane) {
  using T = typename L::T;
  constexpr int step = FWD ? 1 : -1;
  bool changed = false;
  const int own = VERT ? C0 + lane : R0 + lane;      // lane's column / row
  const int side = (lane == 0) ? own - 1 : own + 1;  // outside neighbour line
  const int k0 = VERT ? (FWD ? R0 : R0 + 63) : (FWD ? C0 : C0 + 63);
  auto cell = [&](int k) { return VERT ? G::at(k, own) : G::at(own, k); };
  auto edge_cell = [&](int k) { return VERT ? G::at(k, side) : G::at(side, k); };
  T pc = sf[cell(k0 - step)];
  T fc_n = sf[cell(k0)], dv_n = sd[cell(k0)], e_n = sf[edge_cell(k0 - step)];
  int k = k0;
#pragma unroll 4
  for (int s = 0; s < 64; s++, k += step) {
    const T fc = fc_n, dv = dv_n, edge = e_n;
    if (s < 63) {
      fc_n = sf[cell(k + step)];
      dv_n = sd[cell(k + step)];
      e_n = sf[edge_cell(k)];
    }
    const T a = L::from_below(pc, edge);
    const T b = L::from_above(pc, edge);
    const T nv = L::hi(dv, L::lo(L::lo(a, b), L::lo(pc, fc)));
    const bool upd = nv < fc;
    if (upd) sf[cell(k)] = nv;
    changed |= upd;
    pc = L::lo(nv, fc);
  }
  return changed;
}

template <class G, int MODE>
__global__ __launch_bounds__(G::THREADS) void fill_persistent(FillArgs a) {
  using L = typename std::conditional<MODE == kModeSpill, KeyLevel, FloatLevel>::type;
  using T = typename L::T;
  __shared__ T sf[G::CELLS];
  __shared__ T sd[G::CELLS];
  __shared__ T sper[2 * (G::TX + G::TY)];
  __shared__ unsigned s_bits;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int R0 = (wave / G::QX) * 64;
  const int C0 = (wave % G::QX) * 64;

  auto visit = [&](int t, bool first, unsigned nvis) -> unsigned {
    const unsigned long long c0 = wall_clock64();
    const int tx = t % a.ntx, ty = t / a.ntx;
    const long long x0 = (long long)tx * G::TX, y0 = (long long)ty * G::TY;
    // ---- stage tile + halo
    if constexpr (MODE == kModeSpill) {
      constexpr int E = G::ROWS * (G::TX + 2);
      constexpr int U = 8;
      for (int base = 0; base < E; base += G::THREADS * U) {
        T fv[U], dv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int idx = base + u * G::THREADS + tid;
          const int r = idx / (G::TX + 2) - 1;
          const int c = idx % (G::TX + 2) - 1;
          const long long gx = x0 + c, gy = y0 + r;
          T f = L::kUnreached, d = L::kInert;
          if (idx < E && gx >= 0 && gx < a.nx && gy >= 0 && gy < a.ny) {
            const size_t g = (size_t)gy * a.nx + gx;
            const float raw = a.src[g];
            if (raw == raw) {
              d = KeyLevel::make(raw, (unsigned)g);
              f = first ? ((a.bc[g] == 1) ? d : KeyLevel::kUnreached) : a.keys[g];
            }
          }
          fv[u] = f;
          dv[u] = d;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int idx = base + u * G::THREADS + tid;
          if (idx < E) {
            const int r = idx / (G::TX + 2) - 1;
            const int c = idx % (G::TX + 2) - 1;
            sf[G::at(r, c)] = fv[u];
            sd[G::at(r, c)] = dv[u];
          }
        }
      }
    } else {
      // raw inputs first (bound source in sd; bc flag or stored state in sf),
      // then one LDS pass turns them into (level, bound)
      const StageF32 src{a.src, sd, __builtin_nanf("")};
      if (MODE == kModeFill && first)
        stage_tile<G>(src, StageU8{a.bc, sf, 0.f}, true, x0, y0, a.nx, a.ny);
      else
        stage_tile<G>(src, StageF32{a.state, sf, kInf}, true, x0, y0, a.nx, a.ny);
      __syncthreads();
      for (int k = tid; k < G::CELLS; k += G::THREADS) {
        float d = sd[k], f = sf[k];
        if (MODE == kModeFill) {
          if (first) f = (f == 1.f) ? d : kInf;
        } else {
          d = -d;
          f = fmaxf(d, (f == f) ? -f : kInf);
        }
        if (d != d) {  // NaN input: inert cell
          d = kInf;
          f = kInf;
        }
        sd[k] = d;
        sf[k] = f;
      }
    }
    __syncthreads();
    // halo cells whose owner tile's stored state availability differs from
    // what was staged (recon state is always initialised: the marker)
    if (MODE != kModeRecon) {
      for_stale_halo<G>(first, nvis, [&](int r, int c, bool from_state) {
        const long long gx = x0 + c, gy = y0 + r;
        if (gx < 0 || gx >= a.nx || gy < 0 || gy >= a.ny) return;
        const size_t g = (size_t)gy * a.nx + gx;
        const float raw = a.src[g];
        if (raw != raw) return;  // inert either way
        if constexpr (MODE == kModeSpill) {
          sf[G::at(r, c)] = from_state ? a.keys[g]
                                       : ((a.bc[g] == 1) ? KeyLevel::make(raw, (unsigned)g)
                                                         : KeyLevel::kUnreached);
        } else {
          sf[G::at(r, c)] = from_state ? a.state[g] : ((a.bc[g] == 1) ? raw : kInf);
        }
      });
      __syncthreads();
    }
    save_perimeter<G>(sper, sf);
    const unsigned long long c1 = wall_clock64();
    // ---- relax to the local fixed point
    unsigned it = 0;
    int any, any_change = 0;
    do {
      bool ch = fill_sweep<G, L, true, true>(sf, sd, R0, C0, lane);
      ch |= fill_sweep<G, L, true, false>(sf, sd, R0, C0, lane);
      ch |= fill_sweep<G, L, false, true>(sf, sd, R0, C0, lane);
      ch |= fill_sweep<G, L, false, false>(sf, sd, R0, C0, lane);
      any = __syncthreads_or(ch ? 1 : 0);
      any_change |= any;
      it++;
    } while (any && it < a.max_local);
    const unsigned long long c2 = wall_clock64();
    // ---- write back: the first visit initialises the tile's state; later
    // visits skip tiles whose relaxation changed nothing. Inert cells (NaN
    // input) take the input bits.
    if (first || any_change) {
      for (int idx = tid; idx < G::TY * G::TX; idx += G::THREADS) {
        const int r = idx / G::TX, c = idx % G::TX;
        const long long gx = x0 + c, gy = y0 + r;
        if (gx < a.nx && gy < a.ny) {
          const size_t g = (size_t)gy * a.nx + gx;
          const T v = sf[G::at(r, c)];
          const bool inert = sd[G::at(r, c)] == L::kInert;
          if constexpr (MODE == kModeSpill) {
            // one aligned 64-bit store: a neighbour reading its halo
            // concurrently sees the old or the new key, never a mix
            a.keys[g] = v;
            a.state[g] = inert ? a.src[g] : unord_bits((unsigned)(v >> 32));
            a.spill[g] = inert ? -1 : (int)(unsigned)v;
          } else {
            float out = v;
            if (inert) {
              const float raw = a.src[g];
              if (raw != raw) out = (MODE == kModeFill) ? raw : -raw;
            }
            a.state[g] = (MODE == kModeRecon) ? -out : out;
          }
        }
      }
    }
    // ---- changed perimeter -> neighbour tiles to queue
    const unsigned bits = block_or(changed_perimeter<G, true>(sper, sf), &s_bits);
    if (tid == 0) visit_stats(a.q.sync, it, any != 0, c0, c1, c2);
    return bits;
  };

  run_queue(a.q, a.ntx, a.nty, visit);
}

template <class G, int MODE>
static hipError_t launch_fill(FillArgs a, hipStream_t stream, RunStats* stats) {
  a.ntx = (int)((a.nx + G::TX - 1) / G::TX);
  a.nty = (int)((a.ny + G::TY - 1) / G::TY);
  a.max_local = 1u << 16;
  return launch_persistent(fill_persistent<G, MODE>, G::THREADS, a.ntx * a.nty, a, stream, stats,
                           G::TX, G::TY);
}

template <int MODE>
static hipError_t dispatch_fill(FillArgs a, TileShape ts, hipStream_t stream, RunStats* stats) {
  if constexpr (MODE == kModeSpill) {
    // 64-bit keys: at most two quadrants fit LDS
    if (ts.qx == 1 && ts.qy == 1) return launch_fill<TileGeom<1, 1>, MODE>(a, stream, stats);
    if (ts.qx == 1 && ts.qy == 2) return launch_fill<TileGeom<1, 2>, MODE>(a, stream, stats);
    return launch_fill<TileGeom<2, 1>, MODE>(a, stream, stats);
  } else {
    if (ts.qx == 2 && ts.qy == 2) return launch_fill<TileGeom<2, 2>, MODE>(a, stream, stats);
    if (ts.qx == 2 && ts.qy == 1) return launch_fill<TileGeom<2, 1>, MODE>(a, stream, stats);
    if (ts.qx == 1 && ts.qy == 2) return launch_fill<TileGeom<1, 2>, MODE>(a, stream, stats);
    return launch_fill<TileGeom<1, 1>, MODE>(a, stream, stats);
  }
}

hipError_t run_fill(float* state, const float* src, const uint8_t* bc, int64_t nx, int64_t ny,
                    int mode, TileShape ts, hipStream_t stream, RunStats* stats) {
  if (nx <= 0 || ny <= 0) return hipSuccess;
  FillArgs a{};
  a.state = state;
  a.src = src;
  a.bc = bc;
  a.nx = nx;
  a.ny = ny;
  if (mode == kModeFill) return dispatch_fill<kModeFill>(a, ts, stream, stats);
  return dispatch_fill<kModeRecon>(a, ts, stream, stats);
}

hipError_t run_fill_spill(float* out, int* spill, const float* dem, const uint8_t* bc,
                          int64_t nx, int64_t ny, TileShape ts, hipStream_t stream,
                          RunStats* stats) {
  if (nx <= 0 || ny <= 0) return hipSuccess;
  FillArgs a{};
  a.state = out;
  a.spill = spill;
  a.src = dem;
  a.bc = bc;
  a.nx = nx;
  a.ny = ny;
  // (level, index) keys live in one 64-bit array between visits; the
  // separate level and index outputs are written alongside
  hipError_t e = hipMallocAsync((void**)&a.keys, (size_t)nx * ny * 8, stream);
  if (e != hipSuccess) return e;
  e = dispatch_fill<kModeSpill>(a, ts, stream, stats);
  hipError_t f = hipFreeAsync(a.keys, stream);
  return e != hipSuccess ? e : f;
}

}  // namespace tt
