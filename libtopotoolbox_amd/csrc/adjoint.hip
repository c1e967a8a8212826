// Adjoint kernels (no reference counterpart: the reference has no
// derivatives). Both are deterministic: the same inputs give the same bits
// whatever the schedule.
//
// fillsinks backward. out[p] = dem[spill[p]] (spill from the lexicographic
// forward), so g_dem[q] = sum of g_out[p] over p with spill[p] == q. The
// sum is accumulated in 64-bit fixed point (2^-32 resolution), where
// addition is associative, so atomics in any order give the same result;
// contributions to one spill cell are pre-summed per workgroup in an LDS
// hash table. Range: |sum| < 2^31 per spill cell.
//
// excesstopography_fsm2d backward. At the converged state u, every cell is
// either constrained (u == dem: du/ddem = 1) or eikonal (u = G(u1, u2, fi)
// with fi = cellsize * slope), and its upwind parents hold strictly lower
// values, so L = sum(g_out * u) has adjoint lambda = g_out + J^T lambda,
// a triangular system solved by the same persistent tiled relaxation as
// the forward (pulling lambda from the cells that use a cell as parent).
// Then g_dem = lambda on constrained cells, g_slopes = lambda * du/dfi *
// cellsize on eikonal cells; NaN (no-data) cells get zero gradients. The
// local linearisation follows the selection rules of the FP64 oracle (x parent: i+1 unless i-1 is strictly lower;
// y parent: j+1 unless j-1 is strictly lower; one-sided ties to x).
#include "eikonal.h"
#include "kernels.h"
#include "tile_engine.h"
#include "workspace.h"

namespace tt {

constexpr double kFixScale = 4294967296.0;  // 2^32

// ---------------------------------------------------------------- fill backward
__global__ void fill_bwd_init(unsigned long long* acc, const int* spill, const float* g_out,
                              size_t n, unsigned* err) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x) {
    const float g = g_out[q];
    if (!isfinite(g)) atomicOr(err, 1u);
    const long long v = isfinite(g) ? llrint((double)g * kFixScale) : 0ll;
    acc[q] = (spill[q] == (int)q) ? (unsigned long long)v : 0ull;
  }
}

constexpr int kBwdThreads = 256;
constexpr int kBwdChunk = 4096;
constexpr int kBwdSlots = 1024;

__global__ __launch_bounds__(kBwdThreads) void fill_bwd_scatter(unsigned long long* acc,
                                                                const int* spill,
                                                                const float* g_out, size_t n) {
  __shared__ int keys[kBwdSlots];
  __shared__ unsigned long long vals[kBwdSlots];
  const int tid = threadIdx.x;
  for (size_t base = (size_t)blockIdx.x * kBwdChunk; base < n;
       base += (size_t)gridDim.x * kBwdChunk) {
    for (int k = tid; k < kBwdSlots; k += kBwdThreads) {
      keys[k] = -1;
      vals[k] = 0ull;
    }
    __syncthreads();
    for (int k = tid; k < kBwdChunk; k += kBwdThreads) {
      const size_t p = base + k;
      if (p >= n) break;
      const int s = spill[p];
      if (s < 0 || s == (int)p) continue;
      const float g = g_out[p];
      if (!isfinite(g)) continue;
      const unsigned long long v = (unsigned long long)llrint((double)g * kFixScale);
      unsigned h = ((unsigned)s * 2654435761u) >> 22;  // 10 bits
      bool done = false;
      for (int probe = 0; probe < 16; probe++) {
        const int prev = atomicCAS(&keys[h], -1, s);
        if (prev == -1 || prev == s) {
          atomicAdd(&vals[h], v);
          done = true;
          break;
        }
        h = (h + 1) & (kBwdSlots - 1);
      }
      if (!done) atomicAdd(&acc[s], v);
    }
    __syncthreads();
    for (int k = tid; k < kBwdSlots; k += kBwdThreads) {
      const int s = keys[k];
      if (s >= 0) atomicAdd(&acc[s], vals[k]);
    }
    __syncthreads();
  }
}

__global__ void fill_bwd_finish(float* g_dem, const unsigned long long* acc, size_t n) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x)
    g_dem[q] = (float)((double)(long long)acc[q] / kFixScale);
}

static int grid_for(size_t n, int threads) {
  size_t b = (n + threads - 1) / threads;
  return (int)(b > 8192 ? 8192 : (b < 1 ? 1 : b));
}

hipError_t run_fill_backward(float* g_dem, const int* spill, const float* g_out, int64_t nx,
                             int64_t ny, hipStream_t stream, uint32_t* nonfinite) {
  const size_t n = (size_t)nx * (size_t)ny;
  if (n == 0) return hipSuccess;
  void* ws = nullptr;
  hipError_t e = hipMallocAsync(&ws, n * 8 + 256, stream);
  if (e != hipSuccess) return e;
  unsigned long long* acc = (unsigned long long*)ws;
  unsigned* err = (unsigned*)((char*)ws + n * 8);
  zero_words<<<1, 64, 0, stream>>>(err, 64);
  fill_bwd_init<<<grid_for(n, 256), 256, 0, stream>>>(acc, spill, g_out, n, err);
  size_t chunks = (n + kBwdChunk - 1) / kBwdChunk;
  int gs = (int)(chunks > 4096 ? 4096 : chunks);
  fill_bwd_scatter<<<gs, kBwdThreads, 0, stream>>>(acc, spill, g_out, n);
  fill_bwd_finish<<<grid_for(n, 256), 256, 0, stream>>>(g_dem, acc, n);
  e = hipGetLastError();
  if (nonfinite && e == hipSuccess) {
    e = hipMemcpyAsync(nonfinite, err, 4, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
  }
  hipError_t f = hipFreeAsync(ws, stream);
  return e != hipSuccess ? e : f;
}

// ---------------------------------------------------------------- FSM backward
// Per cell: sx = +w / -w for an x parent at i+1 / i-1 (0: none), sy likewise
// for j+1 / j-1, dfic = du/dfi * cellsize on eikonal cells, NaN on
// constrained cells, -INF on NaN (no-data) cells, whose gradients are 0.
__global__ void fsm_linearize(float* sx, float* sy, float* dfic, const float* u, const float* dem,
                              const float* slopes, float cellsize, long long nx, long long ny) {
  const size_t n = (size_t)nx * ny;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x) {
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
}

__global__ void fsm_adjoint_finish(float* g_dem, float* g_slopes, const float* lam,
                                   const float* dfic, size_t n) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x) {
    const float l = lam[q], dc = dfic[q];
    const bool constrained = dc != dc;
    const bool nodata = dc == -kInf;
    g_dem[q] = constrained ? l : 0.f;
    g_slopes[q] = (constrained || nodata) ? 0.f : l * dc;
  }
}

struct AdjArgs {
  float* lam;
  const float* g;
  const float* sx;
  const float* sy;
  long long nx, ny;
  int ntx, nty;
  TileQueue q;
  unsigned max_local;
};

// Contribution of neighbour lambda l through weight w (0 when w <= 0).
__device__ __forceinline__ float pull(float w, float l) { return w > 0.f ? w * l : 0.f; }

// lambda = g + up + down + left + right, always summed in that order.
template <class G, bool VERT, bool FWD>
__device__ __forceinline__ bool adj_sweep(float* sl, const float* sg, const float* ssx,
                                          const float* ssy, int R0, int C0, int lane) {
  constexpr int step = FWD ? 1 : -1;
  bool changed = false;
  const int own = VERT ? C0 + lane : R0 + lane;
  const int side = (lane == 0) ? own - 1 : own + 1;
  const int k0 = VERT ? (FWD ? R0 : R0 + 63) : (FWD ? C0 : C0 + 63);
  auto cell = [&](int k) { return VERT ? G::at(k, own) : G::at(own, k); };
  auto edge_cell = [&](int k) { return VERT ? G::at(k, side) : G::at(side, k); };
  // weights along the sweep axis: s_along = sy for vertical sweeps, sx for
  // horizontal ones; across the lanes: the other one
  const float* s_al = VERT ? ssy : ssx;
  const float* s_ac = VERT ? ssx : ssy;
  float pl = sl[cell(k0 - step)];    // lambda of the previous cell along the sweep
  float ps = s_al[cell(k0 - step)];  // its weight along the sweep axis
  int k = k0;
#pragma unroll 2
  for (int s = 0; s < 64; s++, k += step) {
    const float cur = sl[cell(k)];
    const float g = sg[cell(k)];
    const float lf = sl[cell(k + step)];
    const float sf = s_al[cell(k + step)];
    const float sa = s_ac[cell(k)];
    const float le = sl[edge_cell(k)];
    const float se = s_ac[edge_cell(k)];
    // across-lane neighbours: lower lane (column C-1 or row R-1), upper lane
    const float l_lo = lane_from_below(cur, le), s_lo = lane_from_below(sa, se);
    const float l_hi = lane_from_above(cur, le), s_hi = lane_from_above(sa, se);
    // along-axis neighbours: "before" (index k-1) and "after" (index k+1)
    const float l_bf = FWD ? pl : lf, s_bf = FWD ? ps : sf;
    const float l_af = FWD ? lf : pl, s_af = FWD ? sf : ps;
    float up, dn, lt, rt;
    if (VERT) {  // along = y: before = up (R-1), after = down (R+1)
      up = pull(s_bf, l_bf);
      dn = pull(-s_af, l_af);
      lt = pull(s_lo, l_lo);
      rt = pull(-s_hi, l_hi);
    } else {  // along = x: before = left, after = right; lanes are rows
      lt = pull(s_bf, l_bf);
      rt = pull(-s_af, l_af);
      up = pull(s_lo, l_lo);
      dn = pull(-s_hi, l_hi);
    }
    const float nv = (((g + up) + dn) + lt) + rt;
    const bool upd = __float_as_uint(nv) != __float_as_uint(cur);
    if (upd) sl[cell(k)] = nv;
    changed |= upd;
    pl = nv;
    ps = s_al[cell(k)];
  }
  return changed;
}

template <class G>
__global__ __launch_bounds__(G::THREADS) void fsm_adjoint_persistent(AdjArgs a) {
  __shared__ float sl[G::CELLS];
  __shared__ float sg[G::CELLS];
  __shared__ float ssx[G::CELLS];
  __shared__ float ssy[G::CELLS];
  __shared__ unsigned sper[2 * (G::TX + G::TY)];
  __shared__ unsigned s_bits;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int R0 = (wave / G::QX) * 64;
  const int C0 = (wave % G::QX) * 64;
  const unsigned* sl_bits = reinterpret_cast<const unsigned*>(sl);

  auto visit = [&](int t, bool first, unsigned nvis) -> unsigned {
    const unsigned long long c0 = wall_clock64();
    const int tx = t % a.ntx, ty = t / a.ntx;
    const long long x0 = (long long)tx * G::TX, y0 = (long long)ty * G::TY;
    // lambda starts at g; outside the grid everything is 0 (no pulls)
    stage_tile<G>(StageF32{first ? a.g : a.lam, sl, 0.f}, StageF32{a.g, sg, 0.f}, true, x0, y0,
                  a.nx, a.ny);
    stage_tile<G>(StageF32{a.sx, ssx, 0.f}, StageF32{a.sy, ssy, 0.f}, true, x0, y0, a.nx, a.ny);
    __syncthreads();
    for_stale_halo<G>(first, nvis, [&](int r, int c, bool from_state) {
      const long long gx = x0 + c, gy = y0 + r;
      if (gx < 0 || gx >= a.nx || gy < 0 || gy >= a.ny) return;
      const size_t q = (size_t)gy * a.nx + gx;
      sl[G::at(r, c)] = from_state ? a.lam[q] : a.g[q];
    });
    __syncthreads();
    save_perimeter<G>(sper, sl_bits);
    const unsigned long long c1 = wall_clock64();
    unsigned it = 0;
    int any, any_change = 0;
    do {
      bool ch = adj_sweep<G, true, true>(sl, sg, ssx, ssy, R0, C0, lane);
      ch |= adj_sweep<G, true, false>(sl, sg, ssx, ssy, R0, C0, lane);
      ch |= adj_sweep<G, false, true>(sl, sg, ssx, ssy, R0, C0, lane);
      ch |= adj_sweep<G, false, false>(sl, sg, ssx, ssy, R0, C0, lane);
      any = __syncthreads_or(ch ? 1 : 0);
      any_change |= any;
      it++;
    } while (any && it < a.max_local);
    const unsigned long long c2 = wall_clock64();
    if (first || any_change) {
      for (int idx = tid; idx < G::TY * G::TX; idx += G::THREADS) {
        const int r = idx / G::TX, c = idx % G::TX;
        const long long gx = x0 + c, gy = y0 + r;
        if (gx < a.nx && gy < a.ny) a.lam[(size_t)gy * a.nx + gx] = sl[G::at(r, c)];
      }
    }
    const unsigned bits = block_or(changed_perimeter<G, false>(sper, sl_bits), &s_bits);
    if (tid == 0) visit_stats(a.q.sync, it, any != 0, c0, c1, c2);
    return bits;
  };

  run_queue(a.q, a.ntx, a.nty, visit);
}

hipError_t run_fsm_backward(float* g_dem, float* g_slopes, const float* u, const float* dem,
                            const float* slopes, float cellsize, const float* g_out, int64_t nx,
                            int64_t ny, hipStream_t stream, RunStats* stats) {
  const size_t n = (size_t)nx * (size_t)ny;
  if (n == 0) return hipSuccess;
  void* ws = nullptr;
  hipError_t e = hipMallocAsync(&ws, n * 16, stream);
  if (e != hipSuccess) return e;
  float* sx = (float*)ws;
  float* sy = sx + n;
  float* dfic = sy + n;
  float* lam = dfic + n;
  fsm_linearize<<<grid_for(n, 256), 256, 0, stream>>>(sx, sy, dfic, u, dem, slopes, cellsize, nx,
                                                      ny);
  e = hipGetLastError();
  if (e == hipSuccess) {
    using G = TileGeom<2, 1>;  // four float arrays: two quadrants fit LDS
    AdjArgs a{};
    a.lam = lam;
    a.g = g_out;
    a.sx = sx;
    a.sy = sy;
    a.nx = nx;
    a.ny = ny;
    a.ntx = (int)((nx + G::TX - 1) / G::TX);
    a.nty = (int)((ny + G::TY - 1) / G::TY);
    a.max_local = 1u << 16;
    e = launch_persistent(fsm_adjoint_persistent<G>, G::THREADS, a.ntx * a.nty, a, stream, stats,
                          G::TX, G::TY);
  }
  if (e == hipSuccess) {
    fsm_adjoint_finish<<<grid_for(n, 256), 256, 0, stream>>>(g_dem, g_slopes, lam, dfic, n);
    e = hipGetLastError();
  }
  hipError_t f = hipFreeAsync(ws, stream);
  return e != hipSuccess ? e : f;
}

}  // namespace tt
