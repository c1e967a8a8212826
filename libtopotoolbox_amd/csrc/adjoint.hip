// Adjoint kernels on one grid (no reference counterpart: the reference has
// no derivatives). Both adjoints are deterministic: the same inputs give
// the same bits whatever the schedule.
//
// fillsinks backward: the spill scatter in 64-bit fixed point with a scale
// chosen per call from max |g_out| (adjoint_visit.h), so no sum overflows;
// contributions to one spill cell are pre-summed per workgroup in an LDS
// hash table; non-finite g_out entries propagate as in an IEEE sum.
//
// excesstopography_fsm2d backward: linearise, relax lambda with the
// persistent work-queue kernel over adj_visit, finish (adjoint_visit.h
// states the system).
#include "adjoint_visit.h"
#include "workspace.h"

namespace tt {

// ---------------------------------------------------------------- fill backward
// hdr[0]: max |g_out| over the finite entries (float bits), hdr[1]: 1 if
// any entry is non-finite. Also clears the accumulator.
__global__ void fill_bwd_prep(unsigned long long* acc, unsigned* hdr, const float* g_out,
                              size_t n) {
  __shared__ unsigned smax, snf;
  if (threadIdx.x == 0) smax = snf = 0u;
  __syncthreads();
  unsigned m = 0u, nf = 0u;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x) {
    const float g = g_out[q];
    m = max(m, abs_bits_if_finite(g));
    nf |= isfinite(g) ? 0u : 1u;
    acc[q] = 0ull;
  }
  if (m) atomicMax(&smax, m);
  if (nf) atomicOr(&snf, 1u);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (smax) atomicMax(&hdr[0], smax);
    if (snf) atomicOr(&hdr[1], 1u);
  }
}

constexpr int kBwdThreads = 256;
constexpr int kBwdChunk = 4096;
constexpr int kBwdSlots = 1024;

// Every finite g_out[p] with spill[p] >= 0 and spill[p] != p onto
// acc[spill[p]], pre-summed per spill cell in an LDS hash table (a cell
// that is its own spill cell, about half of a DEM, adds its own term in
// fill_bwd_finish instead: through the table, each would take a slot).
__global__ __launch_bounds__(kBwdThreads) void fill_bwd_scatter(unsigned long long* acc,
                                                                const unsigned* hdr,
                                                                const int* spill,
                                                                const float* g_out, size_t n) {
  __shared__ int keys[kBwdSlots];
  __shared__ unsigned long long vals[kBwdSlots];
  const int e = fix_exponent(hdr[0], n);
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
      const unsigned long long v = to_fixed(g_out[p], e);
      if (v == 0ull) continue;
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

// g_dem from the accumulator plus the cell's own term when it is its own
// spill cell (the same integer sum in another order).
__global__ void fill_bwd_finish(float* g_dem, const unsigned long long* acc, const unsigned* hdr,
                                const int* spill, const float* g_out, size_t n) {
  const int e = fix_exponent(hdr[0], n);
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x)
    g_dem[q] = from_fixed(acc[q] + (spill[q] == (int)q ? to_fixed(g_out[q], e) : 0ull), e);
}

// Non-finite g_out entries, only when there are any (hdr[1]).
__global__ void fill_bwd_nonfinite(float* g_dem, const unsigned* hdr, const int* spill,
                                   const float* g_out, size_t n) {
  if (hdr[1] == 0u) return;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (size_t)gridDim.x * blockDim.x) {
    const float g = g_out[p];
    const int s = spill[p];
    if (!isfinite(g) && s >= 0) combine_nonfinite(g_dem, (size_t)s, g);
  }
}

static int grid_for(size_t n, int threads) {
  size_t b = (n + threads - 1) / threads;
  return (int)(b > 8192 ? 8192 : (b < 1 ? 1 : b));
}

hipError_t run_fill_backward(float* g_dem, const int* spill, const float* g_out, int64_t nx,
                             int64_t ny, hipStream_t stream) {
  const size_t n = (size_t)nx * (size_t)ny;
  if (n == 0) return hipSuccess;
  void* ws = nullptr;
  hipError_t e = hipMallocAsync(&ws, n * 8 + 256, stream);
  if (e != hipSuccess) return e;
  unsigned long long* acc = (unsigned long long*)ws;
  unsigned* hdr = (unsigned*)((char*)ws + n * 8);
  zero_words<<<1, 64, 0, stream>>>(hdr, 64);
  fill_bwd_prep<<<grid_for(n, 256), 256, 0, stream>>>(acc, hdr, g_out, n);
  size_t chunks = (n + kBwdChunk - 1) / kBwdChunk;
  int gs = (int)(chunks > 4096 ? 4096 : chunks);
  fill_bwd_scatter<<<gs, kBwdThreads, 0, stream>>>(acc, hdr, spill, g_out, n);
  fill_bwd_finish<<<grid_for(n, 256), 256, 0, stream>>>(g_dem, acc, hdr, spill, g_out, n);
  fill_bwd_nonfinite<<<grid_for(n, 256), 256, 0, stream>>>(g_dem, hdr, spill, g_out, n);
  e = hipGetLastError();
  hipError_t f = hipFreeAsync(ws, stream);
  return e != hipSuccess ? e : f;
}

// ---------------------------------------------------------------- FSM backward
__global__ void fsm_linearize(float* sx, float* sy, float* dfic, const float* u, const float* dem,
                              const float* slopes, float cellsize, long long nx, long long ny) {
  const size_t n = (size_t)nx * ny;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x)
    fsm_linearize_cell(q, sx, sy, dfic, u, dem, slopes, cellsize, nx, ny);
}

__global__ void fsm_adjoint_finish(float* g_dem, float* g_slopes, const float* lam,
                                   const float* dfic, size_t n) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x)
    fsm_adjoint_finish_cell(q, g_dem, g_slopes, lam, dfic);
}

struct AdjArgs {
  AdjView v;
  int ntx, nty;
  TileQueue q;
  unsigned max_local;
};

template <class G>
__global__ __launch_bounds__(G::THREADS) void fsm_adjoint_persistent(AdjArgs a) {
  __shared__ AdjLds<G> lds;
  run_queue(a.q, GridTopo{a.ntx, a.nty}, [&](int t, bool first, unsigned nvis) {
    const long long x0 = (long long)(t % a.ntx) * G::TX, y0 = (long long)(t / a.ntx) * G::TY;
    return adj_visit<G>(a.v, x0, y0, first, nvis, a.max_local, a.q.stats_sink(), lds);
  });
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
    AdjArgs a{};
    a.v.lam = lam;
    a.v.g = g_out;
    a.v.sx = sx;
    a.v.sy = sy;
    a.v.nx = nx;
    a.v.ny = ny;
    a.max_local = 1u << 16;
    // four float arrays per cell in LDS: 64x64 tiles fit two workgroups per
    // CU (the backward pass at 8192^2: 9.7 ms, against 11.1 ms with 128x64
    // and 12.3 ms with 64x128 tiles, one workgroup per CU)
    using G = TileGeom<1, 1>;
    a.ntx = (int)((nx + G::TX - 1) / G::TX);
    a.nty = (int)((ny + G::TY - 1) / G::TY);
    e = launch_persistent(fsm_adjoint_persistent<G>, G::THREADS, a.ntx * a.nty, a, stream, stats,
                          G::TX, G::TY, a.ntx, a.nty);
  }
  if (e == hipSuccess) {
    fsm_adjoint_finish<<<grid_for(n, 256), 256, 0, stream>>>(g_dem, g_slopes, lam, dfic, n);
    e = hipGetLastError();
  }
  hipError_t f = hipFreeAsync(ws, stream);
  return e != hipSuccess ? e : f;
}

}  // namespace tt
