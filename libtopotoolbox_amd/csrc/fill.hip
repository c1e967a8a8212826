// Sink filling and grayscale reconstruction on one grid: the persistent
// work-queue kernel over fill_visit (fill_visit.h states the semantics),
// and the spill indices of the fillsinks adjoint (labelling of spill_ccl.h,
// then the spill relaxation from the dirty tiles).
#include <mutex>

#include "prio_queue.h"
#include "spill_ccl.h"
#include "workspace.h"

namespace tt {

struct FillArgs {
  FillView v;
  int ntx, nty;
  TileQueue q;
  unsigned max_local;
  unsigned* dirty;  // kModeSpill: per tile, relax on the first visit
};

template <class G, int MODE>
__global__ __launch_bounds__(G::THREADS) void fill_persistent(FillArgs a) {
  __shared__ FillLds<G, MODE> lds;
  const GridTopo topo{a.ntx, a.nty};
  run_queue(a.q, topo, [&](int t, bool first, unsigned nvis) -> unsigned {
    if constexpr (MODE == kModeSpill)
      if (spill_skip(a.dirty, t, first, &lds.bits)) return 0u;
    const long long x0 = (long long)(t % a.ntx) * G::TX, y0 = (long long)(t / a.ntx) * G::TY;
    const unsigned bits =
        fill_visit<G, MODE>(a.v, x0, y0, first, nvis, a.max_local, a.q.stats_sink(), lds);
    if constexpr (MODE == kModeSpill) spill_mark(a.dirty, topo, t, bits);
    return bits;
  });
}

// ---------------------------------------------------------------- spill labelling
constexpr int kLabelTile = 64;     // labelling tiles: 64 x 64 cells
constexpr int kLabelThreads = 256;

__global__ __launch_bounds__(kLabelThreads) void spill_label_kernel(SpillView v) {
  __shared__ SpillLabelLds<kLabelTile, kLabelTile> lds;
  spill_label_tile<kLabelTile, kLabelTile, kLabelThreads, true>(
      v, (long long)blockIdx.x * kLabelTile, (long long)blockIdx.y * kLabelTile, lds);
}

__global__ __launch_bounds__(kLabelThreads) void spill_border_kernel(SpillView v) {
  spill_border_tile<kLabelTile, kLabelTile, kLabelThreads>(
      v, (long long)blockIdx.x * kLabelTile, (long long)blockIdx.y * kLabelTile);
}

// The per-cell steps: each thread takes four cells' cmin words at once
// (one 16-byte load; cmin is 256-byte aligned), most of which say "nothing
// to do here".
template <void (*CELL)(const SpillView&, size_t, unsigned)>
__global__ void spill_cells_kernel(SpillView v) {
  const size_t n = (size_t)v.nx * v.ny, n4 = n / 4;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const uint4* c4 = reinterpret_cast<const uint4*>(v.cmin);
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n4; q += stride) {
    const uint4 c = c4[q];
    CELL(v, 4 * q, c.x);
    CELL(v, 4 * q + 1, c.y);
    CELL(v, 4 * q + 2, c.z);
    CELL(v, 4 * q + 3, c.w);
  }
  for (size_t g = 4 * n4 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += stride)
    CELL(v, g, v.cmin[g]);
}

template <class G, int MODE>
static hipError_t launch_fill(FillArgs a, hipStream_t stream, RunStats* stats,
                              void* scratch = nullptr, unsigned max_visits = 0) {
  a.ntx = (int)((a.v.nx + G::TX - 1) / G::TX);
  a.nty = (int)((a.v.ny + G::TY - 1) / G::TY);
  a.max_local = 1u << 16;
  return launch_persistent(fill_persistent<G, MODE>, G::THREADS, a.ntx * a.nty, a, stream, stats,
                           G::TX, G::TY, a.ntx, a.nty, scratch, max_visits);
}

// ---------------------------------------------------------------- lowest level first
// kModeFill on one grid with the priority queue of prio_queue.h.
struct FillPrioArgs {
  FillView v;
  int ntx, nty;
  PrioQueue q;
  unsigned max_local;
};

template <class G, int MODE>
__global__ __launch_bounds__(G::THREADS) void fill_prio_persistent(FillPrioArgs a) {
  __shared__ FillLds<G, MODE> lds;
  const GridTopo topo{a.ntx, a.nty};
  run_prio_queue(a.q, topo, [&](int t, bool first, unsigned nvis, unsigned* lev) -> unsigned {
    const long long x0 = (long long)(t % a.ntx) * G::TX, y0 = (long long)(t / a.ntx) * G::TY;
    return fill_visit<G, MODE>(a.v, x0, y0, first, nvis, a.max_local, a.q.stats_sink(), lds, lev);
  });
}

__device__ inline unsigned wave_min(unsigned x) {
  for (int o = 32; o > 0; o >>= 1) x = min(x, (unsigned)__shfl_xor((int)x, o));
  return x;
}
__device__ inline unsigned wave_max(unsigned x) {
  for (int o = 32; o > 0; o >>= 1) x = max(x, (unsigned)__shfl_xor((int)x, o));
  return x;
}

// One workgroup per tile: the key of the tile's lowest seed into q.seed,
// and the finite range of the bounds into the control block (pad[0] =
// ~key(min), pad[1] = key(max), raised from zero). Fill: a seed is a bc
// cell with a value, at its DEM level; reconstruction: any cell with a
// marker, at max(-mask, -marker). Each thread reads four adjacent cells
// per row (16-byte value and 4-byte bc loads where the row allows).
constexpr int kSeedThreads = 256;
template <class G, int MODE>
__global__ __launch_bounds__(kSeedThreads) void prio_seed_kernel(FillView v, int ntx, PrioQueue q) {
  constexpr int kPerRow = G::TX / 4;              // threads per tile row
  constexpr int kRows = kSeedThreads / kPerRow;   // rows per pass
  static_assert(G::TY % kRows == 0, "whole passes");
  const int t = blockIdx.x;
  const long long x0 = (long long)(t % ntx) * G::TX, y0 = (long long)(t / ntx) * G::TY;
  const long long gx = x0 + 4 * (threadIdx.x % kPerRow);
  constexpr bool kRecon = MODE == kModeRecon;
  const bool vec = v.nx % 4 == 0 && gx + 3 < v.nx &&
                   (reinterpret_cast<uintptr_t>(v.src) & 15) == 0 &&
                   (kRecon ? (reinterpret_cast<uintptr_t>(v.state) & 15) == 0
                           : (reinterpret_cast<uintptr_t>(v.bc) & 3) == 0);
  unsigned smin = ~0u, fmin = ~0u, fmax = 0u;
  // fill: d = dem, a seed where bc == 1 at level d; reconstruction (the
  // negated domain of fill_visit): d = -mask, every cell with a marker a
  // seed at level max(d, -marker)
  auto cell = [&](float src, float aux) {
    const float d = kRecon ? -src : src;
    if (d != d) return;
    const unsigned k = ord_bits(d);
    if (kRecon) {
      const float f = fmaxf(d, (aux == aux) ? -aux : kInf);
      if (f < kInf) smin = min(smin, ord_bits(f));
    } else if (aux == 1.f) {
      smin = min(smin, k);
    }
    if (d - d == 0.f) {  // finite
      fmin = min(fmin, k);
      fmax = max(fmax, k);
    }
  };
#pragma unroll 4
  for (int r = threadIdx.x / kPerRow; r < G::TY; r += kRows) {
    const long long gy = y0 + r;
    if (gy >= v.ny) break;
    const size_t g = (size_t)gy * v.nx + gx;
    if (vec) {
      const float4 d = *reinterpret_cast<const float4*>(v.src + g);
      float4 a;
      if (kRecon) {
        a = *reinterpret_cast<const float4*>(v.state + g);
      } else {
        const uchar4 b = *reinterpret_cast<const uchar4*>(v.bc + g);
        a = make_float4(b.x, b.y, b.z, b.w);
      }
      cell(d.x, a.x);
      cell(d.y, a.y);
      cell(d.z, a.z);
      cell(d.w, a.w);
    } else {
      for (int k = 0; k < 4; k++)
        if (gx + k < v.nx) cell(v.src[g + k], kRecon ? v.state[g + k] : (float)v.bc[g + k]);
    }
  }
  __shared__ unsigned red[3][kSeedThreads / 64];
  smin = wave_min(smin);
  fmin = wave_min(fmin);
  fmax = wave_max(fmax);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][wave] = smin;
    red[1][wave] = fmin;
    red[2][wave] = fmax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kSeedThreads / 64; w++) {
      smin = min(smin, red[0][w]);
      fmin = min(fmin, red[1][w]);
      fmax = max(fmax, red[2][w]);
    }
    q.seed[t] = smin;
    // thousands of workgroups raise the same two words: only those that
    // would raise them issue the atomic
    if (fmin != ~0u && ld_relaxed(&q.sync->pad[0]) < ~fmin) atomicMax(&q.sync->pad[0], ~fmin);
    if (fmax != 0u && ld_relaxed(&q.sync->pad[1]) < fmax) atomicMax(&q.sync->pad[1], fmax);
  }
}

// Initial queue: every tile, at the bucket of its lowest seed, or in the
// unreached bucket when it holds none. A wavefront takes 64 consecutive
// tiles (one bitmap word): its unreached tiles, usually most of them, are
// marked with one word update and one count update.
__global__ void prio_queue_kernel(PrioQueue q) {
  const PrioScale sc = prio_scale(q.sync);
  const int lane = threadIdx.x & 63;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    q.sync->pending = (unsigned)q.ntiles;
    q.sync->tail = (unsigned)q.ntiles;
  }
  const int stride = gridDim.x * blockDim.x;
  for (int t0 = (blockIdx.x * blockDim.x + threadIdx.x) & ~63; t0 < q.ntiles; t0 += stride) {
    const int t = t0 + lane;
    const bool in = t < q.ntiles;
    int b = kPrioUnreached;
    if (in) {
      const unsigned k = q.seed[t];
      if (k != ~0u) b = sc.bucket(k);
      q.state[t] = kQueued | ((unsigned)b << 8);
      q.visited[t] = 0u;
    }
    const unsigned long long un = __ballot(in && b == kPrioUnreached);
    if (in && b != kPrioUnreached) pq_mark(q, t, b);
    if (lane == 0 && un) {
      atomicOr(&q.bits[(size_t)kPrioUnreached * q.nwords + (t0 >> 6)], un);
      atomicAdd(&q.count[kPrioUnreached], __popcll(un));
    }
  }
}

static int launch_blocks(size_t n) {
  size_t b = (n + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

template <class G, int MODE>
static hipError_t launch_fill_prio(const FillView& v, hipStream_t stream, RunStats* stats) {
  FillPrioArgs a{};
  a.v = v;
  a.ntx = (int)((v.nx + G::TX - 1) / G::TX);
  a.nty = (int)((v.ny + G::TY - 1) / G::TY);
  a.max_local = 1u << 16;
  const int ntiles = a.ntx * a.nty;
  int grid = 0;
  hipError_t e =
      persistent_grid((const void*)fill_prio_persistent<G, MODE>, G::THREADS, ntiles, &grid);
  if (e != hipSuccess) return e;
  const size_t words = prio_workspace_words(ntiles);
  void* mem = nullptr;
  e = hipMallocAsync(&mem, words * sizeof(unsigned), stream);
  if (e != hipSuccess) return e;
  // finish() reads the control block's statistics and frees the memory
  Workspace ws;
  ws.base = mem;
  ws.owned = true;
  ws.stream = stream;
  a.q = prio_layout(mem, ntiles);
  ws.q.sync = a.q.sync;
  ws.q.ntiles = ntiles;
  a.q.wait_limit = wait_limit_ticks();
  a.q.stats = stats ? 1 : 0;
  // measured at 8192^2: 2x the workgroups 1-2% faster than 1x, 1/4x and
  // 4x slower
  a.q.window = 2 * grid;
  e = sticky_error_word(&a.q.sticky);
  if (e == hipSuccess) {
    zero_words<<<launch_blocks(words), 256, 0, stream>>>(static_cast<unsigned*>(mem), words);
    prio_seed_kernel<G, MODE><<<ntiles, kSeedThreads, 0, stream>>>(v, a.ntx, a.q);
    prio_queue_kernel<<<launch_blocks((size_t)ntiles), 256, 0, stream>>>(a.q);
    fill_prio_persistent<G, MODE><<<grid, G::THREADS, 0, stream>>>(a);
    e = hipGetLastError();
  }
  if (stats) {
    stats->grid = grid;
    stats->tile_x = G::TX;
    stats->tile_y = G::TY;
  }
  hipError_t f = ws.finish(stats);
  return e != hipSuccess ? e : f;
}

template <int MODE>
static hipError_t dispatch_fill(FillArgs a, TileShape ts, hipStream_t stream, RunStats* stats) {
  if constexpr (MODE == kModeSpill) {
    // 64-bit keys and a kind byte per cell: a 2x2 tile fits LDS alone, a
    // 2x1 or 1x2 tile twice per CU
    if (ts.qx == 2 && ts.qy == 2) return launch_fill<TileGeom<2, 2>, MODE>(a, stream, stats);
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

// Schedule of sink filling on one grid: the lowest-level-first queue pays
// off once tiles outnumber the workgroups by kPrioMinTilesPerWorker or
// more (fewer tiles: every tile has a workgroup of its own most of the
// time, and the FIFO queue's cheaper pops win), with at most
// kPrioMaxWorkers workgroups (more poppers contend for the lowest tiles:
// 1024 workgroups of 64 x 64 tiles lose to FIFO on 8192 x 8192 grids).
// set_fill_schedule overrides the choice for the calling thread.
constexpr int kPrioMinTilesPerWorker = 4;
constexpr int kPrioMaxWorkers = 512;

static thread_local FillSchedule g_fill_schedule = kSchedAuto;

void set_fill_schedule(FillSchedule s) { g_fill_schedule = s; }

static FillSchedule fill_schedule() { return g_fill_schedule; }

template <class G, int MODE>
static hipError_t launch_fill_scheduled(FillArgs a, hipStream_t stream, RunStats* stats) {
  const FillSchedule s = fill_schedule();
  if (s == kSchedAuto) {
    if (MODE == kModeRecon) return launch_fill<G, MODE>(a, stream, stats);
    const long long ntiles = ((a.v.nx + G::TX - 1) / G::TX) * ((a.v.ny + G::TY - 1) / G::TY);
    int grid = 0;
    hipError_t e =
        persistent_grid((const void*)fill_prio_persistent<G, MODE>, G::THREADS, 1 << 30, &grid);
    if (e != hipSuccess) return e;
    if (ntiles < (long long)kPrioMinTilesPerWorker * grid || grid > kPrioMaxWorkers)
      return launch_fill<G, MODE>(a, stream, stats);
    return launch_fill_prio<G, MODE>(a.v, stream, stats);
  }
  if (s == kSchedFifo) return launch_fill<G, MODE>(a, stream, stats);
  return launch_fill_prio<G, MODE>(a.v, stream, stats);
}

hipError_t run_fill(float* state, const float* src, const uint8_t* bc, int64_t nx, int64_t ny,
                    int mode, TileShape ts, hipStream_t stream, RunStats* stats) {
  if (nx <= 0 || ny <= 0) return hipSuccess;
  FillArgs a{};
  a.v.state = state;
  a.v.src = src;
  a.v.bc = bc;
  a.v.nx = nx;
  a.v.ny = ny;
  if (mode == kModeFill) {
    if (ts.qx == 2 && ts.qy == 2)
      return launch_fill_scheduled<TileGeom<2, 2>, kModeFill>(a, stream, stats);
    if (ts.qx == 2 && ts.qy == 1)
      return launch_fill_scheduled<TileGeom<2, 1>, kModeFill>(a, stream, stats);
    if (ts.qx == 1 && ts.qy == 2)
      return launch_fill_scheduled<TileGeom<1, 2>, kModeFill>(a, stream, stats);
    return launch_fill_scheduled<TileGeom<1, 1>, kModeFill>(a, stream, stats);
  }
  if (ts.qx == 2 && ts.qy == 2)
    return launch_fill_scheduled<TileGeom<2, 2>, kModeRecon>(a, stream, stats);
  if (ts.qx == 2 && ts.qy == 1)
    return launch_fill_scheduled<TileGeom<2, 1>, kModeRecon>(a, stream, stats);
  if (ts.qx == 1 && ts.qy == 2)
    return launch_fill_scheduled<TileGeom<1, 2>, kModeRecon>(a, stream, stats);
  return launch_fill_scheduled<TileGeom<1, 1>, kModeRecon>(a, stream, stats);
}

hipError_t run_fill_capped(float* state, const float* src, const uint8_t* bc, int64_t nx,
                           int64_t ny, int mode, unsigned max_visits, hipStream_t stream,
                           RunStats* stats, int* converged) {
  *converged = 1;
  if (nx <= 0 || ny <= 0) return hipSuccess;
  FillArgs a{};
  a.v.state = state;
  a.v.src = src;
  a.v.bc = bc;
  a.v.nx = nx;
  a.v.ny = ny;
  using G = TileGeom<2, 2>;
  hipError_t e = mode == kModeFill
                     ? launch_fill<G, kModeFill>(a, stream, stats, nullptr, max_visits)
                     : launch_fill<G, kModeRecon>(a, stream, stats, nullptr, max_visits);
  if (e != hipSuccess) return e;
  *converged = stats->pending == 0 ? 1 : 0;
  stats->error &= ~kStopVisitCap;
  return hipSuccess;
}

hipError_t run_fill_spill(float* out, int* spill, const float* dem, const uint8_t* bc,
                          int64_t nx, int64_t ny, TileShape ts, hipStream_t stream,
                          RunStats* stats) {
  if (nx <= 0 || ny <= 0) return hipSuccess;
  // exact levels first (float relaxation, the fast path)
  hipError_t e = run_fill(out, dem, bc, nx, ny, kModeFill, TileShape{2, 2}, stream, nullptr);
  if (e != hipSuccess) return e;
  // labelling: spill := exact indices or upper bounds, dirty := relaxation
  // tiles holding a flat cell
  const size_t n = (size_t)nx * (size_t)ny;
  const int rtx = 64 * ts.qx, rty = 64 * ts.qy;
  const int rntx = (int)((nx + rtx - 1) / rtx), rnty = (int)((ny + rty - 1) / rty);
  const size_t ntiles = (size_t)rntx * rnty;
  const size_t cmin_bytes = (n * 4 + 255) & ~(size_t)255;
  void* ws = nullptr;
  e = hipMallocAsync(&ws, cmin_bytes + ntiles * 4, stream);
  if (e != hipSuccess) return e;
  SpillView sv{};
  sv.F = out;
  sv.dem = dem;
  sv.bc = bc;
  sv.spill = spill;
  sv.cmin = static_cast<unsigned*>(ws);
  sv.dirty = reinterpret_cast<unsigned*>(static_cast<char*>(ws) + cmin_bytes);
  sv.nx = nx;
  sv.ny = ny;
  sv.rtx = rtx;
  sv.rty = rty;
  sv.rntx = rntx;
  const dim3 tiles((unsigned)((nx + kLabelTile - 1) / kLabelTile),
                   (unsigned)((ny + kLabelTile - 1) / kLabelTile));
  size_t cb = (n / 4 + 255) / 256;
  const int cells_grid = (int)(cb > 8192 ? 8192 : (cb < 1 ? 1 : cb));
  zero_words<<<(int)((ntiles + 255) / 256 > 1024 ? 1024 : (ntiles + 255) / 256), 256, 0, stream>>>(
      sv.dirty, ntiles);
  spill_label_kernel<<<tiles, kLabelThreads, 0, stream>>>(sv);
  spill_border_kernel<<<tiles, kLabelThreads, 0, stream>>>(sv);
  spill_cells_kernel<spill_root_cell><<<cells_grid, 256, 0, stream>>>(sv);
  spill_cells_kernel<spill_final_cell><<<cells_grid, 256, 0, stream>>>(sv);
  spill_cells_kernel<spill_final_root><<<cells_grid, 256, 0, stream>>>(sv);
  e = hipGetLastError();
  if (e == hipSuccess) {
    // relaxation of the indices the labelling left as upper bounds
    FillArgs a{};
    a.v.state = out;
    a.v.spill = spill;
    a.v.src = dem;
    a.v.bc = bc;
    a.v.nx = nx;
    a.v.ny = ny;
    a.dirty = sv.dirty;
    e = dispatch_fill<kModeSpill>(a, ts, stream, stats);
  }
  hipError_t f = hipFreeAsync(ws, stream);
  return e != hipSuccess ? e : f;
}

// ---------------------------------------------------------------- no allocation
// fillsinks for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops): the queue workspace lives in a
// fixed file-scope array instead of a stream-ordered allocation. Each GPU
// holds its own copy of the array until the process ends, and every call
// on that GPU shares it, so calls that overlap in time on one GPU race.
// The init kernel rewrites the whole control block and queue each call, so
// nothing depends on what the previous call left there.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices];  // device address of `scratch`, per GPU
}  // namespace

// Device address of this GPU's `scratch` (the array's name in host code is
// only a stand-in for it), looked up once per device with that device
// current, on the device's first call.
static hipError_t scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return hipSuccess;
}

int64_t noalloc_fill_tiles(int64_t nx, int64_t ny) {
  using G = TileGeom<2, 2>;
  return ((nx + G::TX - 1) / G::TX) * ((ny + G::TY - 1) / G::TY);
}

hipError_t run_fill_noalloc(float* state, const float* src, const uint8_t* bc, int64_t nx,
                            int64_t ny, hipStream_t stream) {
  if (nx <= 0 || ny <= 0) return hipSuccess;
  if (noalloc_fill_tiles(nx, ny) > kNoAllocMaxTiles) return hipErrorInvalidValue;
  void* ws = nullptr;
  hipError_t e = scratch_address(&ws);
  if (e != hipSuccess) return e;
  FillArgs a{};
  a.v.state = state;
  a.v.src = src;
  a.v.bc = bc;
  a.v.nx = nx;
  a.v.ny = ny;
  return launch_fill<TileGeom<2, 2>, kModeFill>(a, stream, nullptr, ws);
}

}  // namespace tt
