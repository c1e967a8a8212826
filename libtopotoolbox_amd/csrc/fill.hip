// Sink filling and grayscale reconstruction on one grid: the persistent
// work-queue kernel over fill_visit (fill_visit.h states the semantics).
#include <mutex>

#include "fill_visit.h"
#include "workspace.h"

namespace tt {

struct FillArgs {
  FillView v;
  int ntx, nty;
  TileQueue q;
  unsigned max_local;
};

template <class G, int MODE>
__global__ __launch_bounds__(G::THREADS) void fill_persistent(FillArgs a) {
  __shared__ FillLds<G, MODE> lds;
  run_queue(a.q, GridTopo{a.ntx, a.nty}, [&](int t, bool first, unsigned nvis) {
    const long long x0 = (long long)(t % a.ntx) * G::TX, y0 = (long long)(t / a.ntx) * G::TY;
    return fill_visit<G, MODE>(a.v, x0, y0, first, nvis, a.max_local, a.q.stats_sink(), lds);
  });
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

hipError_t run_fill(float* state, const float* src, const uint8_t* bc, int64_t nx, int64_t ny,
                    int mode, TileShape ts, hipStream_t stream, RunStats* stats) {
  if (nx <= 0 || ny <= 0) return hipSuccess;
  FillArgs a{};
  a.v.state = state;
  a.v.src = src;
  a.v.bc = bc;
  a.v.nx = nx;
  a.v.ny = ny;
  if (mode == kModeFill) return dispatch_fill<kModeFill>(a, ts, stream, stats);
  return dispatch_fill<kModeRecon>(a, ts, stream, stats);
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
  FillArgs a{};
  a.v.state = out;
  a.v.spill = spill;
  a.v.src = dem;
  a.v.bc = bc;
  a.v.nx = nx;
  a.v.ny = ny;
  // exact levels first (float relaxation, the fast path), then the
  // (level, index) keys from (level, no index): only spill indices move
  hipError_t e = run_fill(out, dem, bc, nx, ny, kModeFill, TileShape{2, 2}, stream, nullptr);
  if (e != hipSuccess) return e;
  a.v.levels = true;
  // (level, index) keys live in one 64-bit array between visits; the
  // index output is written alongside (the levels in out are final)
  e = hipMallocAsync((void**)&a.v.keys, (size_t)nx * ny * 8, stream);
  if (e != hipSuccess) return e;
  e = dispatch_fill<kModeSpill>(a, ts, stream, stats);
  hipError_t f = hipFreeAsync(a.v.keys, stream);
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
