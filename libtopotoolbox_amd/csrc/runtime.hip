// Workspace, occupancy and error plumbing shared by the kernels.
#include <stdarg.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <mutex>

#include "workspace.h"

namespace tt {

// Error bits of every launch on a GPU since the host last cleared them
// (device errors of calls that return before their kernels finish).
static __device__ unsigned sticky_error;

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_sticky_mu;
unsigned* g_sticky_dev[kMaxDevices];
std::atomic<unsigned long long> g_wait_limit{kWaitTimeoutTicks};
}  // namespace

hipError_t sticky_error_word(unsigned** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_sticky_mu);
  if (!g_sticky_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(sticky_error));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_sticky_dev[dev] = static_cast<unsigned*>(p);
  }
  *out = g_sticky_dev[dev];
  return hipSuccess;
}

unsigned long long wait_limit_ticks() { return g_wait_limit.load(); }

void set_wait_limit_ticks(unsigned long long ticks) {
  g_wait_limit.store(ticks ? ticks : kWaitTimeoutTicks);
}

hipError_t take_sticky_error(bool clear, unsigned* bits) {
  *bits = 0;
  unsigned* p = nullptr;
  hipError_t e = sticky_error_word(&p);
  if (e != hipSuccess) return e;
  e = hipDeviceSynchronize();
  if (e != hipSuccess) return e;
  e = hipMemcpy(bits, p, sizeof(unsigned), hipMemcpyDeviceToHost);
  if (e != hipSuccess || !clear || *bits == 0) return e;
  const unsigned zero = 0;
  return hipMemcpy(p, &zero, sizeof(unsigned), hipMemcpyHostToDevice);
}

__global__ void zero_words(unsigned* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0u;
}

// Position of tile t in the initial queue when tiles are ordered from the
// grid boundary inwards (concentric rings of tiles, each ring in row-major
// order): a tile's first visit then tends to find its outer neighbours,
// through which drainage from the edge arrives, already relaxed.
__device__ inline int boundary_first_tile(int k, int ntx, int nty) {
  // ring r holds the tiles whose distance to the boundary is r; count the
  // tiles of the rings before k's ring by peeling rings off the rectangle
  int r = 0, w = ntx, h = nty;
  for (;;) {
    const int n = (w <= 0 || h <= 0) ? 0 : ((w == 1 || h == 1) ? w * h : 2 * w + 2 * h - 4);
    if (k < n || n == 0) break;
    k -= n;
    r++;
    w -= 2;
    h -= 2;
  }
  // k-th tile (row-major) of ring r of the w x h rectangle at offset (r, r)
  int tx, ty;
  if (h == 1 || w == 1) {
    tx = (h == 1) ? k : 0;
    ty = (h == 1) ? 0 : k;
  } else if (k < w) {  // top row
    tx = k;
    ty = 0;
  } else if (k < w + 2 * (h - 2)) {  // side columns, row by row
    const int m = k - w;
    ty = 1 + m / 2;
    tx = (m % 2) ? w - 1 : 0;
  } else {  // bottom row
    tx = k - w - 2 * (h - 2);
    ty = h - 1;
  }
  return (ty + r) * ntx + (tx + r);
}

__global__ void init_queue(Sync* s, unsigned long long* ring, unsigned* state,
                           unsigned* visited, int ntiles, int ntx, int nty) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    unsigned* w = reinterpret_cast<unsigned*>(s);
    for (unsigned k = 0; k < sizeof(Sync) / 4; k++) w[k] = 0u;
    s->tail = (unsigned)ntiles;
    s->pending = (unsigned)ntiles;
  }
  const bool rings = ntx > 0 && nty > 0 && ntx * nty == ntiles;
  for (int k = i; k < ntiles; k += gridDim.x * blockDim.x) {
    const int t = rings ? boundary_first_tile(k, ntx, nty) : k;
    ring[k] = ((unsigned long long)k << 32) | ((unsigned)t + 1u);
    state[k] = kQueued;
    visited[k] = 0u;
  }
}

hipError_t Workspace::alloc(int ntiles, hipStream_t s) { return alloc(ntiles, 0, 0, s); }

hipError_t Workspace::alloc(int ntiles, int ntx, int nty, hipStream_t s) {
  void* mem = nullptr;
  hipError_t e = hipMallocAsync(&mem, queue_workspace_bytes(ntiles), s);
  if (e != hipSuccess) return e;
  e = attach(mem, ntiles, ntx, nty, s);
  owned = true;
  return e;
}

hipError_t Workspace::attach(void* mem, int ntiles, int ntx, int nty, hipStream_t s) {
  stream = s;
  base = mem;
  owned = false;
  // Sync block (rounded to 64 words) + ring (2 words per tile) + state and
  // visited (1 word per tile)
  const size_t head = 64;
  q.sync = reinterpret_cast<Sync*>(base);
  q.ring = reinterpret_cast<unsigned long long*>(reinterpret_cast<unsigned*>(base) + head);
  q.state = reinterpret_cast<unsigned*>(q.ring + ntiles);
  q.visited = q.state + ntiles;
  q.ntiles = ntiles;
  q.wait_limit = wait_limit_ticks();
  hipError_t e = sticky_error_word(&q.sticky);
  if (e != hipSuccess) return e;
  int blocks = (ntiles + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  if (blocks < 1) blocks = 1;
  init_queue<<<blocks, 256, 0, s>>>(q.sync, q.ring, q.state, q.visited, ntiles, ntx, nty);
  return hipGetLastError();
}

hipError_t Workspace::finish(RunStats* st) {
  hipError_t e = hipSuccess;
  if (st) {
    Sync h;
    e = hipMemcpyAsync(&h, q.sync, sizeof(Sync), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e == hipSuccess) {
      st->requeues = (int64_t)h.tail - (int64_t)q.ntiles;
      st->visits = (int64_t)h.visits;
      st->local_iters = (int64_t)h.local_iters;
      st->max_local = h.max_local;
      st->error = h.error;
      st->pending = (int64_t)h.pending;
      st->t_stage = (int64_t)h.t_stage;
      st->t_relax = (int64_t)h.t_relax;
      st->t_write = (int64_t)h.t_write;
      st->t_wait = (int64_t)h.t_wait;
    }
  }
  hipError_t f = owned ? hipFreeAsync(base, stream) : hipSuccess;
  base = nullptr;
  owned = false;
  return e != hipSuccess ? e : f;
}

hipError_t persistent_grid(const void* kernel, int threads, int ntiles, int* grid) {
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0);
  if (e != hipSuccess) return e;
  if (per_cu < 1) return hipErrorInvalidConfiguration;
  long long g = (long long)cus * per_cu;
  if (g > ntiles) g = ntiles;
  if (g < 1) g = 1;
  *grid = (int)g;
  return hipSuccess;
}

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

const char* last_error() { return g_err; }

bool tile_shape_ok(TileShape ts) {
  return (ts.qx == 1 || ts.qx == 2) && (ts.qy == 1 || ts.qy == 2);
}

}  // namespace tt
