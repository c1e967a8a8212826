// C ABI: the reference's drop-in entry points (host pointers) and the
// device-pointer entry points used by the Python layer.
#include <limits.h>
#include <stdio.h>

#include <mutex>

#include "../../include/topotoolbox_amd.h"
#include "kernels.h"
#include "workspace.h"

namespace tt {
const char* last_error();
}

namespace {

// Tile shape set by tt_set_tile_shape; until then each entry point uses
// the shape measured fastest for it at 8192^2 on MI355X.
thread_local tt::TileShape g_shape;
thread_local bool g_shape_set = false;

tt::TileShape shape_for(tt::TileShape fallback) { return g_shape_set ? g_shape : fallback; }

void to_public(const tt::RunStats& s, tt_stats* o) {
  o->requeues = s.requeues;
  o->visits = s.visits;
  o->local_iters = s.local_iters;
  o->max_local = s.max_local;
  o->grid = s.grid;
  o->tile_x = s.tile_x;
  o->tile_y = s.tile_y;
  o->error = s.error;
  o->t_stage = s.t_stage;
  o->t_relax = s.t_relax;
  o->t_write = s.t_write;
  o->t_wait = s.t_wait;
}

int fail(const char* where, hipError_t e) {
  tt::set_error("%s: %s", where, hipGetErrorString(e));
  return (int)e;
}

int device_error(const char* where, uint32_t bits) {
  tt::set_error("%s: device error bits 0x%x (1: a queue wait exceeded its time limit)", where,
                bits);
  return 1000 + (int)bits;
}

// Device memory of the host-pointer drop-ins: one grow-only buffer per GPU,
// reused by every call (no allocation per call once it is large enough)
// and held until tt_host_workspace_release(). A call holds its GPU's lock
// from the first copy in to the last copy out, so concurrent drop-in calls
// on one GPU run one after the other.
constexpr int kMaxDevices = 64;
struct HostCache {
  std::mutex mu;
  void* p = nullptr;
  size_t bytes = 0;
};
HostCache g_host_cache[kMaxDevices];

// Device buffers of one drop-in call, carved from its GPU's cache.
class HostStage {
 public:
  // sizes: bytes of each buffer (each starts 256-byte aligned)
  hipError_t acquire(const size_t* sizes, int count) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    cache_ = &g_host_cache[dev];
    lock_ = std::unique_lock<std::mutex>(cache_->mu);
    size_t total = 0;
    for (int k = 0; k < count; k++) total += (sizes[k] + 255) & ~(size_t)255;
    if (total > cache_->bytes) {
      if (cache_->p) {
        e = hipFree(cache_->p);
        cache_->p = nullptr;
        cache_->bytes = 0;
        if (e != hipSuccess) return e;
      }
      e = hipMalloc(&cache_->p, total ? total : 256);
      if (e != hipSuccess) return e;
      cache_->bytes = total;
    }
    char* b = static_cast<char*>(cache_->p);
    for (int k = 0; k < count; k++) {
      ptr_[k] = b;
      b += (sizes[k] + 255) & ~(size_t)255;
    }
    return hipSuccess;
  }
  void* operator[](int k) const { return ptr_[k]; }

 private:
  HostCache* cache_ = nullptr;
  std::unique_lock<std::mutex> lock_;
  void* ptr_[8] = {};
};

// Device errors of a synchronous drop-in call (its kernels have finished
// when this runs).
int sticky_check(const char* where) {
  unsigned bits = 0;
  hipError_t e = tt::take_sticky_error(true, &bits);
  if (e != hipSuccess) return fail(where, e);
  if (bits) return device_error(where, bits);
  return 0;
}

// Drop-in failures are reported once on stderr: the reference API has no
// error channel, and a silent failure would hand back garbage.
void report(const char* fn) {
  static thread_local int reported = 0;
  if (!reported) {
    fprintf(stderr, "libtopotoolbox_amd: %s failed: %s\n", fn, tt::last_error());
    reported = 1;
  }
}


bool dims_ok(const ptrdiff_t* dims, const char* fn) {
  if (!dims || dims[0] < 0 || dims[1] < 0) {
    tt::set_error("%s: invalid dims", fn);
    return false;
  }
  return true;
}

}  // namespace

extern "C" {

int has_topotoolbox(void) { return 1; }

#ifndef TT_SOURCE_HASH
#define TT_SOURCE_HASH "unstamped"
#endif
const char* tt_build_hash(void) { return TT_SOURCE_HASH; }

const char* tt_last_error(void) { return tt::last_error(); }

int tt_set_tile_shape(int qx, int qy) {
  tt::TileShape s{qx, qy};
  if (!tt::tile_shape_ok(s)) {
    tt::set_error("tt_set_tile_shape: unsupported shape %dx%d", qx, qy);
    return 1;
  }
  g_shape = s;
  g_shape_set = true;
  return 0;
}

int tt_set_fill_schedule(int schedule) {
  if (schedule != TT_FILL_SCHEDULE_AUTO && schedule != TT_FILL_SCHEDULE_FIFO &&
      schedule != TT_FILL_SCHEDULE_PRIORITY) {
    tt::set_error("tt_set_fill_schedule: unknown schedule %d", schedule);
    return 1;
  }
  tt::set_fill_schedule(static_cast<tt::FillSchedule>(schedule));
  return 0;
}

int tt_fillsinks_dev(float* output, const float* dem, const uint8_t* bc, ptrdiff_t nx,
                     ptrdiff_t ny, void* stream, tt_stats* stats) {
  tt::set_error("%s", "");
  if (nx < 0 || ny < 0) {
    tt::set_error("tt_fillsinks_dev: invalid dims");
    return 1;
  }
  tt::RunStats st;
  hipError_t e = tt::run_fill(output, dem, bc, nx, ny, tt::kModeFill, shape_for({2, 2}),
                              (hipStream_t)stream, stats ? &st : nullptr);
  if (e != hipSuccess) return fail("tt_fillsinks_dev", e);
  if (stats) {
    to_public(st, stats);
    if (st.error) return device_error("tt_fillsinks_dev", st.error);
  }
  return 0;
}

int tt_fillsinks_noalloc_dev(float* output, const float* dem, const uint8_t* bc, ptrdiff_t nx,
                             ptrdiff_t ny, void* stream) {
  tt::set_error("%s", "");
  if (nx < 0 || ny < 0) {
    tt::set_error("tt_fillsinks_noalloc_dev: invalid dims");
    return 1;
  }
  if (tt::noalloc_fill_tiles(nx, ny) > tt::kNoAllocMaxTiles) {
    tt::set_error(
        "tt_fillsinks_noalloc_dev: %lld x %lld cells need %lld tiles of 128 x 128; the fixed "
        "workspace holds at most %lld (use tt_fillsinks_dev)",
        (long long)nx, (long long)ny, (long long)tt::noalloc_fill_tiles(nx, ny),
        (long long)tt::kNoAllocMaxTiles);
    return 2;
  }
  hipError_t e = tt::run_fill_noalloc(output, dem, bc, nx, ny, (hipStream_t)stream);
  if (e != hipSuccess) return fail("tt_fillsinks_noalloc_dev", e);
  return 0;
}

int tt_reconstruct_dev(float* marker, const float* mask, ptrdiff_t nx, ptrdiff_t ny,
                       void* stream, tt_stats* stats) {
  tt::set_error("%s", "");
  if (nx < 0 || ny < 0) {
    tt::set_error("tt_reconstruct_dev: invalid dims");
    return 1;
  }
  tt::RunStats st;
  hipError_t e = tt::run_fill(marker, mask, nullptr, nx, ny, tt::kModeRecon, shape_for({2, 2}),
                              (hipStream_t)stream, stats ? &st : nullptr);
  if (e != hipSuccess) return fail("tt_reconstruct_dev", e);
  if (stats) {
    to_public(st, stats);
    if (st.error) return device_error("tt_reconstruct_dev", st.error);
  }
  return 0;
}

static int fill_capped(const char* fn, float* state, const float* src, const uint8_t* bc,
                       ptrdiff_t nx, ptrdiff_t ny, int mode, int64_t max_visits, void* stream,
                       int* converged, tt_stats* stats) {
  tt::set_error("%s", "");
  if (nx < 0 || ny < 0) {
    tt::set_error("%s: invalid dims", fn);
    return 1;
  }
  if (max_visits < 1 || max_visits > (int64_t)UINT32_MAX) {
    tt::set_error("%s: max_visits must be in [1, 2^32 - 1], got %lld", fn, (long long)max_visits);
    return 1;
  }
  if (!converged) {
    tt::set_error("%s: converged must not be NULL", fn);
    return 1;
  }
  tt::RunStats st;
  hipError_t e = tt::run_fill_capped(state, src, bc, nx, ny, mode, (unsigned)max_visits,
                                     (hipStream_t)stream, &st, converged);
  if (e != hipSuccess) return fail(fn, e);
  if (stats) to_public(st, stats);
  if (st.error) return device_error(fn, st.error);
  return 0;
}

int tt_fillsinks_capped_dev(float* output, const float* dem, const uint8_t* bc, ptrdiff_t nx,
                            ptrdiff_t ny, int64_t max_visits, void* stream, int* converged,
                            tt_stats* stats) {
  return fill_capped("tt_fillsinks_capped_dev", output, dem, bc, nx, ny, tt::kModeFill,
                     max_visits, stream, converged, stats);
}

int tt_reconstruct_capped_dev(float* marker, const float* mask, ptrdiff_t nx, ptrdiff_t ny,
                              int64_t max_visits, void* stream, int* converged,
                              tt_stats* stats) {
  return fill_capped("tt_reconstruct_capped_dev", marker, mask, nullptr, nx, ny, tt::kModeRecon,
                     max_visits, stream, converged, stats);
}

int tt_excesstopography_fsm2d_dev(float* excess, const float* dem, const float* slopes,
                                  float cellsize, ptrdiff_t nx, ptrdiff_t ny, void* stream,
                                  tt_stats* stats) {
  tt::set_error("%s", "");
  if (nx < 0 || ny < 0) {
    tt::set_error("tt_excesstopography_fsm2d_dev: invalid dims");
    return 1;
  }
  tt::RunStats st;
  hipError_t e = tt::run_fsm(excess, dem, slopes, cellsize, nx, ny, shape_for({2, 1}),
                             (hipStream_t)stream,
                             stats ? &st : nullptr);
  if (e != hipSuccess) return fail("tt_excesstopography_fsm2d_dev", e);
  if (stats) {
    to_public(st, stats);
    if (st.error) return device_error("tt_excesstopography_fsm2d_dev", st.error);
  }
  return 0;
}

int tt_excesstopography_fmm3d_dev(float* excess, const float* dem, const float* lithstack,
                                  const float* threshold_slopes, float cellsize, ptrdiff_t nx,
                                  ptrdiff_t ny, ptrdiff_t nlayers, void* stream, tt_stats* stats) {
  tt::set_error("%s", "");
  if (nx < 0 || ny < 0 || nlayers < 1 || nlayers > 1024) {
    tt::set_error("tt_excesstopography_fmm3d_dev: dims must be >= 0 and 1 <= nlayers <= 1024");
    return 1;
  }
  tt::RunStats st;
  hipError_t e = tt::run_fmm3d(excess, dem, lithstack, threshold_slopes, cellsize, nx, ny,
                               nlayers, (hipStream_t)stream, stats ? &st : nullptr);
  if (e != hipSuccess) return fail("tt_excesstopography_fmm3d_dev", e);
  if (stats) {
    to_public(st, stats);
    if (st.error) return device_error("tt_excesstopography_fmm3d_dev", st.error);
  }
  return 0;
}

int tt_fillsinks_spill_dev(float* output, int32_t* spill, const float* dem, const uint8_t* bc,
                           ptrdiff_t nx, ptrdiff_t ny, void* stream, tt_stats* stats) {
  tt::set_error("%s", "");
  if (nx < 0 || ny < 0 || (nx > 0 && ny > INT32_MAX / nx)) {
    tt::set_error("tt_fillsinks_spill_dev: dims must be >= 0 with nx * ny <= INT32_MAX");
    return 1;
  }
  tt::RunStats st;
  hipError_t e = tt::run_fill_spill(output, spill, dem, bc, nx, ny, shape_for({2, 2}),
                                    (hipStream_t)stream,
                                    stats ? &st : nullptr);
  if (e != hipSuccess) return fail("tt_fillsinks_spill_dev", e);
  if (stats) {
    to_public(st, stats);
    if (st.error) return device_error("tt_fillsinks_spill_dev", st.error);
  }
  return 0;
}

int tt_fillsinks_backward_dev(float* g_dem, const int32_t* spill, const float* g_out,
                              ptrdiff_t nx, ptrdiff_t ny, void* stream) {
  tt::set_error("%s", "");
  if (nx < 0 || ny < 0 || (nx > 0 && ny > INT32_MAX / nx)) {
    tt::set_error("tt_fillsinks_backward_dev: dims must be >= 0 with nx * ny <= INT32_MAX");
    return 1;
  }
  hipError_t e = tt::run_fill_backward(g_dem, spill, g_out, nx, ny, (hipStream_t)stream);
  if (e != hipSuccess) return fail("tt_fillsinks_backward_dev", e);
  return 0;
}

int tt_excesstopography_fsm2d_backward_dev(float* g_dem, float* g_slopes, const float* excess,
                                           const float* dem, const float* slopes, float cellsize,
                                           const float* g_out, ptrdiff_t nx, ptrdiff_t ny,
                                           void* stream, tt_stats* stats) {
  tt::set_error("%s", "");
  if (nx < 0 || ny < 0) {
    tt::set_error("tt_excesstopography_fsm2d_backward_dev: invalid dims");
    return 1;
  }
  tt::RunStats st;
  hipError_t e = tt::run_fsm_backward(g_dem, g_slopes, excess, dem, slopes, cellsize, g_out, nx,
                                      ny, (hipStream_t)stream, stats ? &st : nullptr);
  if (e != hipSuccess) return fail("tt_excesstopography_fsm2d_backward_dev", e);
  if (stats) {
    to_public(st, stats);
    if (st.error) return device_error("tt_excesstopography_fsm2d_backward_dev", st.error);
  }
  return 0;
}

int tt_identifyflats_dev(int32_t* output, const float* dem, ptrdiff_t nx, ptrdiff_t ny,
                         void* stream, int64_t* count) {
  tt::set_error("%s", "");
  if (nx < 0 || ny < 0) {
    tt::set_error("tt_identifyflats_dev: invalid dims");
    return 1;
  }
  hipError_t e = tt::run_identifyflats(output, dem, nx, ny, (hipStream_t)stream, count);
  if (e != hipSuccess) return fail("tt_identifyflats_dev", e);
  return 0;
}

int tt_gwdt_computecosts_dev(float* costs, int64_t* conncomps, const int32_t* flats,
                             const float* original_dem, const float* filled_dem, ptrdiff_t nx,
                             ptrdiff_t ny, void* stream) {
  tt::set_error("%s", "");
  if (nx < 0 || ny < 0 || (nx > 0 && (unsigned long long)ny >= 0xffffffffull / (unsigned long long)nx)) {
    tt::set_error("tt_gwdt_computecosts_dev: dims must be >= 0 with nx * ny < 2^32 - 1");
    return 1;
  }
  hipError_t e = tt::run_gwdt_computecosts(costs, conncomps, flats, original_dem, filled_dem, nx,
                                           ny, (hipStream_t)stream);
  if (e != hipSuccess) return fail("tt_gwdt_computecosts_dev", e);
  return 0;
}

int tt_value_filter_dev(int op, float* output, const float* dem,
                        const uint8_t* structuring_element, const ptrdiff_t se_dims[3],
                        ptrdiff_t nx, ptrdiff_t ny, void* stream) {
  tt::set_error("%s", "");
  if (nx < 0 || ny < 0 || !se_dims || se_dims[0] < 0 || se_dims[1] < 0 || se_dims[2] < 0 ||
      op == 0) {
    tt::set_error("tt_value_filter_dev: invalid dims or op");
    return 1;
  }
  hipError_t e = tt::run_value_filter(op, output, dem, structuring_element, se_dims, nx, ny,
                                      (hipStream_t)stream);
  if (e != hipSuccess) return fail("tt_value_filter_dev", e);
  return 0;
}

int tt_value_filter_square_dev(int op, float* output, const float* dem, float* tmp, int width,
                               ptrdiff_t nx, ptrdiff_t ny, void* stream) {
  tt::set_error("%s", "");
  if (nx < 0 || ny < 0 || width < 0 || op == 0) {
    tt::set_error("tt_value_filter_square_dev: invalid dims, width or op");
    return 1;
  }
  hipError_t e =
      tt::run_value_filter_square(op, output, dem, tmp, width, nx, ny, (hipStream_t)stream);
  if (e != hipSuccess) return fail("tt_value_filter_square_dev", e);
  return 0;
}

int tt_batch_plan(int op, const tt_batch_item* items, int nitems, int qx, int qy, void* plan,
                  size_t* plan_bytes, size_t* workspace_bytes) {
  tt::set_error("%s", "");
  hipError_t e = tt::batch_plan(op, nitems, items, tt::TileShape{qx, qy}, plan, plan_bytes,
                                workspace_bytes);
  return e == hipSuccess ? 0 : (int)e;
}

int tt_batch_plan_set_sweep_caps(void* plan, int nitems, const uint32_t* caps) {
  tt::set_error("%s", "");
  hipError_t e = tt::batch_plan_set_sweep_caps(plan, nitems, caps);
  return e == hipSuccess ? 0 : (int)e;
}

int tt_batch_run_dev(const void* plan_host, const void* plan_dev,
                     const tt_batch_buffers* buffers, void* workspace, size_t workspace_bytes,
                     unsigned max_local, int flags, void* stream, tt_stats* stats) {
  tt::set_error("%s", "");
  if (!buffers) {
    tt::set_error("tt_batch_run_dev: buffers is NULL");
    return 1;
  }
  tt::RunStats st;
  hipError_t e = tt::run_batch(plan_host, plan_dev, *buffers, workspace, workspace_bytes,
                               max_local, flags, (hipStream_t)stream, stats ? &st : nullptr);
  if (e == hipErrorInvalidValue && tt::last_error()[0]) return (int)e;
  if (e != hipSuccess) return fail("tt_batch_run_dev", e);
  if (stats) {
    to_public(st, stats);
    if (st.error) return device_error("tt_batch_run_dev", st.error);
  }
  return 0;
}

// Host-pointer drop-ins: copy in, run the device entry point on the null
// stream, copy out (hipMemcpy waits for the kernels), then report device
// errors. Device memory comes from the per-GPU cache (HostStage).
#define TT_HOST_TRY(fn, expr)           \
  do {                                  \
    hipError_t e_ = (expr);             \
    if (e_ != hipSuccess) {             \
      fail(fn, e_);                     \
      return report(fn);                \
    }                                   \
  } while (0)

void fillsinks(float* output, float* dem, uint8_t* bc, ptrdiff_t dims[2]) {
  const char* fn = "fillsinks";
  tt::set_error("%s", "");
  if (!dims_ok(dims, fn)) return report(fn);
  const size_t n = (size_t)dims[0] * (size_t)dims[1];
  if (n == 0) return;
  HostStage st;
  const size_t sz[3] = {n * 4, n * 4, n};
  TT_HOST_TRY(fn, st.acquire(sz, 3));
  TT_HOST_TRY(fn, hipMemcpy(st[1], dem, n * 4, hipMemcpyHostToDevice));
  TT_HOST_TRY(fn, hipMemcpy(st[2], bc, n, hipMemcpyHostToDevice));
  if (tt_fillsinks_dev((float*)st[0], (const float*)st[1], (const uint8_t*)st[2], dims[0],
                       dims[1], nullptr, nullptr) != 0)
    return report(fn);
  TT_HOST_TRY(fn, hipMemcpy(output, st[0], n * 4, hipMemcpyDeviceToHost));
  if (sticky_check(fn) != 0) report(fn);
}

void reconstruct(float* marker, float* mask, ptrdiff_t dims[2]) {
  const char* fn = "reconstruct";
  tt::set_error("%s", "");
  if (!dims_ok(dims, fn)) return report(fn);
  const size_t n = (size_t)dims[0] * (size_t)dims[1];
  if (n == 0) return;
  HostStage st;
  const size_t sz[2] = {n * 4, n * 4};
  TT_HOST_TRY(fn, st.acquire(sz, 2));
  TT_HOST_TRY(fn, hipMemcpy(st[0], marker, n * 4, hipMemcpyHostToDevice));
  TT_HOST_TRY(fn, hipMemcpy(st[1], mask, n * 4, hipMemcpyHostToDevice));
  if (tt_reconstruct_dev((float*)st[0], (const float*)st[1], dims[0], dims[1], nullptr,
                         nullptr) != 0)
    return report(fn);
  TT_HOST_TRY(fn, hipMemcpy(marker, st[0], n * 4, hipMemcpyDeviceToHost));
  if (sticky_check(fn) != 0) report(fn);
}

void excesstopography_fsm2d(float* excess, float* dem, float* threshold_slopes, float cellsize,
                            ptrdiff_t dims[2]) {
  const char* fn = "excesstopography_fsm2d";
  tt::set_error("%s", "");
  if (!dims_ok(dims, fn)) return report(fn);
  const size_t n = (size_t)dims[0] * (size_t)dims[1];
  if (n == 0) return;
  HostStage st;
  const size_t sz[3] = {n * 4, n * 4, n * 4};
  TT_HOST_TRY(fn, st.acquire(sz, 3));
  TT_HOST_TRY(fn, hipMemcpy(st[1], dem, n * 4, hipMemcpyHostToDevice));
  TT_HOST_TRY(fn, hipMemcpy(st[2], threshold_slopes, n * 4, hipMemcpyHostToDevice));
  if (tt_excesstopography_fsm2d_dev((float*)st[0], (const float*)st[1], (const float*)st[2],
                                    cellsize, dims[0], dims[1], nullptr, nullptr) != 0)
    return report(fn);
  TT_HOST_TRY(fn, hipMemcpy(excess, st[0], n * 4, hipMemcpyDeviceToHost));
  if (sticky_check(fn) != 0) report(fn);
}

void fillsinks_hybrid(float* output, ptrdiff_t* queue, float* dem, uint8_t* bc,
                      ptrdiff_t dims[2]) {
  (void)queue;
  fillsinks(output, dem, bc, dims);
}

int reconstruct_hybrid(float* marker, ptrdiff_t* queue, float* mask, ptrdiff_t dims[2]) {
  (void)queue;
  reconstruct(marker, mask, dims);
  return tt::last_error()[0] ? 1 : 0;
}

void excesstopography_fmm2d(float* excess, ptrdiff_t* heap, ptrdiff_t* back, float* dem,
                            float* threshold_slopes, float cellsize, ptrdiff_t dims[2]) {
  (void)heap;
  (void)back;
  excesstopography_fsm2d(excess, dem, threshold_slopes, cellsize, dims);
}

void excesstopography_fmm3d(float* excess, ptrdiff_t* heap, ptrdiff_t* back, float* dem,
                            float* lithstack, float* threshold_slopes, float cellsize,
                            ptrdiff_t dims[2], ptrdiff_t nlayers) {
  (void)heap;
  (void)back;
  const char* fn = "excesstopography_fmm3d";
  tt::set_error("%s", "");
  if (!dims_ok(dims, fn)) return report(fn);
  if (nlayers < 1 || nlayers > 1024) {
    tt::set_error("%s: nlayers must be in 1..1024", fn);
    return report(fn);
  }
  const size_t n = (size_t)dims[0] * (size_t)dims[1];
  if (n == 0) return;
  const size_t nl = (size_t)nlayers;
  HostStage st;
  const size_t sz[4] = {n * 4, n * 4, n * nl * 4, nl * 4};
  TT_HOST_TRY(fn, st.acquire(sz, 4));
  TT_HOST_TRY(fn, hipMemcpy(st[1], dem, n * 4, hipMemcpyHostToDevice));
  TT_HOST_TRY(fn, hipMemcpy(st[2], lithstack, n * nl * 4, hipMemcpyHostToDevice));
  TT_HOST_TRY(fn, hipMemcpy(st[3], threshold_slopes, nl * 4, hipMemcpyHostToDevice));
  if (tt_excesstopography_fmm3d_dev((float*)st[0], (const float*)st[1], (const float*)st[2],
                                    (const float*)st[3], cellsize, dims[0], dims[1], nlayers,
                                    nullptr, nullptr) != 0)
    return report(fn);
  TT_HOST_TRY(fn, hipMemcpy(excess, st[0], n * 4, hipMemcpyDeviceToHost));
  if (sticky_check(fn) != 0) report(fn);
}

ptrdiff_t identifyflats(int32_t* output, float* dem, ptrdiff_t dims[2]) {
  const char* fn = "identifyflats";
  tt::set_error("%s", "");
  if (!dims_ok(dims, fn)) return report(fn), 0;
  const size_t n = (size_t)dims[0] * (size_t)dims[1];
  if (n == 0) return 0;
  HostStage st;
  const size_t sz[2] = {n * 4, n * 4};
  hipError_t e = st.acquire(sz, 2);
  if (e == hipSuccess) e = hipMemcpy(st[1], dem, n * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail(fn, e), report(fn), 0;
  int64_t count = 0;
  if (tt_identifyflats_dev((int32_t*)st[0], (const float*)st[1], dims[0], dims[1], nullptr,
                           &count) != 0)
    return report(fn), 0;
  e = hipMemcpy(output, st[0], n * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail(fn, e), report(fn), 0;
  return (ptrdiff_t)count;
}

void gwdt_computecosts(float* costs, ptrdiff_t* conncomps, int32_t* flats, float* original_dem,
                       float* filled_dem, ptrdiff_t dims[2]) {
  const char* fn = "gwdt_computecosts";
  tt::set_error("%s", "");
  if (!dims_ok(dims, fn)) return report(fn);
  const size_t n = (size_t)dims[0] * (size_t)dims[1];
  if (n == 0) return;
  HostStage st;
  const size_t sz[5] = {n * 4, n * 8, n * 4, n * 4, n * 4};
  TT_HOST_TRY(fn, st.acquire(sz, 5));
  TT_HOST_TRY(fn, hipMemcpy(st[2], flats, n * 4, hipMemcpyHostToDevice));
  TT_HOST_TRY(fn, hipMemcpy(st[3], original_dem, n * 4, hipMemcpyHostToDevice));
  TT_HOST_TRY(fn, hipMemcpy(st[4], filled_dem, n * 4, hipMemcpyHostToDevice));
  if (tt_gwdt_computecosts_dev((float*)st[0], (int64_t*)st[1], (const int32_t*)st[2],
                               (const float*)st[3], (const float*)st[4], dims[0], dims[1],
                               nullptr) != 0)
    return report(fn);
  TT_HOST_TRY(fn, hipMemcpy(costs, st[0], n * 4, hipMemcpyDeviceToHost));
  TT_HOST_TRY(fn, hipMemcpy(conncomps, st[1], n * 8, hipMemcpyDeviceToHost));
}

void tt_host_workspace_release(void) {
  for (int d = 0; d < kMaxDevices; d++) {
    HostCache& c = g_host_cache[d];
    std::lock_guard<std::mutex> lock(c.mu);
    if (!c.p) continue;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return;
    if (hipSetDevice(d) == hipSuccess) {
      (void)hipFree(c.p);
      c.p = nullptr;
      c.bytes = 0;
    }
    (void)hipSetDevice(cur);
  }
}

int tt_device_error(int clear, uint32_t* bits) {
  tt::set_error("%s", "");
  unsigned b = 0;
  hipError_t e = tt::take_sticky_error(clear != 0, &b);
  if (e != hipSuccess) return fail("tt_device_error", e);
  if (bits) *bits = b;
  return 0;
}

int tt_set_wait_limit(uint64_t ticks) {
  tt::set_wait_limit_ticks(ticks);
  return 0;
}

}  // extern "C"

namespace {

void value_filter_host(const char* fn, int op, float* output, const float* dem,
                       const uint8_t* se, const ptrdiff_t* io_dims, const ptrdiff_t* se_dims) {
  tt::set_error("%s", "");
  if (!dims_ok(io_dims, fn)) return report(fn);
  if (!se_dims || se_dims[0] < 0 || se_dims[1] < 0 || se_dims[2] < 0) {
    tt::set_error("%s: invalid se_dims", fn);
    return report(fn);
  }
  const size_t n = (size_t)io_dims[0] * (size_t)io_dims[1];
  const size_t ns = (size_t)se_dims[0] * (size_t)se_dims[1] * (size_t)se_dims[2];
  if (n == 0) return;
  HostStage st;
  const size_t sz[3] = {n * 4, n * 4, ns};
  TT_HOST_TRY(fn, st.acquire(sz, 3));
  TT_HOST_TRY(fn, hipMemcpy(st[1], dem, n * 4, hipMemcpyHostToDevice));
  if (ns) TT_HOST_TRY(fn, hipMemcpy(st[2], se, ns, hipMemcpyHostToDevice));
  if (tt_value_filter_dev(op, (float*)st[0], (const float*)st[1], (const uint8_t*)st[2],
                          se_dims, io_dims[0], io_dims[1], nullptr) != 0)
    return report(fn);
  TT_HOST_TRY(fn, hipMemcpy(output, st[0], n * 4, hipMemcpyDeviceToHost));
}

void value_filter_square_host(const char* fn, int op, float* output, const float* dem,
                              float* tmp, int width, const ptrdiff_t* io_dims) {
  tt::set_error("%s", "");
  if (!dims_ok(io_dims, fn)) return report(fn);
  const size_t n = (size_t)io_dims[0] * (size_t)io_dims[1];
  if (n == 0) return;
  HostStage st;
  const size_t sz[3] = {n * 4, n * 4, n * 4};
  TT_HOST_TRY(fn, st.acquire(sz, 3));
  TT_HOST_TRY(fn, hipMemcpy(st[1], dem, n * 4, hipMemcpyHostToDevice));
  if (tt_value_filter_square_dev(op, (float*)st[0], (const float*)st[1], (float*)st[2], width,
                                 io_dims[0], io_dims[1], nullptr) != 0)
    return report(fn);
  TT_HOST_TRY(fn, hipMemcpy(output, st[0], n * 4, hipMemcpyDeviceToHost));
  if (tmp) TT_HOST_TRY(fn, hipMemcpy(tmp, st[2], n * 4, hipMemcpyDeviceToHost));
}

}  // namespace

extern "C" {

void min_filter(float* output, float* dem, uint8_t* structuring_element, ptrdiff_t io_dims[2],
                ptrdiff_t se_dims[3]) {
  value_filter_host("min_filter", -1, output, dem, structuring_element, io_dims, se_dims);
}

void max_filter(float* output, float* dem, uint8_t* structuring_element, ptrdiff_t io_dims[2],
                ptrdiff_t se_dims[3]) {
  value_filter_host("max_filter", 1, output, dem, structuring_element, io_dims, se_dims);
}

void min_filter_square(float* output, float* dem, float* tmp, uint8_t width,
                       ptrdiff_t io_dims[2]) {
  value_filter_square_host("min_filter_square", -1, output, dem, tmp, width, io_dims);
}

void max_filter_square(float* output, float* dem, float* tmp, uint8_t width,
                       ptrdiff_t io_dims[2]) {
  value_filter_square_host("max_filter_square", 1, output, dem, tmp, width, io_dims);
}

}  // extern "C"
