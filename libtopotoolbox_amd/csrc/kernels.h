// Host-side interface of the HIP kernels (C++; the C ABI lives in capi.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/topotoolbox_amd.h"

namespace tt {

struct RunStats {
  int64_t requeues = 0;     // tiles queued again after the initial queue
  int64_t visits = 0;       // tile visits
  int64_t local_iters = 0;  // local sweep iterations summed over visits
  int64_t max_local = 0;    // largest local iteration count of one visit
  int64_t grid = 0;         // workgroups in the persistent grid
  int64_t tile_x = 0, tile_y = 0;
  uint32_t error = 0;       // Sync error bits
  int64_t pending = 0;      // tiles still queued or running at the end (0: converged)
  // wall-clock ticks (100 MHz) summed over workgroups, by phase
  int64_t t_stage = 0, t_relax = 0, t_write = 0, t_wait = 0;
};

// Tile shape in 64x64 quadrants (one wavefront per quadrant).
struct TileShape {
  int qx = 2, qy = 2;
};

enum FillMode : int {
  kModeFill = 0,   // fillsinks: state = output levels, src = dem, bc used
  kModeRecon = 1,  // reconstruct: state = marker (in place), src = mask
};

// Sink filling / reconstruction. `stats` (host) may be null; when set the
// call synchronises the stream to read the control block.
hipError_t run_fill(float* state, const float* src, const uint8_t* bc, int64_t nx, int64_t ny,
                    int mode, TileShape ts, hipStream_t stream, RunStats* stats);

// Work-queue schedule of sink filling on one grid (kModeFill) for later
// calls on this thread: automatic (lowest level first on large grids, FIFO
// otherwise), FIFO, or lowest level first. Both queues reach the same
// result.
enum FillSchedule : int { kSchedAuto = 0, kSchedFifo = 1, kSchedPriority = 2 };
void set_fill_schedule(FillSchedule s);

// run_fill (128 x 128 tiles) that stops after about max_visits (>= 1) tile
// visits: every tile is visited at least once, and the state is a valid
// partial result (between the fixed point and the initial marker), so a
// later reconstruction from it reaches the same fixed point. Synchronises;
// *converged = 1 if the launch finished all work; stats (required) without
// the stop bit.
hipError_t run_fill_capped(float* state, const float* src, const uint8_t* bc, int64_t nx,
                           int64_t ny, int mode, unsigned max_visits, hipStream_t stream,
                           RunStats* stats, int* converged);

// fillsinks (work-queue algorithm, 128 x 128 tiles) with its queue
// workspace in a fixed per-GPU device array instead of an allocation:
// allocates nothing. Grids of more than kNoAllocMaxTiles tiles are
// rejected (hipErrorInvalidValue); calls overlapping on one GPU race.
constexpr size_t kNoAllocScratchBytes = (size_t)4 << 20;  // 4 MiB per GPU
constexpr int64_t kNoAllocMaxTiles = (int64_t)((kNoAllocScratchBytes - 64 * 4) / 16);
int64_t noalloc_fill_tiles(int64_t nx, int64_t ny);
hipError_t run_fill_noalloc(float* state, const float* src, const uint8_t* bc, int64_t nx,
                            int64_t ny, hipStream_t stream);

// Fast-sweeping excess topography: u (output) from dem and slopes.
hipError_t run_fsm(float* u, const float* dem, const float* slopes, float cellsize, int64_t nx,
                   int64_t ny, TileShape ts, hipStream_t stream, RunStats* stats);

// Excess topography with layered lithology (fmm3d.hip): lithstack holds
// nlayers layer tops per cell (cell-major), slopes one threshold slope per
// layer (device memory). 1 <= nlayers <= 1024 (hipErrorInvalidValue else).
hipError_t run_fmm3d(float* u, const float* dem, const float* lithstack, const float* slopes,
                     float cellsize, int64_t nx, int64_t ny, int64_t nlayers, hipStream_t stream,
                     RunStats* stats);

// fillsinks plus the spill cell of every output value (int32, -1 for NaN
// and unreached cells); needs nx * ny <= INT32_MAX.
hipError_t run_fill_spill(float* out, int* spill, const float* dem, const uint8_t* bc, int64_t nx,
                          int64_t ny, TileShape ts, hipStream_t stream, RunStats* stats);

// g_dem[q] = sum of g_out[p] over p with spill[p] == q (deterministic;
// fixed point scaled per call, so no overflow; NaN / INF entries propagate
// as in an IEEE sum).
hipError_t run_fill_backward(float* g_dem, const int* spill, const float* g_out, int64_t nx,
                             int64_t ny, hipStream_t stream);

// Adjoint of excesstopography_fsm2d at its converged output u.
hipError_t run_fsm_backward(float* g_dem, float* g_slopes, const float* u, const float* dem,
                            const float* slopes, float cellsize, const float* g_out, int64_t nx,
                            int64_t ny, hipStream_t stream, RunStats* stats);

// ---------------------------------------------------------------- batches
enum : int {
  kBatchFill = TT_BATCH_FILLSINKS,
  kBatchFsm = TT_BATCH_EXCESSTOPOGRAPHY,
  kBatchForwardOnly = TT_BATCH_FORWARD_ONLY,
};

using BatchItem = tt_batch_item;
using BatchBuffers = tt_batch_buffers;

// One item as the kernel sees it (in the plan, device memory).
struct BatchItemDev {
  long long nx, ny, offset;
  float cellsize;
  int ntx, nty;     // tiles along x, y
  int ntiles;       // ntx * nty
  int tile_base;    // global id of the item's first tile
  unsigned max_local;  // sweep cap of this item's visits (0: the launch's)
};

// Plan layout: header | items (BatchItemDev) | tile -> item (unsigned),
// each part 256-byte aligned. The workspace layout is recorded too.
constexpr unsigned kBatchPlanMagic = 0x54544250u;  // "PBTT"
struct BatchPlanHeader {
  unsigned magic;
  int op, nitems, ntiles, qx, qy;
  long long cells;             // end of the packed cell range
  size_t items_off, tile_item_off, bytes;
  size_t ws_queue, ws_items, ws_scratch, ws_bytes;
};

// Host-only: sizes (plan == null) or writes the plan.
hipError_t batch_plan(int op, int nitems, const BatchItem* items, TileShape ts, void* plan,
                      size_t* plan_bytes, size_t* workspace_bytes);

// Host-only: per-item sweep caps written into a plan (caps[b] = 0: the
// launch's cap). The plan must be copied to the device again afterwards.
hipError_t batch_plan_set_sweep_caps(void* plan, int nitems, const unsigned* caps);

// One persistent launch (plus a small init kernel) over a planned batch;
// no allocation, no synchronisation unless stats is non-null.
hipError_t run_batch(const void* plan_host, const void* plan_dev, const BatchBuffers& buf,
                     void* workspace, size_t workspace_bytes, unsigned max_local, int flags,
                     hipStream_t stream, RunStats* stats);

// Stencils next to the hot path (stencil.hip). count_host, when non-null,
// receives the number of flats after a synchronisation.
hipError_t run_identifyflats(int32_t* out, const float* dem, int64_t nx, int64_t ny,
                             hipStream_t stream, int64_t* count_host);
hipError_t run_value_filter(int op, float* out, const float* dem, const uint8_t* se,
                            const int64_t se_dims[3], int64_t nx, int64_t ny,
                            hipStream_t stream);
hipError_t run_value_filter_square(int op, float* out, const float* dem, float* tmp, int width,
                                   int64_t nx, int64_t ny, hipStream_t stream);

// gwdt_computecosts (gwdt.hip); synchronises the stream (once, or three times
// when components with tied deepest depths need the scan replay).
hipError_t run_gwdt_computecosts(float* costs, int64_t* conncomps, const int32_t* flats,
                                 const float* original, const float* filled, int64_t nx,
                                 int64_t ny, hipStream_t stream);

// Thread-local error message set by the entry points.
void set_error(const char* fmt, ...);

// Checks a tile shape is one of the compiled variants.
bool tile_shape_ok(TileShape ts);

}  // namespace tt
