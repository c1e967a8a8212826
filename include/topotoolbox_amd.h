/*
  topotoolbox_amd.h - C ABI of the MI355X (gfx950) sink-filling and
  excess-topography library.

  Part 1 declares the reference's own entry points with the reference's
  exact signatures (has_topotoolbox, fillsinks, reconstruct,
  excesstopography_fsm2d, their hybrid / fast-marching siblings, and the
  steps next to them: identifyflats, gwdt_computecosts, min_filter,
  max_filter and their square variants), so a C caller of libtopotoolbox
  links against libtopotoolbox_amd.so unchanged. They take host pointers, copy to the
  GPU, run the HIP kernels and copy back before returning.

  Part 2 adds device-pointer entry points (stream-ordered, for ctypes /
  torch callers whose data already lives in HBM), the spill-index forward
  and the adjoints, the batched tile entry, and error / tuning hooks. None
  of these have a counterpart in the reference; each says what it is for.

  Layout everywhere: cell (i, j) is at p = j * dims[0] + i, dims[0]
  changing fastest (reference include/topotoolbox.h:284-290). In the
  device entry points nx = dims[0] and ny = dims[1].

  Errors: the reference functions return void and "always succeed"
  (reference docs/ARCHITECTURE.md:83-98). Here a failure (no GPU, out of
  memory, a kernel error) leaves the outputs unspecified, is reported on
  stderr once, and is readable from tt_last_error(). There is no CPU
  fallback.
*/
#ifndef TOPOTOOLBOX_AMD_H
#define TOPOTOOLBOX_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TOPOTOOLBOX_AMD_VERSION_MAJOR 0
#define TOPOTOOLBOX_AMD_VERSION_MINOR 2

/* The reference's visibility and version macros (reference
   include/topotoolbox.h:10-32), so that a caller switching headers keeps
   them. The bodies equal the reference's on Linux (an empty
   TOPOTOOLBOX_API: the symbols have default ELF visibility), so both
   headers can be included in one translation unit, in either order. The
   version is the reference interface these entry points implement. */
#ifndef TOPOTOOLBOX_API
#define TOPOTOOLBOX_API
#endif
#define TOPOTOOLBOX_VERSION_MAJOR 3
#define TOPOTOOLBOX_VERSION_MINOR 0
#define TOPOTOOLBOX_VERSION_PATCH 0
#ifndef TOPOTOOLBOX_OPENMP_VERSION
#define TOPOTOOLBOX_OPENMP_VERSION 0
#endif

/* ---------------------------------------------------------------------
   Part 1: drop-in replacements (host pointers)
   --------------------------------------------------------------------- */

/* Replaces has_topotoolbox, reference include/topotoolbox.h:47-48
   (src/topotoolbox.c:4-5). Always returns 1. */
TOPOTOOLBOX_API
int has_topotoolbox(void);

/* Replaces fillsinks, reference include/topotoolbox.h:292-293
   (src/fillsinks.c:22-48). `dem` is read only (the reference mutates and
   restores it; the observable contract is the same). */
TOPOTOOLBOX_API
void fillsinks(float *output, float *dem, uint8_t *bc, ptrdiff_t dims[2]);

/* Replaces reconstruct, reference include/topotoolbox.h:1055-1056
   (src/reconstruct.c:270-279). Runs to convergence; the reference stops
   after 1000 iterations, which matters only for inputs that need more. */
TOPOTOOLBOX_API
void reconstruct(float *marker, float *mask, ptrdiff_t dims[2]);

/* Replaces excesstopography_fsm2d, reference
   include/topotoolbox.h:607-609 (src/excesstopography.c:47-111). */
TOPOTOOLBOX_API
void excesstopography_fsm2d(float *excess, float *dem, float *threshold_slopes,
                            float cellsize, ptrdiff_t dims[2]);

/* Replaces fillsinks_hybrid, reference include/topotoolbox.h:351-352
   (src/fillsinks.c:50-77). Same output as fillsinks (the reference's hybrid
   algorithm reaches the same reconstruction); `queue` is scratch the
   reference needs for its FIFO and is not used here. */
TOPOTOOLBOX_API
void fillsinks_hybrid(float *output, ptrdiff_t *queue, float *dem, uint8_t *bc,
                      ptrdiff_t dims[2]);

/* Replaces reconstruct_hybrid, reference include/topotoolbox.h:1069-1071
   (src/reconstruct.c:291-312). Same output as reconstruct; `queue` is not
   used. Returns 0 (converged), or 1 if the call failed (the reference
   returns 1 when its queue overflowed twice without converging). */
TOPOTOOLBOX_API
int reconstruct_hybrid(float *marker, ptrdiff_t *queue, float *mask, ptrdiff_t dims[2]);

/* Replaces excesstopography_fmm2d, reference include/topotoolbox.h:694-697
   (src/excesstopography.c:113-172). Served by the GPU fast sweeping kernel:
   both methods solve the same constrained eikonal equation, and the
   reference's own tests require them to agree within 1e-4
   (test/excesstopography.cpp:69-82,133); `heap` and `back` (priority-queue
   scratch of the reference) are not used. */
TOPOTOOLBOX_API
void excesstopography_fmm2d(float *excess, ptrdiff_t *heap, ptrdiff_t *back, float *dem,
                            float *threshold_slopes, float cellsize, ptrdiff_t dims[2]);

/* Replaces excesstopography_fmm3d, reference include/topotoolbox.h:780-783
   (src/excesstopography.c:174-262): excess topography with layered
   lithology. lithstack holds nlayers layer tops per cell, bottom up
   (lithstack[(j * dims[0] + i) * nlayers + layer]); threshold_slopes one
   slope per layer. A cell takes the Godunov proposal of the lowest layer
   whose proposal lies below that layer's top. Solved by fast sweeping on
   the GPU (`heap` and `back` are not used): the reference's heap order and
   the sweeping order can reach different surfaces where the layer choice
   switches, since the choice is not monotone in the neighbours; both keep
   excess <= dem, and with one layer the result is that of
   excesstopography_fsm2d. Threshold slopes must be >= 0; 1 <= nlayers <=
   1024. */
TOPOTOOLBOX_API
void excesstopography_fmm3d(float *excess, ptrdiff_t *heap, ptrdiff_t *back, float *dem,
                            float *lithstack, float *threshold_slopes, float cellsize,
                            ptrdiff_t dims[2], ptrdiff_t nlayers);

/* Replaces identifyflats, reference include/topotoolbox.h:405
   (src/identifyflats.c:32-121). output[p] & 1: flat, & 2: sill, & 4:
   presill; returns the number of flats. */
TOPOTOOLBOX_API
ptrdiff_t identifyflats(int32_t *output, float *dem, ptrdiff_t dims[2]);

/* Replace min_filter and max_filter, reference include/topotoolbox.h:101-102
   and :155-156 (src/value_filters.c:9-59, 61-110). The structuring element
   is se_dims[0] x se_dims[1] x se_dims[2] bytes (slices folded in turn);
   the result has the reference's bits, including the sign of zeros. */
TOPOTOOLBOX_API
void min_filter(float *output, float *dem, uint8_t *structuring_element, ptrdiff_t io_dims[2],
                ptrdiff_t se_dims[3]);
TOPOTOOLBOX_API
void max_filter(float *output, float *dem, uint8_t *structuring_element, ptrdiff_t io_dims[2],
                ptrdiff_t se_dims[3]);

/* Replace min_filter_square and max_filter_square, reference
   include/topotoolbox.h:204-205 and :253-254 (src/value_filters.c:112-178,
   180-246). `tmp` receives the pass along dims[0], as in the reference. */
TOPOTOOLBOX_API
void min_filter_square(float *output, float *dem, float *tmp, uint8_t width,
                       ptrdiff_t io_dims[2]);
TOPOTOOLBOX_API
void max_filter_square(float *output, float *dem, float *tmp, uint8_t width,
                       ptrdiff_t io_dims[2]);

/* Replaces gwdt_computecosts, reference include/topotoolbox.h:464-466
   (src/gwdt.c:166-274): costs = (largest depth of the flat's 8-connected
   flat component - its own depth) + 0.1, depth = filled - original;
   conncomps = the component's label, the same pixel the reference's scan
   picks (a deepest pixel of the component). Non-flats get 0 and 0. Flats
   are expected as identifyflats makes them (none on the grid border).
   NaN depths (flats of INF - INF) are handled as the reference handles
   them: a NaN depth is never "strictly deeper", so the label follows the
   reference's scan and the cost is NaN. nx * ny must be below 2^32. */
TOPOTOOLBOX_API
void gwdt_computecosts(float *costs, ptrdiff_t *conncomps, int32_t *flats, float *original_dem,
                       float *filled_dem, ptrdiff_t dims[2]);

/* ---------------------------------------------------------------------
   Part 2: device-pointer API (no reference counterpart)
   All pointers are device pointers; `stream` is a hipStream_t (NULL = the
   default stream). Calls are stream-ordered and asynchronous unless
   `stats` is non-NULL, in which case the call waits for completion and
   fills *stats. Return 0 on success, a nonzero code otherwise
   (message in tt_last_error()).
   --------------------------------------------------------------------- */

typedef struct tt_stats {
  int64_t requeues;    /* tiles queued again after the initial queue of all tiles */
  int64_t visits;      /* tile visits (all tiles once, plus the requeues) */
  int64_t local_iters; /* local sweep iterations summed over visits */
  int64_t max_local;   /* largest local iteration count of one visit */
  int64_t grid;        /* workgroups of the persistent launch */
  int64_t tile_x;      /* tile width (cells along dims[0]) */
  int64_t tile_y;      /* tile height (cells along dims[1]) */
  uint32_t error;      /* device error bits (0 = none) */
  /* wall-clock ticks (100 MHz) summed over workgroups, by phase */
  int64_t t_stage;     /* staging tiles into LDS */
  int64_t t_relax;     /* local sweeps */
  int64_t t_write;     /* write-back and neighbour queueing */
  int64_t t_wait;      /* waiting for a tile to pop */
} tt_stats;

/* fillsinks on device buffers (for callers whose DEM is already in HBM). */
int tt_fillsinks_dev(float *output, const float *dem, const uint8_t *bc, ptrdiff_t nx,
                     ptrdiff_t ny, void *stream, tt_stats *stats);

/* fillsinks on device buffers that allocates nothing. No counterpart in
   the reference: for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc). Same result as tt_fillsinks_dev, bit for bit. The work-queue
   state lives in a fixed 4 MiB array per GPU, held until the process ends,
   which limits the grid to 262128 tiles of 128 x 128 cells (about 4.29e9
   cells); larger grids are rejected with status 2. Every call on a GPU
   shares that GPU's array: calls that overlap in time on one GPU race and
   must be serialised by the caller (one stream, or events). */
int tt_fillsinks_noalloc_dev(float *output, const float *dem, const uint8_t *bc, ptrdiff_t nx,
                             ptrdiff_t ny, void *stream);

/* reconstruct on device buffers; `marker` is updated in place. */
int tt_reconstruct_dev(float *marker, const float *mask, ptrdiff_t nx, ptrdiff_t ny,
                       void *stream, tt_stats *stats);

/* fillsinks / reconstruct with an iteration cap, the counterpart of the
   reference's 1000-iteration cap in reconstruct (src/reconstruct.c:273):
   the launch stops taking new work after about `max_visits` tile visits
   (1 <= max_visits < 2^32; 128 x 128 tiles, each visited at least once).
   The output is then a partial result: pointwise between the full result
   and the initial marker (for fillsinks: the full fill <= output, and
   output >= dem), and reconstruct called on it (for fillsinks: marker =
   -output, mask = -dem, negated back) reaches the full result bit for bit.
   The capped state differs from the reference's capped marker (another
   update order). Waits for completion; *converged (required) = 1 when all
   work finished before the cap. */
int tt_fillsinks_capped_dev(float *output, const float *dem, const uint8_t *bc, ptrdiff_t nx,
                            ptrdiff_t ny, int64_t max_visits, void *stream, int *converged,
                            tt_stats *stats);
int tt_reconstruct_capped_dev(float *marker, const float *mask, ptrdiff_t nx, ptrdiff_t ny,
                              int64_t max_visits, void *stream, int *converged,
                              tt_stats *stats);

/* excesstopography_fsm2d on device buffers. */
int tt_excesstopography_fsm2d_dev(float *excess, const float *dem, const float *slopes,
                                  float cellsize, ptrdiff_t nx, ptrdiff_t ny, void *stream,
                                  tt_stats *stats);

/* excesstopography_fmm3d on device buffers (lithstack: nx * ny * nlayers
   floats, threshold_slopes: nlayers floats, both device memory). */
int tt_excesstopography_fmm3d_dev(float *excess, const float *dem, const float *lithstack,
                                  const float *threshold_slopes, float cellsize, ptrdiff_t nx,
                                  ptrdiff_t ny, ptrdiff_t nlayers, void *stream, tt_stats *stats);

/* fillsinks that also returns, for every cell, the index of the cell whose
   DEM value the output copies (its spill cell; -1 for NaN and unreached
   cells). Ties are broken by the lexicographic (level, index) order, so the
   index is unique and deterministic. For the adjoint; nx * ny must not
   exceed INT32_MAX. */
int tt_fillsinks_spill_dev(float *output, int32_t *spill, const float *dem, const uint8_t *bc,
                           ptrdiff_t nx, ptrdiff_t ny, void *stream, tt_stats *stats);

/* Adjoint of fillsinks: g_dem[q] = sum of g_out[p] over p with
   spill[p] == q. Finite entries are accumulated in 64-bit fixed point, so
   the result does not depend on the order of the atomic adds; the scale is
   chosen per call from max |g_out| so that no sum can overflow: with
   max |g_out| < 2^E and nx * ny < 2^B, a term is rounded to a multiple of
   2^(E + B - 62). NaN and INF entries propagate as in an IEEE sum
   (opposite infinities give NaN). */
int tt_fillsinks_backward_dev(float *g_dem, const int32_t *spill, const float *g_out,
                              ptrdiff_t nx, ptrdiff_t ny, void *stream);

/* Adjoint of excesstopography_fsm2d at its converged output `excess`:
   for L = sum(g_out * excess), writes dL/ddem and dL/dthreshold_slopes. */
int tt_excesstopography_fsm2d_backward_dev(float *g_dem, float *g_slopes, const float *excess,
                                           const float *dem, const float *slopes, float cellsize,
                                           const float *g_out, ptrdiff_t nx, ptrdiff_t ny,
                                           void *stream, tt_stats *stats);

/* identifyflats on device buffers. With `count` non-NULL the call waits and
   stores the number of flats there; with NULL it stays asynchronous. */
int tt_identifyflats_dev(int32_t *output, const float *dem, ptrdiff_t nx, ptrdiff_t ny,
                         void *stream, int64_t *count);

/* gwdt_computecosts on device buffers. Waits for completion (the number
   of pixels in components whose deepest depth is tied is read back, and
   when there are any, the number of scan steps to replay). NaN depths are
   handled as in the reference (see gwdt_computecosts). Takes about 42
   bytes of device workspace per pixel. Returns nonzero with a message for
   nx * ny >= 2^32. */
int tt_gwdt_computecosts_dev(float *costs, int64_t *conncomps, const int32_t *flats,
                             const float *original_dem, const float *filled_dem, ptrdiff_t nx,
                             ptrdiff_t ny, void *stream);

/* min_filter (op < 0) or max_filter (op > 0) on device buffers;
   `structuring_element` is device memory of se_dims[0]*se_dims[1]*se_dims[2]
   bytes. */
int tt_value_filter_dev(int op, float *output, const float *dem,
                        const uint8_t *structuring_element, const ptrdiff_t se_dims[3],
                        ptrdiff_t nx, ptrdiff_t ny, void *stream);

/* min_filter_square (op < 0) or max_filter_square (op > 0) on device
   buffers; `tmp` is device memory of nx * ny floats. */
int tt_value_filter_square_dev(int op, float *output, const float *dem, float *tmp, int width,
                               ptrdiff_t nx, ptrdiff_t ny, void *stream);

/* ---------------------------------------------------------------------
   Batched entry (no reference counterpart: the reference handles one grid
   per call). Many grids ("items") of arbitrary, uneven sizes run in ONE
   persistent launch whose tiles share a single work queue; each item goes
   through the forward pass and, unless TT_BATCH_FORWARD_ONLY, the adjoint,
   with no host round trip in between. The launch allocates nothing and
   does not synchronise (unless `stats` is non-NULL), so it can be captured
   in a HIP graph and replayed.

   Usage: tt_batch_plan once with plan = NULL to get the sizes, again with a
   host buffer of plan_bytes; copy the plan to device memory; allocate
   workspace_bytes of device memory; then tt_batch_run_dev as often as
   wanted (the plan and workspace can be reused by calls that do not
   overlap in time). Items are packed in the caller's buffers: item b's
   cell (i, j) is element offset + j * nx + i of every per-cell buffer.
   Items must not overlap.
   --------------------------------------------------------------------- */

#define TT_BATCH_FILLSINKS 0         /* forward: fillsinks + spill index; adjoint: g_dem */
#define TT_BATCH_EXCESSTOPOGRAPHY 1  /* forward: excess; adjoint: g_dem, g_slopes */
#define TT_BATCH_FORWARD_ONLY 1      /* flag: skip the adjoint phases */

typedef struct tt_batch_item {
  ptrdiff_t nx;      /* cells along dims[0] */
  ptrdiff_t ny;      /* cells along dims[1] */
  ptrdiff_t offset;  /* index of cell (0, 0) in the packed buffers */
  float cellsize;    /* excess topography only */
} tt_batch_item;

typedef struct tt_batch_buffers {
  float *output;          /* filled DEM / excess topography */
  int32_t *spill;         /* fillsinks: spill index, item-local (as tt_fillsinks_spill_dev) */
  const float *dem;
  const uint8_t *bc;      /* fillsinks: boundary cells */
  const float *slopes;    /* excess topography: threshold slopes */
  const float *g_out;     /* adjoint seed dL/doutput (unused when forward only) */
  float *g_dem;           /* dL/ddem */
  float *g_slopes;        /* excess topography: dL/dthreshold_slopes */
} tt_batch_buffers;

/* Builds the plan of a batch for tiles of qx x qy 64x64 quadrants
   (qx * qy <= 2). With plan == NULL only *plan_bytes and *workspace_bytes
   are set. Host-only: no GPU call. */
int tt_batch_plan(int op, const tt_batch_item *items, int nitems, int qx, int qy, void *plan,
                  size_t *plan_bytes, size_t *workspace_bytes);

/* Per-item sweep caps (caps[b] caps the sweeps of one visit of item b's
   tiles; 0 = the cap passed to tt_batch_run_dev), written into a host plan
   made by tt_batch_plan, which the caller then copies to the device again.
   An autotuner sets them per item size class (small items converge in one
   visit, large ones pass progress to their neighbours sooner with a low
   cap). Like the launch's cap they change the schedule, never the result.
   Host-only. */
int tt_batch_plan_set_sweep_caps(void *plan, int nitems, const uint32_t *caps);

/* Runs a planned batch. plan_host and plan_dev hold the same plan bytes
   (host and device copies); `workspace` is device memory of at least
   workspace_bytes. max_local caps the sweeps of one tile visit (0 = 65536;
   a capped visit queues its tile again, so the cap changes the schedule,
   never the result). Results equal the single-grid entry points' bit for
   bit (fillsinks, spill, the fillsinks adjoint) or to the same FP32
   rounding behaviour (excess topography and its adjoint). */
int tt_batch_run_dev(const void *plan_host, const void *plan_dev,
                     const tt_batch_buffers *buffers, void *workspace, size_t workspace_bytes,
                     unsigned max_local, int flags, void *stream, tt_stats *stats);

/* Tile shape used by subsequent calls on this thread, in 64x64 quadrants
   (qx, qy in {1, 2}); returns nonzero for an unsupported shape. Until it is
   called, fillsinks, reconstruct and the spill forward use 2x2 and excess
   topography 2x1 (the fastest measured at 8192^2). */
int tt_set_tile_shape(int qx, int qy);

/* Work-queue schedule of single-grid sink filling (tt_fillsinks_dev,
   fillsinks, fillsinks_hybrid, the spill forward's levels) and
   reconstruction (tt_reconstruct_dev, reconstruct, reconstruct_hybrid)
   for subsequent calls on this thread; returns nonzero for an unknown
   value. AUTO (the default) visits queued fill tiles lowest pending level
   first when tiles outnumber the persistent workgroups at least four to
   one and there are at most 512 workgroups, and in FIFO order otherwise;
   reconstruction runs FIFO under AUTO, since lowest-first pays off only
   for sparse markers (28% faster at 8192^2 for a marker set on the edge
   alone, 15-20% slower for a marker everywhere) and the marker is not
   seen before the launch. FIFO and PRIORITY force one queue. Every
   schedule gives the same, bit-exact result. No counterpart in the
   reference (a GPU scheduling choice). */
#define TT_FILL_SCHEDULE_AUTO 0
#define TT_FILL_SCHEDULE_FIFO 1
#define TT_FILL_SCHEDULE_PRIORITY 2
int tt_set_fill_schedule(int schedule);

/* Device errors of asynchronous calls. A call made without `stats` returns
   before its kernels end, so a device error (a bounded queue wait that
   expired: bit 1) cannot be in its return code. Every persistent launch
   also ORs its error bits into a word per GPU, which this call reads for
   the current GPU after synchronising the device (clearing it when
   `clear` is nonzero). No counterpart in the reference, whose functions
   are synchronous. Returns 0 and stores the bits in *bits. The host-pointer
   drop-ins check (and clear) the word themselves before returning. */
int tt_device_error(int clear, uint32_t *bits);

/* Upper bound, in ticks of the 100 MHz wall clock, on one bounded device
   wait of later launches (process-wide; 0 restores the default of 10 s).
   No counterpart in the reference: for tests of the error path, which set
   a tiny bound so that a wait gives up and the launch ends early with its
   error bit set. Always returns 0. */
int tt_set_wait_limit(uint64_t ticks);

/* Frees the device memory the host-pointer drop-ins keep between calls
   (one grow-only buffer per GPU, sized by the largest call so far). No
   counterpart in the reference. */
void tt_host_workspace_release(void);

/* Hash of the sources the library was built from (the Python loader
   compares it with the sources beside it and refuses a stale binary). No
   counterpart in the reference. */
const char *tt_build_hash(void);

/* Message of the last failure on this thread ("" if none). */
const char *tt_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TOPOTOOLBOX_AMD_H */
