// Batched forward (+ adjoint) over many grids of uneven shapes in one
// persistent launch (no reference counterpart: the reference processes one
// grid per call).
//
// Every grid ("item") is cut into tiles; the tiles of all items share one
// work queue (tile_engine.h), so small and large items interleave and the
// whole batch keeps every CU busy until the last tile. Each item walks
// through phases; its next phase is queued, all its tiles at once, when the
// last tile of the current phase is released (a per-item pending count):
//
//   fillsinks:        0 fill levels (fill_visit, kModeFill)
//                     1 spill indices by labelling each tile's submerged
//                       components (spill_ccl.h, tile by tile)
//                     2 spill relaxation from the dirty tiles (fill_visit,
//                       kModeSpill): components across tile edges, flats
//                     3 accumulator := 0; the item's max |g_out| (the
//                       fixed-point scale) and whether g_out has NaN / INF
//                     4 scatter of g_out onto spill cells
//                     5 g_dem from the accumulator
//                     6 NaN / INF entries onto g_dem (queued only for
//                       items that have any)
//   excesstopography: 0 relaxation of u (fsm_visit)
//                     1 linearisation (sx, sy, dfic)
//                     2 relaxation of lambda (adj_visit)
//                     3 g_dem, g_slopes
//
// Forward-only runs stop after the forward phases (fillsinks: 2,
// excess topography: 0). The launch allocates nothing and
// never synchronises: the plan (item table) and the workspace come from
// the caller, so the call can be captured in a graph.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <utility>
#include <vector>

#include "adjoint_visit.h"
#include "fsm_visit.h"
#include "spill_ccl.h"
#include "workspace.h"

namespace tt {

// ---------------------------------------------------------------- device kernels
template <class G, int OP>
struct BatchLds;
template <class G>
struct BatchLds<G, kBatchFill> {
  union {
    FillLds<G, kModeFill> level;
    FillLds<G, kModeSpill> fill;
    SpillLabelLds<G::TX, G::TY> label;
    struct {
      int keys[1024];
      unsigned long long vals[1024];
    } hash;
  };
};
template <class G>
struct BatchLds<G, kBatchFsm> {
  union {
    FsmLds<G> fsm;
    AdjLds<G> adj;
  };
};

struct BatchArgs {
  const BatchItemDev* items;
  const unsigned* tile_item;
  unsigned* item_phase;
  unsigned* item_pending;
  unsigned* item_gmax;       // fillsinks adjoint: max |g_out| (float bits)
  unsigned* item_nonfinite;  // fillsinks adjoint: g_out has NaN / INF
  int nitems, ntiles, last_phase;
  BatchBuffers buf;
  unsigned long long* acc;  // fillsinks: the adjoint's accumulator
  unsigned* dirty;          // fillsinks: per tile, relax on the first spill visit
  float* sx;                // excesstopography: linearisation and lambda
  float* sy;
  float* dfic;
  float* lam;
  TileQueue q;
  unsigned max_local;
};

struct BatchTopo {
  const BatchItemDev* items;
  const unsigned* tile_item;
  unsigned* item_phase;
  unsigned* item_pending;
  const unsigned* item_nonfinite;  // non-null: the last phase is optional
  int last_phase;

  __device__ int neighbour(int t, int k) const {
    const BatchItemDev& it = items[tile_item[t]];
    const int lt = t - it.tile_base;
    const int x = lt % it.ntx + kNbrDx[k], y = lt / it.ntx + kNbrDy[k];
    return (x < 0 || x >= it.ntx || y < 0 || y >= it.nty) ? -1 : it.tile_base + y * it.ntx + x;
  }
  __device__ void on_new(int t) const {
    __hip_atomic_fetch_add(&item_pending[tile_item[t]], 1u, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
  }
  // Thread 0, tile t went idle. When it was the item's last queued or
  // running tile, the item's phase is complete: advance the phase, count
  // all the item's tiles in again and reserve their ring positions; the
  // workgroup then queues them in expand(). Token: pos << 32 | item.
  __device__ unsigned long long on_release(const TileQueue& q, int t) const {
    const unsigned b = tile_item[t];
    if (__hip_atomic_fetch_sub(&item_pending[b], 1u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT) != 1u)
      return kNoExpand;
    // every tile of the phase released its writes before its decrement
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const unsigned ph = ld_relaxed(&item_phase[b]);
    if ((int)ph >= last_phase) return kNoExpand;
    if (item_nonfinite && (int)ph == last_phase - 1 && ld_relaxed(&item_nonfinite[b]) == 0u)
      return kNoExpand;
    const unsigned n = (unsigned)items[b].ntiles;
    __hip_atomic_store(&item_phase[b], ph + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&item_pending[b], n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(&q.sync->pending, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned pos =
        __hip_atomic_fetch_add(&q.sync->tail, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return ((unsigned long long)pos << 32) | b;
  }
  // All threads: queue every tile of the item (all idle: no tile of the
  // item runs, so nothing else touches their state words).
  __device__ void expand(const TileQueue& q, unsigned long long token) const {
    const BatchItemDev& it = items[(unsigned)token];
    const unsigned pos = (unsigned)(token >> 32);
    for (int k = threadIdx.x; k < it.ntiles; k += blockDim.x) {
      __hip_atomic_store(&q.state[it.tile_base + k], kQueued, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&q.visited[it.tile_base + k], 0u, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    // the reset flags and the new phase (thread 0) precede every slot write
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    for (int k = threadIdx.x; k < it.ntiles; k += blockDim.x)
      fill_slot(q, pos + (unsigned)k, it.tile_base + k);
    __syncthreads();
  }
};

// Scatter of the finite g_out of one tile's cells onto their spill cells
// other than themselves, pre-summed per spill cell in an LDS hash table
// (fixed point with scale 2^e: order-independent); a cell that is its own
// spill cell adds its term when g_dem is formed.
template <class G>
__device__ void scatter_tile(unsigned long long* acc, const int* spill, const float* g, int e,
                             long long nx, long long ny, long long x0, long long y0,
                             int* keys, unsigned long long* vals) {
  constexpr int kSlots = 1024;
  const int tid = threadIdx.x;
  for (int k = tid; k < kSlots; k += G::THREADS) {
    keys[k] = -1;
    vals[k] = 0ull;
  }
  __syncthreads();
  for (int idx = tid; idx < G::TY * G::TX; idx += G::THREADS) {
    const long long gx = x0 + idx % G::TX, gy = y0 + idx / G::TX;
    if (gx >= nx || gy >= ny) continue;
    const size_t p = (size_t)gy * nx + gx;
    const int s = spill[p];
    if (s < 0 || s == (int)p) continue;
    const unsigned long long v = to_fixed(g[p], e);
    if (v == 0ull) continue;
    unsigned h = ((unsigned)s * 2654435761u) >> 22;
    bool done = false;
    for (int probe = 0; probe < 16; probe++) {
      const int prev = atomicCAS(&keys[h], -1, s);
      if (prev == -1 || prev == s) {
        atomicAdd(&vals[h], v);
        done = true;
        break;
      }
      h = (h + 1) & (kSlots - 1);
    }
    if (!done) atomicAdd(&acc[s], v);
  }
  __syncthreads();
  for (int k = tid; k < kSlots; k += G::THREADS)
    if (keys[k] >= 0) atomicAdd(&acc[keys[k]], vals[k]);
}

// The spill relaxation's visit as a call of its own. Inlined next to the
// levels' visit, the two made the kernel spill to scratch (their registers
// are allocated together); each alone does not. The call takes its LDS and
// global pointers back into their address spaces, so that its accesses
// stay LDS and global instructions rather than flat ones.
template <class G>
__device__ __attribute__((noinline)) unsigned batch_spill_visit(FillView v, long long x0,
                                                                long long y0, bool first,
                                                                unsigned nvis, unsigned ml,
                                                                Sync* sink,
                                                                FillLds<G, kModeSpill>& lds) {
  using LdsT = FillLds<G, kModeSpill>;
  typedef __attribute__((address_space(3))) char lds_char;
  typedef __attribute__((address_space(1))) char glb_char;
  LdsT& l = *reinterpret_cast<LdsT*>((char*)(lds_char*)(char*)&lds);
  v.state = reinterpret_cast<float*>((char*)(glb_char*)(char*)v.state);
  v.spill = reinterpret_cast<int*>((char*)(glb_char*)(char*)v.spill);
  v.src = reinterpret_cast<const float*>((char*)(glb_char*)(char*)v.src);
  v.bc = reinterpret_cast<const uint8_t*>((char*)(glb_char*)(char*)v.bc);
  return fill_visit<G, kModeSpill>(v, x0, y0, first, nvis, ml, sink, l);
}

template <class G, int OP>
__global__ __launch_bounds__(G::THREADS) void batch_persistent(BatchArgs a) {
  __shared__ BatchLds<G, OP> lds;
  __shared__ unsigned tile_gmax, tile_nonfinite;
  const BatchTopo topo{a.items, a.tile_item, a.item_phase, a.item_pending, a.item_nonfinite,
                       a.last_phase};
  run_queue(a.q, topo, [&](int t, bool first, unsigned nvis) -> unsigned {
    const unsigned b = a.tile_item[t];
    const BatchItemDev it = a.items[b];
    const int lt = t - it.tile_base;
    const long long x0 = (long long)(lt % it.ntx) * G::TX, y0 = (long long)(lt / it.ntx) * G::TY;
    const long long nx = it.nx, ny = it.ny;
    const size_t off = (size_t)it.offset;
    const unsigned ph = ld_relaxed(&a.item_phase[b]);
    const unsigned ml = it.max_local ? it.max_local : a.max_local;
    // per-cell phases walk the tile's cells directly
    auto for_cells = [&](auto fn) {
      for (int idx = threadIdx.x; idx < G::TY * G::TX; idx += G::THREADS) {
        const long long gx = x0 + idx % G::TX, gy = y0 + idx / G::TX;
        if (gx < nx && gy < ny) fn((size_t)gy * nx + gx);
      }
    };
    if constexpr (OP == kBatchFill) {
      unsigned long long* acc = a.acc + off;
      const FillView v{a.buf.output + off, a.buf.spill + off, a.buf.dem + off, a.buf.bc + off,
                       nx, ny};
      if (ph == 0)
        return fill_visit<G, kModeFill>(v, x0, y0, first, nvis, ml, a.q.stats_sink(), lds.level);
      if (ph == 1) {
        // tile-by-tile labelling (label tile = relaxation tile): each tile
        // writes its dirty flag
        const SpillView sv{a.buf.output + off, a.buf.dem + off, a.buf.bc + off,
                           a.buf.spill + off, nullptr, a.dirty + it.tile_base, nx, ny,
                           G::TX, G::TY, it.ntx};
        spill_label_tile<G::TX, G::TY, G::THREADS, false>(sv, x0, y0, lds.label);
        return 0u;
      }
      if (ph == 2) {
        if (spill_skip(a.dirty, t, first, &lds.fill.bits)) return 0u;
        const unsigned bits =
            batch_spill_visit<G>(v, x0, y0, first, nvis, ml, a.q.stats_sink(), lds.fill);
        spill_mark(a.dirty, topo, t, bits);
        return bits;
      }
      const int* spill = a.buf.spill + off;
      const float* g = a.buf.g_out + off;
      const unsigned long long ncell = (unsigned long long)nx * (unsigned long long)ny;
      float* g_dem = a.buf.g_dem + off;
      if (ph == 3) {
        if (threadIdx.x == 0) tile_gmax = tile_nonfinite = 0u;
        __syncthreads();
        unsigned m = 0u, nf = 0u;
        for_cells([&](size_t p) {
          const float gp = g[p];
          m = max(m, abs_bits_if_finite(gp));
          nf |= isfinite(gp) ? 0u : 1u;
          acc[p] = 0ull;
        });
        if (m) atomicMax(&tile_gmax, m);
        if (nf) atomicOr(&tile_nonfinite, 1u);
        __syncthreads();
        if (threadIdx.x == 0) {
          if (tile_gmax) atomicMax(&a.item_gmax[b], tile_gmax);
          if (tile_nonfinite) atomicOr(&a.item_nonfinite[b], 1u);
        }
      } else if (ph == 4) {
        const int e = fix_exponent(ld_relaxed(&a.item_gmax[b]), ncell);
        scatter_tile<G>(acc, spill, g, e, nx, ny, x0, y0, lds.hash.keys, lds.hash.vals);
      } else if (ph == 5) {
        const int e = fix_exponent(ld_relaxed(&a.item_gmax[b]), ncell);
        for_cells([&](size_t p) {
          g_dem[p] = from_fixed(acc[p] + (spill[p] == (int)p ? to_fixed(g[p], e) : 0ull), e);
        });
      } else {
        for_cells([&](size_t p) {
          const float gp = g[p];
          const int s = spill[p];
          if (!isfinite(gp) && s >= 0) combine_nonfinite(g_dem, (size_t)s, gp);
        });
      }
      return 0u;
    } else {
      if (ph == 0) {
        const FsmView v{a.buf.output + off, a.buf.dem + off, a.buf.slopes + off, it.cellsize,
                        nx, ny};
        return fsm_visit<G>(v, x0, y0, first, nvis, ml, a.q.stats_sink(), lds.fsm);
      }
      if (ph == 1) {
        for_cells([&](size_t p) {
          fsm_linearize_cell(p, a.sx + off, a.sy + off, a.dfic + off, a.buf.output + off,
                             a.buf.dem + off, a.buf.slopes + off, it.cellsize, nx, ny);
        });
        return 0u;
      }
      if (ph == 2) {
        const AdjView v{a.lam + off, a.buf.g_out + off, a.sx + off, a.sy + off, nx, ny};
        return adj_visit<G>(v, x0, y0, first, nvis, ml, a.q.stats_sink(), lds.adj);
      }
      for_cells([&](size_t p) {
        fsm_adjoint_finish_cell(p, a.buf.g_dem + off, a.buf.g_slopes + off, a.lam + off,
                                a.dfic + off);
      });
      return 0u;
    }
  });
}

__global__ void batch_init(TileQueue q, const BatchItemDev* items, int nitems,
                           unsigned* item_phase, unsigned* item_pending, unsigned* item_gmax,
                           unsigned* item_nonfinite) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    unsigned* w = reinterpret_cast<unsigned*>(q.sync);
    for (unsigned k = 0; k < sizeof(Sync) / 4; k++) w[k] = 0u;
    q.sync->tail = (unsigned)q.ntiles;
    q.sync->pending = (unsigned)q.ntiles;
  }
  const int stride = gridDim.x * blockDim.x;
  for (int k = i; k < q.ntiles; k += stride) {
    q.ring[k] = ((unsigned long long)k << 32) | ((unsigned)k + 1u);
    q.state[k] = kQueued;
    q.visited[k] = 0u;
  }
  for (int b = i; b < nitems; b += stride) {
    item_phase[b] = 0u;
    item_pending[b] = (unsigned)items[b].ntiles;
    item_gmax[b] = 0u;
    item_nonfinite[b] = 0u;
  }
}

// ---------------------------------------------------------------- host side
static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static bool batch_shape_ok(TileShape ts) {
  // 64-bit spill keys and the adjoint's four arrays: at most two quadrants
  return tile_shape_ok(ts) && ts.qx * ts.qy <= 2;
}

hipError_t batch_plan(int op, int nitems, const BatchItem* items, TileShape ts, void* plan,
                      size_t* plan_bytes, size_t* workspace_bytes) {
  if (nitems <= 0 || !items || !plan_bytes || !workspace_bytes) {
    set_error("tt_batch_plan: need nitems > 0, items and both size outputs");
    return hipErrorInvalidValue;
  }
  if (op != kBatchFill && op != kBatchFsm) {
    set_error("tt_batch_plan: unknown op %d", op);
    return hipErrorInvalidValue;
  }
  if (!batch_shape_ok(ts)) {
    set_error("tt_batch_plan: tile shape %dx%d quadrants not supported (1x1, 2x1, 1x2)", ts.qx,
              ts.qy);
    return hipErrorInvalidValue;
  }
  const int tx = 64 * ts.qx, ty = 64 * ts.qy;
  std::vector<BatchItemDev> dev(nitems);
  std::vector<std::pair<long long, long long>> ranges(nitems);
  long long ntiles = 0, cells_end = 0;
  for (int b = 0; b < nitems; b++) {
    const BatchItem& it = items[b];
    if (it.nx <= 0 || it.ny <= 0 || it.offset < 0 ||
        (long long)it.nx * it.ny > (long long)INT32_MAX) {
      set_error("tt_batch_plan: item %d has nx=%lld ny=%lld offset=%lld (need positive sizes, "
                "offset >= 0, nx*ny <= 2^31-1)",
                b, (long long)it.nx, (long long)it.ny, (long long)it.offset);
      return hipErrorInvalidValue;
    }
    BatchItemDev& d = dev[b];
    d.nx = it.nx;
    d.ny = it.ny;
    d.offset = it.offset;
    d.cellsize = it.cellsize;
    d.ntx = (int)((it.nx + tx - 1) / tx);
    d.nty = (int)((it.ny + ty - 1) / ty);
    d.ntiles = d.ntx * d.nty;
    d.tile_base = (int)ntiles;
    ntiles += d.ntiles;
    if (ntiles > INT32_MAX / 4) {
      set_error("tt_batch_plan: too many tiles");
      return hipErrorInvalidValue;
    }
    ranges[b] = {it.offset, it.offset + (long long)it.nx * it.ny};
    if (ranges[b].second > cells_end) cells_end = ranges[b].second;
  }
  std::sort(ranges.begin(), ranges.end());
  for (int b = 1; b < nitems; b++)
    if (ranges[b].first < ranges[b - 1].second) {
      set_error("tt_batch_plan: items overlap in the packed buffers");
      return hipErrorInvalidValue;
    }
  BatchPlanHeader h{};
  h.magic = kBatchPlanMagic;
  h.op = op;
  h.nitems = nitems;
  h.ntiles = (int)ntiles;
  h.qx = ts.qx;
  h.qy = ts.qy;
  h.cells = cells_end;
  h.items_off = align256(sizeof(BatchPlanHeader));
  h.tile_item_off = align256(h.items_off + sizeof(BatchItemDev) * nitems);
  h.bytes = h.tile_item_off + sizeof(unsigned) * ntiles;
  // workspace: control block + ring (2 words per tile) + state + visited +
  // dirty, item phase, pending count, max |g_out| and non-finite flag, then
  // per-cell scratch (fillsinks: the adjoint's 64-bit accumulator; excess
  // topography: sx, sy, dfic, lambda)
  h.ws_queue = 0;
  h.ws_items = align256((64 + 5 * (size_t)ntiles) * 4);
  h.ws_scratch = align256(h.ws_items + 16 * (size_t)nitems);
  h.ws_bytes = h.ws_scratch + (size_t)cells_end * (op == kBatchFill ? 8 : 16);
  *plan_bytes = h.bytes;
  *workspace_bytes = h.ws_bytes;
  if (!plan) return hipSuccess;
  char* p = static_cast<char*>(plan);
  memset(p, 0, h.bytes);
  memcpy(p, &h, sizeof h);
  memcpy(p + h.items_off, dev.data(), sizeof(BatchItemDev) * nitems);
  unsigned* ti = reinterpret_cast<unsigned*>(p + h.tile_item_off);
  for (int b = 0; b < nitems; b++)
    for (int k = 0; k < dev[b].ntiles; k++) ti[dev[b].tile_base + k] = (unsigned)b;
  return hipSuccess;
}

hipError_t batch_plan_set_sweep_caps(void* plan, int nitems, const unsigned* caps) {
  if (!plan || !caps) {
    set_error("tt_batch_plan_set_sweep_caps: plan and caps are required");
    return hipErrorInvalidValue;
  }
  BatchPlanHeader h;
  memcpy(&h, plan, sizeof h);
  if (h.magic != kBatchPlanMagic || h.nitems != nitems) {
    set_error("tt_batch_plan_set_sweep_caps: not a plan of %d items made by tt_batch_plan",
              nitems);
    return hipErrorInvalidValue;
  }
  char* p = static_cast<char*>(plan) + h.items_off;
  for (int b = 0; b < nitems; b++) {
    BatchItemDev d;
    memcpy(&d, p + sizeof(BatchItemDev) * b, sizeof d);
    d.max_local = caps[b];
    memcpy(p + sizeof(BatchItemDev) * b, &d, sizeof d);
  }
  return hipSuccess;
}

template <class G, int OP>
static hipError_t launch_batch(BatchArgs a, hipStream_t stream, RunStats* stats) {
  // the occupancy query is host-side only (capturable); cache it per kernel
  static thread_local int cached = 0;
  if (cached == 0) {
    hipError_t e = persistent_grid((const void*)batch_persistent<G, OP>, G::THREADS, INT32_MAX,
                                   &cached);
    if (e != hipSuccess) return e;
  }
  const int grid = cached < a.ntiles ? cached : a.ntiles;
  int blocks = (a.ntiles + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  batch_init<<<blocks, 256, 0, stream>>>(a.q, a.items, a.nitems, a.item_phase, a.item_pending,
                                         a.item_gmax, a.item_gmax + a.nitems);
  batch_persistent<G, OP><<<grid, G::THREADS, 0, stream>>>(a);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && stats) {
    Sync h;
    e = hipMemcpyAsync(&h, a.q.sync, sizeof(Sync), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e == hipSuccess) {
      stats->requeues = (int64_t)h.tail - a.ntiles;
      stats->visits = (int64_t)h.visits;
      stats->local_iters = (int64_t)h.local_iters;
      stats->max_local = h.max_local;
      stats->error = h.error;
      stats->t_stage = (int64_t)h.t_stage;
      stats->t_relax = (int64_t)h.t_relax;
      stats->t_write = (int64_t)h.t_write;
      stats->t_wait = (int64_t)h.t_wait;
      stats->grid = grid;
      stats->tile_x = G::TX;
      stats->tile_y = G::TY;
    }
  }
  return e;
}

template <int OP>
static hipError_t dispatch_batch(BatchArgs a, TileShape ts, hipStream_t s, RunStats* st) {
  if (ts.qx == 2 && ts.qy == 1) return launch_batch<TileGeom<2, 1>, OP>(a, s, st);
  if (ts.qx == 1 && ts.qy == 2) return launch_batch<TileGeom<1, 2>, OP>(a, s, st);
  return launch_batch<TileGeom<1, 1>, OP>(a, s, st);
}

hipError_t run_batch(const void* plan_host, const void* plan_dev, const BatchBuffers& buf,
                     void* workspace, size_t workspace_bytes, unsigned max_local, int flags,
                     hipStream_t stream, RunStats* stats) {
  if (!plan_host || !plan_dev || !workspace) {
    set_error("tt_batch_run_dev: plan_host, plan_dev and workspace are required");
    return hipErrorInvalidValue;
  }
  BatchPlanHeader h;
  memcpy(&h, plan_host, sizeof h);
  if (h.magic != kBatchPlanMagic || !batch_shape_ok(TileShape{h.qx, h.qy})) {
    set_error("tt_batch_run_dev: plan_host is not a plan made by tt_batch_plan");
    return hipErrorInvalidValue;
  }
  if (workspace_bytes < h.ws_bytes) {
    set_error("tt_batch_run_dev: workspace of %zu bytes, the plan needs %zu", workspace_bytes,
              h.ws_bytes);
    return hipErrorInvalidValue;
  }
  const bool fwd_only = (flags & kBatchForwardOnly) != 0;
  bool ok = buf.output && buf.dem;
  if (h.op == kBatchFill) ok = ok && buf.spill && buf.bc;
  else ok = ok && buf.slopes;
  if (!fwd_only) ok = ok && buf.g_out && buf.g_dem && (h.op == kBatchFill || buf.g_slopes);
  if (!ok) {
    set_error("tt_batch_run_dev: a buffer the %s %s needs is NULL",
              h.op == kBatchFill ? "fillsinks" : "excess topography",
              fwd_only ? "forward" : "forward + adjoint");
    return hipErrorInvalidValue;
  }
  const char* pd = static_cast<const char*>(plan_dev);
  char* ws = static_cast<char*>(workspace);
  BatchArgs a{};
  a.items = reinterpret_cast<const BatchItemDev*>(pd + h.items_off);
  a.tile_item = reinterpret_cast<const unsigned*>(pd + h.tile_item_off);
  a.item_phase = reinterpret_cast<unsigned*>(ws + h.ws_items);
  a.item_pending = a.item_phase + h.nitems;
  a.item_gmax = a.item_pending + h.nitems;
  a.item_nonfinite = (h.op == kBatchFill && !fwd_only) ? a.item_gmax + h.nitems : nullptr;
  a.nitems = h.nitems;
  a.ntiles = h.ntiles;
  a.last_phase = (h.op == kBatchFill) ? (fwd_only ? 2 : 6) : (fwd_only ? 0 : 3);
  a.buf = buf;
  a.max_local = max_local ? max_local : (1u << 16);
  a.q.sync = reinterpret_cast<Sync*>(ws + h.ws_queue);
  a.q.ring = reinterpret_cast<unsigned long long*>(reinterpret_cast<unsigned*>(a.q.sync) + 64);
  a.q.state = reinterpret_cast<unsigned*>(a.q.ring + h.ntiles);
  a.q.visited = a.q.state + h.ntiles;
  a.dirty = a.q.visited + h.ntiles;
  a.q.ntiles = h.ntiles;
  a.q.wait_limit = wait_limit_ticks();
  {
    const hipError_t se = sticky_error_word(&a.q.sticky);
    if (se != hipSuccess) return se;
  }
  a.q.stats = stats ? 1 : 0;
  const size_t cells = (size_t)h.cells;
  if (h.op == kBatchFill) {
    a.acc = reinterpret_cast<unsigned long long*>(ws + h.ws_scratch);
    return dispatch_batch<kBatchFill>(a, TileShape{h.qx, h.qy}, stream, stats);
  }
  a.sx = reinterpret_cast<float*>(ws + h.ws_scratch);
  a.sy = a.sx + cells;
  a.dfic = a.sy + cells;
  a.lam = a.dfic + cells;
  return dispatch_batch<kBatchFsm>(a, TileShape{h.qx, h.qy}, stream, stats);
}

}  // namespace tt
