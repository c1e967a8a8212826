// Excess topography with layered lithology (the reference's
// excesstopography_fmm3d, src/excesstopography.c:174-262) by fast sweeping
// on the persistent tile engine of fsm_visit.h.
//
// Each cell has nlayers layer tops (lithstack, cell-major: top of layer l of
// cell p at lithstack[p * nlayers + l], bottom up) and each layer one
// threshold slope. A cell's proposal is the Godunov update of the lowest
// layer whose proposal lies below that layer's top at the cell, with fi =
// cellsize * slope of that layer (src/excesstopography.c:204-214); no
// proposal if no layer qualifies. As in the reference, a cell accepts a
// proposal only when it is strictly lower than its current value, starting
// from the DEM, and NaN cells never change.
//
// The layer choice is not monotone in the neighbours (a lower neighbour
// can make a lower layer qualify, whose slope may be steeper), so the
// relaxation's fixed point depends on the update order: the reference's
// binary-heap order and this tiled sweeping order can reach different
// surfaces. Both satisfy excess <= dem; with one layer (or equal slopes)
// the rule is the plain excess topography and the result is that of
// excesstopography_fsm2d.
#include "fsm_visit.h"
#include "workspace.h"

namespace tt {

namespace {

constexpr int kMaxLayers = 1024;

struct Fmm3dView {
  float* u;
  const float* dem;
  const float* lith;    // nx * ny * nlayers layer tops, cell-major
  const float* slopes;  // nlayers threshold slopes
  float cellsize;
  long long nx, ny;
  int nlayers;
};

// Proposal of tile cell (r, c): the first layer, bottom up, whose Godunov
// proposal lies below the layer's top at the cell (NaN: none, never
// accepted). Cells beyond the grid have none.
struct LayerProp {
  const float* lith;
  const float* fi;  // LDS: cellsize * slope per layer
  long long x0, y0, nx, ny;
  int nl;
  __device__ __forceinline__ float operator()(int r, int c, float u1, float u2) const {
    const long long gx = x0 + c, gy = y0 + r;
    if (gx >= nx || gy >= ny) return __builtin_nanf("");
    const float* top = lith + ((size_t)gy * (size_t)nx + (size_t)gx) * (size_t)nl;
    for (int l = 0; l < nl; l++) {
      const float p = godunov(u1, u2, fi[l]);
      if (p < top[l]) return p;
    }
    return __builtin_nanf("");
  }
};

template <class G>
struct Fmm3dLds {
  float su[G::CELLS];
  float fi[kMaxLayers];
  unsigned sper[2 * (G::TX + G::TY)];
  unsigned bits;
};

template <class G>
__device__ inline unsigned fmm3d_visit(const Fmm3dView& v, long long x0, long long y0, bool first,
                                       unsigned nvis, unsigned max_local, Sync* sync,
                                       Fmm3dLds<G>& lds) {
  float* su = lds.su;
  const unsigned* su_bits = reinterpret_cast<const unsigned*>(su);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int R0 = (wave / G::QX) * 64;
  const int C0 = (wave % G::QX) * 64;
  const unsigned long long c0 = wall_clock64();
  // u from the DEM on the first visit, from the stored state afterwards
  // (+INF outside the grid)
  stage_tile<G>(StageF32{first ? v.dem : v.u, su, kInf}, StageF32{nullptr, nullptr, 0.f}, false,
                x0, y0, v.nx, v.ny);
  for (int l = tid; l < v.nlayers; l += G::THREADS) lds.fi[l] = v.cellsize * v.slopes[l];
  __syncthreads();
  for_stale_halo<G>(first, nvis, [&](int r, int c, bool from_state) {
    const long long gx = x0 + c, gy = y0 + r;
    if (gx < 0 || gx >= v.nx || gy < 0 || gy >= v.ny) return;
    const size_t g = (size_t)gy * v.nx + gx;
    su[G::at(r, c)] = from_state ? v.u[g] : v.dem[g];
  });
  __syncthreads();
  save_perimeter<G>(lds.sper, su_bits);
  const unsigned long long c1 = wall_clock64();
  int any = 0, any_change = 0;
  const LayerProp prop{v.lith, lds.fi, x0, y0, v.nx, v.ny, v.nlayers};
  const unsigned it = fsm_relax_regs<G>(su, prop, R0, C0, lane, max_local, &any, &any_change);
  const unsigned long long c2 = wall_clock64();
  if (first || any_change) write_tile_f32<G>(v.u, su, x0, y0, v.nx, v.ny);
  // every proposal through p is at least u_p (fi >= 0), so a neighbour cell
  // q can only be lowered by p if u_p < u_q
  unsigned bits = block_or(
      pushes_from_perimeter<G, false>(lds.sper, su_bits, [&](int p, int q) { return su[p] < su[q]; }),
      &lds.bits);
  if (any) bits |= kSelf;
  if (tid == 0) visit_stats(sync, it, c0, c1, c2);
  return bits;
}

struct Fmm3dArgs {
  Fmm3dView v;
  int ntx, nty;
  TileQueue q;
  unsigned max_local;
};

}  // namespace

template <class G>
__global__ __launch_bounds__(G::THREADS) void fmm3d_persistent(Fmm3dArgs a) {
  __shared__ Fmm3dLds<G> lds;
  run_queue(a.q, GridTopo{a.ntx, a.nty}, [&](int t, bool first, unsigned nvis) {
    const long long x0 = (long long)(t % a.ntx) * G::TX, y0 = (long long)(t / a.ntx) * G::TY;
    return fmm3d_visit<G>(a.v, x0, y0, first, nvis, a.max_local, a.q.stats_sink(), lds);
  });
}

hipError_t run_fmm3d(float* u, const float* dem, const float* lithstack, const float* slopes,
                     float cellsize, int64_t nx, int64_t ny, int64_t nlayers, hipStream_t stream,
                     RunStats* stats) {
  if (nx <= 0 || ny <= 0) return hipSuccess;
  if (nlayers < 1 || nlayers > kMaxLayers) return hipErrorInvalidValue;
  using G = TileGeom<2, 1>;
  Fmm3dArgs a{};
  a.v.u = u;
  a.v.dem = dem;
  a.v.lith = lithstack;
  a.v.slopes = slopes;
  a.v.cellsize = cellsize;
  a.v.nx = nx;
  a.v.ny = ny;
  a.v.nlayers = (int)nlayers;
  a.ntx = (int)((nx + G::TX - 1) / G::TX);
  a.nty = (int)((ny + G::TY - 1) / G::TY);
  a.max_local = 1u << 16;
  return launch_persistent(fmm3d_persistent<G>, G::THREADS, a.ntx * a.nty, a, stream, stats,
                           G::TX, G::TY, a.ntx, a.nty);
}

}  // namespace tt
