// Excess topography on one grid: the persistent work-queue kernel over
// fsm_visit (fsm_visit.h states the semantics).
#include "fsm_visit.h"
#include "workspace.h"

namespace tt {

struct FsmArgs {
  FsmView v;
  int ntx, nty;
  TileQueue q;
  unsigned max_local;
};

template <class G>
__global__ __launch_bounds__(G::THREADS) void fsm_persistent(FsmArgs a) {
  __shared__ FsmLds<G> lds;
  run_queue(a.q, GridTopo{a.ntx, a.nty}, [&](int t, bool first, unsigned nvis) {
    const long long x0 = (long long)(t % a.ntx) * G::TX, y0 = (long long)(t / a.ntx) * G::TY;
    return fsm_visit<G>(a.v, x0, y0, first, nvis, a.max_local, a.q.stats_sink(), lds);
  });
}

template <class G>
static hipError_t launch_fsm(FsmArgs a, hipStream_t stream, RunStats* stats) {
  a.ntx = (int)((a.v.nx + G::TX - 1) / G::TX);
  a.nty = (int)((a.v.ny + G::TY - 1) / G::TY);
  a.max_local = 1u << 16;
  return launch_persistent(fsm_persistent<G>, G::THREADS, a.ntx * a.nty, a, stream, stats, G::TX,
                           G::TY, a.ntx, a.nty);
}

hipError_t run_fsm(float* u, const float* dem, const float* slopes, float cellsize, int64_t nx,
                   int64_t ny, TileShape ts, hipStream_t stream, RunStats* stats) {
  if (nx <= 0 || ny <= 0) return hipSuccess;
  FsmArgs a{};
  a.v.u = u;
  a.v.dem = dem;
  a.v.slopes = slopes;
  a.v.cellsize = cellsize;
  a.v.nx = nx;
  a.v.ny = ny;
  if (ts.qx == 2 && ts.qy == 2) return launch_fsm<TileGeom<2, 2>>(a, stream, stats);
  if (ts.qx == 2 && ts.qy == 1) return launch_fsm<TileGeom<2, 1>>(a, stream, stats);
  if (ts.qx == 1 && ts.qy == 2) return launch_fsm<TileGeom<1, 2>>(a, stream, stats);
  return launch_fsm<TileGeom<1, 1>>(a, stream, stats);
}

}  // namespace tt
