// Single-pass stencils that consume a DEM next to the hot path: flat / sill
// / presill classification and the minimum / maximum value filters. Each
// output cell depends only on a fixed neighbourhood of the input, so these
// are one read and one write per cell at HBM rate: a workgroup stages a
// window (its tile plus the stencil's halo) into LDS with coalesced row
// loads, and every cell folds its neighbourhood out of LDS.
//
// Reference semantics:
//   identifyflats      src/identifyflats.c:32-121
//   min_filter,        src/value_filters.c:9-110: a sequential fold over
//   max_filter         the structuring element (slices, rows, columns in
//                      that order), the running value replaced unless it is
//                      strictly better, so the result's bits (e.g. -0 vs
//                      +0) follow the same order here
//   *_filter_square    src/value_filters.c:112-246: the same fold along
//                      dims[0] into tmp, then along dims[1]
//
// Cells outside the grid are staged as NaN: the filters skip NaN
// neighbours and out-of-grid neighbours alike, and identifyflats never
// matches a NaN, so one marker serves both.
#include "kernels.h"

namespace tt {

// Tiles are kTW cells wide (one column per lane and kCols columns per lane)
// and kTH rows tall; four waves share the rows.
constexpr int kCols = 4, kTW = 64 * kCols, kTH = 16, kWaves = 4, kThreads = 64 * kWaves;

__device__ __forceinline__ float nan_marker() { return __builtin_nanf(""); }

// A window of up to RPW rows per wave and kChunks x 64 columns, loaded
// into registers (all loads of a thread in flight together) and then
// written to LDS. The tiled kernels load the next tile's window while they
// compute the current one, so HBM latency overlaps the LDS work.
constexpr int kChunks = 5;

template <int RPW>
struct WindowLoader {
  float v[RPW][kChunks];

  // rows [wy0, wy0 + H) x columns [wx0, wx0 + W) of src; NaN outside the grid
  __device__ __forceinline__ void load(const float* src, long long nx, long long ny,
                                       long long wx0, long long wy0, int W, int H) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int u = 0; u < RPW; u++) {
      const int r = wave + kWaves * u;
      const long long gy = wy0 + r;
      const bool row_ok = r < H && gy >= 0 && gy < ny;
#pragma unroll
      for (int ch = 0; ch < kChunks; ch++) {
        const int c = lane + 64 * ch;
        const long long gx = wx0 + c;
        v[u][ch] = (row_ok && c < W && gx >= 0 && gx < nx) ? src[gy * nx + gx] : nan_marker();
      }
    }
  }

  // into LDS s with row pitch P
  __device__ __forceinline__ void store(float* s, int P, int W, int H) const {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int u = 0; u < RPW; u++) {
      const int r = wave + kWaves * u;
#pragma unroll
      for (int ch = 0; ch < kChunks; ch++) {
        const int c = lane + 64 * ch;
        if (r < H && c < W) s[r * P + c] = v[u][ch];
      }
    }
  }
};

// Tiles of the grid in row-major order, split into contiguous chunks per
// XCD (workgroups are dispatched round-robin over the 8 XCDs, each with
// its own L2): vertically and horizontally adjacent tiles, which share
// halo rows and cache lines, then tend to run on the same XCD.
struct TileWalk {
  long long ntx, ntiles, per_xcd, begin, end, stride;
  __device__ TileWalk(long long nx, long long ny) {
    ntx = (nx + kTW - 1) / kTW;
    ntiles = ntx * ((ny + kTH - 1) / kTH);
    const int xcds = gridDim.x >= 8 ? 8 : 1;
    const long long xcd = blockIdx.x % xcds, rank = blockIdx.x / xcds;
    const long long groups = gridDim.x / xcds;  // workgroups per XCD
    per_xcd = (ntiles + xcds - 1) / xcds;
    begin = xcd * per_xcd + rank;
    end = (xcd + 1) * per_xcd < ntiles ? (xcd + 1) * per_xcd : ntiles;
    stride = groups;
    // workgroups beyond groups * xcds (grid not a multiple of 8) idle
    if (rank >= groups) begin = end;
  }
};

// ---------------------------------------------------------------- identifyflats
// Window: the tile plus 2 cells on every side (a sill or presill of a tile
// cell depends on whether its neighbours are flats, which depends on their
// neighbours). Flat bits are computed for the tile plus 1 cell.
constexpr int kFlatW = kTW + 4, kFlatH = kTH + 4;
constexpr int kFlagW = kTW + 2, kFlagH = kTH + 2;
static_assert(kThreads == kTW, "identifyflats walks one tile column per thread");
static_assert(2 * kFlagH <= kThreads, "the side flag columns take one thread per cell");

__global__ __launch_bounds__(kThreads, 5) void identifyflats_kernel(
    int32_t* out, const float* dem, long long nx, long long ny, unsigned long long* partial) {
  __shared__ float sd[kFlatH * kFlatW];
  __shared__ unsigned char sflat[kFlagH * kFlagW];
  __shared__ unsigned long long wave_count[kWaves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const TileWalk w(nx, ny);
  unsigned long long flats = 0;
  WindowLoader<(kFlatH + kWaves - 1) / kWaves> win;
  if (w.begin < w.end)
    win.load(dem, nx, ny, (w.begin % w.ntx) * kTW - 2, (w.begin / w.ntx) * kTH - 2, kFlatW,
             kFlatH);
  for (long long t = w.begin; t < w.end; t += w.stride) {
    const long long x0 = (t % w.ntx) * kTW, y0 = (t / w.ntx) * kTH;
    win.store(sd, kFlatW, kFlatW, kFlatH);
    __syncthreads();
    const long long tn = t + w.stride;
    if (tn < w.end)
      win.load(dem, nx, ny, (tn % w.ntx) * kTW - 2, (tn / w.ntx) * kTH - 2, kFlatW, kFlatH);
    // flats: interior cells equal to the minimum of their 3x3 window, NaN
    // neighbours counting as -INF (a NaN cell is never equal to anything).
    // Thread c owns tile column c and walks the flag rows with a 3x3 window
    // in registers (3 LDS reads per cell); flag column fc is window column
    // fc + 1, flag row fr window row fr + 1.
    {
      const int c = threadIdx.x;  // kThreads == kTW: one column per thread
      const long long gx = x0 + c;
      const bool col_in = gx > 0 && gx < nx - 1;
      float t0 = sd[c + 1], t1 = sd[c + 2], t2 = sd[c + 3];
      float m0 = sd[kFlatW + c + 1], m1 = sd[kFlatW + c + 2], m2 = sd[kFlatW + c + 3];
      for (int fr = 0; fr < kFlagH; fr++) {
        const float* rb = sd + (fr + 2) * kFlatW + c + 1;
        const float b0 = rb[0], b1 = rb[1], b2 = rb[2];
        const long long gy = y0 - 1 + fr;
        bool flat = false;
        if (col_in && gy > 0 && gy < ny - 1) {
          float lo = m1;
          lo = fminf(lo, t0 != t0 ? -__builtin_inff() : t0);
          lo = fminf(lo, t1 != t1 ? -__builtin_inff() : t1);
          lo = fminf(lo, t2 != t2 ? -__builtin_inff() : t2);
          lo = fminf(lo, m0 != m0 ? -__builtin_inff() : m0);
          lo = fminf(lo, m2 != m2 ? -__builtin_inff() : m2);
          lo = fminf(lo, b0 != b0 ? -__builtin_inff() : b0);
          lo = fminf(lo, b1 != b1 ? -__builtin_inff() : b1);
          lo = fminf(lo, b2 != b2 ? -__builtin_inff() : b2);
          flat = m1 == lo;
        }
        sflat[fr * kFlagW + c + 1] = flat;
        t0 = m0; t1 = m1; t2 = m2;
        m0 = b0; m1 = b1; m2 = b2;
      }
      // the two flag columns beside the tile (x0 - 1 and x0 + kTW)
      if (c < 2 * kFlagH) {
        const int fr = c % kFlagH, fc = c < kFlagH ? 0 : kFlagW - 1;
        const long long hx = x0 - 1 + fc, hy = y0 - 1 + fr;
        bool flat = false;
        if (hx > 0 && hx < nx - 1 && hy > 0 && hy < ny - 1) {
          const float* q = sd + fr * kFlatW + fc;  // window cell (fr + 1, fc + 1) is the centre
          const float h = q[kFlatW + 1];
          float lo = h;
          for (int dr = 0; dr < 3; dr++)
            for (int dc = 0; dc < 3; dc++) {
              if (dr == 1 && dc == 1) continue;
              const float v = q[dr * kFlatW + dc];
              lo = fminf(lo, v != v ? -__builtin_inff() : v);
            }
          flat = h == lo;
        }
        sflat[fr * kFlagW + fc] = flat;
      }
    }
    __syncthreads();
    // sills and presills, thread c down tile column c with values and flags
    // of the 3x3 neighbourhood in registers
    {
      const int c = threadIdx.x;
      const long long gx = x0 + c;
      const bool col_in = gx < nx;
      const bool has_w = gx > 0, has_e = gx + 1 < nx;
      // rows r - 1 and r of the window (values) and of the flags
      float vt0 = sd[1 * kFlatW + c + 1], vt1 = sd[1 * kFlatW + c + 2], vt2 = sd[1 * kFlatW + c + 3];
      float vm0 = sd[2 * kFlatW + c + 1], vm1 = sd[2 * kFlatW + c + 2], vm2 = sd[2 * kFlatW + c + 3];
      bool ft0 = sflat[c], ft1 = sflat[c + 1], ft2 = sflat[c + 2];
      bool fm0 = sflat[kFlagW + c], fm1 = sflat[kFlagW + c + 1], fm2 = sflat[kFlagW + c + 2];
      for (int r = 0; r < kTH; r++) {
        const float* vb = sd + (r + 3) * kFlatW + c + 1;
        const unsigned char* fb = sflat + (r + 2) * kFlagW + c;
        const float vb0 = vb[0], vb1 = vb[1], vb2 = vb[2];
        const bool fb0 = fb[0], fb1 = fb[1], fb2 = fb[2];
        const long long gy = y0 + r;
        if (col_in && gy < ny) {
          const bool has_n = gy > 0, has_s = gy + 1 < ny;
          const float h = vm1;
          const bool flat = fm1;
          int32_t bits = flat;
          // a neighbour outside the grid is NaN in the window: never equal
          auto look = [&](bool inside, float v, bool qflat) {
            if (!inside || v != h) return;
            if (!flat && qflat) bits |= 2;  // sill: touches an equal flat
            if (flat && !qflat) bits |= 4;  // presill: a flat touching an equal sill
          };
          look(has_n && has_w, vt0, ft0);
          look(has_n, vt1, ft1);
          look(has_n && has_e, vt2, ft2);
          look(has_w, vm0, fm0);
          look(has_e, vm2, fm2);
          look(has_s && has_w, vb0, fb0);
          look(has_s, vb1, fb1);
          look(has_s && has_e, vb2, fb2);
          out[gy * nx + gx] = bits;
          flats += flat;
        }
        vt0 = vm0; vt1 = vm1; vt2 = vm2;
        vm0 = vb0; vm1 = vb1; vm2 = vb2;
        ft0 = fm0; ft1 = fm1; ft2 = fm2;
        fm0 = fb0; fm1 = fb1; fm2 = fb2;
      }
    }
    __syncthreads();  // the next tile's store overwrites sd and sflat
  }
  if (partial) {
    for (int o = 32; o > 0; o >>= 1) flats += __shfl_xor(flats, o);
    if (lane == 0) wave_count[wave] = flats;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long s = 0;
      for (int i = 0; i < kWaves; i++) s += wave_count[i];
      partial[blockIdx.x] = s;
    }
  }
}

// identifyflats streaming through registers (nx % 4 == 0, 16-byte aligned
// arrays): each lane owns four columns, and each wavefront walks a strip of
// kStripF rows down the grid with a rolling window. It keeps the values of
// the last four rows (columns -2 .. 5: side columns by DPP, the
// wavefront's outside ones by a load in its edge lanes), the order keys of
// the last three and the flat bits of the last three, with the loads of
// the next kAheadF rows in flight. A row's flat bits need its 3x3 windows
// and a row's flags the flat bits of the rows around it, so the flags
// trail the loads by two rows; the rules are the tiled kernel's. Each row
// is read once per strip, plus four halo rows per strip.
constexpr int kStripF = 32, kAheadF = 4;

struct RawRowF {
  float4 c;
  float2 lw, rw;
};

__device__ __forceinline__ RawRowF load_row_f(const float* dem, long long nx, long long ny,
                                              long long x, long long y, int lane, bool xin) {
  const float nan = __builtin_nanf("");
  RawRowF r;
  r.c = make_float4(nan, nan, nan, nan);
  r.lw = make_float2(nan, nan);
  r.rw = make_float2(nan, nan);
  const bool yin = y >= 0 && y < ny;
  if (xin && yin) r.c = *reinterpret_cast<const float4*>(dem + y * nx + x);
  if (lane == 0 && yin && x >= 2 && x - 2 < nx)
    r.lw = *reinterpret_cast<const float2*>(dem + y * nx + x - 2);
  if (lane == 63 && yin && x + 4 < nx)
    r.rw = *reinterpret_cast<const float2*>(dem + y * nx + x + 4);
  return r;
}

// Order-preserving integer key of a value, NaN as -INF and -0 as +0 (the
// float comparison finds -0 and +0 equal).
__device__ __forceinline__ unsigned flat_key(float q) {
  const unsigned u = __float_as_uint(q + 0.0f);  // -0 + 0 = +0
  return q != q ? 0x007fffffu : (u ^ ((u >> 31) ? 0xffffffffu : 0x80000000u));
}

__global__ __launch_bounds__(256) void identifyflats_stream(int32_t* out, const float* dem,
                                                            long long nx, long long ny,
                                                            unsigned long long* partial) {
  __shared__ unsigned long long wave_count[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // per-XCD runs of strips, as in value_filter_3x3
  const long long gxb = gridDim.x, total = gxb * gridDim.y;
  long long L = (long long)blockIdx.y * gxb + blockIdx.x;
  const long long block_id = L;
  if ((total & 7) == 0) L = (L & 7) * (total >> 3) + (L >> 3);
  const long long bx = L % gxb, by = L / gxb;
  const long long x = (bx * 64 + lane) * 4;
  const long long ys = (by * 4 + wave) * kStripF;
  const bool xin = x < nx;
  const float nan = __builtin_nanf("");
  // interior columns of the six flat columns (bit fc: x - 1 + fc)
  unsigned colin = 0u;
#pragma unroll
  for (int fc = 0; fc < 6; fc++) colin |= (unsigned)(x - 1 + fc > 0 && x - 1 + fc < nx - 1) << fc;
  auto shr = [](float old, float src) {  // lane i <- lane i - 1
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), 0x138, 0xf, 0xf, false));
  };
  auto shl = [](float old, float src) {  // lane i <- lane i + 1
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), 0x130, 0xf, 0xf, false));
  };
  float V[4][8];     // values of rows y - 3 .. y (V[3] newest)
  unsigned K[3][8];  // keys of rows y - 2 .. y
  unsigned F[3];     // flat bits of rows y - 3 .. y - 1 (bit fc: column x - 1 + fc)
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 8; j++) V[i][j] = nan;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    F[i] = 0u;
#pragma unroll
    for (int j = 0; j < 8; j++) K[i][j] = 0x007fffffu;
  }
  unsigned long long flats = 0;
  const bool active = ys < ny;
  constexpr int T = kStripF + 4;  // rows ys - 2 .. ys + kStripF + 1
  static_assert(T % kAheadF == 0, "whole batches of rows");
  RawRowF q[kAheadF];
#pragma unroll
  for (int d = 0; d < kAheadF; d++)
    q[d] = load_row_f(dem, nx, active ? ny : 0, x, ys - 2 + d, lane, xin);
  for (int t0 = 0; t0 < T; t0 += kAheadF) {
#pragma unroll
    for (int d = 0; d < kAheadF; d++) {
      const int t = t0 + d;
      const long long y = ys - 2 + t;
      const RawRowF cur = q[d];
      if (t + kAheadF < T) q[d] = load_row_f(dem, nx, active ? ny : 0, x, y + kAheadF, lane, xin);
      float v[8];
      v[0] = shr(cur.lw.x, cur.c.z);
      v[1] = shr(cur.lw.y, cur.c.w);
      v[2] = cur.c.x;
      v[3] = cur.c.y;
      v[4] = cur.c.z;
      v[5] = cur.c.w;
      v[6] = shl(cur.rw.x, cur.c.x);
      v[7] = shl(cur.rw.y, cur.c.y);
#pragma unroll
      for (int j = 0; j < 8; j++) {
        V[0][j] = V[1][j];
        V[1][j] = V[2][j];
        V[2][j] = V[3][j];
        V[3][j] = v[j];
        K[0][j] = K[1][j];
        K[1][j] = K[2][j];
        K[2][j] = flat_key(v[j]);
      }
      // flat bits of row y - 1: an interior cell equal to the minimum of
      // its 3x3 window, NaN neighbours counting as -INF; as keys, a non-NaN
      // cell is flat exactly when the window's minimum key is its own
      unsigned fnew = 0u;
      {
        unsigned cm[8];
#pragma unroll
        for (int j = 0; j < 8; j++) cm[j] = min(min(K[0][j], K[1][j]), K[2][j]);
#pragma unroll
        for (int fc = 0; fc < 6; fc++) {
          const float h = V[2][fc + 1];
          const unsigned m9 = min(min(cm[fc], cm[fc + 1]), cm[fc + 2]);
          fnew |= (unsigned)(h == h && m9 == K[1][fc + 1]) << fc;
        }
        const long long gy = y - 1;
        fnew = (gy > 0 && gy < ny - 1) ? (fnew & colin) : 0u;
      }
      F[0] = F[1];
      F[1] = F[2];
      F[2] = fnew;
      // flags of row y - 2 (own cells). A neighbour outside the grid is NaN
      // in the window, never equal to anything. Neighbour n of a cell
      // (n = 0..7: rows -1, 0, +1, columns -1, 0, +1 in order, the cell
      // itself skipped): eq bit n if its value equals the cell's, nf bit n
      // if it is flat.
      const long long gy = y - 2;
      if (active && xin && gy >= ys && gy < ys + kStripF && gy < ny) {
        int o[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const float h = V[1][k + 2];
          const unsigned flat = (F[1] >> (k + 1)) & 1u;
          const unsigned nf = ((F[0] >> k) & 7u) | (((F[1] >> k) & 1u) << 3) |
                              (((F[1] >> (k + 2)) & 1u) << 4) | (((F[2] >> k) & 7u) << 5);
          unsigned eq = 0u;
          int n = 0;
#pragma unroll
          for (int dr = 0; dr < 3; dr++)
#pragma unroll
            for (int dc = -1; dc <= 1; dc++) {
              if (dr == 1 && dc == 0) continue;
              eq |= (unsigned)(V[dr][k + 2 + dc] == h) << n;
              n++;
            }
          // sill: not flat, touches an equal flat; presill: flat, touches an
          // equal non-flat cell (a sill)
          const unsigned sill = !flat && (eq & nf) != 0u;
          const unsigned presill = flat && (eq & ~nf) != 0u;
          o[k] = (int)(flat | (sill << 1) | (presill << 2));
        }
        flats += __popc((F[1] >> 1) & 15u);
        typedef int v4i __attribute__((ext_vector_type(4)));
        const v4i ov = {o[0], o[1], o[2], o[3]};
        __builtin_nontemporal_store(ov, reinterpret_cast<v4i*>(out + gy * nx + x));
      }
    }
  }
  if (partial) {
    for (int off = 32; off > 0; off >>= 1) flats += __shfl_xor(flats, off);
    if (lane == 0) wave_count[wave] = flats;
    __syncthreads();
    if (threadIdx.x == 0)
      partial[block_id] = wave_count[0] + wave_count[1] + wave_count[2] + wave_count[3];
  }
}

__global__ void sum_partials(unsigned long long* partial, int n) {
  __shared__ unsigned long long s[256];
  unsigned long long v = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) v += partial[i];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[n] = s[0];
}

// ---------------------------------------------------------------- value filters
// op < 0: minimum, op > 0: maximum. The running value v stays only if it
// beats h strictly (the reference's "continue" test), so ties take h; NaN
// (and out-of-grid) cells are skipped.
template <int OP>
__device__ __forceinline__ float fold(float v, float h) {
  return (h != h || (OP < 0 ? (v < h) : (v > h))) ? v : h;
}

// Structuring elements handled through LDS: at most kSeMax set entries and
// a halo of at most kHaloMax cells on each axis; larger ones take the
// direct kernel below.
constexpr int kSeMax = 1024, kHaloMax = 16;

// The structuring element's set entries in the reference's visiting order
// (slice, row, column), packed as (row << 16 | column); returns the count.
__device__ int compact_se(unsigned* list, const uint8_t* se, long long sx, long long sy,
                          long long ss) {
  __shared__ unsigned wave_base[kWaves];
  __shared__ unsigned total;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long n = sx * sy * ss;
  if (threadIdx.x == 0) total = 0;
  __syncthreads();
  for (long long k0 = 0; k0 < n; k0 += kThreads) {
    const long long k = k0 + threadIdx.x;
    const bool set = k < n && se[k];
    const unsigned long long m = __ballot(set);
    const unsigned below = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wave_base[wave] = (unsigned)__popcll(m);
    __syncthreads();
    unsigned off = total;
    for (int i = 0; i < wave; i++) off += wave_base[i];
    if (set) {
      const long long in_slice = k % (sx * sy);
      list[off + below] = (unsigned)((in_slice / sx) << 16 | (in_slice % sx));
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int i = 0; i < kWaves; i++) total += wave_base[i];
    __syncthreads();
  }
  return (int)total;
}

template <int OP, int RPW>
__global__ __launch_bounds__(kThreads) void value_filter_tiled(
    float* out, const float* dem, const uint8_t* se, long long sx, long long sy, long long ss,
    long long nx, long long ny) {
  extern __shared__ float smem[];
  __shared__ unsigned list[kSeMax];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cx = (int)((sx + 1) / 2 - 1), cy = (int)((sy + 1) / 2 - 1);
  const int W = kTW + (int)sx - 1, H = kTH + (int)sy - 1;
  const TileWalk w(nx, ny);
  WindowLoader<RPW> win;
  if (w.begin < w.end)
    win.load(dem, nx, ny, (w.begin % w.ntx) * kTW - cx, (w.begin / w.ntx) * kTH - cy, W, H);
  const int nlist = compact_se(list, se, sx, sy, ss);
  for (long long t = w.begin; t < w.end; t += w.stride) {
    const long long x0 = (t % w.ntx) * kTW, y0 = (t / w.ntx) * kTH;
    win.store(smem, W, W, H);
    __syncthreads();
    const long long tn = t + w.stride;
    if (tn < w.end) win.load(dem, nx, ny, (tn % w.ntx) * kTW - cx, (tn / w.ntx) * kTH - cy, W, H);
    for (int r = wave; r < kTH; r += kWaves) {
      const long long gy = y0 + r;
      if (gy >= ny) break;
      float v[kCols];
#pragma unroll
      for (int k = 0; k < kCols; k++) v[k] = OP < 0 ? __builtin_inff() : -__builtin_inff();
      for (int e = 0; e < nlist; e++) {
        const unsigned ab = list[e];
        const float* row = smem + (r + (int)(ab >> 16)) * W + (int)(ab & 0xffffu) + lane;
#pragma unroll
        for (int k = 0; k < kCols; k++) v[k] = fold<OP>(v[k], row[64 * k]);
      }
#pragma unroll
      for (int k = 0; k < kCols; k++) {
        const int c = lane + 64 * k;
        const long long gx = x0 + c;
        if (gx >= nx) continue;
        const float self = smem[(r + cy) * W + c + cx];
        out[gy * nx + gx] = self != self ? nan_marker() : v[k];
      }
    }
    __syncthreads();
  }
}

// Any structuring element: each cell folds straight from global memory.
template <int OP>
__global__ void value_filter_direct(float* out, const float* dem, const uint8_t* se,
                                    long long sx, long long sy, long long ss, long long nx,
                                    long long ny) {
  const long long cx = (sx + 1) / 2 - 1, cy = (sy + 1) / 2 - 1;
  for (long long j = blockIdx.y; j < ny; j += gridDim.y)
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nx;
         i += (long long)gridDim.x * blockDim.x) {
      const long long p = j * nx + i;
      if (dem[p] != dem[p]) {
        out[p] = nan_marker();
        continue;
      }
      float v = OP < 0 ? __builtin_inff() : -__builtin_inff();
      for (long long s = 0; s < ss; s++)
        for (long long b = 0; b < sy; b++) {
          const long long jj = j + b - cy;
          if (jj < 0 || jj >= ny) continue;
          const uint8_t* row = se + (s * sy + b) * sx;
          for (long long a = 0; a < sx; a++) {
            const long long ii = i + a - cx;
            if (!row[a] || ii < 0 || ii >= nx) continue;
            v = fold<OP>(v, dem[jj * nx + ii]);
          }
        }
      out[p] = v;
    }
}

// Square filter in one pass: each window row is folded along x and the
// result written over the row's first kTW entries in LDS (one wave owns a
// row, and its lanes read the row before any of them writes it). Rows
// outside the grid are marked NaN, which the vertical fold skips; tmp
// itself is never NaN, since each fold starts at +-INF and skips NaN. The
// tile's own rows also go to `tmp`; whether a tile cell is NaN is kept
// aside before its row is overwritten.
constexpr int kSquareMax = 33;

template <int OP, int RPW>
__global__ __launch_bounds__(kThreads) void value_filter_square_tiled(
    float* out, float* tmp, const float* dem, int width, long long nx, long long ny) {
  extern __shared__ float smem[];
  __shared__ unsigned char self_nan[kTH * kTW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = (width + 1) / 2 - 1;
  const int W = kTW + width - 1, H = kTH + width - 1;
  const TileWalk w(nx, ny);
  WindowLoader<RPW> win;
  if (w.begin < w.end)
    win.load(dem, nx, ny, (w.begin % w.ntx) * kTW - c, (w.begin / w.ntx) * kTH - c, W, H);
  for (long long t = w.begin; t < w.end; t += w.stride) {
    const long long x0 = (t % w.ntx) * kTW, y0 = (t / w.ntx) * kTH;
    win.store(smem, W, W, H);
    __syncthreads();
    const long long tn = t + w.stride;
    if (tn < w.end) win.load(dem, nx, ny, (tn % w.ntx) * kTW - c, (tn / w.ntx) * kTH - c, W, H);
    for (int r = wave; r < H; r += kWaves) {
      const long long gy = y0 - c + r;
      const bool row_in = gy >= 0 && gy < ny;
      const bool own = r >= c && r < c + kTH;
      float* row = smem + r * W;
      float v[kCols];
#pragma unroll
      for (int k = 0; k < kCols; k++) v[k] = OP < 0 ? __builtin_inff() : -__builtin_inff();
      for (int a = 0; a < width; a++)
#pragma unroll
        for (int k = 0; k < kCols; k++) v[k] = fold<OP>(v[k], row[lane + a + 64 * k]);
      if (own) {
#pragma unroll
        for (int k = 0; k < kCols; k++) {
          const float self = row[lane + 64 * k + c];
          self_nan[(r - c) * kTW + lane + 64 * k] = self != self;
        }
      }
#pragma unroll
      for (int k = 0; k < kCols; k++) {
        const int col = lane + 64 * k;
        row[col] = row_in ? v[k] : nan_marker();
        if (own && row_in && x0 + col < nx) tmp[gy * nx + x0 + col] = v[k];
      }
    }
    __syncthreads();
    for (int r = wave; r < kTH; r += kWaves) {
      const long long gy = y0 + r;
      if (gy >= ny) break;
      float v[kCols];
#pragma unroll
      for (int k = 0; k < kCols; k++) v[k] = OP < 0 ? __builtin_inff() : -__builtin_inff();
      for (int b = 0; b < width; b++)
#pragma unroll
        for (int k = 0; k < kCols; k++) v[k] = fold<OP>(v[k], smem[(r + b) * W + lane + 64 * k]);
#pragma unroll
      for (int k = 0; k < kCols; k++) {
        const int col = lane + 64 * k;
        if (x0 + col >= nx) continue;
        out[gy * nx + x0 + col] = self_nan[r * kTW + col] ? nan_marker() : v[k];
      }
    }
    __syncthreads();
  }
}

// One axis of the square filter for any width: along x (AXIS 0, from `in`
// into `out`, no NaN pass-through) or along y (AXIS 1, NaN cells of `dem`
// stay NaN).
template <int OP, int AXIS>
__global__ void value_filter_axis(float* out, const float* in, const float* dem, int width,
                                  long long nx, long long ny) {
  const long long c = (width + 1) / 2 - 1;
  for (long long j = blockIdx.y; j < ny; j += gridDim.y)
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nx;
         i += (long long)gridDim.x * blockDim.x) {
      const long long p = j * nx + i;
      if (AXIS == 1 && dem[p] != dem[p]) {
        out[p] = nan_marker();
        continue;
      }
      float v = OP < 0 ? __builtin_inff() : -__builtin_inff();
      for (int k = 0; k < width; k++) {
        if (AXIS == 0) {
          const long long ii = i + k - c;
          if (ii >= 0 && ii < nx) v = fold<OP>(v, in[j * nx + ii]);
        } else {
          const long long jj = j + k - c;
          if (jj >= 0 && jj < ny) v = fold<OP>(v, in[jj * nx + i]);
        }
      }
      out[p] = v;
    }
}

// ---------------------------------------------------------------- host side
// Workgroups for a tiled kernel: enough to fill every CU at the kernel's
// occupancy, a multiple of the 8 XCDs, no more than the tiles.
static hipError_t tiled_grid(const void* kernel, size_t lds, long long ntiles, int* grid) {
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kThreads, lds);
  if (e != hipSuccess) return e;
  if (per_cu < 1) return hipErrorInvalidConfiguration;
  long long g = (long long)cus * per_cu * 2;  // two rounds of tiles in flight
  if (g > ntiles) g = ntiles;
  if (g >= 8) g -= g % 8;
  *grid = (int)(g < 1 ? 1 : g);
  return hipSuccess;
}

static long long tile_count(long long nx, long long ny) {
  return ((nx + kTW - 1) / kTW) * ((ny + kTH - 1) / kTH);
}

static dim3 direct_grid(long long nx, long long ny) {
  long long bx = (nx + 255) / 256;
  if (bx > 4096) bx = 4096;
  return dim3((unsigned)bx, (unsigned)(ny < 65535 ? ny : 65535), 1);
}

hipError_t run_identifyflats(int32_t* out, const float* dem, int64_t nx, int64_t ny,
                             hipStream_t stream, int64_t* count_host) {
  if (nx <= 0 || ny <= 0) {
    if (count_host) *count_host = 0;
    return hipSuccess;
  }
  const auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  const long long sgx = (nx + 255) / 256, sgy = (ny + 4 * kStripF - 1) / (4 * kStripF);
  const bool stream_path = (nx & 3) == 0 && al(out) && al(dem) && sgy <= 65535;
  int grid = 1;
  hipError_t e = hipSuccess;
  if (stream_path) {
    grid = (int)(sgx * sgy);
  } else {
    e = tiled_grid((const void*)identifyflats_kernel, 0, tile_count(nx, ny), &grid);
    if (e != hipSuccess) return e;
  }
  unsigned long long* partial = nullptr;
  if (count_host) {
    e = hipMallocAsync((void**)&partial, sizeof(*partial) * (grid + 1), stream);
    if (e != hipSuccess) return e;
  }
  if (stream_path)
    identifyflats_stream<<<dim3((unsigned)sgx, (unsigned)sgy), 256, 0, stream>>>(out, dem, nx, ny,
                                                                                 partial);
  else
    identifyflats_kernel<<<grid, kThreads, 0, stream>>>(out, dem, nx, ny, partial);
  e = hipGetLastError();
  if (count_host) {
    if (e == hipSuccess) {
      sum_partials<<<1, 256, 0, stream>>>(partial, grid);
      e = hipGetLastError();
    }
    unsigned long long h = 0;
    if (e == hipSuccess)
      e = hipMemcpyAsync(&h, partial + grid, sizeof(h), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    *count_host = (int64_t)h;
    hipError_t f = hipFreeAsync(partial, stream);
    if (e == hipSuccess) e = f;
  }
  return e;
}

// 3x3 structuring elements of one slice, streaming through registers:
// each lane owns four adjacent columns of a kRows3-row strip and holds the
// strip's rows plus the one above and below (16-byte loads, all in flight
// together); the columns beside its four come from the neighbouring lanes
// by DPP, the wavefront's outside columns by one extra load in its edge
// lanes. Every cell folds the set entries in the reference's visiting
// order (row, then column), as the tiled kernel does. Needs nx % 4 == 0
// and 16-byte aligned arrays. Short strips (many wavefronts in flight),
// nontemporal stores and per-XCD runs of strips measured fastest at
// 8192^2: 0.096 ms against 0.170 ms for the tiled kernel.
constexpr int kRows3 = 4;

template <int OP>
__device__ __forceinline__ void fold3x3_row(float (&acc)[4], const float (&w)[6], unsigned bits) {
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int dx = 0; dx < 3; dx++)
      if (bits & (1u << dx)) acc[k] = fold<OP>(acc[k], w[k + dx]);
}

template <int OP>
__global__ __launch_bounds__(256) void value_filter_3x3(float* out, const float* dem,
                                                        const uint8_t* se, long long nx,
                                                        long long ny) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // workgroups go round-robin over the 8 XCDs; when the grid divides by 8,
  // renumber them so that each XCD takes a contiguous run of strips, whose
  // shared halo rows then meet in its L2
  const long long gx = gridDim.x, total = gx * gridDim.y;
  long long L = (long long)blockIdx.y * gx + blockIdx.x;
  if ((total & 7) == 0) L = (L & 7) * (total >> 3) + (L >> 3);
  const long long bx = L % gx, by = L / gx;
  const long long x = (bx * 64 + lane) * 4;
  const long long y0 = (by * 4 + wave) * kRows3;
  if (y0 >= ny) return;  // whole wavefront
  unsigned mask = 0u;  // bit 3 * row + column of the element
#pragma unroll
  for (int k = 0; k < 9; k++) mask |= (se[k] ? 1u : 0u) << k;
  const float nan = __builtin_nanf("");
  const bool xin = x < nx;
  float4 c[kRows3 + 2];
  float lft[kRows3 + 2], rgt[kRows3 + 2];
#pragma unroll
  for (int i = 0; i < kRows3 + 2; i++) {
    const long long y = y0 - 1 + i;
    const bool yin = y >= 0 && y < ny;
    c[i] = make_float4(nan, nan, nan, nan);
    lft[i] = nan;
    rgt[i] = nan;
    if (xin && yin) c[i] = *reinterpret_cast<const float4*>(dem + y * nx + x);
    if (lane == 0 && yin && x >= 1 && x - 1 < nx) lft[i] = dem[y * nx + x - 1];
    if (lane == 63 && yin && x + 4 < nx) rgt[i] = dem[y * nx + x + 4];
  }
  // row i's six columns x - 1 .. x + 4
  float w[kRows3 + 2][6];
#pragma unroll
  for (int i = 0; i < kRows3 + 2; i++) {
    w[i][0] = __int_as_float(__builtin_amdgcn_update_dpp(
        __float_as_int(lft[i]), __float_as_int(c[i].w), 0x138, 0xf, 0xf, false));
    w[i][1] = c[i].x;
    w[i][2] = c[i].y;
    w[i][3] = c[i].z;
    w[i][4] = c[i].w;
    w[i][5] = __int_as_float(__builtin_amdgcn_update_dpp(
        __float_as_int(rgt[i]), __float_as_int(c[i].x), 0x130, 0xf, 0xf, false));
  }
  const float init = OP < 0 ? __builtin_inff() : -__builtin_inff();
#pragma unroll
  for (int j = 0; j < kRows3; j++) {
    const long long y = y0 + j;
    if (!xin || y >= ny) continue;
    float acc[4] = {init, init, init, init};
    fold3x3_row<OP>(acc, w[j], mask);
    fold3x3_row<OP>(acc, w[j + 1], mask >> 3);
    fold3x3_row<OP>(acc, w[j + 2], mask >> 6);
    float o[4];
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = w[j + 1][k + 1] != w[j + 1][k + 1] ? nan : acc[k];
    typedef float v4f __attribute__((ext_vector_type(4)));
    const v4f ov = {o[0], o[1], o[2], o[3]};
    __builtin_nontemporal_store(ov, reinterpret_cast<v4f*>(out + y * nx + x));
  }
}

template <int OP>
static hipError_t value_filter_launch(float* out, const float* dem, const uint8_t* se,
                                      const int64_t d[3], int64_t nx, int64_t ny,
                                      hipStream_t stream) {
  const auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  if (d[0] == 3 && d[1] == 3 && d[2] == 1 && (nx & 3) == 0 && al(out) && al(dem) &&
      (ny + 4 * kRows3 - 1) / (4 * kRows3) <= 65535) {
    const dim3 g((unsigned)((nx + 255) / 256), (unsigned)((ny + 4 * kRows3 - 1) / (4 * kRows3)));
    value_filter_3x3<OP><<<g, 256, 0, stream>>>(out, dem, se, nx, ny);
    return hipGetLastError();
  }
  const bool tiled = d[0] >= 1 && d[1] >= 1 && d[2] >= 1 && d[0] <= 2 * kHaloMax + 1 &&
                     d[1] <= 2 * kHaloMax + 1 && d[0] * d[1] * d[2] <= kSeMax;
  if (tiled) {
    const size_t lds = sizeof(float) * (size_t)(kTW + d[0] - 1) * (size_t)(kTH + d[1] - 1);
    const int rows = kTH + (int)d[1] - 1;
    auto go = [&](auto kernel) {
      int grid = 1;
      hipError_t e = tiled_grid((const void*)kernel, lds, tile_count(nx, ny), &grid);
      if (e != hipSuccess) return e;
      kernel<<<grid, kThreads, lds, stream>>>(out, dem, se, d[0], d[1], d[2], nx, ny);
      return hipGetLastError();
    };
    if (rows <= 5 * kWaves) return go(value_filter_tiled<OP, 5>);
    if (rows <= 8 * kWaves) return go(value_filter_tiled<OP, 8>);
    return go(value_filter_tiled<OP, 12>);
  } else {
    value_filter_direct<OP><<<direct_grid(nx, ny), 256, 0, stream>>>(out, dem, se, d[0], d[1],
                                                                     d[2], nx, ny);
  }
  return hipGetLastError();
}

hipError_t run_value_filter(int op, float* out, const float* dem, const uint8_t* se,
                            const int64_t se_dims[3], int64_t nx, int64_t ny,
                            hipStream_t stream) {
  if (nx <= 0 || ny <= 0) return hipSuccess;
  return op < 0 ? value_filter_launch<-1>(out, dem, se, se_dims, nx, ny, stream)
                : value_filter_launch<1>(out, dem, se, se_dims, nx, ny, stream);
}

// Square elements of width 1..5 streaming through registers (nx % 4 ==
// 0, 16-byte aligned arrays), laid out as identifyflats_stream: each lane
// owns four columns of a kRowsS-row strip and holds the rows its window
// needs, columns -2 .. 5 around its own. Each row is folded along x
// (written to tmp for the strip's own rows; rows outside the grid become
// NaN, which the vertical fold skips), then each cell along y, both in the
// reference's order; a NaN input cell gives NaN.
constexpr int kRowsS = 4;

template <int OP, int WD>
__global__ __launch_bounds__(256) void value_filter_square_stream(float* out, float* tmp,
                                                                  const float* dem, long long nx,
                                                                  long long ny) {
  constexpr int C = (WD + 1) / 2 - 1;  // cells before the centre on each axis
  constexpr int NR = kRowsS + WD - 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // per-XCD runs of strips, as in value_filter_3x3
  const long long gxb = gridDim.x, total = gxb * gridDim.y;
  long long L = (long long)blockIdx.y * gxb + blockIdx.x;
  if ((total & 7) == 0) L = (L & 7) * (total >> 3) + (L >> 3);
  const long long bx = L % gxb, by = L / gxb;
  const long long x = (bx * 64 + lane) * 4;
  const long long y0 = (by * 4 + wave) * kRowsS;
  if (y0 >= ny) return;  // whole wavefront
  const bool xin = x < nx;
  const float nan = __builtin_nanf("");
  const float init = OP < 0 ? __builtin_inff() : -__builtin_inff();
  typedef float v4f __attribute__((ext_vector_type(4)));
  float h[NR][4];          // row y0 - C + i folded along x
  unsigned self_nan[kRowsS];  // bit k: input cell (y0 + r, x + k) is NaN
#pragma unroll
  for (int i = 0; i < NR; i++) {
    const long long y = y0 - C + i;
    const bool yin = y >= 0 && y < ny;
    float4 c = make_float4(nan, nan, nan, nan);
    float2 lw = make_float2(nan, nan), rw = make_float2(nan, nan);
    if (xin && yin) c = *reinterpret_cast<const float4*>(dem + y * nx + x);
    if (lane == 0 && yin && x >= 2 && x - 2 < nx)
      lw = *reinterpret_cast<const float2*>(dem + y * nx + x - 2);
    if (lane == 63 && yin && x + 4 < nx)
      rw = *reinterpret_cast<const float2*>(dem + y * nx + x + 4);
    const float w[8] = {  // columns x - 2 .. x + 5; lane i - 1 / i + 1 by DPP
        __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(lw.x), __float_as_int(c.z),
                                                   0x138, 0xf, 0xf, false)),
        __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(lw.y), __float_as_int(c.w),
                                                   0x138, 0xf, 0xf, false)),
        c.x, c.y, c.z, c.w,
        __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(rw.x), __float_as_int(c.x),
                                                   0x130, 0xf, 0xf, false)),
        __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(rw.y), __float_as_int(c.y),
                                                   0x130, 0xf, 0xf, false))};
#pragma unroll
    for (int k = 0; k < 4; k++) {
      float acc = init;
#pragma unroll
      for (int a = 0; a < WD; a++) acc = fold<OP>(acc, w[k - C + a + 2]);
      h[i][k] = yin ? acc : nan;
    }
    if (i >= C && i < C + kRowsS) {
      self_nan[i - C] = (unsigned)(c.x != c.x) | ((unsigned)(c.y != c.y) << 1) |
                        ((unsigned)(c.z != c.z) << 2) | ((unsigned)(c.w != c.w) << 3);
      if (xin && yin) {
        const v4f tv = {h[i][0], h[i][1], h[i][2], h[i][3]};
        __builtin_nontemporal_store(tv, reinterpret_cast<v4f*>(tmp + y * nx + x));
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kRowsS; r++) {
    const long long y = y0 + r;
    if (!xin || y >= ny) continue;
    float o[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      float acc = init;
#pragma unroll
      for (int a = 0; a < WD; a++) acc = fold<OP>(acc, h[r + a][k]);
      o[k] = ((self_nan[r] >> k) & 1u) ? nan : acc;
    }
    const v4f ov = {o[0], o[1], o[2], o[3]};
    __builtin_nontemporal_store(ov, reinterpret_cast<v4f*>(out + y * nx + x));
  }
}

template <int OP>
static hipError_t square_launch(float* out, const float* dem, float* tmp, int width, int64_t nx,
                                int64_t ny, hipStream_t stream) {
  const auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  if (width >= 1 && width <= 5 && (nx & 3) == 0 && al(out) && al(tmp) && al(dem) &&
      (ny + 4 * kRowsS - 1) / (4 * kRowsS) <= 65535) {
    const dim3 g((unsigned)((nx + 255) / 256), (unsigned)((ny + 4 * kRowsS - 1) / (4 * kRowsS)));
    switch (width) {
      case 1: value_filter_square_stream<OP, 1><<<g, 256, 0, stream>>>(out, tmp, dem, nx, ny); break;
      case 2: value_filter_square_stream<OP, 2><<<g, 256, 0, stream>>>(out, tmp, dem, nx, ny); break;
      case 3: value_filter_square_stream<OP, 3><<<g, 256, 0, stream>>>(out, tmp, dem, nx, ny); break;
      case 4: value_filter_square_stream<OP, 4><<<g, 256, 0, stream>>>(out, tmp, dem, nx, ny); break;
      default: value_filter_square_stream<OP, 5><<<g, 256, 0, stream>>>(out, tmp, dem, nx, ny); break;
    }
    return hipGetLastError();
  }
  if (width >= 1 && width <= kSquareMax) {
    const size_t lds = sizeof(float) * (size_t)(kTH + width - 1) * (size_t)(kTW + width - 1);
    const int rows = kTH + width - 1;
    auto go = [&](auto kernel) {
      int grid = 1;
      hipError_t e = tiled_grid((const void*)kernel, lds, tile_count(nx, ny), &grid);
      if (e != hipSuccess) return e;
      kernel<<<grid, kThreads, lds, stream>>>(out, tmp, dem, width, nx, ny);
      return hipGetLastError();
    };
    if (rows <= 5 * kWaves) return go(value_filter_square_tiled<OP, 5>);
    if (rows <= 8 * kWaves) return go(value_filter_square_tiled<OP, 8>);
    return go(value_filter_square_tiled<OP, 12>);
  } else {
    const dim3 g = direct_grid(nx, ny);
    value_filter_axis<OP, 0><<<g, 256, 0, stream>>>(tmp, dem, dem, width, nx, ny);
    value_filter_axis<OP, 1><<<g, 256, 0, stream>>>(out, tmp, dem, width, nx, ny);
  }
  return hipGetLastError();
}

hipError_t run_value_filter_square(int op, float* out, const float* dem, float* tmp, int width,
                                   int64_t nx, int64_t ny, hipStream_t stream) {
  if (nx <= 0 || ny <= 0) return hipSuccess;
  return op < 0 ? square_launch<-1>(out, dem, tmp, width, nx, ny, stream)
                : square_launch<1>(out, dem, tmp, width, nx, ny, stream);
}

}  // namespace tt
