// Single-pass stencils that consume a DEM next to the hot path: flat / sill
// / presill classification and the minimum / maximum value filters. Each
// output cell depends only on a fixed neighbourhood of the input, so these
// are one read and one write per cell, streamed at HBM rate.
//
// Reference semantics:
//   identifyflats      src/identifyflats.c:32-121
//   min_filter,        src/value_filters.c:9-113: a sequential fold over
//   max_filter         the structuring element (slices, rows, columns in
//                      that order), the running value replaced unless it is
//                      strictly better, so the result's bits (e.g. -0 vs
//                      +0) follow the same order here
//   *_filter_square    src/value_filters.c:115-246: the same fold along
//                      dims[0] into tmp, then along dims[1]
#include "kernels.h"

namespace tt {

// ---------------------------------------------------------------- identifyflats
constexpr int kFlatTW = 64, kFlatTH = 8, kFlatThreads = 256;

// Tile of kFlatTW x kFlatTH cells; LDS holds the DEM with a 2-cell halo
// and the flat bits with a 1-cell halo (a sill or presill of a tile cell
// depends on its neighbours being flats).
__global__ __launch_bounds__(kFlatThreads) void identifyflats_kernel(
    int32_t* out, const float* dem, long long nx, long long ny, unsigned long long* count) {
  constexpr int DW = kFlatTW + 4, DH = kFlatTH + 4;  // DEM window
  constexpr int FW = kFlatTW + 2, FH = kFlatTH + 2;  // flat-bit window
  __shared__ float sd[DH][DW];
  __shared__ unsigned char sflat[FH][FW];
  __shared__ unsigned block_flats;
  const int tid = threadIdx.x;
  const long long ntx = (nx + kFlatTW - 1) / kFlatTW, nty = (ny + kFlatTH - 1) / kFlatTH;
  for (long long t = blockIdx.x; t < ntx * nty; t += gridDim.x) {
    const long long x0 = (t % ntx) * kFlatTW, y0 = (t / ntx) * kFlatTH;
    if (tid == 0) block_flats = 0u;
    for (int k = tid; k < DW * DH; k += kFlatThreads) {
      const int r = k / DW, c = k % DW;
      const long long gx = x0 + c - 2, gy = y0 + r - 2;
      sd[r][c] = (gx >= 0 && gx < nx && gy >= 0 && gy < ny) ? dem[gy * nx + gx]
                                                            : __builtin_nanf("");
    }
    __syncthreads();
    // flats: interior cells equal to the minimum of their 3x3 window, NaN
    // neighbours counting as -INF
    for (int k = tid; k < FW * FH; k += kFlatThreads) {
      const int r = k / FW, c = k % FW;  // DEM window cell (r + 1, c + 1)
      const long long gx = x0 + c - 1, gy = y0 + r - 1;
      bool flat = false;
      if (gx > 0 && gx < nx - 1 && gy > 0 && gy < ny - 1) {
        const float h = sd[r + 1][c + 1];
        float lowest = h;
        for (int dr = 0; dr < 3; dr++)
          for (int dc = 0; dc < 3; dc++) {
            if (dr == 1 && dc == 1) continue;
            const float q = sd[r + dr][c + dc];
            lowest = fminf(lowest, q != q ? -__builtin_inff() : q);
          }
        flat = h == lowest;
      }
      sflat[r][c] = flat ? 1 : 0;
    }
    __syncthreads();
    unsigned flats = 0;
    for (int k = tid; k < kFlatTW * kFlatTH; k += kFlatThreads) {
      const int r = k / kFlatTW, c = k % kFlatTW;
      const long long gx = x0 + c, gy = y0 + r;
      if (gx >= nx || gy >= ny) continue;
      const bool flat = sflat[r + 1][c + 1];
      const float h = sd[r + 2][c + 2];
      int32_t bits = flat ? 1 : 0;
      for (int dr = -1; dr <= 1; dr++)
        for (int dc = -1; dc <= 1; dc++) {
          if (!dr && !dc) continue;
          const long long qx = gx + dc, qy = gy + dr;
          if (qx < 0 || qx >= nx || qy < 0 || qy >= ny) continue;
          if (sd[r + 2 + dr][c + 2 + dc] != h) continue;  // NaN never equal
          const bool qflat = sflat[r + 1 + dr][c + 1 + dc];
          if (!flat && qflat) bits |= 2;  // sill: touches an equal flat
          if (flat && !qflat) bits |= 4;  // presill: a flat touching an equal sill
        }
      out[gy * nx + gx] = bits;
      flats += flat;
    }
    // wave sum, then one atomic per wave into the tile's count
    for (int o = 32; o > 0; o >>= 1) flats += __shfl_xor(flats, o);
    if ((tid & 63) == 0 && flats) atomicAdd(&block_flats, flats);
    __syncthreads();
    if (tid == 0 && block_flats) atomicAdd(count, (unsigned long long)block_flats);
    __syncthreads();
  }
}

// ---------------------------------------------------------------- value filters
// op < 0: minimum, op > 0: maximum. `better(v, h)`: v stays unless h beats
// it strictly (the reference's "continue" test), so ties take h.
template <int OP>
__device__ __forceinline__ float fold(float v, float h) {
  if (h != h) return v;
  return (OP < 0 ? (v < h) : (v > h)) ? v : h;
}

template <int OP>
__global__ void value_filter_kernel(float* out, const float* dem, const uint8_t* se, long long sx,
                                    long long sy, long long ss, long long nx, long long ny) {
  const long long cx = (sx + 1) / 2 - 1, cy = (sy + 1) / 2 - 1;
  const long long n = nx * ny;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (long long)gridDim.x * blockDim.x) {
    const long long i = p % nx, j = p / nx;
    const float self = dem[p];
    if (self != self) {
      out[p] = __builtin_nanf("");
      continue;
    }
    float v = OP < 0 ? __builtin_inff() : -__builtin_inff();
    for (long long s = 0; s < ss; s++)
      for (long long b = 0; b < sy; b++) {
        const long long jj = j + b - cy;
        if (jj < 0 || jj >= ny) continue;
        const uint8_t* row = se + (s * sy + b) * sx;
        for (long long a = 0; a < sx; a++) {
          if (!row[a]) continue;
          const long long ii = i + a - cx;
          if (ii < 0 || ii >= nx) continue;
          v = fold<OP>(v, dem[jj * nx + ii]);
        }
      }
    out[p] = v;
  }
}

// One axis of the square filter: along x (AXIS 0, from `in` into `out`,
// no NaN pass-through) or along y (AXIS 1, NaN cells of `dem` stay NaN).
template <int OP, int AXIS>
__global__ void value_filter_axis(float* out, const float* in, const float* dem, int width,
                                  long long nx, long long ny) {
  const long long c = (width + 1) / 2 - 1;
  const long long n = nx * ny;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (long long)gridDim.x * blockDim.x) {
    const long long i = p % nx, j = p / nx;
    if (AXIS == 1 && dem[p] != dem[p]) {
      out[p] = __builtin_nanf("");
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

__global__ void zero_u64(unsigned long long* p) { *p = 0ull; }

static int stream_grid(long long n, int threads) {
  const long long b = (n + threads - 1) / threads;
  return (int)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

hipError_t run_identifyflats(int32_t* out, const float* dem, int64_t nx, int64_t ny,
                             hipStream_t stream, int64_t* count_host) {
  if (nx <= 0 || ny <= 0) {
    if (count_host) *count_host = 0;
    return hipSuccess;
  }
  unsigned long long* count = nullptr;
  hipError_t e = hipMallocAsync((void**)&count, sizeof(*count), stream);
  if (e != hipSuccess) return e;
  zero_u64<<<1, 1, 0, stream>>>(count);
  const long long tiles = ((nx + kFlatTW - 1) / kFlatTW) * ((ny + kFlatTH - 1) / kFlatTH);
  const int grid = (int)(tiles < 16384 ? tiles : 16384);
  identifyflats_kernel<<<grid, kFlatThreads, 0, stream>>>(out, dem, nx, ny, count);
  e = hipGetLastError();
  if (e == hipSuccess && count_host) {
    unsigned long long h = 0;
    e = hipMemcpyAsync(&h, count, sizeof(h), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    *count_host = (int64_t)h;
  }
  hipError_t f = hipFreeAsync(count, stream);
  return e != hipSuccess ? e : f;
}

hipError_t run_value_filter(int op, float* out, const float* dem, const uint8_t* se,
                            const int64_t se_dims[3], int64_t nx, int64_t ny,
                            hipStream_t stream) {
  if (nx <= 0 || ny <= 0) return hipSuccess;
  const int grid = stream_grid(nx * ny, 256);
  if (op < 0)
    value_filter_kernel<-1><<<grid, 256, 0, stream>>>(out, dem, se, se_dims[0], se_dims[1],
                                                      se_dims[2], nx, ny);
  else
    value_filter_kernel<1><<<grid, 256, 0, stream>>>(out, dem, se, se_dims[0], se_dims[1],
                                                     se_dims[2], nx, ny);
  return hipGetLastError();
}

hipError_t run_value_filter_square(int op, float* out, const float* dem, float* tmp, int width,
                                   int64_t nx, int64_t ny, hipStream_t stream) {
  if (nx <= 0 || ny <= 0) return hipSuccess;
  const int grid = stream_grid(nx * ny, 256);
  if (op < 0) {
    value_filter_axis<-1, 0><<<grid, 256, 0, stream>>>(tmp, dem, dem, width, nx, ny);
    value_filter_axis<-1, 1><<<grid, 256, 0, stream>>>(out, tmp, dem, width, nx, ny);
  } else {
    value_filter_axis<1, 0><<<grid, 256, 0, stream>>>(tmp, dem, dem, width, nx, ny);
    value_filter_axis<1, 1><<<grid, 256, 0, stream>>>(out, tmp, dem, width, nx, ny);
  }
  return hipGetLastError();
}

}  // namespace tt
