// Costs of the gray-weighted distance transform (gwdt_computecosts):
// connected components of the flats (8-connected) with the largest depth
// (filled - original) of each, every flat's cost
// (max depth - own depth) + 0.1 and the component's label.
//
// Reference semantics, src/gwdt.c:83-274: one raster scan builds a
// union-find over the flats, unifying each flat with its flat neighbours
// among (i-1,j-1), (i,j-1), (i+1,j-1), (i-1,j) in that order; a merge keeps
// the neighbour's root unless the current pixel's root is strictly deeper.
// The label is that final root: a pixel of largest depth, and when several
// pixels share it, the one the scan's merge order picks.
//
// Here the components and their largest depth come from a parallel
// union-find. A component with one pixel of largest depth is labelled with
// it. Otherwise its label is replayed from the few scan steps that can
// change which deepest pixel is the root, in scan order, one component per
// thread:
//   birth     a flat with no flat scanned neighbour starts a set;
//   merge     a flat whose scanned neighbours lie in two different sets
//             (only possible when (i,j-1) is not flat, (i+1,j-1) is, and
//             (i-1,j-1) or (i-1,j) is); the set met second in the
//             neighbour order keeps its deepest root, else the first,
//             else the pixel itself if it is one of the deepest;
//   first     the first deepest pixel that joins a set grown from one
//             birth makes it a set with a deepest root, if it has none.
// All other scan steps leave every deepest root where it is, since only a
// strictly deeper pixel can displace a root. Sets are tracked by their
// birth pixel: each flat points at one flat scanned neighbour, and the
// chains end at births.
//
// NaN depths (e.g. flats of -INF cells, where filled - original is NaN):
// the scan's "strictly deeper" test is false whenever a NaN is involved, so
// a component whose depths are all NaN behaves as one whose depths all tie,
// and is replayed the same way with every pixel a deepest one. A component
// mixing NaN and numbers has no such order; its scan is replayed in full,
// pixel by pixel, with the reference's merge rule.
#include <hipcub/hipcub.hpp>

#include "kernels.h"

namespace tt {
namespace {

constexpr unsigned kNone = 0xffffffffu;
// component kinds (bits): depths that are numbers, depths that are NaN
constexpr unsigned kHasNumber = 1u, kHasNaN = 2u, kMixed = 3u;

__device__ __forceinline__ unsigned gw_ord(float f) {
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float gw_unord(unsigned u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__device__ __forceinline__ unsigned ld(const unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st(unsigned* p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving; every pointer points at a smaller or equal
// index, so concurrent halving and linking only ever store ancestors.
__device__ unsigned find_halving(unsigned* P, unsigned x) {
  for (;;) {
    const unsigned px = ld(P + x);
    if (px == x) return x;
    const unsigned ppx = ld(P + px);
    if (ppx != px) st(P + x, ppx);
    x = px;
  }
}

// Links the sets of a and b, the larger root under the smaller.
__device__ void unite(unsigned* P, unsigned a, unsigned b) {
  for (;;) {
    a = find_halving(P, a);
    b = find_halving(P, b);
    if (a == b) return;
    if (a < b) {
      const unsigned t = a;
      a = b;
      b = t;
    }
    if (atomicCAS(P + a, a, b) == a) return;
  }
}

struct Grid {
  const int32_t* flats;
  long long nx, ny;
  __device__ bool flat(long long i, long long j) const {
    return i >= 0 && i < nx && j >= 0 && j < ny && (flats[j * nx + i] & 1);
  }
};

// Merge candidate at (i, j): the neighbour of the first set (A: (i-1,j-1)
// if flat, else (i-1,j)) and whether (i-1,j-1) is flat, which decides the
// order in which the scan meets the two sets. Returns false if the pixel
// cannot join two sets.
__device__ __forceinline__ bool merge_site(const Grid& g, long long i, long long j,
                                           long long* a_rep, bool* nw) {
  if (g.flat(i, j - 1) || !g.flat(i + 1, j - 1)) return false;
  *nw = g.flat(i - 1, j - 1);
  if (!*nw && !g.flat(i - 1, j)) return false;
  *a_rep = *nw ? (j - 1) * g.nx + (i - 1) : j * g.nx + (i - 1);
  return true;
}

struct Ws {
  unsigned* A;    // per flat: run start / chain-group union-find, then birth
  unsigned* C;    // per birth: component union-find, then component
  unsigned* W;    // per component: largest depth (order-preserving bits)
  unsigned* Mc;   // per component: number of deepest pixels
  unsigned* Mf;   // per component: first deepest pixel, then the label
  unsigned* D;    // per birth: first deepest pixel of its chain group
  unsigned* F;    // per birth: union-find of the replayed scan (first: roots of A)
  unsigned* H;    // per replayed set: its deepest root (first: roots of C)
  unsigned* K;    // per component: kind bits (kHasNumber, kHasNaN)
  unsigned* err;
};

#define GW_FOR_PIXELS                                                             \
  for (long long j = blockIdx.y; j < g.ny; j += gridDim.y)                        \
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < g.nx; \
         i += (long long)gridDim.x * blockDim.x)

// One flat scanned neighbour of (i, j), or the pixel itself if it has none
// (a birth). The trees these links form are the chain groups; their
// smallest pixel, the root, is their birth.
__device__ __forceinline__ long long scanned_parent(const Grid& g, long long i, long long j) {
  const long long p = j * g.nx + i;
  if (g.flat(i - 1, j)) return p - 1;
  if (g.flat(i, j - 1)) return p - g.nx;
  if (g.flat(i - 1, j - 1)) return p - g.nx - 1;
  if (g.flat(i + 1, j - 1)) return p - g.nx + 1;
  return p;
}

// Every flat's scanned parent is its (i-1, j) neighbour when that is flat,
// so the chain groups are horizontal runs of flats joined through their
// first pixels. A[p] = the first pixel of p's run: one block per row walks
// the row in chunks of 256, the last non-flat column carried along. The
// births (run starts with no flat in the row above) initialise their
// per-birth entries.
__global__ __launch_bounds__(256) void gw_runs(Grid g, Ws w) {
  __shared__ long long wave_last[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (long long j = blockIdx.y; j < g.ny; j += gridDim.y) {
    long long carry = -1;  // last non-flat column left of the chunk
    for (long long i0 = 0; i0 < g.nx; i0 += 256) {
      const long long i = i0 + threadIdx.x;
      const bool flat = i < g.nx && (g.flats[j * g.nx + i] & 1);
      // lanes past the row's end count as non-flat; they follow every pixel
      const unsigned long long nonflat = __ballot(!flat);
      const unsigned long long upto =
          nonflat & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
      const long long base = i0 + 64 * wave;
      if (lane == 0) wave_last[wave] = nonflat ? base + 63 - __clzll(nonflat) : -1;
      __syncthreads();
      long long last = upto ? base + 63 - __clzll(upto) : -1;
      for (int v = wave - 1; last < 0 && v >= 0; v--) last = wave_last[v];
      if (last < 0) last = carry;
      if (flat) {
        const long long p = j * g.nx + i;
        w.A[p] = (unsigned)(j * g.nx + last + 1);
        if (last == i - 1 && !g.flat(i - 1, j - 1) && !g.flat(i, j - 1) &&
            !g.flat(i + 1, j - 1)) {  // a birth
          w.C[p] = (unsigned)p;
          w.W[p] = 0u;  // below every ordered float
          w.Mc[p] = 0u;
          w.Mf[p] = kNone;
          w.D[p] = kNone;
          w.K[p] = 0u;
        }
      }
      for (int v = 3; v >= 0; v--)
        if (wave_last[v] >= 0) {
          carry = wave_last[v];
          break;
        }
      __syncthreads();
    }
  }
}

// Joins each run to the run of its first pixel's scanned parent (in the
// row above); the smaller root always becomes the parent, so each group's
// root is its smallest pixel, its birth.
__global__ void gw_link(Grid g, Ws w) {
  GW_FOR_PIXELS {
    const long long p = j * g.nx + i;
    if (!(g.flats[p] & 1) || w.A[p] != (unsigned)p) continue;  // run starts only
    const long long par = scanned_parent(g, i, j);
    if (par != p) unite(w.A, (unsigned)p, (unsigned)par);
  }
}

// Roots are written to a second array and copied back by the next kernel:
// a root stored into A[p] itself could be overwritten by another thread's
// path halving that read A[p]'s old parent just before, leaving a
// non-root behind.
// Invariant after gw_link: every entry of A points at a run start. The
// initial entries do (a flat points at the start of its run), and links and
// path halving only ever store roots and grandparents, which are entries
// of A and therefore run starts. A non-start flat can still be the `par`
// of a unite() in gw_link, whose find_halving then rewrites its entry, but
// only with a run start. So the roots are found from the run starts alone,
// every other flat takes its run start's root, and F[A[p]] is defined for
// every flat (F is filled for every run start).
__global__ void gw_births(Grid g, Ws w) {
  GW_FOR_PIXELS {
    const long long p = j * g.nx + i;
    if ((g.flats[p] & 1) && !g.flat(i - 1, j)) w.F[p] = find_halving(w.A, (unsigned)p);
  }
}

__global__ void gw_births_store(Grid g, Ws w) {
  GW_FOR_PIXELS {
    const long long p = j * g.nx + i;
    if (g.flats[p] & 1) w.A[p] = w.F[g.flat(i - 1, j) ? w.A[p] : (unsigned)p];
  }
}

__global__ void gw_components(Grid g, Ws w) {
  GW_FOR_PIXELS {
    const long long p = j * g.nx + i;
    long long a_rep;
    bool nw;
    if (!(g.flats[p] & 1) || !merge_site(g, i, j, &a_rep, &nw)) continue;
    const unsigned a = w.A[a_rep], b = w.A[p - g.nx + 1];
    if (a != b) unite(w.C, a, b);
  }
}

// Component root of every birth, through H and back into C (as above).
__global__ void gw_component_roots(Grid g, Ws w) {
  GW_FOR_PIXELS {
    const long long p = j * g.nx + i;
    if ((g.flats[p] & 1) && w.A[p] == (unsigned)p) w.H[p] = find_halving(w.C, (unsigned)p);
  }
}

__global__ void gw_component_store(Grid g, Ws w) {
  GW_FOR_PIXELS {
    const long long p = j * g.nx + i;
    if ((g.flats[p] & 1) && w.A[p] == (unsigned)p) w.C[p] = w.H[p];
  }
}

__device__ __forceinline__ float depth_at(const float* original, const float* filled,
                                          long long p) {
  return filled[p] - original[p];
}

// Whether p is one of the deepest pixels of its component of kind k.
__device__ __forceinline__ bool deepest_pixel(unsigned k, float d, unsigned wbits) {
  return k == kHasNaN ? d != d : d == gw_unord(wbits);  // float ==: -0 and +0 tie
}

// Per component: kind bits and largest depth (gw_depth_max_runs); the
// number of deepest pixels and the first of them, and per birth the first
// deepest pixel of its chain group (gw_deepest_runs). One wavefront per
// block walks kRunChunks consecutive 64-pixel chunks of a row and keeps a
// running entry (wave-uniform) per target array, flushed with the atomics
// only when the key changes. Flats come in long runs of one component, so
// a lake's waves no longer all hit its words once per 64 pixels; the
// atomics are associative and commutative, so the words end the same.
constexpr int kRunChunks = 16;

__global__ __launch_bounds__(64) void gw_depth_max_runs(Grid g, Ws w, const float* original,
                                                        const float* filled) {
  const int lane = threadIdx.x;
  for (long long j = blockIdx.y; j < g.ny; j += gridDim.y) {
    unsigned ck = kNone, co = 0u, cm = 0u;
    auto flush = [&]() {
      if (ck != kNone && lane == 0) {
        if (co) atomicOr(w.K + ck, co);
        if (cm) atomicMax(w.W + ck, cm);
      }
    };
    const long long ibase = (long long)blockIdx.x * 64 * kRunChunks;
    for (int t = 0; t < kRunChunks; t++) {
      const long long i0 = ibase + 64LL * t;
      if (i0 >= g.nx) break;
      const long long i = i0 + lane;
      const long long p = j * g.nx + i;
      unsigned c = kNone, kind = 0u, wb = 0u;
      if (i < g.nx && (g.flats[p] & 1)) {
        const float d = depth_at(original, filled, p);
        c = w.C[w.A[p]];
        kind = d != d ? kHasNaN : kHasNumber;
        wb = d != d ? 0u : gw_ord(d);
      }
      unsigned long long pending = __ballot(c != kNone);
      while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const unsigned k0 = __shfl(c, leader);
        const bool mine = c == k0;
        unsigned o = mine ? kind : 0u, m = mine ? wb : 0u;
        for (int dd = 32; dd > 0; dd >>= 1) {
          o |= __shfl_xor(o, dd);
          m = max(m, __shfl_xor(m, dd));
        }
        if (k0 == ck) {
          co |= o;
          cm = max(cm, m);
        } else {
          flush();
          ck = k0;
          co = o;
          cm = m;
        }
        pending &= ~__ballot(mine);
      }
    }
    flush();
  }
}

__global__ __launch_bounds__(64) void gw_deepest_runs(Grid g, Ws w, const float* original,
                                                      const float* filled) {
  const int lane = threadIdx.x;
  for (long long j = blockIdx.y; j < g.ny; j += gridDim.y) {
    // running entries: component (count of deepest pixels, first one) and
    // birth (first deepest pixel of its chain group)
    unsigned ck = kNone, ccount = 0u, cmin = kNone, bk = kNone, bmin = kNone;
    auto flush_c = [&]() {
      if (ck != kNone && lane == 0) {
        atomicAdd(w.Mc + ck, ccount);
        atomicMin(w.Mf + ck, cmin);
      }
    };
    auto flush_b = [&]() {
      if (bk != kNone && lane == 0) atomicMin(w.D + bk, bmin);
    };
    const long long ibase = (long long)blockIdx.x * 64 * kRunChunks;
    for (int t = 0; t < kRunChunks; t++) {
      const long long i0 = ibase + 64LL * t;
      if (i0 >= g.nx) break;
      const long long i = i0 + lane;
      const long long p = j * g.nx + i;
      unsigned c = kNone, b = kNone;
      if (i < g.nx && (g.flats[p] & 1)) {
        b = w.A[p];
        const unsigned cc = w.C[b], k = w.K[cc];
        if (k != kMixed && deepest_pixel(k, depth_at(original, filled, p), w.W[cc])) c = cc;
        else b = kNone;
      }
      unsigned long long pending = __ballot(c != kNone);
      while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const unsigned k0 = __shfl(c, leader);
        const unsigned long long group = __ballot(c == k0);
        unsigned m = c == k0 ? (unsigned)p : kNone;
        for (int dd = 32; dd > 0; dd >>= 1) m = min(m, __shfl_xor(m, dd));
        const unsigned n = (unsigned)__popcll(group);
        if (k0 == ck) {
          ccount += n;
          cmin = min(cmin, m);
        } else {
          flush_c();
          ck = k0;
          ccount = n;
          cmin = m;
        }
        pending &= ~group;
      }
      pending = __ballot(b != kNone);
      while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const unsigned k0 = __shfl(b, leader);
        const unsigned long long group = __ballot(b == k0);
        unsigned m = b == k0 ? (unsigned)p : kNone;
        for (int dd = 32; dd > 0; dd >>= 1) m = min(m, __shfl_xor(m, dd));
        if (k0 == bk) {
          bmin = min(bmin, m);
        } else {
          flush_b();
          bk = k0;
          bmin = m;
        }
        pending &= ~group;
      }
    }
    flush_c();
    flush_b();
  }
}

// Scan steps to replay (components with two or more deepest pixels; every
// pixel of a mixed component), as keys component << 32 | pixel: counted
// when keys == null, else written.
__global__ void gw_events(Grid g, Ws w, unsigned long long* keys, unsigned* count) {
  GW_FOR_PIXELS {
    const long long p = j * g.nx + i;
    if (!(g.flats[p] & 1)) continue;
    const unsigned b = w.A[p], c = w.C[b];
    const bool mixed = w.K[c] == kMixed;
    if (!mixed && w.Mc[c] < 2u) continue;
    bool ev = mixed || b == (unsigned)p || w.D[b] == (unsigned)p;
    long long a_rep;
    bool nw;
    if (!ev && merge_site(g, i, j, &a_rep, &nw)) ev = w.A[a_rep] != w.A[p - g.nx + 1];
    if (!ev) continue;
    const unsigned k = atomicAdd(count, 1u);
    if (keys) keys[k] = ((unsigned long long)c << 32) | (unsigned long long)p;
  }
}

__device__ unsigned find_seq(unsigned* P, unsigned x) {
  while (P[x] != x) {
    const unsigned px = P[x];
    P[x] = P[px];
    x = px;
  }
  return x;
}

// The reference's scan over every pixel of a mixed component, in scan
// order: each pixel starts a set and is merged with its flat scanned
// neighbours in order; a merge keeps the neighbour's root unless the
// current root is strictly deeper. F is the union-find, H the root (label)
// of each set.
__device__ void replay_full(const Grid& g, const Ws& w, const unsigned long long* keys,
                            unsigned k, unsigned e, const float* original, const float* filled) {
  const unsigned c = (unsigned)(keys[k] >> 32);
  for (unsigned q = k; q < e; q++) {
    const unsigned p = (unsigned)keys[q];
    const long long i = (long long)p % g.nx, j = (long long)p / g.nx;
    w.F[p] = p;
    w.H[p] = p;
    const long long nb[4][2] = {{i - 1, j - 1}, {i, j - 1}, {i + 1, j - 1}, {i - 1, j}};
    for (int t = 0; t < 4; t++) {
      if (!g.flat(nb[t][0], nb[t][1])) continue;
      const unsigned a = find_seq(w.F, (unsigned)(nb[t][1] * g.nx + nb[t][0]));
      const unsigned s = find_seq(w.F, p);
      if (a == s) continue;
      const unsigned ra = w.H[a], rs = w.H[s];
      const bool current_deeper = depth_at(original, filled, ra) < depth_at(original, filled, rs);
      w.F[s] = a;
      w.H[a] = current_deeper ? rs : ra;
    }
  }
  w.Mf[c] = w.H[find_seq(w.F, c)];
}

// One thread per component: the replay of its scan steps, in scan order.
__global__ void gw_replay(Grid g, Ws w, const unsigned long long* keys, unsigned n,
                          const float* original, const float* filled) {
  for (unsigned k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const unsigned c = (unsigned)(keys[k] >> 32);
    if (k > 0 && (unsigned)(keys[k - 1] >> 32) == c) continue;  // not the first of its component
    unsigned e = k;
    while (e < n && (unsigned)(keys[e] >> 32) == c) e++;
    const unsigned kind = w.K[c];
    if (kind == kMixed) {
      replay_full(g, w, keys, k, e, original, filled);
      continue;
    }
    const unsigned wbits = w.W[c];
    for (unsigned q = k; q < e; q++) {
      const unsigned p = (unsigned)keys[q];
      const long long i = (long long)p % g.nx, j = (long long)p / g.nx;
      const bool pm = deepest_pixel(kind, depth_at(original, filled, p), wbits);
      const unsigned b = w.A[p];
      if (b == p) {
        w.F[p] = p;
        w.H[p] = pm ? p : kNone;
        continue;  // a birth is the first deepest pixel of its own group if it is one
      }
      long long a_rep;
      bool nw;
      if (merge_site(g, i, j, &a_rep, &nw)) {
        unsigned a = w.A[a_rep], bb = w.A[p - g.nx + 1];
        if (a != bb) {
          a = find_seq(w.F, a);
          bb = find_seq(w.F, bb);
          if (a != bb) {
            const unsigned first = nw ? a : bb, second = nw ? bb : a;
            const unsigned r = w.H[second] != kNone ? w.H[second]
                               : w.H[first] != kNone ? w.H[first]
                               : pm                  ? p
                                                     : kNone;
            w.F[first] = second;
            w.H[second] = r;
          }
        }
      }
      if (w.D[b] == p) {
        const unsigned s = find_seq(w.F, b);
        if (w.H[s] == kNone) w.H[s] = p;
      }
    }
    w.Mf[c] = w.H[find_seq(w.F, c)];
  }
}

// Every pixel's outputs: 0 and 0 for non-flats; for flats the component's
// label and (root depth - own depth, in float) + 0.1 rounded from double,
// as the reference.
__global__ void gw_final(Grid g, Ws w, const float* original, const float* filled, float* costs,
                         int64_t* conncomps) {
  GW_FOR_PIXELS {
    const long long p = j * g.nx + i;
    if (!(g.flats[p] & 1)) {
      costs[p] = 0.0f;
      conncomps[p] = 0;
      continue;
    }
    const unsigned label = w.Mf[w.C[w.A[p]]];
    const float gap = depth_at(original, filled, label) - depth_at(original, filled, p);
    costs[p] = (float)((double)gap + 0.1);
    conncomps[p] = (int64_t)label;
  }
}

__global__ void gw_zero(unsigned* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

dim3 pixel_grid(long long nx, long long ny) {
  long long bx = (nx + 255) / 256;
  if (bx > 4096) bx = 4096;
  return dim3((unsigned)bx, (unsigned)(ny < 65535 ? ny : 65535), 1);
}

}  // namespace

hipError_t run_gwdt_computecosts(float* costs, int64_t* conncomps, const int32_t* flats,
                                 const float* original, const float* filled, int64_t nx,
                                 int64_t ny, hipStream_t s) {
  if (nx <= 0 || ny <= 0) return hipSuccess;
  const size_t n = (size_t)nx * (size_t)ny;
  // per-pixel arrays in one workspace (the outputs are written only by the
  // last kernel)
  Ws w;
  unsigned* ws = nullptr;
  hipError_t e = hipMallocAsync((void**)&ws, (9 * n + 64) * sizeof(unsigned), s);
  if (e != hipSuccess) return e;
  w.A = ws;
  w.W = ws + n;
  w.Mc = ws + 2 * n;
  w.Mf = ws + 3 * n;
  w.D = ws + 4 * n;
  w.F = ws + 5 * n;
  w.H = ws + 6 * n;
  w.K = ws + 7 * n;
  w.C = ws + 8 * n;
  w.err = ws + 9 * n;  // [0] unused, [1] replay count
  Grid g{flats, nx, ny};
  const dim3 grid = pixel_grid(nx, ny);
  unsigned long long* keys = nullptr;
  unsigned long long* sorted = nullptr;
  void* temp = nullptr;
  unsigned nev = 0;
  auto done = [&](hipError_t r) {
    if (keys) (void)hipFreeAsync(keys, s);
    if (sorted) (void)hipFreeAsync(sorted, s);
    if (temp) (void)hipFreeAsync(temp, s);
    hipError_t f = hipFreeAsync(ws, s);
    return r != hipSuccess ? r : f;
  };
  gw_zero<<<1, 64, 0, s>>>(w.err, 2);
  gw_runs<<<dim3(1, (unsigned)(ny < 65535 ? ny : 65535)), 256, 0, s>>>(g, w);
  gw_link<<<grid, 256, 0, s>>>(g, w);
  gw_births<<<grid, 256, 0, s>>>(g, w);
  gw_births_store<<<grid, 256, 0, s>>>(g, w);
  gw_components<<<grid, 256, 0, s>>>(g, w);
  gw_component_roots<<<grid, 256, 0, s>>>(g, w);
  gw_component_store<<<grid, 256, 0, s>>>(g, w);
  {
    const long long span = 64LL * kRunChunks;
    const dim3 rgrid((unsigned)((nx + span - 1) / span), (unsigned)(ny < 65535 ? ny : 65535));
    gw_depth_max_runs<<<rgrid, 64, 0, s>>>(g, w, original, filled);
    gw_deepest_runs<<<rgrid, 64, 0, s>>>(g, w, original, filled);
  }
  unsigned* d_count = w.err + 1;
  gw_events<<<grid, 256, 0, s>>>(g, w, nullptr, d_count);
  if ((e = hipGetLastError()) != hipSuccess) return done(e);
  unsigned host[2] = {0, 0};
  e = hipMemcpyAsync(host, w.err, sizeof(host), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return done(e);
  nev = host[1];
  if (nev) {
    e = hipMallocAsync((void**)&keys, nev * sizeof(unsigned long long), s);
    if (e == hipSuccess) e = hipMallocAsync((void**)&sorted, nev * sizeof(unsigned long long), s);
    if (e != hipSuccess) return done(e);
    gw_zero<<<1, 64, 0, s>>>(d_count, 1);
    gw_events<<<grid, 256, 0, s>>>(g, w, keys, d_count);
    size_t temp_bytes = 0;
    e = hipcub::DeviceRadixSort::SortKeys(nullptr, temp_bytes, keys, sorted, (int)nev, 0, 64, s);
    if (e == hipSuccess) e = hipMallocAsync(&temp, temp_bytes ? temp_bytes : 1, s);
    if (e == hipSuccess)
      e = hipcub::DeviceRadixSort::SortKeys(temp, temp_bytes, keys, sorted, (int)nev, 0, 64, s);
    if (e != hipSuccess) return done(e);
    unsigned blocks = (nev + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    gw_replay<<<blocks, 256, 0, s>>>(g, w, sorted, nev, original, filled);
  }
  gw_final<<<grid, 256, 0, s>>>(g, w, original, filled, costs, conncomps);
  return done(hipGetLastError());
}

}  // namespace tt
