/*
  CPU oracle for the sink-filling and excess-topography path.

  TEST INFRASTRUCTURE ONLY. Nothing in the libtopotoolbox_amd product
  links, loads or calls this file; only tests/, __graft_entry__.smoke()
  and bench.py's cpu_baseline leg use it, and only as the checker or as
  the timed CPU baseline.

  Contents (each function cites the reference lines it restates):

    oracle_pcg4d                 test/utils.c:6-28
    oracle_grid_random           test/grid.c:136-168
    oracle_reconstruct           src/reconstruct.c:68-112,134-202,270-279
    oracle_fillsinks             src/fillsinks.c:22-48
    oracle_fillsinks_pf          textbook Priority-Flood (Barnes et al. 2014,
                                 no epsilon), independent cross-check
    oracle_fillsinks_spill       Priority-Flood over the lexicographic
                                 (value, index) order: fill level plus the
                                 deterministic spill-cell index
    oracle_fillsinks_adjoint_f64 scatter of g_out onto spill cells, FP64
    oracle_fsm2d_f32             src/excesstopography.c:16-41,47-111
    oracle_fsm2d_f64             same scheme in double precision
    oracle_fmm2d_f64             src/excesstopography.c:113-172 (fast
                                 marching), double precision cross-check
    oracle_fsm2d_adjoint_f64     adjoint of the converged Godunov system
    oracle_identifyflats         src/identifyflats.c:32-121
    oracle_value_filter          src/value_filters.c:9-113 (min_filter,
                                 max_filter: op = -1 / +1)
    oracle_value_filter_square   src/value_filters.c:115-246
                                 (min_filter_square, max_filter_square)
    oracle_gwdt_computecosts     src/gwdt.c:83-274 (union-find over the
                                 raster scan, heaviest root kept)

  Pinning (which check fixes which function) is stated in the tests that
  exercise them: tests/test_oracle.py.

  Layout everywhere: p = j * dims[0] + i with dims[0] the fastest
  dimension (include/topotoolbox.h:284-290 in the reference).
*/
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Test-input generators                                              */
/* ------------------------------------------------------------------ */

/* PCG4D hash of Jarzynski & Olano (2020); test/utils.c:6-28. */
float oracle_pcg4d(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
  uint32_t x = a * 1664525u + 1013904223u;
  uint32_t y = b * 1664525u + 1013904223u;
  uint32_t z = c * 1664525u + 1013904223u;
  uint32_t w = d * 1664525u + 1013904223u;
  x += y * w;
  y += z * x;
  z += x * y;
  w += y * z;
  x ^= x >> 16;
  y ^= y >> 16;
  z ^= z >> 16;
  w ^= w >> 16;
  x += y * w;
  y += z * x;
  z += x * y;
  w += y * z;
  return (float)(w >> 8) / (1 << 24);
}

/* White-noise DEM of test/grid.c:136-168: 100 * pcg4d(i, j, hi, lo). */
void oracle_grid_random(float *out, uint64_t seed, const ptrdiff_t dims[2]) {
  uint32_t c = (uint32_t)(seed >> 32);
  uint32_t d = (uint32_t)(seed & 0xffffffffu);
  for (ptrdiff_t j = 0; j < dims[1]; j++)
    for (ptrdiff_t i = 0; i < dims[0]; i++)
      out[j * dims[0] + i] = 100.0f * oracle_pcg4d((uint32_t)i, (uint32_t)j, c, d);
}

/* ------------------------------------------------------------------ */
/* Grayscale reconstruction (restatement of the reference algorithm)  */
/* ------------------------------------------------------------------ */

/* One raster scan, j and i increasing; src/reconstruct.c:68-112.
   Causal neighbours (i+1,j-1), (i,j-1), (i-1,j-1), (i-1,j). */
static ptrdiff_t scan_forward(float *marker, const float *mask,
                              const ptrdiff_t dims[2]) {
  static const ptrdiff_t dj[4] = {-1, -1, -1, 0};
  static const ptrdiff_t di[4] = {1, 0, -1, -1};
  ptrdiff_t changed = 0;
  for (ptrdiff_t j = 0; j < dims[1]; j++) {
    for (ptrdiff_t i = 0; i < dims[0]; i++) {
      ptrdiff_t p = j * dims[0] + i;
      float top = marker[p];
      for (int k = 0; k < 4; k++) {
        ptrdiff_t ni = i + di[k], nj = j + dj[k];
        if (ni < 0 || ni >= dims[0] || nj < 0 || nj >= dims[1]) continue;
        top = fmaxf(top, marker[nj * dims[0] + ni]);
      }
      /* a NaN mask makes the comparison false and selects the mask */
      float z = top < mask[p] ? top : mask[p];
      if (z > marker[p]) changed++;
      marker[p] = z;
    }
  }
  return changed;
}

/* One anti-raster scan, j and i decreasing; src/reconstruct.c:134-202
   (the path without a queue). */
static ptrdiff_t scan_backward(float *marker, const float *mask,
                               const ptrdiff_t dims[2]) {
  static const ptrdiff_t dj[4] = {1, 1, 1, 0};
  static const ptrdiff_t di[4] = {-1, 0, 1, 1};
  ptrdiff_t changed = 0;
  for (ptrdiff_t j = dims[1] - 1; j >= 0; j--) {
    for (ptrdiff_t i = dims[0] - 1; i >= 0; i--) {
      ptrdiff_t p = j * dims[0] + i;
      float top = marker[p];
      for (int k = 0; k < 4; k++) {
        ptrdiff_t ni = i + di[k], nj = j + dj[k];
        if (ni < 0 || ni >= dims[0] || nj < 0 || nj >= dims[1]) continue;
        top = fmaxf(top, marker[nj * dims[0] + ni]);
      }
      float z = top < mask[p] ? top : mask[p];
      if (z > marker[p]) changed++;
      marker[p] = z;
    }
  }
  return changed;
}

/* Reconstruction by dilation of `mask` from `marker`, in place;
   src/reconstruct.c:270-279. The reference caps the loop at 1000
   iterations; `max_iterations` makes the cap a parameter (1000 reproduces
   the reference). Returns the number of iterations performed; *converged
   is set to 1 when the last iteration changed nothing. */
int64_t oracle_reconstruct(float *marker, const float *mask,
                           const ptrdiff_t dims[2], int32_t max_iterations,
                           int32_t *converged) {
  ptrdiff_t n = dims[0] * dims[1];
  int32_t it = 0;
  for (; it < max_iterations && n > 0; it++) {
    n = scan_forward(marker, mask, dims);
    n += scan_backward(marker, mask, dims);
  }
  if (converged) *converged = (n == 0);
  return it;
}

/* fillsinks, src/fillsinks.c:22-48: negate the DEM, seed the marker with
   the negated DEM on bc==1 cells and -INF elsewhere, reconstruct, negate
   both arrays back. `dem` is modified during the call and restored. */
int64_t oracle_fillsinks(float *output, float *dem, const uint8_t *bc,
                         const ptrdiff_t dims[2], int32_t max_iterations,
                         int32_t *converged) {
  ptrdiff_t n = dims[0] * dims[1];
  for (ptrdiff_t p = 0; p < n; p++) {
    dem[p] = -dem[p];
    output[p] = (bc[p] == 1) ? dem[p] : -INFINITY;
  }
  int64_t it = oracle_reconstruct(output, dem, dims, max_iterations, converged);
  for (ptrdiff_t p = 0; p < n; p++) {
    dem[p] = -dem[p];
    output[p] = -output[p];
  }
  return it;
}

/* ------------------------------------------------------------------ */
/* Priority-Flood over (value, index)                                 */
/* ------------------------------------------------------------------ */

typedef struct {
  float v;
  int64_t idx; /* tie-break key (lexicographic order) */
  int64_t cell;
} PFItem;

typedef struct {
  PFItem *a;
  int64_t n, cap;
} PFHeap;

static int pf_less(const PFItem *x, const PFItem *y) {
  if (x->v < y->v) return 1;
  if (x->v > y->v) return 0;
  return x->idx < y->idx;
}

static int pf_push(PFHeap *h, PFItem it) {
  if (h->n == h->cap) {
    int64_t nc = h->cap ? h->cap * 2 : 1024;
    PFItem *na = (PFItem *)realloc(h->a, (size_t)nc * sizeof(PFItem));
    if (!na) return -1;
    h->a = na;
    h->cap = nc;
  }
  int64_t k = h->n++;
  while (k > 0) {
    int64_t par = (k - 1) / 2;
    if (!pf_less(&it, &h->a[par])) break;
    h->a[k] = h->a[par];
    k = par;
  }
  h->a[k] = it;
  return 0;
}

static PFItem pf_pop(PFHeap *h) {
  PFItem top = h->a[0];
  PFItem last = h->a[--h->n];
  int64_t k = 0;
  for (;;) {
    int64_t c = 2 * k + 1;
    if (c >= h->n) break;
    if (c + 1 < h->n && pf_less(&h->a[c + 1], &h->a[c])) c++;
    if (!pf_less(&h->a[c], &last)) break;
    h->a[k] = h->a[c];
    k = c;
  }
  if (h->n > 0) h->a[k] = last;
  return top;
}

/* Fill level and spill index by Priority-Flood on the lexicographic
   (value, index) order, seeded at every bc==1 cell whose DEM is not NaN.
   The fill level of p is the min over 8-connected paths from p to a seed
   of the max DEM along the path (the greatest fixed point the reference
   reconstruction converges to); spill[p] is the index of the cell whose
   (dem, index) pair is the lexicographic max on the lexicographically
   minimal path. NaN cells are impassable and keep their NaN; cells no
   seed reaches get +INF and spill -1. `spill` may be NULL. */
int oracle_fillsinks_spill(float *output, int64_t *spill, const float *dem,
                           const uint8_t *bc, const ptrdiff_t dims[2]) {
  ptrdiff_t nx = dims[0], ny = dims[1], n = nx * ny;
  uint8_t *done = (uint8_t *)calloc((size_t)(n > 0 ? n : 1), 1);
  if (!done) return -1;
  PFHeap h = {0};
  for (ptrdiff_t p = 0; p < n; p++) {
    if (isnan(dem[p])) {
      output[p] = dem[p];
      if (spill) spill[p] = -1;
      done[p] = 1;
    } else {
      output[p] = INFINITY;
      if (spill) spill[p] = -1;
    }
  }
  for (ptrdiff_t p = 0; p < n; p++) {
    if (bc[p] == 1 && !isnan(dem[p])) {
      PFItem it = {dem[p], p, p};
      done[p] = 1;
      output[p] = dem[p];
      if (spill) spill[p] = p;
      if (pf_push(&h, it)) goto fail;
    }
  }
  while (h.n > 0) {
    PFItem c = pf_pop(&h);
    ptrdiff_t ci = c.cell % nx, cj = c.cell / nx;
    for (int dj = -1; dj <= 1; dj++) {
      for (int di = -1; di <= 1; di++) {
        if (!di && !dj) continue;
        ptrdiff_t qi = ci + di, qj = cj + dj;
        if (qi < 0 || qi >= nx || qj < 0 || qj >= ny) continue;
        ptrdiff_t q = qj * nx + qi;
        if (done[q]) continue;
        done[q] = 1;
        PFItem own = {dem[q], q, q};
        PFItem it = pf_less(&c, &own) ? own : c; /* lexicographic max */
        it.cell = q;
        output[q] = it.v;
        if (spill) spill[q] = it.idx;
        if (pf_push(&h, it)) goto fail;
      }
    }
  }
  free(h.a);
  free(done);
  return 0;
fail:
  free(h.a);
  free(done);
  return -1;
}

/* Plain Priority-Flood fill (no spill index). */
int oracle_fillsinks_pf(float *output, const float *dem, const uint8_t *bc,
                        const ptrdiff_t dims[2]) {
  return oracle_fillsinks_spill(output, NULL, dem, bc, dims);
}

/* FP64 adjoint of fillsinks: output[p] = dem[spill[p]], so
   g_dem[q] = sum over p with spill[p] == q of g_out[p], summed in
   increasing p. Cells with spill < 0 contribute nothing. */
void oracle_fillsinks_adjoint_f64(double *g_dem, const int64_t *spill,
                                  const double *g_out, const ptrdiff_t dims[2]) {
  ptrdiff_t n = dims[0] * dims[1];
  for (ptrdiff_t p = 0; p < n; p++) g_dem[p] = 0.0;
  for (ptrdiff_t p = 0; p < n; p++)
    if (spill[p] >= 0) g_dem[spill[p]] += g_out[p];
}

/* ------------------------------------------------------------------ */
/* Excess topography: fast sweeping (restatement)                     */
/* ------------------------------------------------------------------ */

/* Godunov upwind proposal for cell (i,j); src/excesstopography.c:16-41.
   Out-of-grid neighbours are +INF; fminf ignores NaN neighbours. */
static float godunov_f32(const float *u, float fi, ptrdiff_t i, ptrdiff_t j,
                         const ptrdiff_t dims[2]) {
  float s = i < dims[0] - 1 ? u[j * dims[0] + i + 1] : INFINITY;
  float nn = i > 0 ? u[j * dims[0] + i - 1] : INFINITY;
  float u1 = fminf(s, nn);
  float e = j < dims[1] - 1 ? u[(j + 1) * dims[0] + i] : INFINITY;
  float w = j > 0 ? u[(j - 1) * dims[0] + i] : INFINITY;
  float u2 = fminf(e, w);
  if (fabsf(u1 - u2) < fi) {
    return (u1 + u2) / 2 + sqrtf(-(u1 - u2) * (u1 - u2) + 2 * (fi * fi)) / 2;
  }
  return fminf(u1 + fi, u2 + fi);
}

#define FSM_VISIT(T, SOLVER)                                    \
  do {                                                          \
    ptrdiff_t p = j * dims[0] + i;                              \
    T fi = cellsize * slopes[p];                                \
    T prop = SOLVER(u, fi, i, j, dims);                         \
    if (prop < u[p]) {                                          \
      u[p] = prop;                                              \
      count++;                                                  \
    }                                                           \
  } while (0)

/* excesstopography_fsm2d, src/excesstopography.c:47-111: u = dem, then
   four alternating Gauss-Seidel sweeps (j+ i+, j- i+, j- i-, j+ i-) until a
   full iteration accepts no proposal. The reference has no iteration cap;
   `max_iterations` <= 0 means none. Returns the iterations performed. */
int64_t oracle_fsm2d_f32(float *u, const float *dem, const float *slopes,
                         float cellsize, const ptrdiff_t dims[2],
                         int64_t max_iterations) {
  ptrdiff_t n = dims[0] * dims[1];
  for (ptrdiff_t p = 0; p < n; p++) u[p] = dem[p];
  ptrdiff_t count = n;
  int64_t it = 0;
  while (count > 0 && (max_iterations <= 0 || it < max_iterations)) {
    count = 0;
    it++;
    for (ptrdiff_t j = 0; j < dims[1]; j++)
      for (ptrdiff_t i = 0; i < dims[0]; i++) FSM_VISIT(float, godunov_f32);
    for (ptrdiff_t j = dims[1] - 1; j >= 0; j--)
      for (ptrdiff_t i = 0; i < dims[0]; i++) FSM_VISIT(float, godunov_f32);
    for (ptrdiff_t j = dims[1] - 1; j >= 0; j--)
      for (ptrdiff_t i = dims[0] - 1; i >= 0; i--) FSM_VISIT(float, godunov_f32);
    for (ptrdiff_t j = 0; j < dims[1]; j++)
      for (ptrdiff_t i = dims[0] - 1; i >= 0; i--) FSM_VISIT(float, godunov_f32);
  }
  return it;
}

/* Same Godunov update in double precision. fmin ignores NaN exactly as
   fminf does. */
static double godunov_f64(const double *u, double fi, ptrdiff_t i, ptrdiff_t j,
                          const ptrdiff_t dims[2]) {
  double s = i < dims[0] - 1 ? u[j * dims[0] + i + 1] : INFINITY;
  double nn = i > 0 ? u[j * dims[0] + i - 1] : INFINITY;
  double u1 = fmin(s, nn);
  double e = j < dims[1] - 1 ? u[(j + 1) * dims[0] + i] : INFINITY;
  double w = j > 0 ? u[(j - 1) * dims[0] + i] : INFINITY;
  double u2 = fmin(e, w);
  if (fabs(u1 - u2) < fi) {
    return (u1 + u2) / 2 + sqrt(-(u1 - u2) * (u1 - u2) + 2 * (fi * fi)) / 2;
  }
  return fmin(u1 + fi, u2 + fi);
}

/* FP64 fast sweeping oracle. `slopes` and `dem` are double arrays;
   fi = cellsize * slope is formed in double. */
int64_t oracle_fsm2d_f64(double *u, const double *dem, const double *slopes,
                         double cellsize, const ptrdiff_t dims[2],
                         int64_t max_iterations) {
  ptrdiff_t n = dims[0] * dims[1];
  for (ptrdiff_t p = 0; p < n; p++) u[p] = dem[p];
  ptrdiff_t count = n;
  int64_t it = 0;
  while (count > 0 && (max_iterations <= 0 || it < max_iterations)) {
    count = 0;
    it++;
    for (ptrdiff_t j = 0; j < dims[1]; j++)
      for (ptrdiff_t i = 0; i < dims[0]; i++) FSM_VISIT(double, godunov_f64);
    for (ptrdiff_t j = dims[1] - 1; j >= 0; j--)
      for (ptrdiff_t i = 0; i < dims[0]; i++) FSM_VISIT(double, godunov_f64);
    for (ptrdiff_t j = dims[1] - 1; j >= 0; j--)
      for (ptrdiff_t i = dims[0] - 1; i >= 0; i--) FSM_VISIT(double, godunov_f64);
    for (ptrdiff_t j = 0; j < dims[1]; j++)
      for (ptrdiff_t i = dims[0] - 1; i >= 0; i--) FSM_VISIT(double, godunov_f64);
  }
  return it;
}

/* ------------------------------------------------------------------ */
/* Excess topography: fast marching cross-check (FP64)                */
/* ------------------------------------------------------------------ */

typedef struct {
  int64_t *heap; /* heap position -> cell */
  int64_t *pos;  /* cell -> heap position, -1 once removed */
  double *key;
  int64_t n;
} DHeap;

static int dh_less(const DHeap *h, int64_t a, int64_t b) {
  /* NaN keys sort last, as in src/helpers/priority_queue.c:39-41,57-71 */
  double x = h->key[a], y = h->key[b];
  if (isnan(x)) return 0;
  if (isnan(y)) return 1;
  return x < y || (x == y && a < b);
}
static void dh_swap(DHeap *h, int64_t i, int64_t j) {
  int64_t a = h->heap[i], b = h->heap[j];
  h->heap[i] = b;
  h->heap[j] = a;
  h->pos[a] = j;
  h->pos[b] = i;
}
static void dh_up(DHeap *h, int64_t k) {
  while (k > 0) {
    int64_t par = (k - 1) / 2;
    if (!dh_less(h, h->heap[k], h->heap[par])) break;
    dh_swap(h, k, par);
    k = par;
  }
}
static void dh_down(DHeap *h, int64_t k) {
  for (;;) {
    int64_t c = 2 * k + 1;
    if (c >= h->n) break;
    if (c + 1 < h->n && dh_less(h, h->heap[c + 1], h->heap[c])) c++;
    if (!dh_less(h, h->heap[c], h->heap[k])) break;
    dh_swap(h, k, c);
    k = c;
  }
}

/* Fast marching (Sethian 1996) as in src/excesstopography.c:113-172: every
   cell enters the heap at its DEM value; the minimum is finalised and its
   4-neighbours that are not yet finalised receive the Godunov proposal
   computed from the current values, accepted when lower. */
int oracle_fmm2d_f64(double *u, const double *dem, const double *slopes,
                     double cellsize, const ptrdiff_t dims[2]) {
  ptrdiff_t nx = dims[0], ny = dims[1], n = nx * ny;
  DHeap h;
  h.heap = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
  h.pos = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
  h.key = u;
  h.n = n;
  if (!h.heap || !h.pos) {
    free(h.heap);
    free(h.pos);
    return -1;
  }
  for (ptrdiff_t p = 0; p < n; p++) {
    u[p] = dem[p];
    h.heap[p] = p;
    h.pos[p] = p;
  }
  for (ptrdiff_t k = n / 2 - 1; k >= 0; k--) dh_down(&h, k);
  while (h.n > 0) {
    int64_t c = h.heap[0];
    dh_swap(&h, 0, h.n - 1);
    h.n--;
    h.pos[c] = -1;
    if (h.n > 0) dh_down(&h, 0);
    ptrdiff_t ci = c % nx, cj = c / nx;
    const int di[4] = {1, -1, 0, 0};
    const int dj[4] = {0, 0, 1, -1};
    for (int k = 0; k < 4; k++) {
      ptrdiff_t qi = ci + di[k], qj = cj + dj[k];
      if (qi < 0 || qi >= nx || qj < 0 || qj >= ny) continue;
      ptrdiff_t q = qj * nx + qi;
      if (h.pos[q] < 0) continue;
      double prop = godunov_f64(u, cellsize * slopes[q], qi, qj, dims);
      if (prop < u[q]) {
        u[q] = prop;
        dh_up(&h, h.pos[q]);
      }
    }
  }
  free(h.heap);
  free(h.pos);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Excess topography adjoint (FP64)                                   */
/* ------------------------------------------------------------------ */

/* Local linearisation of the converged state at cell p. Writes the
   upwind neighbour cells (or -1) and the partials of u_p with respect to
   them and to fi. Returns 0 for a constrained cell (u_p == dem_p, so
   du_p/ddem_p = 1) or a NaN cell, 1 for an eikonal cell.

   Selection rules (shared with the GPU adjoint):
     u1 = min(S, N), S = (i+1, j), N = (i-1, j); tie -> S
     u2 = min(E, W), E = (i, j+1), W = (i, j-1); tie -> E
     two-sided when |u1 - u2| < fi, else one-sided on min(u1, u2)
     (tie -> u1). Out-of-grid and NaN neighbours never win. */
static int fsm_linearise(const double *u, const double *dem, double fi,
                         ptrdiff_t i, ptrdiff_t j, const ptrdiff_t dims[2],
                         int64_t nb[2], double w[2], double *dfi) {
  ptrdiff_t nx = dims[0], ny = dims[1], p = j * nx + i;
  nb[0] = nb[1] = -1;
  w[0] = w[1] = 0.0;
  *dfi = 0.0;
  if (isnan(u[p]) || !(u[p] < dem[p])) return 0;
  double s = i < nx - 1 ? u[p + 1] : INFINITY;
  double nn = i > 0 ? u[p - 1] : INFINITY;
  double e = j < ny - 1 ? u[p + nx] : INFINITY;
  double ww = j > 0 ? u[p - nx] : INFINITY;
  int64_t a = -1, b = -1;
  double u1 = INFINITY, u2 = INFINITY;
  if (!isnan(s) && !(nn < s)) { u1 = s; a = (s < INFINITY) ? p + 1 : -1; }
  else if (!isnan(nn)) { u1 = nn; a = (nn < INFINITY) ? p - 1 : -1; }
  if (!isnan(e) && !(ww < e)) { u2 = e; b = (e < INFINITY) ? p + nx : -1; }
  else if (!isnan(ww)) { u2 = ww; b = (ww < INFINITY) ? p - nx : -1; }
  double d = u1 - u2;
  if (fabs(d) < fi) {
    double D = sqrt(2.0 * fi * fi - d * d);
    nb[0] = a;
    nb[1] = b;
    w[0] = (D - d) / (2.0 * D);
    w[1] = (D + d) / (2.0 * D);
    *dfi = fi / D;
  } else {
    if (u1 <= u2) { nb[0] = a; w[0] = 1.0; }
    else { nb[0] = b; w[0] = 1.0; }
    *dfi = 1.0;
  }
  return 1;
}

typedef struct {
  double u;
  int64_t p;
} UCell;

static int ucell_desc(const void *x, const void *y) {
  const UCell *a = (const UCell *)x, *b = (const UCell *)y;
  if (a->u > b->u) return -1;
  if (a->u < b->u) return 1;
  return (a->p < b->p) ? -1 : (a->p > b->p);
}

/* Adjoint of u = FSM(dem, slopes) at the converged state u. With
   L = sum g_out * u, solves lambda = g_out + J^T lambda in decreasing u
   (every upwind dependency points to a strictly lower value), then
   g_dem = lambda on constrained cells and
   g_slopes = lambda * du/dfi * cellsize on eikonal cells. */
int oracle_fsm2d_adjoint_f64(double *g_dem, double *g_slopes, const double *u,
                             const double *dem, const double *slopes,
                             double cellsize, const double *g_out,
                             const ptrdiff_t dims[2]) {
  ptrdiff_t nx = dims[0], ny = dims[1], n = nx * ny;
  UCell *order = (UCell *)malloc((size_t)(n > 0 ? n : 1) * sizeof(UCell));
  double *lam = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  if (!order || !lam) {
    free(order);
    free(lam);
    return -1;
  }
  ptrdiff_t m = 0;
  for (ptrdiff_t p = 0; p < n; p++) {
    lam[p] = g_out[p];
    g_dem[p] = 0.0;
    g_slopes[p] = 0.0;
    if (!isnan(u[p])) {
      order[m].u = u[p];
      order[m].p = p;
      m++;
    }
  }
  qsort(order, (size_t)m, sizeof(UCell), ucell_desc);
  for (ptrdiff_t k = 0; k < m; k++) {
    int64_t p = order[k].p;
    ptrdiff_t i = p % nx, j = p / nx;
    int64_t nb[2];
    double w[2], dfi;
    double fi = cellsize * slopes[p];
    if (fsm_linearise(u, dem, fi, i, j, dims, nb, w, &dfi)) {
      for (int t = 0; t < 2; t++)
        if (nb[t] >= 0) lam[nb[t]] += w[t] * lam[p];
      g_slopes[p] = lam[p] * dfi * cellsize;
    } else {
      g_dem[p] = lam[p];
    }
  }
  free(order);
  free(lam);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Flats, sills, presills; value filters                              */
/* ------------------------------------------------------------------ */

/* Bit 0: flat (an interior cell equal to the minimum of its 3x3
   neighbourhood, NaN neighbours counting as -INF); bit 1: sill (a non-flat
   cell next to a flat of equal height); bit 2: presill (such a flat).
   Returns the number of flats. src/identifyflats.c:32-121. */
ptrdiff_t oracle_identifyflats(int32_t *out, const float *dem, const ptrdiff_t dims[2]) {
  const ptrdiff_t nx = dims[0], ny = dims[1];
  ptrdiff_t flats = 0;
  for (ptrdiff_t j = 0; j < ny; j++)
    for (ptrdiff_t i = 0; i < nx; i++) {
      const ptrdiff_t p = j * nx + i;
      out[p] = 0;
      if (i == 0 || j == 0 || i == nx - 1 || j == ny - 1) continue;
      float lowest = dem[p];
      for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++) {
          if (!di && !dj) continue;
          const float h = dem[p + dj * nx + di];
          lowest = fminf(lowest, isnan(h) ? -INFINITY : h);
        }
      if (dem[p] == lowest) {
        out[p] |= 1;
        flats++;
      }
    }
  for (ptrdiff_t j = 0; j < ny; j++)
    for (ptrdiff_t i = 0; i < nx; i++) {
      const ptrdiff_t p = j * nx + i;
      if (out[p] & 1) continue;
      for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++) {
          if (!di && !dj) continue;
          const ptrdiff_t ii = i + di, jj = j + dj;
          if (ii < 0 || ii >= nx || jj < 0 || jj >= ny) continue;
          const ptrdiff_t q = jj * nx + ii;
          if ((out[q] & 1) && dem[q] == dem[p]) {
            out[p] |= 2;
            out[q] |= 4;
          }
        }
    }
  return flats;
}

/* Minimum (op < 0) or maximum (op > 0) over the cells under a structuring
   element of se_dims[0] x se_dims[1] x se_dims[2] (slices visited in turn)
   centred at ((se_dims[0] + 1) / 2 - 1, (se_dims[1] + 1) / 2 - 1); NaN
   cells stay NaN, NaN neighbours and cells outside the grid are skipped,
   the running value is replaced unless it is strictly better (so equal
   values, e.g. -0 and +0, take the later one). src/value_filters.c:9-113. */
void oracle_value_filter(float *out, const float *dem, const uint8_t *se,
                         const ptrdiff_t dims[2], const ptrdiff_t se_dims[3], int op) {
  const ptrdiff_t nx = dims[0], ny = dims[1];
  const ptrdiff_t cx = (se_dims[0] + 1) / 2 - 1, cy = (se_dims[1] + 1) / 2 - 1;
  for (ptrdiff_t j = 0; j < ny; j++)
    for (ptrdiff_t i = 0; i < nx; i++) {
      const ptrdiff_t p = j * nx + i;
      if (isnan(dem[p])) {
        out[p] = NAN;
        continue;
      }
      float v = op < 0 ? INFINITY : -INFINITY;
      for (ptrdiff_t s = 0; s < se_dims[2]; s++)
        for (ptrdiff_t b = 0; b < se_dims[1]; b++) {
          const ptrdiff_t jj = j + b - cy;
          if (jj < 0 || jj >= ny) continue;
          for (ptrdiff_t a = 0; a < se_dims[0]; a++) {
            if (!se[a + b * se_dims[0] + s * se_dims[0] * se_dims[1]]) continue;
            const ptrdiff_t ii = i + a - cx;
            if (ii < 0 || ii >= nx) continue;
            const float h = dem[jj * nx + ii];
            if (isnan(h)) continue;
            if (op < 0 ? (v < h) : (v > h)) continue;
            v = h;
          }
        }
      out[p] = v;
    }
}

/* Square, full structuring element of the given width, separably: first
   along dims[0] into tmp (NaN inputs skipped), then along dims[1]
   (NaN cells of the input stay NaN). src/value_filters.c:115-246. */
void oracle_value_filter_square(float *out, const float *dem, float *tmp, int width,
                                const ptrdiff_t dims[2], int op) {
  const ptrdiff_t nx = dims[0], ny = dims[1], c = (width + 1) / 2 - 1;
  for (ptrdiff_t j = 0; j < ny; j++)
    for (ptrdiff_t i = 0; i < nx; i++) {
      float v = op < 0 ? INFINITY : -INFINITY;
      for (ptrdiff_t k = 0; k < width; k++) {
        const ptrdiff_t ii = i + k - c;
        if (ii < 0 || ii >= nx) continue;
        const float h = dem[j * nx + ii];
        if (isnan(h)) continue;
        if (op < 0 ? (v < h) : (v > h)) continue;
        v = h;
      }
      tmp[j * nx + i] = v;
    }
  for (ptrdiff_t j = 0; j < ny; j++)
    for (ptrdiff_t i = 0; i < nx; i++) {
      const ptrdiff_t p = j * nx + i;
      if (isnan(dem[p])) {
        out[p] = NAN;
        continue;
      }
      float v = op < 0 ? INFINITY : -INFINITY;
      for (ptrdiff_t k = 0; k < width; k++) {
        const ptrdiff_t jj = j + k - c;
        if (jj < 0 || jj >= ny) continue;
        const float h = tmp[jj * nx + i];
        if (isnan(h)) continue;
        if (op < 0 ? (v < h) : (v > h)) continue;
        v = h;
      }
      out[p] = v;
    }
}


/* ------------------------------------------------------------------ */
/* Costs of the gray-weighted distance transform                      */
/* ------------------------------------------------------------------ */

/* Union-find stored in (label, weight) arrays: label[p] == p marks a root,
   whose weight is the largest depth of its set. src/gwdt.c:83-152. */
static ptrdiff_t gw_root(const ptrdiff_t *label, ptrdiff_t p) {
  while (label[p] != p) p = label[p];
  return p;
}

/* Points every node on the path from p to its root at `root`. */
static void gw_compress(ptrdiff_t *label, float *weight, ptrdiff_t p, ptrdiff_t root,
                        float w) {
  for (;;) {
    const ptrdiff_t next = label[p];
    label[p] = root;
    weight[p] = w;
    if (next == p) break;
    p = next;
  }
}

/* Joins the sets of a (already scanned neighbour) and p (current pixel):
   the neighbour's root stays unless p's root is strictly heavier. */
static void gw_unify(ptrdiff_t *label, float *weight, ptrdiff_t a, ptrdiff_t p) {
  ptrdiff_t r = gw_root(label, a);
  float rw = weight[r];
  if (r != p) {
    const ptrdiff_t s = gw_root(label, p);
    if (rw < weight[s]) {
      r = s;
      rw = weight[s];
    }
    gw_compress(label, weight, p, r, rw);
  }
  gw_compress(label, weight, a, r, rw);
}

/* Cost of every flat: (max depth of its 8-connected flat component - its
   own depth) + 0.1, the sum taken in double and rounded to float; depth =
   filled - original in float. conncomps holds the component's root, the
   pixel whose depth is the largest, chosen by the scan order and tie rule
   above. Non-flat cells get 0 and 0. src/gwdt.c:166-274. */
void oracle_gwdt_computecosts(float *costs, ptrdiff_t *conncomps, const int32_t *flats,
                              const float *original, const float *filled,
                              const ptrdiff_t dims[2]) {
  const ptrdiff_t nx = dims[0], ny = dims[1];
  /* scanned neighbours, in the order the scan unifies them */
  static const int dj[4] = {-1, -1, -1, 0};
  static const int di[4] = {-1, 0, 1, -1};
  for (ptrdiff_t j = 0; j < ny; j++)
    for (ptrdiff_t i = 0; i < nx; i++) {
      const ptrdiff_t p = j * nx + i;
      if (!(flats[p] & 1)) {
        costs[p] = 0.0f;
        conncomps[p] = 0;
        continue;
      }
      conncomps[p] = p;
      costs[p] = filled[p] - original[p];
      for (int k = 0; k < 4; k++) {
        const ptrdiff_t jj = j + dj[k], ii = i + di[k];
        if (jj < 0 || jj >= ny || ii < 0 || ii >= nx) continue;
        const ptrdiff_t q = jj * nx + ii;
        if (flats[q] & 1) gw_unify(conncomps, costs, q, p);
      }
    }
  for (ptrdiff_t j = 1; j < ny - 1; j++)
    for (ptrdiff_t i = 1; i < nx - 1; i++) {
      const ptrdiff_t p = j * nx + i;
      if (!(flats[p] & 1)) continue;
      const ptrdiff_t r = gw_root(conncomps, p);
      conncomps[p] = r;
      costs[p] = costs[r];
    }
  for (ptrdiff_t j = 1; j < ny - 1; j++)
    for (ptrdiff_t i = 1; i < nx - 1; i++) {
      const ptrdiff_t p = j * nx + i;
      if (!(flats[p] & 1)) continue;
      const float depth = filled[p] - original[p];
      costs[p] = (float)((double)(costs[p] - depth) + 0.1);
    }
}
