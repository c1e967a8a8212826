/* A C99 caller of the drop-in entry points, written as a caller of the
   reference writes it: host buffers, ptrdiff_t dims with dims[0] fastest,
   void calls. Usage: dropin_caller nx ny in_prefix out_prefix. Reads
   in_prefix.dem (float32), .bc (uint8) and .slopes (float32); writes
   out_prefix.fill, .recon and .excess (float32). Exit status 0 when every
   call succeeded (tt_last_error() empty). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "topotoolbox_amd.h"

static int read_all(const char *prefix, const char *ext, void *buf, size_t bytes) {
  char path[1024];
  snprintf(path, sizeof path, "%s.%s", prefix, ext);
  FILE *f = fopen(path, "rb");
  if (!f) return 1;
  size_t got = fread(buf, 1, bytes, f);
  fclose(f);
  return got != bytes;
}

static int write_all(const char *prefix, const char *ext, const void *buf, size_t bytes) {
  char path[1024];
  snprintf(path, sizeof path, "%s.%s", prefix, ext);
  FILE *f = fopen(path, "wb");
  if (!f) return 1;
  size_t put = fwrite(buf, 1, bytes, f);
  fclose(f);
  return put != bytes;
}

int main(int argc, char **argv) {
  if (argc != 5) return 2;
  ptrdiff_t dims[2] = {atol(argv[1]), atol(argv[2])};
  size_t n = (size_t)dims[0] * (size_t)dims[1];
  float *dem = malloc(n * 4), *slopes = malloc(n * 4), *fill = malloc(n * 4);
  float *recon = malloc(n * 4), *excess = malloc(n * 4), *mask = malloc(n * 4);
  uint8_t *bc = malloc(n);
  if (!dem || !slopes || !fill || !recon || !excess || !mask || !bc) return 3;
  if (read_all(argv[3], "dem", dem, n * 4) || read_all(argv[3], "bc", bc, n) ||
      read_all(argv[3], "slopes", slopes, n * 4))
    return 4;
  if (has_topotoolbox() != 1) return 5;
  fillsinks(fill, dem, bc, dims);
  /* reconstruct by dilation of the DEM from a marker that is the DEM
     lowered by 10 m (the reference's reconstruct: marker updated in place) */
  for (size_t p = 0; p < n; p++) {
    recon[p] = dem[p] - 10.0f;
    mask[p] = dem[p];
  }
  reconstruct(recon, mask, dims);
  excesstopography_fsm2d(excess, dem, slopes, 30.0f, dims);
  if (tt_last_error()[0]) {
    fprintf(stderr, "dropin_caller: %s\n", tt_last_error());
    return 6;
  }
  if (write_all(argv[4], "fill", fill, n * 4) || write_all(argv[4], "recon", recon, n * 4) ||
      write_all(argv[4], "excess", excess, n * 4))
    return 7;
  printf("dropin_caller: %ld x %ld ok\n", (long)dims[0], (long)dims[1]);
  return 0;
}
