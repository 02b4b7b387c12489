/*
 * Shared I/O for the reference harnesses (test infrastructure only).
 *
 * The harness programs drive the reference's own library and kernels,
 * compiled from where they lie under the reference tree, on the binary mesh
 * and Jacobi files written by oplus2_matlab_amd/mesh.py, and dump the results
 * in a flat binary layout for the fixture script.
 */
#ifndef HARNESS_COMMON_H
#define HARNESS_COMMON_H
#include <stdio.h>
#include <stdlib.h>

static void *h_read(FILE *f, size_t bytes) {
  void *p = malloc(bytes ? bytes : 1);
  if (bytes && fread(p, 1, bytes, f) != bytes) {
    printf("harness: short read\n");
    exit(2);
  }
  return p;
}

static FILE *h_open(const char *path, const char *mode) {
  FILE *f = fopen(path, mode);
  if (!f) {
    printf("harness: cannot open %s\n", path);
    exit(2);
  }
  return f;
}

static void h_write_f64(FILE *f, const double *v, size_t n) {
  if (fwrite(v, sizeof(double), n, f) != n) exit(3);
}

template <class T>
static void h_write_as_f64(FILE *f, const T *v, size_t n) {
  for (size_t i = 0; i < n; i++) {
    double d = (double)v[i];
    if (fwrite(&d, sizeof d, 1, f) != 1) exit(3);
  }
}
#endif
