/*
 * Airfoil time march on the reference OP2 library (test infrastructure only).
 *
 * Usage: harness_airfoil mesh.bin consts.bin niter out.bin [OP_PART_SIZE=n]
 *
 * The loop body is the reference's (airfoil/airfoil.cpp:195-257 for the
 * sequential build, airfoil/airfoil_op.cpp for the OpenMP build); the mesh
 * and constants come from files instead of new_grid.dat so that every
 * arithmetic form starts from identical bits.  HARNESS_F64 compiles the
 * reference's unchanged kernel headers with float spelled double.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef HARNESS_F64
typedef double real;
#define REAL_NAME "double"
#else
typedef float real;
#define REAL_NAME "float"
#endif

real gam, gm1, cfl, eps, mach, alpha, qinf[4];

#ifdef HARNESS_SEQ
#include "op_seq.h"
#ifdef HARNESS_F64
#define float double
#endif
#include "save_soln.h"
#include "adt_calc.h"
#include "res_calc.h"
#include "bres_calc.h"
#include "update.h"
#ifdef HARNESS_F64
#undef float
#endif
#define LOOP(fn, name, set, ...) op_par_loop(fn, name, set, __VA_ARGS__)
#else
#include "op_lib.h"
void op_par_loop_save_soln(char const *, op_set, op_arg, op_arg);
void op_par_loop_adt_calc(char const *, op_set, op_arg, op_arg, op_arg, op_arg,
                          op_arg, op_arg);
void op_par_loop_res_calc(char const *, op_set, op_arg, op_arg, op_arg, op_arg,
                          op_arg, op_arg, op_arg, op_arg);
void op_par_loop_bres_calc(char const *, op_set, op_arg, op_arg, op_arg,
                           op_arg, op_arg, op_arg);
void op_par_loop_update(char const *, op_set, op_arg, op_arg, op_arg, op_arg,
                        op_arg);
#define LOOP(fn, name, set, ...) op_par_loop_##fn(name, set, __VA_ARGS__)
#endif

#include "harness_common.h"

int main(int argc, char **argv) {
  if (argc < 5) {
    printf("usage: %s mesh.bin consts.bin niter out.bin [OP_PART_SIZE=n]\n", argv[0]);
    return 2;
  }
  FILE *fm = h_open(argv[1], "rb");
  int *hdr = (int *)h_read(fm, 8 * sizeof(int));
  const int nnode = hdr[2], ncell = hdr[3], nedge = hdr[4], nbedge = hdr[5];
  double *x64 = (double *)h_read(fm, sizeof(double) * 2 * nnode);
  int *cell = (int *)h_read(fm, sizeof(int) * 4 * ncell);
  int *edge = (int *)h_read(fm, sizeof(int) * 2 * nedge);
  int *ecell = (int *)h_read(fm, sizeof(int) * 2 * nedge);
  int *bedge = (int *)h_read(fm, sizeof(int) * 2 * nbedge);
  int *becell = (int *)h_read(fm, sizeof(int) * nbedge);
  int *bound = (int *)h_read(fm, sizeof(int) * nbedge);
  fclose(fm);
  FILE *fc = h_open(argv[2], "rb");
  double *c = (double *)h_read(fc, 10 * sizeof(double));
  fclose(fc);
  const int niter = atoi(argv[3]);

  gam = (real)c[0]; gm1 = (real)c[1]; cfl = (real)c[2]; eps = (real)c[3];
  mach = (real)c[4]; alpha = (real)c[5];
  for (int i = 0; i < 4; i++) qinf[i] = (real)c[6 + i];

  real *x = (real *)malloc(sizeof(real) * 2 * nnode);
  real *q = (real *)malloc(sizeof(real) * 4 * ncell);
  real *qold = (real *)malloc(sizeof(real) * 4 * ncell);
  real *res = (real *)malloc(sizeof(real) * 4 * ncell);
  real *adt = (real *)malloc(sizeof(real) * ncell);
  for (int n = 0; n < 2 * nnode; n++) x[n] = (real)x64[n];
  for (int n = 0; n < ncell; n++)
    for (int m = 0; m < 4; m++) {
      q[4 * n + m] = qinf[m];
      res[4 * n + m] = 0;
      qold[4 * n + m] = 0;
    }
  for (int n = 0; n < ncell; n++) adt[n] = 0;

  /* argv[5..] are handed to op_init (OP_PART_SIZE=..., OP_BLOCK_SIZE=...) */
  op_init(argc - 4, argv + 4, 1);

  op_set nodes = op_decl_set(nnode, "nodes");
  op_set edges = op_decl_set(nedge, "edges");
  op_set bedges = op_decl_set(nbedge, "bedges");
  op_set cells = op_decl_set(ncell, "cells");
  op_map pedge = op_decl_map(edges, nodes, 2, edge, "pedge");
  op_map pecell = op_decl_map(edges, cells, 2, ecell, "pecell");
  op_map pbedge = op_decl_map(bedges, nodes, 2, bedge, "pbedge");
  op_map pbecell = op_decl_map(bedges, cells, 1, becell, "pbecell");
  op_map pcell = op_decl_map(cells, nodes, 4, cell, "pcell");
  op_dat p_bound = op_decl_dat(bedges, 1, "int", bound, "p_bound");
  op_dat p_x = op_decl_dat(nodes, 2, REAL_NAME, x, "p_x");
  op_dat p_q = op_decl_dat(cells, 4, REAL_NAME, q, "p_q");
  op_dat p_qold = op_decl_dat(cells, 4, REAL_NAME, qold, "p_qold");
  op_dat p_adt = op_decl_dat(cells, 1, REAL_NAME, adt, "p_adt");
  op_dat p_res = op_decl_dat(cells, 4, REAL_NAME, res, "p_res");

  double *hist = (double *)malloc(sizeof(double) * (niter > 0 ? niter : 1));
  real rms;
  for (int iter = 1; iter <= niter; iter++) {
    LOOP(save_soln, "save_soln", cells,
         op_arg_dat(p_q, -1, OP_ID, 4, REAL_NAME, OP_READ),
         op_arg_dat(p_qold, -1, OP_ID, 4, REAL_NAME, OP_WRITE));
    for (int k = 0; k < 2; k++) {
      LOOP(adt_calc, "adt_calc", cells,
           op_arg_dat(p_x, 0, pcell, 2, REAL_NAME, OP_READ),
           op_arg_dat(p_x, 1, pcell, 2, REAL_NAME, OP_READ),
           op_arg_dat(p_x, 2, pcell, 2, REAL_NAME, OP_READ),
           op_arg_dat(p_x, 3, pcell, 2, REAL_NAME, OP_READ),
           op_arg_dat(p_q, -1, OP_ID, 4, REAL_NAME, OP_READ),
           op_arg_dat(p_adt, -1, OP_ID, 1, REAL_NAME, OP_WRITE));
      LOOP(res_calc, "res_calc", edges,
           op_arg_dat(p_x, 0, pedge, 2, REAL_NAME, OP_READ),
           op_arg_dat(p_x, 1, pedge, 2, REAL_NAME, OP_READ),
           op_arg_dat(p_q, 0, pecell, 4, REAL_NAME, OP_READ),
           op_arg_dat(p_q, 1, pecell, 4, REAL_NAME, OP_READ),
           op_arg_dat(p_adt, 0, pecell, 1, REAL_NAME, OP_READ),
           op_arg_dat(p_adt, 1, pecell, 1, REAL_NAME, OP_READ),
           op_arg_dat(p_res, 0, pecell, 4, REAL_NAME, OP_INC),
           op_arg_dat(p_res, 1, pecell, 4, REAL_NAME, OP_INC));
      LOOP(bres_calc, "bres_calc", bedges,
           op_arg_dat(p_x, 0, pbedge, 2, REAL_NAME, OP_READ),
           op_arg_dat(p_x, 1, pbedge, 2, REAL_NAME, OP_READ),
           op_arg_dat(p_q, 0, pbecell, 4, REAL_NAME, OP_READ),
           op_arg_dat(p_adt, 0, pbecell, 1, REAL_NAME, OP_READ),
           op_arg_dat(p_res, 0, pbecell, 4, REAL_NAME, OP_INC),
           op_arg_dat(p_bound, -1, OP_ID, 1, "int", OP_READ));
      rms = 0.0;
      LOOP(update, "update", cells,
           op_arg_dat(p_qold, -1, OP_ID, 4, REAL_NAME, OP_READ),
           op_arg_dat(p_q, -1, OP_ID, 4, REAL_NAME, OP_WRITE),
           op_arg_dat(p_res, -1, OP_ID, 4, REAL_NAME, OP_RW),
           op_arg_dat(p_adt, -1, OP_ID, 1, REAL_NAME, OP_READ),
           op_arg_gbl(&rms, 1, REAL_NAME, OP_INC));
    }
    rms = sqrt(rms / (real)ncell);
    hist[iter - 1] = (double)rms;
  }

  FILE *fo = h_open(argv[4], "wb");
  int ohdr[4] = {niter, ncell, (int)sizeof(real), 0};
  fwrite(ohdr, sizeof(int), 4, fo);
  h_write_f64(fo, hist, niter);
  h_write_as_f64(fo, q, 4 * (size_t)ncell);
  h_write_as_f64(fo, qold, 4 * (size_t)ncell);
  h_write_as_f64(fo, adt, (size_t)ncell);
  h_write_as_f64(fo, res, 4 * (size_t)ncell);
  fclose(fo);
  return 0;
}
