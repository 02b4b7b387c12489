/*
 * Jacobi iteration on the reference OP2 library (test infrastructure only).
 *
 * Usage: harness_jac problem.bin alpha beta niter out.bin [OP_PART_SIZE=n]
 *
 * The loop body is the reference's (jac1/jac.cpp:145-163, or
 * jac1/jac_op.cpp for the OpenMP build) on a problem read from the binary
 * file written by oplus2_matlab_amd/mesh.py.
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

real alpha;

#ifdef HARNESS_SEQ
#include "op_seq.h"
#ifdef HARNESS_F64
#define float double
#endif
#include "res.h"
#include "update.h"
#ifdef HARNESS_F64
#undef float
#endif
#define LOOP(fn, name, set, ...) op_par_loop(fn, name, set, __VA_ARGS__)
#else
#include "op_lib.h"
void op_par_loop_res(char const *, op_set, op_arg, op_arg, op_arg, op_arg);
void op_par_loop_update(char const *, op_set, op_arg, op_arg, op_arg, op_arg,
                        op_arg);
#define LOOP(fn, name, set, ...) op_par_loop_##fn(name, set, __VA_ARGS__)
#endif

#include "harness_common.h"

int main(int argc, char **argv) {
  if (argc < 6) {
    printf("usage: %s problem.bin alpha beta niter out.bin [OP_PART_SIZE=n]\n", argv[0]);
    return 2;
  }
  FILE *fp = h_open(argv[1], "rb");
  int *hdr = (int *)h_read(fp, 4 * sizeof(int));
  const int nnode = hdr[2], nedge = hdr[3];
  int *pp = (int *)h_read(fp, sizeof(int) * 2 * nedge);
  double *A64 = (double *)h_read(fp, sizeof(double) * nedge);
  double *r64 = (double *)h_read(fp, sizeof(double) * nnode);
  double *u64 = (double *)h_read(fp, sizeof(double) * nnode);
  double *du64 = (double *)h_read(fp, sizeof(double) * nnode);
  fclose(fp);
  alpha = (real)atof(argv[2]);
  real beta = (real)atof(argv[3]);
  const int niter = atoi(argv[4]);

  real *A = (real *)malloc(sizeof(real) * nedge);
  real *r = (real *)malloc(sizeof(real) * nnode);
  real *u = (real *)malloc(sizeof(real) * nnode);
  real *du = (real *)malloc(sizeof(real) * nnode);
  for (int e = 0; e < nedge; e++) A[e] = (real)A64[e];
  for (int n = 0; n < nnode; n++) {
    r[n] = (real)r64[n];
    u[n] = (real)u64[n];
    du[n] = (real)du64[n];
  }

  op_init(argc - 5, argv + 5, 1);
  op_set nodes = op_decl_set(nnode, "nodes");
  op_set edges = op_decl_set(nedge, "edges");
  op_map ppedge = op_decl_map(edges, nodes, 2, pp, "ppedge");
  op_dat p_A = op_decl_dat(edges, 1, REAL_NAME, A, "p_A");
  op_dat p_r = op_decl_dat(nodes, 1, REAL_NAME, r, "p_r");
  op_dat p_u = op_decl_dat(nodes, 1, REAL_NAME, u, "p_u");
  op_dat p_du = op_decl_dat(nodes, 1, REAL_NAME, du, "p_du");

  double *hs = (double *)malloc(sizeof(double) * (niter > 0 ? niter : 1));
  double *hm = (double *)malloc(sizeof(double) * (niter > 0 ? niter : 1));
  real u_sum, u_max;
  for (int iter = 0; iter < niter; iter++) {
    LOOP(res, "res", edges,
         op_arg_dat(p_A, -1, OP_ID, 1, REAL_NAME, OP_READ),
         op_arg_dat(p_u, 1, ppedge, 1, REAL_NAME, OP_READ),
         op_arg_dat(p_du, 0, ppedge, 1, REAL_NAME, OP_INC),
         op_arg_gbl(&beta, 1, REAL_NAME, OP_READ));
    u_sum = 0.0f;
    u_max = 0.0f;
    LOOP(update, "update", nodes,
         op_arg_dat(p_r, -1, OP_ID, 1, REAL_NAME, OP_READ),
         op_arg_dat(p_du, -1, OP_ID, 1, REAL_NAME, OP_RW),
         op_arg_dat(p_u, -1, OP_ID, 1, REAL_NAME, OP_INC),
         op_arg_gbl(&u_sum, 1, REAL_NAME, OP_INC),
         op_arg_gbl(&u_max, 1, REAL_NAME, OP_MAX));
    hs[iter] = (double)u_sum;
    hm[iter] = (double)u_max;
  }

  FILE *fo = h_open(argv[5], "wb");
  int ohdr[4] = {niter, nnode, (int)sizeof(real), 0};
  fwrite(ohdr, sizeof(int), 4, fo);
  h_write_f64(fo, hs, niter);
  h_write_f64(fo, hm, niter);
  h_write_as_f64(fo, u, (size_t)nnode);
  h_write_as_f64(fo, du, (size_t)nnode);
  fclose(fo);
  return 0;
}
