/*
 * CPU oracle for the OP2 airfoil / Jacobi loops.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only the tests, the smoke check and the benchmark's CPU
 * baseline leg do, and only as the reference results are compared against or
 * as the timed CPU baseline.
 *
 * It restates, in plain C, the reference's sequential execution order
 * (common/op_seq.h:11-19,339-393), its execution-plan builder
 * (common/op_lib_core.c:627-1089) and the OpenMP schedule that runs on that
 * plan (airfoil/<loop>_kernel.cpp, jac1/<loop>_kernel.cpp).  The user-kernel arithmetic
 * follows airfoil/{save_soln,adt_calc,res_calc,bres_calc,update}.h and
 * jac1/{res,update}.h operation by operation.
 *
 * Each arithmetic form is a separate symbol set:
 *   _f64  state and arithmetic in double
 *   _f32  state and arithmetic in float (the reference's own program)
 *   _mix  state and flux arithmetic in float; update's qold - res/adt and the
 *         rms / u_sum / u_max reductions in double
 */
#ifndef OP2_ORACLE_H
#define OP2_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXARGS 24

enum { ORC_READ = 0, ORC_WRITE = 1, ORC_RW = 2, ORC_INC = 3, ORC_MIN = 4, ORC_MAX = 5 };

/* constants: gam, gm1, cfl, eps, mach, alpha, qinf[0..3] (always doubles;
 * the float forms convert them once) */
#define ORC_NCONST 10

/* ---- execution plan (restates common/op_lib_core.c:627-1089) ---- */

typedef struct {
  int set_size;
  int nargs;
  int ninds;
  int part_size;            /* 0 = adaptive size, op_lib_core.c:671-672 */
  /* per argument */
  const int *map[ORC_MAXARGS];  /* NULL for direct / global arguments */
  int map_dim[ORC_MAXARGS];
  int idx[ORC_MAXARGS];
  int to_size[ORC_MAXARGS];     /* size of the set the map points to */
  int dat_size[ORC_MAXARGS];    /* bytes per element of the dat, 0 if none */
  int dat_dim[ORC_MAXARGS];
  int acc[ORC_MAXARGS];
  int inds[ORC_MAXARGS];        /* indirect dataset number, -1 if direct */
} orc_plan_in;

typedef struct {
  int bsize, nblocks, ncolors, nshared, ninds, nargs, set_size;
  int *nthrcol;     /* [nblocks] */
  int *thrcol;      /* [set_size] */
  int *offset;      /* [nblocks] */
  int *nelems;      /* [nblocks] */
  int *ncolblk;     /* [ncolors] */
  int *blkmap;      /* [nblocks] */
  int *blkcol;      /* [nblocks] colour of each block */
  int *ind_offs;    /* [nblocks*ninds] */
  int *ind_sizes;   /* [nblocks*ninds] */
  int nindirect[ORC_MAXARGS];
  int *ind_maps[ORC_MAXARGS];   /* [nindirect[m]] */
  short *loc_maps[ORC_MAXARGS]; /* [set_size] or NULL */
  float transfer, transfer2, transfer3;
} orc_plan;

int  orc_plan_build(const orc_plan_in *in, orc_plan *out);
void orc_plan_free(orc_plan *p);
/* returns the number of errors found by the reference's plan checks
 * (op_lib_core.c:469-612) plus an exact colouring check (no two elements of
 * one thread colour in one block, and no two blocks of one colour, share an
 * OP_INC target). */
int  orc_plan_check(const orc_plan_in *in, const orc_plan *p);

/* ---- per-form entry points ---- */

#define ORC_DECLARE(SFX, REAL, UPD)                                            \
  void orc_save_soln_##SFX(int ncell, const REAL *q, REAL *qold);              \
  void orc_adt_calc_##SFX(int ncell, const int *pcell, const REAL *x,          \
                          const REAL *q, REAL *adt, const double *c);          \
  void orc_res_calc_##SFX(int nedge, const int *pedge, const int *pecell,      \
                          const REAL *x, const REAL *q, const REAL *adt,       \
                          REAL *res, const double *c);                         \
  void orc_bres_calc_##SFX(int nbedge, const int *pbedge, const int *pbecell,  \
                           const int *bound, const REAL *x, const REAL *q,     \
                           const REAL *adt, REAL *res, const double *c);       \
  void orc_update_##SFX(int ncell, const REAL *qold, REAL *q, REAL *res,       \
                        const REAL *adt, UPD *rms);                            \
  int orc_airfoil_seq_##SFX(int nnode, int ncell, int nedge, int nbedge,       \
                            const int *pcell, const int *pedge,                \
                            const int *pecell, const int *pbedge,              \
                            const int *pbecell, const int *bound,              \
                            const REAL *x, REAL *q, REAL *qold, REAL *adt,     \
                            REAL *res, const double *c, int niter,             \
                            double *rms_hist);                                 \
  int orc_airfoil_omp_##SFX(int nnode, int ncell, int nedge, int nbedge,       \
                            const int *pcell, const int *pedge,                \
                            const int *pecell, const int *pbedge,              \
                            const int *pbecell, const int *bound,              \
                            const REAL *x, REAL *q, REAL *qold, REAL *adt,     \
                            REAL *res, const double *c, int niter,             \
                            double *rms_hist, int part_size, int nthreads,     \
                            double *loop_seconds);                             \
  void orc_jac_res_##SFX(int nedge, const int *pp, const REAL *A,              \
                         const REAL *u, REAL *du, REAL beta);                  \
  void orc_jac_update_##SFX(int nnode, const REAL *r, REAL *du, REAL *u,       \
                            UPD *u_sum, UPD *u_max, REAL alpha);               \
  int orc_jac_seq_##SFX(int nnode, int nedge, const int *pp, const REAL *A,    \
                        const REAL *r, REAL *u, REAL *du, REAL alpha,          \
                        REAL beta, int niter, double *sum_hist,                \
                        double *max_hist);                                     \
  int orc_jac_omp_##SFX(int nnode, int nedge, const int *pp, const REAL *A,    \
                        const REAL *r, REAL *u, REAL *du, REAL alpha,          \
                        REAL beta, int niter, double *sum_hist,                \
                        double *max_hist, int part_size, int nthreads);

ORC_DECLARE(f64, double, double)
ORC_DECLARE(f32, float, float)
ORC_DECLARE(mix, float, double)

#ifdef __cplusplus
}
#endif

#endif
