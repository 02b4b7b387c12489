/*
 * OP2 core types for the MI355X runtime (libop2hip.so).
 *
 * Source-compatible with the reference's common/op_lib_core.h:34-176: the
 * same typedefs, access enum, handle types and macros, so that programs
 * written against the reference (airfoil/airfoil.cpp, jac1/jac.cpp) compile
 * unchanged.  Handle structs keep the reference's fields first and add the
 * device-side fields this runtime needs after them.
 */
#ifndef OP2_HIP_OP_LIB_CORE_H
#define OP2_HIP_OP_LIB_CORE_H

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

typedef unsigned int uint;
typedef long long ll;
typedef unsigned long long ull;

#ifdef __cplusplus
extern "C" {
#endif

/* diagnostics level (op_lib_core.h:52-61): 0 none, >=1 argument checks,
 * >=2 plan info, >=3 loop trace, >=4 plan reuse, >=7 positive plan checks */
extern int OP_diags;

/* op_lib_core.h:67,69 */
typedef enum { OP_READ, OP_WRITE, OP_RW, OP_INC, OP_MIN, OP_MAX } op_access;
typedef enum { OP_ARG_GBL, OP_ARG_DAT } op_arg_type;

/* op_lib_core.h:75-81 */
typedef struct {
  int index;
  int size;
  char const *name;
} op_set_core;
typedef op_set_core *op_set;

/* op_lib_core.h:83-92; map_d is the device copy of map, map_i the
 * runtime's renumbered host copy (NULL unless the sets were renumbered) */
typedef struct {
  int index;
  op_set from, to;
  int dim, *map;
  char const *name;
  int *map_d, *map_i;
} op_map_core;
typedef op_map_core *op_map;

/* op_lib_core.h:94-105; data_d is the device copy of data */
typedef struct {
  int index;
  op_set set;
  int dim, size;
  char *data, *data_d;
  char const *type, *name;
} op_dat_core;
typedef op_dat_core *op_dat;

/* op_lib_core.h:107-119 */
typedef struct {
  int index;
  op_dat dat;
  op_map map;
  int dim, idx, size;
  char *data, *data_d;
  char const *type;
  op_access acc;
  op_arg_type argtype;
} op_arg;

/* op_lib_core.h:122-150: the reference plan fields, filled from this
 * runtime's own plan (see op2_hip.h, op2_plan_info) for programs that read
 * them; device arrays are not exposed through it. */
typedef struct {
  char const *name;
  op_set set;
  int nargs, ninds, part_size;
  op_map *maps;
  op_dat *dats;
  int *idxs;
  op_access *accs;
  int *nthrcol, *thrcol, *offset;
  int **ind_maps;
  int *ind_offs, *ind_sizes, *nindirect;
  short **loc_maps;
  int nblocks;
  int *nelems;
  int ncolors;
  int *ncolblk, *blkmap;
  int nshared;
  float transfer, transfer2;
  int count;
} op_plan;

/* op_lib_core.h:152-158 */
typedef struct {
  char const *name;
  int count;
  float time;
  float transfer;
  float transfer2;
} op_kernel;

#ifdef __cplusplus
}
#endif

#ifndef MIN
#define MIN(a, b) ((a < b) ? (a) : (b))
#endif
#ifndef MAX
#define MAX(a, b) ((a > b) ? (a) : (b))
#endif

/* rounds up to a multiple of 16 bytes (op_lib_core.h:176) */
#define ROUND_UP(bytes) (((bytes) + 15) & ~15)

#endif
