/*
 * OP2 C++ API over the MI355X runtime (header only).
 *
 * Source-compatible with the reference's common/op_lib.h: run-time type
 * checks (:45-51), ZERO_<type> constants (:65-71), OP_ID / OP_GBL (:76-77),
 * the C++-linkage functions op_init, op_decl_dat_char, op_decl_const_char,
 * op_fetch_data, op_plan_get and op_exit (:103-154) and the templates
 * op_decl_dat, op_decl_const2, op_decl_const and op_arg_gbl (:160-190).
 * Every function forwards to the C ABI of op2_hip.h.
 */
#ifndef OP2_HIP_OP_LIB_H
#define OP2_HIP_OP_LIB_H

#include "op2_hip.h"

#ifdef __cplusplus

#include <cctype>

inline int type_error(const double *, const char *type) { return strcmp(type, "double"); }
inline int type_error(const float *, const char *type) { return strcmp(type, "float"); }
inline int type_error(const int *, const char *type) { return strcmp(type, "int"); }
inline int type_error(const uint *, const char *type) { return strcmp(type, "uint"); }
inline int type_error(const ll *, const char *type) { return strcmp(type, "ll"); }
inline int type_error(const ull *, const char *type) { return strcmp(type, "ull"); }
inline int type_error(const bool *, const char *type) { return strcmp(type, "bool"); }

#ifdef OP_USER_DATATYPES
#include <OP_USER_DATATYPES>
#endif

#define ZERO_double 0.0;
#define ZERO_float 0.0f;
#define ZERO_int 0;
#define ZERO_uint 0;
#define ZERO_ll 0;
#define ZERO_ull 0;
#define ZERO_bool 0;

#define OP_ID (op_map) NULL
#define OP_GBL (op_map) NULL

extern int OP_diags;

// The program's own -DOP_BLOCK_SIZE / -DOP_PART_SIZE apply first and argv's
// OP_BLOCK_SIZE= / OP_PART_SIZE= after them (common/op_lib_core.c:67-92).
// The sets are renumbered for locality at the program's first indirect
// increment loop (op2_renumber), unless OP2_RENUMBER=0 is in the
// environment; fetched data keeps the program's numbering.
inline void op_init(int argc, char **argv, int diags) {
  op2_set_auto_renumber(1);
#ifdef OP_BLOCK_SIZE
  OP_block_size = OP_BLOCK_SIZE;
#endif
#ifdef OP_PART_SIZE
  OP_part_size = OP_PART_SIZE;
#endif
  op2_init(argc, argv, diags);
}

inline op_dat op_decl_dat_char(op_set set, int dim, char const *type, int size,
                               char *data, char const *name) {
  return op2_decl_dat_char(set, dim, type, size, data, name);
}

inline void op_decl_const_char(int dim, char const *type, int size, char *data,
                               char const *name) {
  op2_decl_const_char(dim, type, size, data, name);
}

inline void op_fetch_data(op_dat dat) { op2_fetch_data(dat); }

inline op_plan *op_plan_get(char const *name, op_set set, int part_size, int nargs,
                            op_arg *args, int ninds, int *inds) {
  return op2_plan_get(name, set, part_size, nargs, args, ninds, inds);
}

inline void op_exit() { op2_exit(); }

template <class T>
op_dat op_decl_dat(op_set set, int dim, char const *type, T *data, char const *name) {
  if (type_error(data, type)) {
    printf("incorrect type specified for dataset \"%s\" \n", name);
    exit(1);
  }
  return op_decl_dat_char(set, dim, type, sizeof(T), (char *)data, name);
}

template <class T>
void op_decl_const2(char const *name, int dim, char const *type, T *data) {
  if (type_error(data, type)) {
    printf("incorrect type specified for constant \"%s\" \n", name);
    exit(1);
  }
  op_decl_const_char(dim, type, sizeof(T), (char *)data, name);
}

/* The reference's op_decl_const(dim, type, data) has no name; its CPU
 * library only type-checks (common/op_lib.h:169-175) and its CUDA build
 * relies on op2.m rewriting the call into op_decl_const2 with the variable's
 * name.  Here the macro below passes the source text of the data argument
 * ("&gam" -> "gam", "qinf" -> "qinf") so that the unchanged call registers
 * the constant under the variable's name. */
template <class T>
void op_decl_const_expr(char const *expr, int dim, char const *type, T *data) {
  if (type_error(data, type)) {
    printf("incorrect type specified for constant in op_decl_const");
    exit(1);
  }
  char name[128];
  int n = 0;
  for (const char *c = expr; *c && n < 127; c++)
    if (std::isalnum((unsigned char)*c) || *c == '_') name[n++] = *c;
  name[n] = 0;
  op_decl_const_char(dim, type, sizeof(T), (char *)data, name);
}
#define op_decl_const(dim, type, data) op_decl_const_expr(#data, dim, type, data)

template <class T>
op_arg op_arg_gbl(T *data, int dim, char const *type, op_access acc) {
  if (type_error(data, type))
    return op_arg_gbl_core((char *)data, dim, "error", acc);
  return op_arg_gbl_core((char *)data, dim, type, acc);
}

#endif /* __cplusplus */
#endif
