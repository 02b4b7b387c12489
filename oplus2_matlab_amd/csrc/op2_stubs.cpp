// C++-linkage per-loop entry points of the reference's generated programs
// (airfoil/airfoil_op.cpp:64-99, jac1/jac_op.cpp): each forwards to the
// device executor.  airfoil's and jac1's op_par_loop_update share one
// symbol (same signature); the executor tells them apart by their
// arguments.
#include "op2_hip.h"

void op_par_loop_save_soln(char const *name, op_set set, op_arg a0, op_arg a1) {
  op_arg a[] = {a0, a1};
  op2_par_loop(name, set, 2, a);
}

void op_par_loop_adt_calc(char const *name, op_set set, op_arg a0, op_arg a1, op_arg a2,
                          op_arg a3, op_arg a4, op_arg a5) {
  op_arg a[] = {a0, a1, a2, a3, a4, a5};
  op2_par_loop(name, set, 6, a);
}

void op_par_loop_res_calc(char const *name, op_set set, op_arg a0, op_arg a1, op_arg a2,
                          op_arg a3, op_arg a4, op_arg a5, op_arg a6, op_arg a7) {
  op_arg a[] = {a0, a1, a2, a3, a4, a5, a6, a7};
  op2_par_loop(name, set, 8, a);
}

void op_par_loop_bres_calc(char const *name, op_set set, op_arg a0, op_arg a1, op_arg a2,
                           op_arg a3, op_arg a4, op_arg a5) {
  op_arg a[] = {a0, a1, a2, a3, a4, a5};
  op2_par_loop(name, set, 6, a);
}

void op_par_loop_update(char const *name, op_set set, op_arg a0, op_arg a1, op_arg a2,
                        op_arg a3, op_arg a4) {
  op_arg a[] = {a0, a1, a2, a3, a4};
  op2_par_loop(name, set, 5, a);
}

void op_par_loop_res(char const *name, op_set set, op_arg a0, op_arg a1, op_arg a2,
                     op_arg a3) {
  op_arg a[] = {a0, a1, a2, a3};
  op2_par_loop(name, set, 4, a);
}
