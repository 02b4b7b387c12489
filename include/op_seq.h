/*
 * op_par_loop for programs written against the reference's common/op_seq.h
 * (header only).
 *
 * The reference's op_par_loop templates (common/op_seq.h:26-1340, 1 to 20
 * arguments) call the user kernel on the host for each element.  Here one
 * variadic template packs the arguments and hands the loop to the device
 * executor (op2_par_loop), which runs the runtime's gfx950 kernel for the
 * loop named `name`; the host function pointer only fixes the argument
 * types, as in the reference's CUDA build (op2.m rewrote op_par_loop calls
 * into per-loop stubs).  The generated-stub entry points
 * op_par_loop_<name>(name, set, args...) of airfoil/airfoil_op.cpp and
 * jac1/jac_op.cpp are provided for the programs that call them.
 */
#ifndef OP2_HIP_OP_SEQ_H
#define OP2_HIP_OP_SEQ_H

#include "op_lib.h"

#ifdef __cplusplus

template <class... T, class... A>
void op_par_loop(void (*kernel)(T *...), char const *name, op_set set, A... args) {
  static_assert(sizeof...(T) == sizeof...(A), "one op_arg per kernel argument");
  static_assert(sizeof...(A) >= 1 && sizeof...(A) <= 20, "1 to 20 arguments");
  (void)kernel;
  op_arg arr[sizeof...(A)] = {args...};
  op2_par_loop(name, set, (int)sizeof...(A), arr);
}

#endif /* __cplusplus */
#endif
