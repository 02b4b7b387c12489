/*
 * The reference's OpenMP Jacobi kernels (jac1/<loop>_kernel.cpp), built with
 * float spelled double, in the way jac1/jac_kernels.cpp builds them for
 * float.  Test infrastructure only.
 */
#include "op_lib.h"

void __syncthreads() {}

#define float double
extern float alpha;
#include "res_kernel.cpp"
#include "update_kernel.cpp"
#undef float
