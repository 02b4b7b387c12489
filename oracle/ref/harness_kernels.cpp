/*
 * The reference's OpenMP airfoil kernels (airfoil/<loop>_kernel.cpp), built
 * with float spelled double, in the way airfoil/airfoil_kernels.cpp builds
 * them for float.  Test infrastructure only.
 */
#include "op_lib.h"

void __syncthreads() {}

#define float double
extern float gam, gm1, cfl, eps, mach, alpha, qinf[4];
#include "save_soln_kernel.cpp"
#include "adt_calc_kernel.cpp"
#include "res_calc_kernel.cpp"
#include "bres_calc_kernel.cpp"
#include "update_kernel.cpp"
#undef float
