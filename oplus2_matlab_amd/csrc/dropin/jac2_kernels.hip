// Device side of the reference's jac2 program (jac2/jac.cpp, jac2/jac_op.cpp)
// on the generic loop path: what the reference's translator writes as
// jac2/jac_kernels.cu, res_kernel.cu and update_kernel.cu.  The unchanged
// user kernels jac2/res.h and jac2/update.h are compiled as device functions
// (the csrc Makefile's dropin target adds -I<reference>/jac2) and registered
// by name, so both op_par_loop (jac.cpp) and the op_par_loop_<name> stubs
// (jac_op.cpp) run them.
#include "op_hip_loop.h"

static __constant__ float alpha;  // jac2/jac.cpp:45, set by op_decl_const
OP2_DEVICE_CONST(alpha)

__device__
#include "res.h"

__device__
#include "update.h"

OP2_DEVICE_LOOP(res)
OP2_DEVICE_LOOP(update)
