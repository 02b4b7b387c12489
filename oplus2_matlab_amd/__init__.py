"""MI355X-native OP2 runtime and loop executor (airfoil and Jacobi loops).

The product is the HIP library oplus2_matlab_amd/lib/libop2hip.so (sources in
csrc/, C ABI in include/op2_hip.h); op2.py binds its C ABI, airfoil.py runs
the reference programs' loops through it, and mesh.py builds the inputs.
"""
