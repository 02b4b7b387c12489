/*
 * CPU oracle: the three arithmetic forms of the airfoil / Jacobi loops.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Built with -ffp-contract=off so
 * that no multiply-add is fused, as in the reference's x86-64 build.
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* plain FP64 */
#define REAL double
#define UPD double
#define SFX f64
#define RSQRT sqrt
#define RFABS fabs
#define USQRT sqrt
#include "form.inc"
#undef REAL
#undef UPD
#undef SFX
#undef RSQRT
#undef RFABS
#undef USQRT

/* the reference's own single-precision program */
#define REAL float
#define UPD float
#define SFX f32
#define RSQRT sqrtf
#define RFABS fabsf
#define USQRT sqrtf
#include "form.inc"
#undef REAL
#undef UPD
#undef SFX
#undef RSQRT
#undef RFABS
#undef USQRT

/* precision form: float state and flux, double update and reductions */
#define REAL float
#define UPD double
#define SFX mix
#define RSQRT sqrtf
#define RFABS fabsf
#define USQRT sqrt
#include "form.inc"
