// Loops run by the generic loop path in tests/test_gpu_generic.py: device
// user functions written in the style of the reference's kernel headers,
// covering every argument kind the path supports.  Built into
// oplus2_matlab_amd/lib/libop2_testloops.so by the csrc Makefile.
#include "op_hip_loop.h"

static __constant__ double gen_scale;
OP2_DEVICE_CONST(gen_scale)

// direct loop: mixed types, dims 3 and 2, every OP_GBL access
__device__ inline void gen_direct(const double *a, float *b, const double *s, float *sum,
                                  int *imax, double *dmin) {
  b[0] = (float)(s[0] * a[0] + a[1]);
  b[1] = (float)(s[1] * a[2] * gen_scale);
  sum[0] += b[0];
  sum[1] += b[1] * b[1];
  *imax = MAX(*imax, (int)(a[0] * 100.0));
  *dmin = MIN(*dmin, a[1] - a[2]);
}

// edges: increments of a 3-component float node dat through both indices
// of an edge -> node map, a double node dat read through it, direct edge dats
__device__ inline void gen_edge(const double *x1, const double *x2, float *n1, float *n2,
                                const float *w, double *len) {
  const double dx = x2[0] - x1[0], dy = x2[1] - x1[1];
  *len = sqrt(dx * dx + dy * dy);
  n1[0] += *w;
  n1[1] += (float)dx;
  n1[2] += 1.0f;
  n2[0] -= *w;
  n2[1] -= (float)dy;
  n2[2] += 1.0f;
}

// cells: OP_RW through four indices of a cell -> node map and OP_MAX / OP_MIN
// through a second map (five (map, index) pairs), an OP_WRITE through a third
__device__ inline void gen_cell(const float *c, double *v0, double *v1, double *v2, double *v3,
                                int *hi, int *lo, float *w) {
  const double t = (double)*c;
  *v0 = MAX(*v0, t);
  *v1 = MAX(*v1, t + 1.0);
  *v2 = MAX(*v2, t + 2.0);
  *v3 = MAX(*v3, t + 3.0);
  *hi = MAX(*hi, (int)(t * 10.0));
  *lo = MIN(*lo, (int)(t * 10.0));
  *w = *c;
}

OP2_DEVICE_LOOP(gen_direct)
OP2_DEVICE_LOOP(gen_edge)
OP2_DEVICE_LOOP(gen_cell)
