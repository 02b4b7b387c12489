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

// an indirect loop with global reductions: total and longest edge length,
// edges counted per node (OP_INC through the map)
__device__ inline void gen_edge_sum(const double *x1, const double *x2, int *c1, int *c2,
                                    double *total, double *longest) {
  const double dx = x2[0] - x1[0], dy = x2[1] - x1[1];
  const double l = sqrt(dx * dx + dy * dy);
  *total += l;
  *longest = MAX(*longest, l);
  *c1 += 1;
  *c2 += 1;
}

// airfoil/res_calc.h's arithmetic as a generic loop (8 arguments, 2 maps),
// with its own device constants: the generic path against the hand-written
// res_calc kernels
static __constant__ double gm1;
static __constant__ double eps;
OP2_DEVICE_CONST(gm1)
OP2_DEVICE_CONST(eps)

__device__ inline void gen_res_calc(const double *x1, const double *x2, const double *q1,
                                    const double *q2, const double *adt1, const double *adt2,
                                    double *res1, double *res2) {
  double dx, dy, mu, ri, p1, vol1, p2, vol2, f;
  dx = x1[0] - x2[0];
  dy = x1[1] - x2[1];
  ri = 1.0 / q1[0];
  p1 = gm1 * (q1[3] - 0.5 * ri * (q1[1] * q1[1] + q1[2] * q1[2]));
  vol1 = ri * (q1[1] * dy - q1[2] * dx);
  ri = 1.0 / q2[0];
  p2 = gm1 * (q2[3] - 0.5 * ri * (q2[1] * q2[1] + q2[2] * q2[2]));
  vol2 = ri * (q2[1] * dy - q2[2] * dx);
  mu = 0.5 * ((*adt1) + (*adt2)) * eps;
  f = 0.5 * (vol1 * q1[0] + vol2 * q2[0]) + mu * (q1[0] - q2[0]);
  res1[0] += f;
  res2[0] -= f;
  f = 0.5 * (vol1 * q1[1] + p1 * dy + vol2 * q2[1] + p2 * dy) + mu * (q1[1] - q2[1]);
  res1[1] += f;
  res2[1] -= f;
  f = 0.5 * (vol1 * q1[2] - p1 * dx + vol2 * q2[2] - p2 * dx) + mu * (q1[2] - q2[2]);
  res1[2] += f;
  res2[2] -= f;
  f = 0.5 * (vol1 * (q1[3] + p1) + vol2 * (q2[3] + p2)) + mu * (q1[3] - q2[3]);
  res1[3] += f;
  res2[3] -= f;
}

// the owner-tile form: every written argument an OP_INC through the edge ->
// node map (a float dim-3 dat and an int count, each through both indices)
// or an OP_GBL OP_MAX; the rest read only
__device__ inline void gen_edge_owner(const double *x1, const double *x2, float *n1, float *n2,
                                      int *c1, int *c2, const float *w, double *longest) {
  const double dx = x2[0] - x1[0], dy = x2[1] - x1[1];
  n1[0] += *w;
  n1[1] += (float)dx;
  n1[2] += 1.0f;
  n2[0] -= *w;
  n2[1] -= (float)dy;
  n2[2] += 1.0f;
  *c1 += 1;
  *c2 += 2;
  *longest = MAX(*longest, sqrt(dx * dx + dy * dy));
}

OP2_DEVICE_LOOP(gen_edge_sum)
OP2_DEVICE_LOOP(gen_edge_owner)
OP2_DEVICE_LOOP(gen_res_calc)

// airfoil update's signature (airfoil/update.h) with another body: half the
// move.  Registered as "update" only while a test asks for it, to show that
// a registered loop runs instead of the built-in kernel of the same name and
// argument signature.
__device__ inline void half_update(const double *qold, double *q, double *res, const double *adt,
                                   double *rms) {
  const double adti = 0.5 / *adt;
  for (int n = 0; n < 4; n++) {
    const double del = adti * res[n];
    q[n] = qold[n] - del;
    res[n] = 0.0;
    *rms += del * del;
  }
}

extern "C" int testloops_register_half_update(int on) {
  return op2_register_loop("update", on ? &op2_loop::entry<&half_update> : nullptr);
}
