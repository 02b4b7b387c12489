// Device building blocks shared by the loop kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace op2 {

constexpr int kWave = 64;
constexpr int kXcds = 8;  // MI355X: 8 XCDs, workgroups dealt round-robin

// Workgroup id -> logical block such that each XCD receives a contiguous
// range of logical blocks (neighbouring blocks then share that XCD's L2).
__device__ __forceinline__ int xcd_block(int wg, int n) {
  const int x = wg % kXcds, i = wg / kXcds;
  const int base = n / kXcds, rem = n % kXcds;
  return x * base + (x < rem ? x : rem) + i;
}

// 2- and 4-component element access with 16-byte memory operations.
template <typename T>
struct Vec;

template <>
struct Vec<double> {
  __device__ __forceinline__ static void load2(const double *p, double v[2]) {
    double2 a = *reinterpret_cast<const double2 *>(p);
    v[0] = a.x;
    v[1] = a.y;
  }
  __device__ __forceinline__ static void load4(const double *p, double v[4]) {
    double2 a = reinterpret_cast<const double2 *>(p)[0];
    double2 b = reinterpret_cast<const double2 *>(p)[1];
    v[0] = a.x;
    v[1] = a.y;
    v[2] = b.x;
    v[3] = b.y;
  }
  __device__ __forceinline__ static void store4(double *p, const double v[4]) {
    reinterpret_cast<double2 *>(p)[0] = make_double2(v[0], v[1]);
    reinterpret_cast<double2 *>(p)[1] = make_double2(v[2], v[3]);
  }
  __device__ __forceinline__ static void store2(double *p, const double v[2]) {
    *reinterpret_cast<double2 *>(p) = make_double2(v[0], v[1]);
  }
};

template <>
struct Vec<float> {
  __device__ __forceinline__ static void load2(const float *p, float v[2]) {
    float2 a = *reinterpret_cast<const float2 *>(p);
    v[0] = a.x;
    v[1] = a.y;
  }
  __device__ __forceinline__ static void load4(const float *p, float v[4]) {
    float4 a = *reinterpret_cast<const float4 *>(p);
    v[0] = a.x;
    v[1] = a.y;
    v[2] = a.z;
    v[3] = a.w;
  }
  __device__ __forceinline__ static void store4(float *p, const float v[4]) {
    *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
  }
  __device__ __forceinline__ static void store2(float *p, const float v[2]) {
    *reinterpret_cast<float2 *>(p) = make_float2(v[0], v[1]);
  }
};

// Constants of the airfoil loops (airfoil/airfoil.cpp:48).
template <typename T>
struct AirfoilConsts {
  T gam, gm1, cfl, eps;
  T qinf[4];
};

// ---- user kernels (arithmetic in the order the reference headers write it) --

// airfoil/adt_calc.h:1-26
template <typename T>
__device__ __forceinline__ T adt_of(const T x1[2], const T x2[2], const T x3[2],
                                    const T x4[2], const T q[4],
                                    const AirfoilConsts<T> &k) {
  T ri = T(1) / q[0];
  T u = ri * q[1];
  T v = ri * q[2];
  T c = sqrt(k.gam * k.gm1 * (ri * q[3] - T(0.5) * (u * u + v * v)));
  T dx = x2[0] - x1[0];
  T dy = x2[1] - x1[1];
  T s = fabs(u * dy - v * dx) + c * sqrt(dx * dx + dy * dy);
  dx = x3[0] - x2[0];
  dy = x3[1] - x2[1];
  s += fabs(u * dy - v * dx) + c * sqrt(dx * dx + dy * dy);
  dx = x4[0] - x3[0];
  dy = x4[1] - x3[1];
  s += fabs(u * dy - v * dx) + c * sqrt(dx * dx + dy * dy);
  dx = x1[0] - x4[0];
  dy = x1[1] - x4[1];
  s += fabs(u * dy - v * dx) + c * sqrt(dx * dx + dy * dy);
  return s / k.cfl;
}

// adt_of with the cell's four side vectors given, side i from node i to node
// i+1 (mod 4), each possibly reversed: a reversed (dx, dy) is the exact
// negation (x_a - x_b = -(x_b - x_a) in IEEE arithmetic), which leaves
// fabs(u dy - v dx) and dx dx + dy dy, and so every bit of the result,
// unchanged.
template <typename T>
__device__ __forceinline__ T adt_of_sides(const T dxs[4], const T dys[4], const T q[4],
                                          const AirfoilConsts<T> &k) {
  T ri = T(1) / q[0];
  T u = ri * q[1];
  T v = ri * q[2];
  T c = sqrt(k.gam * k.gm1 * (ri * q[3] - T(0.5) * (u * u + v * v)));
  T s = fabs(u * dys[0] - v * dxs[0]) + c * sqrt(dxs[0] * dxs[0] + dys[0] * dys[0]);
#pragma unroll
  for (int i = 1; i < 4; i++)
    s += fabs(u * dys[i] - v * dxs[i]) + c * sqrt(dxs[i] * dxs[i] + dys[i] * dys[i]);
  return s / k.cfl;
}

// airfoil/res_calc.h:1-30: f is added to the first cell, subtracted from the
// second.
template <typename T>
__device__ __forceinline__ void edge_flux(const T x1[2], const T x2[2], const T q1[4],
                                          const T q2[4], T adt1, T adt2,
                                          const AirfoilConsts<T> &k, T f[4]) {
  T dx = x1[0] - x2[0];
  T dy = x1[1] - x2[1];
  T ri = T(1) / q1[0];
  T p1 = k.gm1 * (q1[3] - T(0.5) * ri * (q1[1] * q1[1] + q1[2] * q1[2]));
  T vol1 = ri * (q1[1] * dy - q1[2] * dx);
  ri = T(1) / q2[0];
  T p2 = k.gm1 * (q2[3] - T(0.5) * ri * (q2[1] * q2[1] + q2[2] * q2[2]));
  T vol2 = ri * (q2[1] * dy - q2[2] * dx);
  T mu = T(0.5) * (adt1 + adt2) * k.eps;
  f[0] = T(0.5) * (vol1 * q1[0] + vol2 * q2[0]) + mu * (q1[0] - q2[0]);
  f[1] = T(0.5) * (vol1 * q1[1] + p1 * dy + vol2 * q2[1] + p2 * dy) + mu * (q1[1] - q2[1]);
  f[2] = T(0.5) * (vol1 * q1[2] - p1 * dx + vol2 * q2[2] - p2 * dx) + mu * (q1[2] - q2[2]);
  f[3] = T(0.5) * (vol1 * (q1[3] + p1) + vol2 * (q2[3] + p2)) + mu * (q1[3] - q2[3]);
}

// airfoil/bres_calc.h:1-33: the increment of the boundary cell; qi is the
// free stream (the qinf constant, or the boundary edge's mesh's in a batch)
template <typename T>
__device__ __forceinline__ void boundary_flux(const T x1[2], const T x2[2], const T q1[4],
                                              T adt1, int bound, const T *qi,
                                              const AirfoilConsts<T> &k, T f[4]) {
  T dx = x1[0] - x2[0];
  T dy = x1[1] - x2[1];
  T ri = T(1) / q1[0];
  T p1 = k.gm1 * (q1[3] - T(0.5) * ri * (q1[1] * q1[1] + q1[2] * q1[2]));
  if (bound == 1) {
    f[0] = T(0);
    f[1] = +p1 * dy;
    f[2] = -p1 * dx;
    f[3] = T(0);
    return;
  }
  T vol1 = ri * (q1[1] * dy - q1[2] * dx);
  ri = T(1) / qi[0];
  T p2 = k.gm1 * (qi[3] - T(0.5) * ri * (qi[1] * qi[1] + qi[2] * qi[2]));
  T vol2 = ri * (qi[1] * dy - qi[2] * dx);
  T mu = adt1 * k.eps;
  f[0] = T(0.5) * (vol1 * q1[0] + vol2 * qi[0]) + mu * (q1[0] - qi[0]);
  f[1] = T(0.5) * (vol1 * q1[1] + p1 * dy + vol2 * qi[1] + p2 * dy) + mu * (q1[1] - qi[1]);
  f[2] = T(0.5) * (vol1 * q1[2] - p1 * dx + vol2 * qi[2] - p2 * dx) + mu * (q1[2] - qi[2]);
  f[3] = T(0.5) * (vol1 * (q1[3] + p1) + vol2 * (qi[3] + p2)) + mu * (q1[3] - qi[3]);
}

// ---- reductions ----------------------------------------------------------

enum RedOp { kRedSum = 0, kRedMax = 1 };

template <int OP>
__device__ __forceinline__ double red_apply(double a, double b) {
  if (OP == kRedSum) return a + b;
  return (a > b) ? a : b;
}

template <int OP>
__device__ __forceinline__ double red_identity() {
  return OP == kRedSum ? 0.0 : -__builtin_huge_val();
}

// Cross-lane butterfly inside a 64-wide wave, then through LDS across the
// waves of the workgroup.  Every lane of wave 0 ends with the result.
template <int OP, int NT>
__device__ __forceinline__ double block_reduce(double v, double *lds) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v = red_apply<OP>(v, __shfl_xor(v, off, kWave));
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  if (w == 0) {
    v = lane < NT / kWave ? lds[lane] : red_identity<OP>();
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = red_apply<OP>(v, __shfl_xor(v, off, kWave));
  }
  return v;
}

}  // namespace op2
