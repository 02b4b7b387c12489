/*
 * Generic loop path of the MI355X OP2 runtime (header only; compile with
 * hipcc --offload-arch=gfx950).
 *
 * The reference turns each op_par_loop into a CUDA kernel with its MATLAB
 * translator (common/op2.m:142-1141; e.g. jac2/res_kernel.cu): the user
 * function is compiled as a __device__ function and the generated kernel
 * hands it one pointer per argument for every element.  Here that kernel is
 * a template instantiated from the user function's own signature, so no
 * generator is needed.  A translation unit of the user's program (the role
 * of the reference's <app>_kernels.cu) does:
 *
 *     #include "op_hip_loop.h"
 *     static __constant__ float alpha;   // device copy of a global constant
 *     OP2_DEVICE_CONST(alpha)             // filled by op_decl_const(...)
 *     __device__
 *     #include "res.h"                    // the unchanged user kernel
 *     OP2_DEVICE_LOOP(res)                // run by op_par_loop(..."res"...)
 *
 * and the runtime runs the registered loop whenever no built-in kernel
 * matches the loop's name and arguments (op2_register_loop).
 *
 * Execution: every element of the set calls the user function with, per
 * argument, a pointer to the element's record (direct dat), to the record
 * its map points at (indirect dat), or to a thread-local copy (OP_GBL).
 * Loops whose indirect writes are all OP_INC through one map run on owner
 * tiles (below); other loops that write through a map (OP_INC, OP_RW,
 * OP_WRITE, OP_MIN, OP_MAX) run on the runtime's generic plan: blocks of one
 * block colour per launch, and inside a block one thread colour at a time,
 * so no two elements running together write the same target and the user
 * function may update global memory directly.  Other loops run as a
 * grid-stride pass.  OP_GBL
 * reductions (OP_INC, OP_MIN, OP_MAX) are combined per workgroup (64-lane
 * shuffles, then LDS, in a fixed order, in double) and folded by one more
 * kernel; the result is added to (INC) or replaces (MIN, MAX, whose start
 * value is the caller's) the caller's variable before op_par_loop returns,
 * as the reference's op_reduction does (common/op_lib.cu:268-332).
 */
#ifndef OP2_HIP_LOOP_H
#define OP2_HIP_LOOP_H

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <tuple>
#include <type_traits>
#include <utility>

#include "op_lib.h"

namespace op2_loop {

constexpr int kThreads = 256;
constexpr int kGblBytes = 64;      // largest OP_GBL argument (dim x type size)
constexpr int kMaxRed = 64;        // reduction components of one loop
constexpr int kDirectGrid = 2048;  // workgroups of a loop without indirect writes

enum Kind { kDat = 0, kGblRead = 1, kGblInc = 2, kGblMin = 3, kGblMax = 4 };

struct Arg {
  char *data;       // dat: device data (the runtime's numbering)
  int const *map;   // dat through a map: the device map, else null
  int idx, mapdim;  // map index and map dimension
  int bytes;        // dat: bytes per element (dim x type size)
  int kind, dim;    // Kind; OP_GBL: number of components
  int red;          // reduction: first component slot
  // owner tiles: an OP_INC argument's (map, index) among the loop's OP_INC
  // pairs, its first argument on the same dat, and the byte offset of its
  // increments in LDS (-1: not an OP_INC argument); an indirect argument's
  // (map, index) in the plan's per-position target rows (-1: direct)
  int inc, leader, lds, slot;
  alignas(16) unsigned char gbl[kGblBytes];  // OP_GBL READ value, MIN/MAX start value
};

template <int N>
struct Args {
  Arg a[N];
};

struct RedKinds {
  int n;
  unsigned char kind[kMaxRed];
};

template <class T>
using elem_t = std::remove_cv_t<T>;

// thread-local storage of one OP_GBL argument
template <class T>
struct Local {
  elem_t<T> v[kGblBytes / sizeof(elem_t<T>)];
};

// Argument classes known at compile time: with G >= 0 the last G arguments
// are OP_GBL and the others dats (the usual order; G up to 4), so dat
// pointers stay global and the OP_GBL copies can live in registers; G < 0:
// any order, decided per argument at run time (flat pointers, copies in
// private memory).
template <int I, int N, int G>
struct ArgClass {
  static constexpr bool dat = G >= 0 && I < N - G;  // certainly a dat
  static constexpr bool gbl = G >= 0 && I >= N - G; // certainly OP_GBL
};


template <class T>
__device__ __forceinline__ void local_init(const Arg &d, Local<T> &l) {
  if (d.kind == kDat) return;
  constexpr int n = kGblBytes / sizeof(elem_t<T>);
  const elem_t<T> *g = reinterpret_cast<const elem_t<T> *>(d.gbl);
  // the whole copy, with constant indices (the gbl bytes past dim are zero)
#pragma unroll
  for (int c = 0; c < n; c++) l.v[c] = d.kind == kGblInc ? elem_t<T>(0) : g[c];
}

template <class T, bool DAT, bool GBL>
__device__ __forceinline__ T *arg_ptr(const Arg &d, int e, Local<T> &l) {
  if (GBL || (!DAT && d.kind != kDat)) return l.v;
  const int el = d.map ? d.map[(size_t)e * d.mapdim + d.idx] : e;
  return reinterpret_cast<T *>(d.data + (size_t)el * d.bytes);
}

__device__ __forceinline__ double red_op(int kind, double a, double b) {
  if (kind == kGblInc) return a + b;
  if (kind == kGblMin) return b < a ? b : a;
  return b > a ? b : a;
}

__device__ __forceinline__ double red_identity(int kind) {
  return kind == kGblInc ? 0.0 : kind == kGblMin ? __builtin_huge_val() : -__builtin_huge_val();
}

// block-wide reduction in a fixed order; every thread gets the result
__device__ __forceinline__ double block_red(int kind, double v, double *lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = red_op(kind, v, __shfl_xor(v, off, 64));
  __syncthreads();  // lds is reused by consecutive calls
  if (threadIdx.x % 64 == 0) lds[threadIdx.x / 64] = v;
  __syncthreads();
  double r = lds[0];
#pragma unroll
  for (int i = 1; i < kThreads / 64; i++) r = red_op(kind, r, lds[i]);
  return r;
}

template <class T, bool DAT>
__device__ __forceinline__ void reduce_arg(const Arg &d, const Local<T> &l, double *partials,
                                           int slot, int nred, double *lds) {
  if (DAT || d.kind < kGblInc) return;
  constexpr int n = kGblBytes / sizeof(elem_t<T>);
#pragma unroll
  for (int c = 0; c < n; c++) {
    if (c < d.dim) {  // uniform
      const double r = block_red(d.kind, (double)l.v[c], lds);
      if (threadIdx.x == 0) partials[(size_t)slot * nred + d.red + c] = r;
    }
  }
}

template <auto F, int G, class... T, size_t... I>
__device__ __forceinline__ void call(const Args<sizeof...(T)> &A, int e,
                                     std::tuple<Local<T>...> &loc, std::index_sequence<I...>) {
  constexpr int N = sizeof...(T);
  F(arg_ptr<T, ArgClass<I, N, G>::dat, ArgClass<I, N, G>::gbl>(A.a[I], e, std::get<I>(loc))...);
}

template <int G, class... T, size_t... I>
__device__ __forceinline__ void init_all(const Args<sizeof...(T)> &A,
                                         std::tuple<Local<T>...> &loc, std::index_sequence<I...>) {
  constexpr int N = sizeof...(T);
  ((ArgClass<I, N, G>::dat ? (void)0 : local_init<T>(A.a[I], std::get<I>(loc))), ...);
}

template <int G, class... T, size_t... I>
__device__ __forceinline__ void reduce_all(const Args<sizeof...(T)> &A,
                                           std::tuple<Local<T>...> &loc, double *partials,
                                           int slot, int nred, double *lds,
                                           std::index_sequence<I...>) {
  constexpr int N = sizeof...(T);
  (reduce_arg<T, ArgClass<I, N, G>::dat>(A.a[I], std::get<I>(loc), partials, slot, nred, lds),
   ...);
}

// One pass of the loop.  With a plan (blk_off != null) the workgroup runs
// plan block blkmap[blk_first + blockIdx.x] one thread colour at a time;
// without, a grid-stride pass over the set.  Reduction partials go to
// partials[slot * nred + component], slot = plan block or workgroup.
template <auto F, int G, class... T>
__global__ __launch_bounds__(kThreads) void k_loop(Args<sizeof...(T)> A, int n, int nred,
                                                   int const *blk_off, int const *perm,
                                                   int const *col_off, int const *col_base,
                                                   int const *blk_ncol, int const *blkmap,
                                                   int blk_first, double *partials) {
  __shared__ double lds[kThreads / 64];
  constexpr auto seq = std::index_sequence_for<T...>{};
  std::tuple<Local<T>...> loc;
  init_all<G, T...>(A, loc, seq);
  int slot;
  if (blk_off) {
    const int b = blkmap[blk_first + blockIdx.x];
    slot = b;
    const int cb = col_base[b], ncol = blk_ncol[b];
    for (int c = 0; c < ncol; c++) {
      const int hi = col_off[cb + c + 1];
      for (int p = col_off[cb + c] + (int)threadIdx.x; p < hi; p += kThreads)
        call<F, G, T...>(A, perm ? perm[p] : p, loc, seq);
      __syncthreads();  // the next colour sees this colour's writes
    }
  } else {
    slot = blockIdx.x;
    for (int e = blockIdx.x * kThreads + threadIdx.x; e < n; e += gridDim.x * kThreads)
      call<F, G, T...>(A, e, loc, seq);
  }
  if (nred > 0) reduce_all<G, T...>(A, loc, partials, slot, nred, lds, seq);
}

// ---- owner tiles -------------------------------------------------------------
// Loops whose indirect writes are all OP_INC through one map, and whose
// written arguments (non-const pointers of the user function) are exactly
// those OP_INC dats and OP_GBL MIN/MAX, run on the runtime's owner plan
// (op2_generic_owner_plan_get) instead of colours.  Each workgroup takes a
// tile: it evaluates every element that increments one of the tile's
// targets, the user function incrementing zeroed copies held in registers
// whose values go to LDS (component-major rows), then each owned target
// adds its increments in element order and is written once.  No colours,
// no atomics.  Elements on a tile boundary are evaluated by both of their
// tiles (each keeps only its own targets' increments), so OP_GBL OP_INC,
// which would count them twice, keeps the coloured plan.

struct OwnerArrays {
  int const *blk_off, *perm, *ids, *owned_off, *owned, *csr_off;
  unsigned short const *csr, *ell;
  int slot_stride;  // ints per position in ids
  int stride;       // length of an LDS row: the plan's largest number of positions
};

template <class T>
constexpr bool written = !std::is_const<T>::value;

// tile of workgroup wg: each XCD (workgroups wg = x mod 8) takes a
// contiguous range of tiles, so neighbouring tiles share its L2
__device__ __forceinline__ int xcd_tile(int wg, int n) {
  const int x = wg % 8, i = wg / 8, base = n / 8, rem = n % 8;
  return x * base + (x < rem ? x : rem) + i;
}

// the record of position pos (global) for the user function: an indirect
// dat's target from the plan's per-position rows, a direct dat's element
// through perm, the zeroed copy of an OP_INC dat, or the OP_GBL copy
template <class T, bool DAT, bool GBL>
__device__ __forceinline__ T *owner_ptr(const Arg &d, const OwnerArrays &P, int pos,
                                        Local<T> &gl, Local<T> &inc) {
  if (GBL) return gl.v;
  if (DAT && written<T>) return inc.v;
  const int el = d.slot >= 0 ? P.ids[(size_t)pos * P.slot_stride + d.slot] : P.perm[pos];
  return reinterpret_cast<T *>(d.data + (size_t)el * d.bytes);
}

template <class T, bool DAT>
__device__ __forceinline__ void owner_zero(Local<T> &inc) {
  if (!(DAT && written<T>)) return;
  constexpr int n = kGblBytes / sizeof(elem_t<T>);
#pragma unroll
  for (int c = 0; c < n; c++) inc.v[c] = elem_t<T>(0);
}

template <class T, bool DAT>
__device__ __forceinline__ void owner_store(const Arg &d, const Local<T> &inc, unsigned char *lds,
                                            int stride, int p) {
  if (!(DAT && written<T>)) return;
  constexpr int n = kGblBytes / sizeof(elem_t<T>);
  elem_t<T> *row = reinterpret_cast<elem_t<T> *>(lds + d.lds);
#pragma unroll
  for (int c = 0; c < n; c++)
    if (c < d.dim) row[c * stride + p] = inc.v[c];
}

template <auto F, int G, class... T, size_t... I>
__device__ __forceinline__ void owner_call(const Args<sizeof...(T)> &A, const OwnerArrays &P,
                                           int pos, std::tuple<Local<T>...> &loc,
                                           std::tuple<Local<T>...> &inc, unsigned char *lds,
                                           int p, std::index_sequence<I...>) {
  constexpr int N = sizeof...(T);
  (owner_zero<T, ArgClass<I, N, G>::dat>(std::get<I>(inc)), ...);
  F(owner_ptr<T, ArgClass<I, N, G>::dat, ArgClass<I, N, G>::gbl>(A.a[I], P, pos,
                                                                  std::get<I>(loc),
                                                                  std::get<I>(inc))...);
  (owner_store<T, ArgClass<I, N, G>::dat>(A.a[I], std::get<I>(inc), lds, P.stride, p), ...);
}

// adds argument J's increment at position pos to v when J writes the dat of
// argument `leader` through (map, index) pair k
template <class TI, class TJ, bool DATJ>
__device__ __forceinline__ void owner_add(const Arg &dj, int leader, int k, int pos,
                                          const unsigned char *lds, int stride, elem_t<TI> *v,
                                          int dim) {
  if (!(DATJ && written<TJ>)) return;
  if (dj.leader != leader || dj.inc != k) return;
  constexpr int n = kGblBytes / sizeof(elem_t<TI>);
  const elem_t<TI> *row = reinterpret_cast<const elem_t<TI> *>(lds + dj.lds);
#pragma unroll
  for (int c = 0; c < n; c++)
    if (c < dim) v[c] += row[c * stride + pos];
}

// target t of the dat of argument I (the first argument on that dat): its
// value plus its increments from every argument on the dat, in the order
// of the list (element order)
template <int I, int G, class Tup, size_t... J>
__device__ __forceinline__ void owner_apply(const Args<std::tuple_size<Tup>::value> &A, int t,
                                            const unsigned short *list, int cnt,
                                            const unsigned char *lds, int stride,
                                            std::index_sequence<J...>) {
  constexpr int N = std::tuple_size<Tup>::value;
  using TI = std::tuple_element_t<I, Tup>;
  if (!(ArgClass<I, N, G>::dat && written<TI>)) return;
  const Arg &d = A.a[I];
  if (d.leader != I) return;
  constexpr int n = kGblBytes / sizeof(elem_t<TI>);
  elem_t<TI> *dst = reinterpret_cast<elem_t<TI> *>(d.data + (size_t)t * d.bytes);
  elem_t<TI> v[n];
#pragma unroll
  for (int c = 0; c < n; c++)
    if (c < d.dim) v[c] = dst[c];
  for (int j = 0; j < cnt; j++) {
    const unsigned e = list[j];
    if (e == 0xffffu) break;  // (a padded record)
    const int k = (int)(e & 3u), pos = (int)(e >> 2);
    (owner_add<TI, std::tuple_element_t<J, Tup>, ArgClass<J, N, G>::dat>(A.a[J], I, k, pos, lds,
                                                                        stride, v, d.dim),
     ...);
  }
#pragma unroll
  for (int c = 0; c < n; c++)
    if (c < d.dim) dst[c] = v[c];
}

template <int G, class... T, size_t... I>
__device__ __forceinline__ void owner_apply_all(const Args<sizeof...(T)> &A, int t,
                                                const unsigned short *list, int cnt,
                                                const unsigned char *lds, int stride,
                                                std::index_sequence<I...> seq) {
  (owner_apply<I, G, std::tuple<T...>>(A, t, list, cnt, lds, stride, seq), ...);
}

template <auto F, int G, class... T>
__global__ __launch_bounds__(kThreads) void k_loop_owner(Args<sizeof...(T)> A, int ntiles,
                                                         int nred, OwnerArrays P,
                                                         double *partials) {
  extern __shared__ __align__(16) unsigned char inc_lds[];
  __shared__ double lds[kThreads / 64];
  constexpr auto seq = std::index_sequence_for<T...>{};
  std::tuple<Local<T>...> loc;
  init_all<G, T...>(A, loc, seq);
  const int b = xcd_tile(blockIdx.x, ntiles);
  const int p0 = P.blk_off[b], np = P.blk_off[b + 1] - p0;
  for (int p = threadIdx.x; p < np; p += kThreads) {
    std::tuple<Local<T>...> inc;
    owner_call<F, G, T...>(A, P, p0 + p, loc, inc, inc_lds, p, seq);
  }
  __syncthreads();
  const int o0 = P.owned_off[b], no = P.owned_off[b + 1] - o0;
  for (int i = threadIdx.x; i < no; i += kThreads) {
    const int t = P.owned ? P.owned[o0 + i] : o0 + i;
    if (P.ell) {  // the increments of the target as one 8-byte record
      const ushort4 r = reinterpret_cast<const ushort4 *>(P.ell)[o0 + i];
      const unsigned short es[4] = {r.x, r.y, r.z, r.w};
      owner_apply_all<G, T...>(A, t, es, 4, inc_lds, P.stride, seq);
    } else {
      const int j0 = P.csr_off[o0 + i];
      owner_apply_all<G, T...>(A, t, P.csr + j0, P.csr_off[o0 + i + 1] - j0, inc_lds, P.stride,
                               seq);
    }
  }
  if (nred > 0) reduce_all<G, T...>(A, loc, partials, b, nred, lds, seq);
}

// Folds the partials of every reduction component (one workgroup each).
template <int D = 0>
__global__ __launch_bounds__(kThreads) void k_loop_fold(const double *partials, int nslots,
                                                        RedKinds rk, double *out) {
  __shared__ double lds[kThreads / 64];
  const int c = blockIdx.x, kind = rk.kind[c];
  double v = red_identity(kind);
  for (int s = threadIdx.x; s < nslots; s += kThreads)
    v = red_op(kind, v, partials[(size_t)s * rk.n + c]);
  v = block_red(kind, v, lds);
  if (threadIdx.x == 0) out[c] = v;
}

// Deferred reductions: one workgroup folds every component in turn, each
// exactly as k_loop_fold does, and appends the results to the runtime's
// device log in component order (a loop's sum for OP_INC; for OP_MIN/OP_MAX
// the result including the caller's value at launch, as the caller would get).
template <int D = 0>
__global__ __launch_bounds__(kThreads) void k_loop_fold_log(const double *partials, int nslots,
                                                            RedKinds rk, double *log,
                                                            unsigned *count, unsigned mask) {
  __shared__ double lds[kThreads / 64];
  __shared__ unsigned base;
  if (threadIdx.x == 0) base = atomicAdd(count, (unsigned)rk.n);
  __syncthreads();
  for (int c = 0; c < rk.n; c++) {
    const int kind = rk.kind[c];
    double v = red_identity(kind);
    for (int s = threadIdx.x; s < nslots; s += kThreads)
      v = red_op(kind, v, partials[(size_t)s * rk.n + c]);
    v = block_red(kind, v, lds);
    if (threadIdx.x == 0) log[(base + c) & mask] = v;
    __syncthreads();
  }
}

inline int error(char const *name, char const *what) {
  char buf[512];
  snprintf(buf, sizeof buf, "op_par_loop error -- loop \"%s\": %s", name, what);
  return op2_report_error(buf);
}

template <class T>
inline bool type_ok(char const *type) {
  return type && type_error((elem_t<T> *)nullptr, type) == 0;
}

template <class T>
inline int fill(char const *name, const op_arg &g, Arg &d, RedKinds &rk, bool &plan) {
  if (g.argtype == OP_ARG_DAT) {
    if (!g.dat || !type_ok<T>(g.dat->type)) return error(name, "dat type does not match the kernel");
    if (g.dat->size % (int)sizeof(elem_t<T>)) return error(name, "dat size does not match the kernel");
    d.kind = kDat;
    d.data = (char *)op2_dat_device_ptr(g.dat);
    d.bytes = g.dat->size;
    d.dim = g.dat->dim;
    d.inc = d.leader = d.lds = d.slot = -1;
    if (!d.data) return error(name, "dat has no device data");
    if (g.map) {
      d.map = op2_map_device_ptr(g.map);
      if (!d.map) return -1;
      d.idx = g.idx;
      d.mapdim = g.map->dim;
      if (g.idx < 0 || g.idx >= g.map->dim) return error(name, "map index out of range");
      if (g.acc != OP_READ) plan = true;
    }
    return 0;
  }
  if (!type_ok<T>(g.type)) return error(name, "global argument type does not match the kernel");
  const int bytes = g.dim * (int)sizeof(elem_t<T>);
  if (g.dim <= 0 || bytes > kGblBytes) return error(name, "global argument larger than 64 bytes");
  d.dim = g.dim;
  switch (g.acc) {
    case OP_READ: d.kind = kGblRead; break;
    case OP_INC: d.kind = kGblInc; break;
    case OP_MIN: d.kind = kGblMin; break;
    case OP_MAX: d.kind = kGblMax; break;
    default: return error(name, "global arguments are OP_READ, OP_INC, OP_MIN or OP_MAX");
  }
  if (d.kind != kGblInc) memcpy(d.gbl, g.data, bytes);
  if (d.kind >= kGblInc) {
    if (rk.n + g.dim > kMaxRed) return error(name, "more than 64 reduction components");
    d.red = rk.n;
    for (int c = 0; c < g.dim; c++) rk.kind[rk.n + c] = (unsigned char)d.kind;
    rk.n += g.dim;
  }
  return 0;
}

template <class T>
inline void combine(const op_arg &g, const Arg &d, const double *h) {
  if (d.kind < kGblInc) return;
  elem_t<T> *user = reinterpret_cast<elem_t<T> *>(g.data);
  for (int c = 0; c < d.dim; c++) {
    if (d.kind == kGblInc)
      user[c] += (elem_t<T>)h[d.red + c];
    else
      user[c] = (elem_t<T>)h[d.red + c];
  }
}

template <class... T, size_t... I>
inline int fill_all(char const *name, op_arg *args, Args<sizeof...(T)> &A, RedKinds &rk,
                    bool &plan, std::index_sequence<I...>) {
  int rc = 0;
  ((rc = rc ? rc : fill<T>(name, args[I], A.a[I], rk, plan)), ...);
  return rc;
}

template <class... T, size_t... I>
inline void combine_all(op_arg *args, const Args<sizeof...(T)> &A, const double *h,
                        std::index_sequence<I...>) {
  (combine<T>(args[I], A.a[I], h), ...);
}

template <auto F, class... T>
int run(char const *name, op_set set, int nargs, op_arg *args, std::tuple<T...> *) {
  constexpr int N = sizeof...(T);
  constexpr auto seq = std::index_sequence_for<T...>{};
  if (nargs != N) return error(name, "number of arguments does not match the kernel");
  Args<N> A;
  memset(&A, 0, sizeof A);
  RedKinds rk;
  memset(&rk, 0, sizeof rk);
  bool plan = false;
  const int rc = fill_all<T...>(name, args, A, rk, plan, seq);
  if (rc) return rc;
  // the kernel specialised for the argument classes: G trailing OP_GBL
  // arguments (0-3) and dats before them, or the run-time form
  int G = 0;
  while (G < N && args[N - 1 - G].argtype == OP_ARG_GBL) G++;
  bool trailing = G <= 4;
  for (int i = 0; i < N - G; i++) trailing = trailing && args[i].argtype == OP_ARG_DAT;
  auto kern = !trailing ? k_loop<F, -1, T...>
              : G == 0  ? k_loop<F, 0, T...>
              : G == 1  ? k_loop<F, (N >= 1 ? 1 : 0), T...>
              : G == 2  ? k_loop<F, (N >= 2 ? 2 : 0), T...>
              : G == 3  ? k_loop<F, (N >= 3 ? 3 : 0), T...>
                        : k_loop<F, (N >= 4 ? 4 : 0), T...>;
  hipStream_t st = (hipStream_t)op2_stream();
  const int n = set->size;
  int nslots = 0;
  double *partials = nullptr;
  auto scratch = [&](int slots) -> int {
    nslots = slots;
    if (rk.n == 0) return 0;
    partials = (double *)op2_device_scratch(sizeof(double) * ((size_t)slots * rk.n + rk.n));
    return partials ? 0 : -1;
  };
  // owner tiles: argument classes known at compile time, and every written
  // argument an indirect OP_INC dat of at most kGblBytes per element or an
  // OP_GBL MIN/MAX; every other argument read only
  bool owner = plan && trailing && n > 0;
  {
    constexpr bool wr[N] = {written<T>...};
    for (int i = 0; i < N && owner; i++) {
      const op_arg &g = args[i];
      if (g.argtype == OP_ARG_DAT)
        owner = wr[i] ? (g.map && g.acc == OP_INC && g.dat->size <= kGblBytes) : g.acc == OP_READ;
      else
        owner = wr[i] ? (g.acc == OP_MIN || g.acc == OP_MAX) : g.acc == OP_READ;
    }
  }
  op2_generic_owner_plan op;
  int info[2 * N];  // per argument: its OP_INC pair, its place in a target row
  if (owner) owner = op2_generic_owner_plan_get(name, set, nargs, args, 0, &op, info) == 0;
  size_t owner_lds = 0;
  if (owner) {
    for (int i = 0; i < N; i++) {
      Arg &d = A.a[i];
      if (args[i].argtype != OP_ARG_DAT) continue;
      d.slot = info[N + i];
      if (info[i] < 0) continue;
      d.inc = info[i];
      d.leader = i;
      for (int j = 0; j < i && d.leader == i; j++)
        if (info[j] >= 0 && A.a[j].data == d.data) d.leader = j;
      d.lds = (int)owner_lds;
      owner_lds += ((size_t)d.bytes * op.max_positions + 15) / 16 * 16;
    }
    owner = owner_lds <= 160 * 1024;
  }
  if (owner) {
    auto okern = G == 0 ? k_loop_owner<F, 0, T...>
                 : G == 1 ? k_loop_owner<F, (N >= 1 ? 1 : 0), T...>
                 : G == 2 ? k_loop_owner<F, (N >= 2 ? 2 : 0), T...>
                 : G == 3 ? k_loop_owner<F, (N >= 3 ? 3 : 0), T...>
                          : k_loop_owner<F, (N >= 4 ? 4 : 0), T...>;
    if (scratch(op.nblocks)) return -1;
    const OwnerArrays P{op.blk_off, op.perm, op.ids,         op.owned_off,    op.owned,
                        op.csr_off, op.csr,  op.ell,         op.slot_stride, op.max_positions};
    if (owner_lds > 65536 &&
        hipFuncSetAttribute(reinterpret_cast<const void *>(okern),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)owner_lds) != hipSuccess)
      return error(name, "cannot raise the LDS limit of the owner-tile kernel");
    if (op.nblocks > 0)
      hipLaunchKernelGGL(okern, dim3(op.nblocks), dim3(kThreads), owner_lds, st, A, op.nblocks,
                         rk.n, P, partials);
  } else if (plan) {
    op2_generic_plan gp;
    if (op2_generic_plan_get(name, set, nargs, args, 0, &gp) != 0) return -1;
    if (scratch(gp.nblocks)) return -1;
    for (int c = 0; c < gp.ncolors; c++) {
      const int first = gp.colblk_off[c], nb = gp.colblk_off[c + 1] - first;
      if (nb == 0) continue;
      hipLaunchKernelGGL(kern, dim3(nb), dim3(kThreads), 0, st, A, n, rk.n,
                         gp.blk_off, gp.perm, gp.col_off, gp.col_base, gp.blk_ncol, gp.blkmap,
                         first, partials);
    }
  } else {
    int grid = (n + kThreads - 1) / kThreads;
    grid = grid < 1 ? 1 : grid > kDirectGrid ? kDirectGrid : grid;
    if (scratch(grid)) return -1;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreads), 0, st, A, n, rk.n,
                       (int const *)nullptr, (int const *)nullptr, (int const *)nullptr,
                       (int const *)nullptr, (int const *)nullptr, (int const *)nullptr, 0,
                       partials);
  }
  if (hipGetLastError() != hipSuccess) return error(name, "kernel launch failed");
  double *log = nullptr;
  unsigned *log_count = nullptr, log_cap = 0;
  if (rk.n > 0 && op2_reduction_log_device(&log, &log_count, &log_cap) == 1) {
    // deferred: the results go to the device log, the caller's variables are
    // left alone, nothing waits (an empty set logs each reduction's identity)
    hipLaunchKernelGGL(k_loop_fold_log<0>, dim3(1), dim3(kThreads), 0, st,
                       (const double *)partials, nslots, rk, log, log_count, log_cap - 1);
    if (hipGetLastError() != hipSuccess) return error(name, "reduction fold launch failed");
    return 0;
  }
  if (rk.n > 0 && nslots > 0) {  // (no blocks: no element, nothing to combine)
    double *out = partials + (size_t)nslots * rk.n;
    hipLaunchKernelGGL(k_loop_fold<0>, dim3(rk.n), dim3(kThreads), 0, st,
                       (const double *)partials, nslots, rk, out);
    double h[kMaxRed];
    if (hipMemcpyAsync(h, out, sizeof(double) * rk.n, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return error(name, "reduction result copy failed");
    combine_all<T...>(args, A, h, seq);
  }
  return 0;
}

template <class F>
struct Sig;
template <class... T>
struct Sig<void (*)(T *...)> {
  using args = std::tuple<T...>;
};

template <auto F>
int entry(char const *name, op_set set, int nargs, op_arg *args) {
  return run<F>(name, set, nargs, args, (typename Sig<decltype(F)>::args *)nullptr);
}

}  // namespace op2_loop

// Registers the __device__ user function `fn` as the loop named "fn".
#define OP2_DEVICE_LOOP(fn) \
  static const int op2_loop_registered_##fn = op2_register_loop(#fn, &op2_loop::entry<&fn>);

// Registers the __constant__ / __device__ variable `name` as the device copy
// of the program's constant of that name (op_decl_const, op_decl_const2).
#define OP2_DEVICE_CONST(name)                                                            \
  static void op2_const_copy_##name(void const *data, size_t bytes, void *stream) {       \
    if (bytes > sizeof(name)) {                                                           \
      op2_report_error("op_decl_const error -- constant " #name                           \
                       " is larger than its device copy");                                \
      return;                                                                             \
    }                                                                                     \
    if (hipMemcpyToSymbolAsync(HIP_SYMBOL(name), data, bytes, 0, hipMemcpyHostToDevice,   \
                               (hipStream_t)stream) != hipSuccess ||                      \
        hipStreamSynchronize((hipStream_t)stream) != hipSuccess)                          \
      op2_report_error("op_decl_const error -- copy of constant " #name " failed");       \
  }                                                                                       \
  static const int op2_const_registered_##name = op2_register_const(#name, op2_const_copy_##name);

#endif
