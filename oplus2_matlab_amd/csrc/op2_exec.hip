// Loop executor: recognises each op_par_loop by name and argument signature
// and runs its hand-written gfx950 kernel.
//
// Loops (reference user kernel / generated CUDA stub they replace):
//   save_soln  airfoil/save_soln.h,  airfoil/save_soln_kernel.cu:13-117
//   adt_calc   airfoil/adt_calc.h,   airfoil/adt_calc_kernel.cu:13-164
//   res_calc   airfoil/res_calc.h,   airfoil/res_calc_kernel.cu:13-245
//   bres_calc  airfoil/bres_calc.h,  airfoil/bres_calc_kernel.cu:13-231
//   update     airfoil/update.h,     airfoil/update_kernel.cu:13-182
//   res        jac1/res.h,           jac1/res_kernel.cu:13-217
//   update     jac1/update.h,        jac1/update_kernel.cu:13-153
// Dats of type "double" run the FP64 kernels.  Dats of type "float" run the
// precision form: float state and flux arithmetic, with update's
// qold - res/adt and every global reduction carried in double.
#include <algorithm>
#include <cstring>
#include <set>
#include <type_traits>

#include "op2_internal.h"
#include "op2_kernels.h"

namespace op2 {

// Threads per workgroup of the kernels whose size is fixed (the folds of
// reduction partials and the copies of graph buffers).  Every other kernel
// takes its workgroup size NT as a template parameter (128, 256 or 512: the
// loop's block size), and the host code launching it passes the same NT.
constexpr int kFixedBlock = 256;

// ============================================================================
// kernels
// ============================================================================

// save_soln: qold = q, as 16-byte chunks
typedef unsigned op2_v4u __attribute__((ext_vector_type(4)));

// save_soln: 16-byte copy.  The copy (qold) is next read by update, after
// adt_calc and res_calc: non-temporal stores keep it from displacing their
// data in the caches.
template <int NT>
__global__ __launch_bounds__(NT) void k_copy16(const uint4 *__restrict__ src,
                                               uint4 *__restrict__ dst, long long n) {
  long long i = (long long)blockIdx.x * NT + threadIdx.x;
  const long long stride = (long long)gridDim.x * NT;
  for (; i < n; i += stride) {
    const op2_v4u v = *reinterpret_cast<const op2_v4u *>(src + i);
    __builtin_nontemporal_store(v, reinterpret_cast<op2_v4u *>(dst + i));
  }
}

// The nodes of cell c gathered through pcell (measured: the non-temporal
// map load pays in FP64 only).
template <typename T>
__device__ __forceinline__ void cell_nodes(const int4 *__restrict__ pcell,
                                           const T *__restrict__ x, int c, T xs[4][2]) {
  int4 nd;
  if (sizeof(T) == 8) {  // FP64: the map is read once, keep it out of the caches
    typedef int v4 __attribute__((ext_vector_type(4)));
    const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4 *>(pcell + c));
    nd = make_int4(v.x, v.y, v.z, v.w);
  } else {
    nd = pcell[c];
  }
  Vec<T>::load2(x + 2 * (size_t)nd.x, xs[0]);
  Vec<T>::load2(x + 2 * (size_t)nd.y, xs[1]);
  Vec<T>::load2(x + 2 * (size_t)nd.z, xs[2]);
  Vec<T>::load2(x + 2 * (size_t)nd.w, xs[3]);
}

// adt of cell c with state qc
template <typename T>
__device__ __forceinline__ T cell_adt(const int4 *__restrict__ pcell, const T *__restrict__ x,
                                      int c, const T qc[4], const AirfoilConsts<T> &k) {
  T xs[4][2];
  cell_nodes(pcell, x, c, xs);
  return adt_of(xs[0], xs[1], xs[2], xs[3], qc, k);
}

// adt_calc: one cell per thread
template <int NT, typename T>
__global__ __launch_bounds__(NT) void k_adt_calc(const int4 *__restrict__ pcell,
                                                 const T *__restrict__ x,
                                                 const T *__restrict__ q, T *__restrict__ adt,
                                                 int ncell, AirfoilConsts<T> k) {
  // each XCD takes a contiguous range of cells, so the nodes that
  // neighbouring cells share are fetched into one L2
  const int c = xcd_block(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  if (c >= ncell) return;
  T qc[4];
  Vec<T>::load4(q + 4 * (size_t)c, qc);
  adt[c] = cell_adt(pcell, x, c, qc, k);
}

// save_soln followed by adt_calc on the same state, as one pass (the two
// loops of a captured graph fused): q is read once, copied to qold (bit for
// bit, non-temporal as in k_copy16) and used for the cell's adt.
template <int NT, typename T>
__global__ __launch_bounds__(NT) void k_save_adt(const int4 *__restrict__ pcell,
                                                 const T *__restrict__ x,
                                                 const T *__restrict__ q, T *__restrict__ qold,
                                                 T *__restrict__ adt, int ncell,
                                                 AirfoilConsts<T> k) {
  const int c = xcd_block(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  if (c >= ncell) return;
  T xs[4][2];
  cell_nodes(pcell, x, c, xs);  // issued first: the longest chain (map, then nodes)
  constexpr int NV = 4 * sizeof(T) / 16;  // 16-byte chunks of one cell's state
  const op2_v4u *src = reinterpret_cast<const op2_v4u *>(q + 4 * (size_t)c);
  op2_v4u *dst = reinterpret_cast<op2_v4u *>(qold + 4 * (size_t)c);
  op2_v4u v[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) v[i] = src[i];
#pragma unroll
  for (int i = 0; i < NV; i++) __builtin_nontemporal_store(v[i], dst + i);
  T qc[4];
  __builtin_memcpy(qc, v, sizeof qc);
  adt[c] = adt_of(xs[0], xs[1], xs[2], xs[3], qc, k);
}

// Sum of p[i] for i = first, first + stride, ... below end, added in that
// order from +0.0; eight loads are issued before their adds, so that their
// latencies overlap (the same bits as the plain loop).
__device__ __forceinline__ double strided_sum(const double *__restrict__ p, int first, int end,
                                              int stride) {
  double acc = 0.0;
  int i = first;
  for (; i + 7 * stride < end; i += 8 * stride) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = p[i + u * stride];
#pragma unroll
    for (int u = 0; u < 8; u++) acc += v[u];
  }
  for (; i < end; i += stride) acc += p[i];
  return acc;
}

// A fold of reduction partials carried by the next kernel of a captured
// graph instead of a launch of its own.  Single form: workgroup 0 of that
// kernel sums partials[0 .. n) exactly as k_fold<1> does, into *out and the
// log.  Batched form (nseg > 0): workgroup m < nseg sums mesh m's partials
// exactly as k_fold_seg does.
struct CarriedFold {
  const double *partials;
  int n;  // single form: number of partials (0: none)
  double *out, *log;
  unsigned *log_count;
  const int *pslot;  // batched form: partial slots per mesh
  int nseg;          // number of meshes (0: single form)
  double *dat;       // the per-mesh dat incremented
  const unsigned *log_base;
};

__device__ __forceinline__ bool carries_fold(const CarriedFold &f) {
  return f.nseg > 0 ? (int)blockIdx.x < f.nseg : (f.n > 0 && blockIdx.x == 0);
}

template <int NT>
__device__ __forceinline__ void carried_fold(const CarriedFold &f, double *lds) {
  double acc = 0.0;
  if (f.nseg > 0) {
    const int m = blockIdx.x;
    acc = strided_sum(f.partials, f.pslot[m] + threadIdx.x, f.pslot[m + 1], NT);
    acc = block_reduce<kRedSum, NT>(acc, lds);
    if (threadIdx.x == 0) {
      f.dat[m] += acc;
      if (f.log) f.log[(*f.log_base + m) & (kLogCap - 1)] = acc;
    }
  } else {
    acc = strided_sum(f.partials, threadIdx.x, f.n, NT);
    acc = block_reduce<kRedSum, NT>(acc, lds);
    if (threadIdx.x == 0) {
      *f.out = acc;
      if (f.log) f.log[atomicAdd(f.log_count, 1u) & (kLogCap - 1)] = acc;
    }
  }
  __syncthreads();  // lds is reused by the caller
}

// Staged edge loop (res_calc).  One workgroup per plan block:
//  1. stage x of the block's nodes and q, adt of its cells in LDS;
//  2. compute the flux of each edge (EPT edges per thread) into registers;
//  3. reuse the q area as the block's res accumulator:
//     COLOURED - thread colours in turn, one barrier per colour;
//     otherwise - LDS atomic adds;
//  4. add the accumulator into res: plain read-modify-write when blocks of
//     one launch share no cell (block colours), global atomics otherwise.
template <int NT, typename T, int EPT, bool COLOURED>
__global__ __launch_bounds__(NT) void k_res_calc(
    const int *__restrict__ blk_off, const int *__restrict__ blkmap, int blk_first, int nblk,
    const int *__restrict__ node_off, const int *__restrict__ node_list,
    const int *__restrict__ cell_off, const int *__restrict__ cell_list,
    const ushort4 *__restrict__ loc, const int *__restrict__ col_off,
    const int *__restrict__ col_base, const int *__restrict__ blk_ncol, const T *__restrict__ x,
    const T *__restrict__ q, const T *__restrict__ adt, T *__restrict__ res,
    AirfoilConsts<T> k, CarriedFold fold) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (carries_fold(fold)) carried_fold<NT>(fold, reinterpret_cast<double *>(smem));
  const int b = blkmap[blk_first + xcd_block(blockIdx.x, nblk)];
  const int e0 = blk_off[b], ne = blk_off[b + 1] - e0;
  const int n0 = node_off[b], nn = node_off[b + 1] - n0;
  const int c0 = cell_off[b], nc = cell_off[b + 1] - c0;
  T *xs = reinterpret_cast<T *>(smem);                   // [nn][2]
  T *qs = xs + 2 * ((nn + 1) & ~1);                      // [nc][4], 16-byte aligned
  T *as = qs + 4 * nc;                                   // [nc]
  const int tid = threadIdx.x;

  for (int i = tid; i < nn; i += NT) {
    T v[2];
    Vec<T>::load2(x + 2 * (size_t)node_list[n0 + i], v);
    Vec<T>::store2(xs + 2 * i, v);
  }
  for (int i = tid; i < nc; i += NT) {
    const int g = cell_list[c0 + i];
    T v[4];
    Vec<T>::load4(q + 4 * (size_t)g, v);
    Vec<T>::store4(qs + 4 * i, v);
    as[i] = adt[g];
  }
  __syncthreads();

  T f[EPT][4];
  int l1[EPT], l2[EPT];
#pragma unroll
  for (int j = 0; j < EPT; j++) {
    const int e = tid + j * NT;
    l1[j] = -1;
    if (e < ne) {
      const ushort4 L = loc[e0 + e];
      T x1[2], x2[2], q1[4], q2[4];
      Vec<T>::load2(xs + 2 * L.x, x1);
      Vec<T>::load2(xs + 2 * L.y, x2);
      Vec<T>::load4(qs + 4 * L.z, q1);
      Vec<T>::load4(qs + 4 * L.w, q2);
      edge_flux(x1, x2, q1, q2, as[L.z], as[L.w], k, f[j]);
      l1[j] = L.z;
      l2[j] = L.w;
    }
  }
  __syncthreads();
  T *rs = qs;
  for (int i = tid; i < 4 * nc; i += NT) rs[i] = T(0);
  __syncthreads();

  if (COLOURED) {
    const int cb = col_base[b], ncol = blk_ncol[b];
    for (int c = 0; c < ncol; c++) {
      const int cs = col_off[cb + c] - e0, ce = col_off[cb + c + 1] - e0;
#pragma unroll
      for (int j = 0; j < EPT; j++) {
        const int e = tid + j * NT;
        if (e >= cs && e < ce) {
#pragma unroll
          for (int d = 0; d < 4; d++) rs[4 * l1[j] + d] += f[j][d];
#pragma unroll
          for (int d = 0; d < 4; d++) rs[4 * l2[j] + d] -= f[j][d];
        }
      }
      __syncthreads();
    }
    for (int i = tid; i < nc; i += NT) {
      const int g = cell_list[c0 + i];
      T r[4], a[4];
      Vec<T>::load4(res + 4 * (size_t)g, r);
      Vec<T>::load4(rs + 4 * i, a);
#pragma unroll
      for (int d = 0; d < 4; d++) r[d] += a[d];
      Vec<T>::store4(res + 4 * (size_t)g, r);
    }
  } else {
#pragma unroll
    for (int j = 0; j < EPT; j++)
      if (l1[j] >= 0) {
#pragma unroll
        for (int d = 0; d < 4; d++) atomicAdd(rs + 4 * l1[j] + d, f[j][d]);
#pragma unroll
        for (int d = 0; d < 4; d++) atomicAdd(rs + 4 * l2[j] + d, -f[j][d]);
      }
    __syncthreads();
    for (int i = tid; i < 4 * nc; i += NT)
      atomicAdd(res + 4 * (size_t)cell_list[c0 + (i >> 2)] + (i & 3), rs[i]);
  }
}

// Owner-compute edge loop (res_calc).  One workgroup per tile of cells:
//  1. stage x of the tile's nodes and q, adt of its cells (owned + halo);
//  2. one flux per element position into LDS (elements between two tiles are
//     evaluated by both);
//  3. one thread per owned cell sums its increments in the plan's fixed
//     order and adds them to res with a plain read-modify-write.
// No atomics, no colours, one launch, deterministic.
template <int NT, typename T, int OWN, bool ELL>
__global__ __launch_bounds__(NT) void k_res_calc_owner(
    const int *__restrict__ blk_off, int nblk, const int *__restrict__ node_off,
    const int *__restrict__ node_list, const int *__restrict__ cell_off,
    const int *__restrict__ cell_list, const ushort4 *__restrict__ loc,
    const int *__restrict__ owned_off, const int *__restrict__ owned,
    const int *__restrict__ csr_off, const uint16_t *__restrict__ csr,
    const ushort4 *__restrict__ ell, const T *__restrict__ x,
    const T *__restrict__ q, const T *__restrict__ adt, T *__restrict__ res,
    AirfoilConsts<T> k, CarriedFold fold, bool zero_res) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (carries_fold(fold)) carried_fold<NT>(fold, reinterpret_cast<double *>(smem));
  constexpr int pad = 16 / sizeof(T);
  const int b = xcd_block(blockIdx.x, nblk);
  const int p0 = blk_off[b], np = blk_off[b + 1] - p0;
  const int n0 = node_off[b], nn = node_off[b + 1] - n0;
  const int c0 = cell_off[b], nc = cell_off[b + 1] - c0;
  const int o0 = owned_off[b], no = owned_off[b + 1] - o0;
  T *xs = reinterpret_cast<T *>(smem);
  T *qs = xs + 2 * ((nn + 1) & ~1);
  T *as = qs + 4 * nc;
  T *fs = as + ((nc + pad - 1) / pad) * pad;
  const int tid = threadIdx.x;
  // the res of this thread's owned cells is read now, so that its latency
  // overlaps the staging and the flux phases (OWN owned cells per thread)
  T rpre[OWN][4];
  int gown[OWN];
  ushort4 e4[OWN];
#pragma unroll
  for (int j = 0; j < OWN; j++) {
    const int i = tid + j * NT;
    gown[j] = i < no ? owned[o0 + i] : -1;
    e4[j] = ELL && i < no ? ell[o0 + i] : make_ushort4(0xffff, 0xffff, 0xffff, 0xffff);
    if (zero_res) {  // res known to hold zeros (as in k_res_calc_owner_direct)
#pragma unroll
      for (int d = 0; d < 4; d++) rpre[j][d] = T(0);
    } else if (gown[j] >= 0) {
      Vec<T>::load4(res + 4 * (size_t)gown[j], rpre[j]);
    }
  }
  for (int i = tid; i < nn; i += NT) {
    T v[2];
    Vec<T>::load2(x + 2 * (size_t)node_list[n0 + i], v);
    Vec<T>::store2(xs + 2 * i, v);
  }
  for (int i = tid; i < nc; i += NT) {
    const int g = cell_list[c0 + i];
    T v[4];
    Vec<T>::load4(q + 4 * (size_t)g, v);
    Vec<T>::store4(qs + 4 * i, v);
    as[i] = adt[g];
  }
  __syncthreads();
  for (int p = tid; p < np; p += NT) {
    const ushort4 L = loc[p0 + p];
    T x1[2], x2[2], q1[4], q2[4], f[4];
    Vec<T>::load2(xs + 2 * L.x, x1);
    Vec<T>::load2(xs + 2 * L.y, x2);
    Vec<T>::load4(qs + 4 * L.z, q1);
    Vec<T>::load4(qs + 4 * L.w, q2);
    edge_flux(x1, x2, q1, q2, as[L.z], as[L.w], k, f);
    Vec<T>::store4(fs + 4 * p, f);
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < OWN; j++) {
    const int i = tid + j * NT;
    if (gown[j] < 0) continue;
    T acc[4] = {T(0), T(0), T(0), T(0)};
    const unsigned short es[4] = {e4[j].x, e4[j].y, e4[j].z, e4[j].w};
    const int c0 = ELL ? 0 : csr_off[o0 + i], j1 = ELL ? 4 : csr_off[o0 + i + 1];
    for (int c = c0; c < j1; c++) {
      const unsigned e = ELL ? es[c] : csr[c];
      if (ELL && e == 0xffffu) break;
      T f[4];
      Vec<T>::load4(fs + 4 * (e >> 2), f);
      if (e & 3u) {
#pragma unroll
        for (int d = 0; d < 4; d++) acc[d] -= f[d];
      } else {
#pragma unroll
        for (int d = 0; d < 4; d++) acc[d] += f[d];
      }
    }
#pragma unroll
    for (int d = 0; d < 4; d++) rpre[j][d] += acc[d];
    Vec<T>::store4(res + 4 * (size_t)gown[j], rpre[j]);
  }
}

// Sum of the fluxes that increment one owned cell, in the plan's fixed
// order: entries (position << 2 | inc slot) of its ELL record, slot 0 adds
// the flux, slot 1 subtracts it (airfoil/res_calc.h:25-29).  fs is
// component-major, fs[d * np + position].
template <typename T>
__device__ __forceinline__ void ell_sum(const T *fs, int np, ushort4 e4, T acc[4]) {
  const unsigned short es[4] = {e4.x, e4.y, e4.z, e4.w};
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const unsigned e = es[s];
    if (e == 0xffffu) break;
    const int pos = e >> 2;
    if (e & 3u) {
#pragma unroll
      for (int d = 0; d < 4; d++) acc[d] -= fs[d * np + pos];
    } else {
#pragma unroll
      for (int d = 0; d < 4; d++) acc[d] += fs[d * np + pos];
    }
  }
}

// Owner-compute edge loop without staging (res_calc): the same tiles, but
// each element position carries the global ids of its nodes and cells, and
// x, q, adt are gathered straight from global memory (the tile's records
// stay in the L1/L2).  Fluxes meet in LDS, component-major (8-byte lanes,
// no bank conflicts), for the fixed-order sums of the owned cells.
// ELL: the owned cells' increment lists come as one 8-byte record each,
// loaded at the start independently of everything else, instead of
// csr_off/csr loads after the barrier.  OWNID: the tiles own consecutive
// cells (renumbered sets), so no owned list is read.  The position ids and
// ELL records are read once: non-temporal loads keep them from evicting
// the gathered x, q and adt from the caches.
typedef int op2_v4i __attribute__((ext_vector_type(4)));
typedef unsigned short op2_v4us __attribute__((ext_vector_type(4)));

// A target of a 16-bit position record: offset h & 0x7fff in the first of
// the tile's two windows of its map (base w0), or in the second (base w0 + d)
// when bit 15 is set.  w0 and d are uniform (scalar registers).
__device__ __forceinline__ int window_id(unsigned h, int w0, int d) {
  return w0 + (int)(h & 0x7fffu) + (((int)(h << 16) >> 31) & d);
}

// The node and cell ids of a position: SHORT, its 8-byte record of 16-bit
// window offsets (w: the tile's {node base, node step, cell base, cell
// step}), else its 16-byte record of global ids.
template <bool SHORT>
__device__ __forceinline__ op2_v4i position_ids(const int4 *__restrict__ ids,
                                                const uint16_t *__restrict__ ids16, int4 w,
                                                int pos) {
  if (SHORT) {
    const op2_v4us h = __builtin_nontemporal_load((const op2_v4us *)(ids16 + 4 * (size_t)pos));
    return op2_v4i{window_id(h.x, w.x, w.y), window_id(h.y, w.x, w.y),
                   window_id(h.z, w.z, w.w), window_id(h.w, w.z, w.w)};
  }
  return __builtin_nontemporal_load((const op2_v4i *)(ids + pos));
}

// One edge of a gather tile: gathers, flux into LDS (component-major) and
// the edge's (x1 - x2).  Each gather instruction takes the lower-numbered
// node (cell) of every position, the next one the higher-numbered:
// positions follow their smaller cell, so neighbouring lanes then read
// neighbouring records (fewer L1 lines per instruction than the map's slot
// order, whose orientation is arbitrary).  The pairs are swapped back in
// registers, so the flux sees the map's order.
template <bool SHORT, typename T>
__device__ __forceinline__ void tile_edge(const int4 *__restrict__ ids,
                                          const uint16_t *__restrict__ ids16, int4 w,
                                          const T *__restrict__ x, const T *__restrict__ q,
                                          const T *__restrict__ adt, const AirfoilConsts<T> &k,
                                          T *fs, int np, int p0, int p, T side[2]) {
  const op2_v4i g = position_ids<SHORT>(ids, ids16, w, p0 + p);
  const int na = min(g.x, g.y), nb = max(g.x, g.y);
  const int ca = min(g.z, g.w), cb = max(g.z, g.w);
  T xa[2], xb[2], qa[4], qb[4], f[4];
  Vec<T>::load2(x + 2 * (size_t)na, xa);
  Vec<T>::load2(x + 2 * (size_t)nb, xb);
  Vec<T>::load4(q + 4 * (size_t)ca, qa);
  Vec<T>::load4(q + 4 * (size_t)cb, qb);
  const T aa = adt[ca], ab = adt[cb];
  const bool sn = g.x != na, sc = g.z != ca;
  T x1[2], x2[2], q1[4], q2[4];
#pragma unroll
  for (int d = 0; d < 2; d++) {
    x1[d] = sn ? xb[d] : xa[d];
    x2[d] = sn ? xa[d] : xb[d];
  }
#pragma unroll
  for (int d = 0; d < 4; d++) {
    q1[d] = sc ? qb[d] : qa[d];
    q2[d] = sc ? qa[d] : qb[d];
  }
  edge_flux(x1, x2, q1, q2, sc ? ab : aa, sc ? aa : ab, k, f);
#pragma unroll
  for (int d = 0; d < 4; d++) fs[d * np + p] = f[d];
  side[0] = x1[0] - x2[0];
  side[1] = x1[1] - x2[1];
}

template <int NT, typename T, int OWN, bool ELL, bool OWNID>
__global__ __launch_bounds__(NT) void k_res_calc_owner_direct(
    const int *__restrict__ blk_off, int nblk, const int4 *__restrict__ ids,
    const uint16_t *__restrict__ ids16, const int4 *__restrict__ id_win,
    const int *__restrict__ owned_off, const int *__restrict__ owned,
    const int *__restrict__ csr_off, const uint16_t *__restrict__ csr,
    const ushort4 *__restrict__ ell, const T *__restrict__ x, const T *__restrict__ q,
    const T *__restrict__ adt, T *__restrict__ res, AirfoilConsts<T> k, CarriedFold fold,
    bool zero_res) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (carries_fold(fold)) carried_fold<NT>(fold, reinterpret_cast<double *>(smem));
  const int b = xcd_block(blockIdx.x, nblk);
  const int p0 = blk_off[b], np = blk_off[b + 1] - p0;
  const int o0 = owned_off[b], no = owned_off[b + 1] - o0;
  T *fs = reinterpret_cast<T *>(smem);  // [4][np]
  const int tid = threadIdx.x;
  int gown[OWN];
  ushort4 e4[OWN];
#pragma unroll
  for (int j = 0; j < OWN; j++) {
    const int i = tid + j * NT;
    gown[j] = i < no ? (OWNID ? o0 + i : owned[o0 + i]) : -1;
    if (ELL) {
      if (i < no) {
        const op2_v4us v = __builtin_nontemporal_load((const op2_v4us *)(ell + o0 + i));
        e4[j] = make_ushort4(v.x, v.y, v.z, v.w);
      } else {
        e4[j] = make_ushort4(0xffff, 0xffff, 0xffff, 0xffff);
      }
    }
  }
  // zero_res: res is known to hold zeros (a captured graph's update wrote
  // them), so its read is skipped; the sum still starts from +0
  T rpre[OWN][4];
#pragma unroll
  for (int j = 0; j < OWN; j++) {
    if (zero_res) {
#pragma unroll
      for (int d = 0; d < 4; d++) rpre[j][d] = T(0);
    } else if (gown[j] >= 0) {
      Vec<T>::load4(res + 4 * (size_t)gown[j], rpre[j]);
    }
  }
  // FP32: tiles with 16-bit records read those (a uniform branch, so each
  // form's gathers schedule as straight-line code; in FP64 the decode's
  // registers would cost a wave per SIMD)
  bool short_done = false;
  if constexpr (std::is_same<T, float>::value) {
    if (ids16) {
      const int4 w = id_win[b];
      if (w.x >= 0) {
        for (int p = tid; p < np; p += NT) {
          T unused[2];
          tile_edge<true>(ids, ids16, w, x, q, adt, k, fs, np, p0, p, unused);
        }
        short_done = true;
      }
    }
  }
  if (!short_done)
    for (int p = tid; p < np; p += NT) {
      T unused[2];
      tile_edge<false>(ids, ids16, int4{}, x, q, adt, k, fs, np, p0, p, unused);
    }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < OWN; j++) {
    const int i = tid + j * NT;
    if (gown[j] < 0) continue;
    T acc[4] = {T(0), T(0), T(0), T(0)};
    if (ELL) {
      ell_sum(fs, np, e4[j], acc);
    } else {
      const int j1 = csr_off[o0 + i + 1];
      for (int c = csr_off[o0 + i]; c < j1; c++) {
        const unsigned e = csr[c];
        const int pos = e >> 2;
        if (e & 3u) {
#pragma unroll
          for (int d = 0; d < 4; d++) acc[d] -= fs[d * np + pos];
        } else {
#pragma unroll
          for (int d = 0; d < 4; d++) acc[d] += fs[d * np + pos];
        }
      }
    }
#pragma unroll
    for (int d = 0; d < 4; d++) rpre[j][d] += acc[d];
    Vec<T>::store4(res + 4 * (size_t)gown[j], rpre[j]);
  }
}

// Owner-compute boundary loop (bres_calc): one increment per element.
template <int NT, typename T>
__global__ __launch_bounds__(NT) void k_bres_calc_owner(
    const int *__restrict__ blk_off, int nblk, const int *__restrict__ perm,
    const int *__restrict__ node_off, const int *__restrict__ node_list,
    const int *__restrict__ cell_off, const int *__restrict__ cell_list,
    const ushort4 *__restrict__ loc, const int *__restrict__ owned_off,
    const int *__restrict__ owned, const int *__restrict__ csr_off,
    const uint16_t *__restrict__ csr, const T *__restrict__ x, const T *__restrict__ q,
    const T *__restrict__ adt, T *__restrict__ res, const int *__restrict__ bound,
    const T *__restrict__ qinf_dat, const int *__restrict__ bmap, AirfoilConsts<T> k) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int pad = 16 / sizeof(T);
  const int b = xcd_block(blockIdx.x, nblk);
  const int p0 = blk_off[b], np = blk_off[b + 1] - p0;
  const int n0 = node_off[b], nn = node_off[b + 1] - n0;
  const int c0 = cell_off[b], nc = cell_off[b + 1] - c0;
  const int o0 = owned_off[b], no = owned_off[b + 1] - o0;
  T *xs = reinterpret_cast<T *>(smem);
  T *qs = xs + 2 * ((nn + 1) & ~1);
  T *as = qs + 4 * nc;
  T *fs = as + ((nc + pad - 1) / pad) * pad;
  const int tid = threadIdx.x;
  for (int i = tid; i < nn; i += NT) {
    T v[2];
    Vec<T>::load2(x + 2 * (size_t)node_list[n0 + i], v);
    Vec<T>::store2(xs + 2 * i, v);
  }
  for (int i = tid; i < nc; i += NT) {
    const int g = cell_list[c0 + i];
    T v[4];
    Vec<T>::load4(q + 4 * (size_t)g, v);
    Vec<T>::store4(qs + 4 * i, v);
    as[i] = adt[g];
  }
  __syncthreads();
  for (int p = tid; p < np; p += NT) {
    const ushort4 L = loc[p0 + p];
    T x1[2], x2[2], q1[4], f[4];
    Vec<T>::load2(xs + 2 * L.x, x1);
    Vec<T>::load2(xs + 2 * L.y, x2);
    Vec<T>::load4(qs + 4 * L.z, q1);
    const int el = perm[p0 + p];
    const T *qi = bmap ? qinf_dat + 4 * (size_t)bmap[el] : k.qinf;
    boundary_flux(x1, x2, q1, as[L.z], bound[el], qi, k, f);
    Vec<T>::store4(fs + 4 * p, f);
  }
  __syncthreads();
  for (int i = tid; i < no; i += NT) {
    T acc[4] = {T(0), T(0), T(0), T(0)};
    const int j1 = csr_off[o0 + i + 1];
    for (int j = csr_off[o0 + i]; j < j1; j++) {
      T f[4];
      Vec<T>::load4(fs + 4 * (csr[j] >> 2), f);
#pragma unroll
      for (int d = 0; d < 4; d++) acc[d] += f[d];
    }
    const size_t g = 4 * (size_t)owned[o0 + i];
    T r[4];
    Vec<T>::load4(res + g, r);
#pragma unroll
    for (int d = 0; d < 4; d++) r[d] += acc[d];
    Vec<T>::store4(res + g, r);
  }
}

// Direct edge loop (res_calc, no plan): each thread reads its edge's maps,
// gathers x, q and adt straight from global memory and adds its flux to
// both cells with global atomics.  No barriers and no LDS, so every wave
// keeps its loads in flight; the L2 supplies the reuse between edges.
template <int NT, typename T>
__global__ __launch_bounds__(NT) void k_res_calc_direct(const int2 *__restrict__ pedge,
                                                        const int2 *__restrict__ pecell,
                                                        const T *__restrict__ x,
                                                        const T *__restrict__ q,
                                                        const T *__restrict__ adt,
                                                        T *__restrict__ res, int nedge,
                                                        AirfoilConsts<T> k, CarriedFold fold) {
  if (carries_fold(fold)) {
    __shared__ double lds[NT / kWave];
    carried_fold<NT>(fold, lds);
  }
  const int e = xcd_block(blockIdx.x, gridDim.x) * NT + threadIdx.x;
  if (e >= nedge) return;
  const int2 n = pedge[e];
  const int2 c = pecell[e];
  T x1[2], x2[2], q1[4], q2[4], f[4];
  Vec<T>::load2(x + 2 * (size_t)n.x, x1);
  Vec<T>::load2(x + 2 * (size_t)n.y, x2);
  Vec<T>::load4(q + 4 * (size_t)c.x, q1);
  Vec<T>::load4(q + 4 * (size_t)c.y, q2);
  const T a1 = adt[c.x], a2 = adt[c.y];
  edge_flux(x1, x2, q1, q2, a1, a2, k, f);
#pragma unroll
  for (int d = 0; d < 4; d++) atomicAdd(res + 4 * (size_t)c.x + d, f[d]);
#pragma unroll
  for (int d = 0; d < 4; d++) atomicAdd(res + 4 * (size_t)c.y + d, -f[d]);
}

// Boundary-edge loop without a plan (bres_calc): one thread per boundary
// edge, global atomics.  The loop is small (one edge per boundary cell on
// the O-mesh), so its time is latency: a single pass of dependent loads.
// qinf_dat / bmap: the batched form's free stream of the edge's mesh.
template <int NT, typename T>
__global__ __launch_bounds__(NT) void k_bres_calc_direct(
    const int2 *__restrict__ pbedge, const int *__restrict__ pbecell, const T *__restrict__ x,
    const T *__restrict__ q, const T *__restrict__ adt, T *__restrict__ res,
    const int *__restrict__ bound, const T *__restrict__ qinf_dat, const int *__restrict__ bmap,
    int nbedge, AirfoilConsts<T> k) {
  const int e = blockIdx.x * NT + threadIdx.x;
  if (e >= nbedge) return;
  const int2 n = pbedge[e];
  const int c = pbecell[e];
  T x1[2], x2[2], q1[4], f[4];
  Vec<T>::load2(x + 2 * (size_t)n.x, x1);
  Vec<T>::load2(x + 2 * (size_t)n.y, x2);
  Vec<T>::load4(q + 4 * (size_t)c, q1);
  const T *qi = qinf_dat ? qinf_dat + 4 * (size_t)bmap[e] : k.qinf;
  boundary_flux(x1, x2, q1, adt[c], bound[e], qi, k, f);
#pragma unroll
  for (int d = 0; d < 4; d++) atomicAdd(res + 4 * (size_t)c + d, f[d]);
}

// Staged boundary-edge loop (bres_calc): as k_res_calc with one increment per
// element and the direct `bound` read through the plan's element order.
template <int NT, typename T, bool COLOURED>
__global__ __launch_bounds__(NT) void k_bres_calc(
    const int *__restrict__ blk_off, const int *__restrict__ blkmap, int blk_first, int nblk,
    const int *__restrict__ perm, const int *__restrict__ node_off,
    const int *__restrict__ node_list, const int *__restrict__ cell_off,
    const int *__restrict__ cell_list, const ushort4 *__restrict__ loc,
    const int *__restrict__ col_off, const int *__restrict__ col_base,
    const int *__restrict__ blk_ncol, const T *__restrict__ x, const T *__restrict__ q,
    const T *__restrict__ adt, T *__restrict__ res, const int *__restrict__ bound,
    const T *__restrict__ qinf_dat, const int *__restrict__ bmap, AirfoilConsts<T> k) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int b = blkmap[blk_first + xcd_block(blockIdx.x, nblk)];
  const int e0 = blk_off[b], ne = blk_off[b + 1] - e0;
  const int n0 = node_off[b], nn = node_off[b + 1] - n0;
  const int c0 = cell_off[b], nc = cell_off[b + 1] - c0;
  T *xs = reinterpret_cast<T *>(smem);
  T *qs = xs + 2 * ((nn + 1) & ~1);
  T *as = qs + 4 * nc;
  T *rs = as + ((nc + 3) & ~3);
  const int tid = threadIdx.x;
  for (int i = tid; i < nn; i += NT) {
    T v[2];
    Vec<T>::load2(x + 2 * (size_t)node_list[n0 + i], v);
    Vec<T>::store2(xs + 2 * i, v);
  }
  for (int i = tid; i < nc; i += NT) {
    const int g = cell_list[c0 + i];
    T v[4];
    Vec<T>::load4(q + 4 * (size_t)g, v);
    Vec<T>::store4(qs + 4 * i, v);
    as[i] = adt[g];
  }
  for (int i = tid; i < 4 * nc; i += NT) rs[i] = T(0);
  __syncthreads();
  const int cb = COLOURED ? col_base[b] : 0;
  const int ncol = COLOURED ? blk_ncol[b] : 1;
  for (int c = 0; c < ncol; c++) {
    const int cs = COLOURED ? col_off[cb + c] - e0 : 0;
    const int ce = COLOURED ? col_off[cb + c + 1] - e0 : ne;
    for (int e = cs + tid; e < ce; e += NT) {
      const ushort4 L = loc[e0 + e];
      const int el = perm ? perm[e0 + e] : e0 + e;
      T x1[2], x2[2], q1[4], f[4];
      Vec<T>::load2(xs + 2 * L.x, x1);
      Vec<T>::load2(xs + 2 * L.y, x2);
      Vec<T>::load4(qs + 4 * L.z, q1);
      const T *qi = bmap ? qinf_dat + 4 * (size_t)bmap[el] : k.qinf;
      boundary_flux(x1, x2, q1, as[L.z], bound[el], qi, k, f);
      if (COLOURED) {
#pragma unroll
        for (int d = 0; d < 4; d++) rs[4 * L.z + d] += f[d];
      } else {
#pragma unroll
        for (int d = 0; d < 4; d++) atomicAdd(rs + 4 * L.z + d, f[d]);
      }
    }
    __syncthreads();
  }
  if (COLOURED) {
    for (int i = tid; i < nc; i += NT) {
      const int g = cell_list[c0 + i];
      T r[4], a[4];
      Vec<T>::load4(res + 4 * (size_t)g, r);
      Vec<T>::load4(rs + 4 * i, a);
#pragma unroll
      for (int d = 0; d < 4; d++) r[d] += a[d];
      Vec<T>::store4(res + 4 * (size_t)g, r);
    }
  } else {
    for (int i = tid; i < 4 * nc; i += NT)
      atomicAdd(res + 4 * (size_t)cell_list[c0 + (i >> 2)] + (i & 3), rs[i]);
  }
}

// update: q = qold - res/adt, res = 0, rms += sum of squared moves; the
// move and the new q are formed in double and q rounded once to T.  Each
// workgroup leaves its partial sum in partials[blockIdx.x]; k_fold adds them.
// Loops of a captured graph fused into one pass: SAVE - a save_soln of the
// q update writes follows (the next iteration's first loop): the new q is
// stored to qold_out as well; ADT - an adt_calc on that q follows: the
// cell's adt is computed from the q just formed.  qold_out may be qold and
// adt_out may be adt (each cell reads its old values before writing).
// BRES - the bres_calc that precedes update (after res_calc) is applied by
// it: a boundary cell first adds the boundary flux of each of its boundary
// edges to res, in edge order, exactly as bres_calc's increments would, then
// updates.  A cell finds its boundary edges through one 8-byte record per
// 32 cells (a bit mask of the cells with boundary edges and the number of
// such cells in earlier records), then two 16-byte records per boundary
// cell (nodes, edge and mesh of up to two boundary edges, copied from the
// maps; edge -1 when absent): two dependent loads before the gathers.
template <typename T>
struct BresIdx {
  const uint2 *word;  // per 32 cells: (mask, boundary cells before)
  const int4 *rec;    // per boundary cell: 2 x (node 1, node 2, edge, mesh)
  const T *x, *q;
  const int *bound;
  const T *qinf_dat;  // batched form: the free stream of each mesh
  // the zeros update leaves in res are not stored: the res_calc that follows
  // in the graph overwrites every cell's res without reading it
  bool keep_res;
};

template <typename T, bool ADT>
struct UpdIn {
  T qo[4], r[4], a, xs[4][2];
  int bidx;
};

template <typename T, bool ADT, bool BRES>
__device__ __forceinline__ void upd_load(int c, const T *qold, const T *res, const T *adt,
                                         const int4 *pcell, const T *x, const BresIdx<T> &bi,
                                         UpdIn<T, ADT> &in) {
  if (ADT) cell_nodes(pcell, x, c, in.xs);  // the longest chain: issued first
  if (BRES) {
    const uint2 w = bi.word[c >> 5];
    const unsigned bit = 1u << (c & 31);
    in.bidx = (w.x & bit) ? (int)(w.y + __popc(w.x & (bit - 1u))) : -1;
  }
  Vec<T>::load4(qold + 4 * (size_t)c, in.qo);
  Vec<T>::load4(res + 4 * (size_t)c, in.r);
  in.a = adt[c];
}

template <typename T, bool ADT, bool SAVE, bool BRES>
__device__ __forceinline__ void upd_store(int c, const UpdIn<T, ADT> &in, T *q, T *res,
                                          T *adt_out, T *qold_out, const BresIdx<T> &bi,
                                          const AirfoilConsts<T> &k, double &acc) {
  T qn[4], r[4] = {in.r[0], in.r[1], in.r[2], in.r[3]};
  if (BRES && in.bidx >= 0) {
    T qc[4];
    Vec<T>::load4(bi.q + 4 * (size_t)c, qc);
    const int4 rec[2] = {bi.rec[2 * in.bidx], bi.rec[2 * in.bidx + 1]};
#pragma unroll
    for (int j = 0; j < 2; j++) {
      if (rec[j].z < 0) break;
      T x1[2], x2[2], f[4];
      Vec<T>::load2(bi.x + 2 * (size_t)rec[j].x, x1);
      Vec<T>::load2(bi.x + 2 * (size_t)rec[j].y, x2);
      const T *qi = bi.qinf_dat ? bi.qinf_dat + 4 * (size_t)rec[j].w : k.qinf;
      boundary_flux(x1, x2, qc, in.a, bi.bound[rec[j].z], qi, k, f);
#pragma unroll
      for (int d = 0; d < 4; d++) r[d] += f[d];
    }
  }
  const double adti = 1.0 / (double)in.a;
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const double del = adti * (double)r[d];
    qn[d] = (T)((double)in.qo[d] - del);
    acc += del * del;
    r[d] = T(0);
  }
  Vec<T>::store4(q + 4 * (size_t)c, qn);
  if (!bi.keep_res) Vec<T>::store4(res + 4 * (size_t)c, r);
  if (SAVE) Vec<T>::store4(qold_out + 4 * (size_t)c, qn);
  if (ADT) adt_out[c] = adt_of(in.xs[0], in.xs[1], in.xs[2], in.xs[3], qn, k);
}

// One pass of update (with BRES, SAVE and ADT as above).  Workgroups
// stride over the cells.  XCD: each XCD takes a contiguous range of the
// cells, as the edge and cell loops do, so the next loops find a cell's new
// state in the L2 that wrote it; used when one pass of the grid covers the
// set (measured: 180k cells, +27% cell-updates/s) and for float state
// (2.88M cells, FP32 form +2.5%).  Large FP64 sets keep the plain
// interleaved order, which streams faster there (2.88M cells: 1.6-4%).
template <int NT, typename T, bool ADT, bool SAVE, bool XCD, bool BRES>
__global__ __launch_bounds__(NT) void k_update(const T *qold, T *q, T *__restrict__ res,
                                               const T *adt, int ncell,
                                               double *__restrict__ partials,
                                               const int4 *__restrict__ pcell,
                                               const T *__restrict__ x, T *adt_out,
                                               T *qold_out, BresIdx<T> bi, AirfoilConsts<T> k) {
  __shared__ double lds[NT / kWave];
  double acc = 0.0;
  if (XCD) {
    // each XCD strides over its own contiguous range of cell chunks (the
    // ranges the edge and cell loops give the XCDs); with one pass of the
    // grid this is xcd_block's mapping
    const int xd = blockIdx.x % kXcds, i = blockIdx.x / kXcds;
    const int nwg = gridDim.x / kXcds + (xd < (int)gridDim.x % kXcds ? 1 : 0);
    const int nch = (ncell + NT - 1) / NT, base = nch / kXcds, rem = nch % kXcds;
    const int lo = xd * base + min(xd, rem), hi = lo + base + (xd < rem ? 1 : 0);
    for (int ch = lo + i; ch < hi; ch += nwg) {
      const int c = ch * NT + threadIdx.x;
      if (c >= ncell) continue;
      UpdIn<T, ADT> in;
      upd_load<T, ADT, BRES>(c, qold, res, adt, pcell, x, bi, in);
      upd_store<T, ADT, SAVE, BRES>(c, in, q, res, adt_out, qold_out, bi, k, acc);
    }
  } else {
    for (int c = blockIdx.x * NT + threadIdx.x; c < ncell; c += gridDim.x * NT) {
      UpdIn<T, ADT> in;
      upd_load<T, ADT, BRES>(c, qold, res, adt, pcell, x, bi, in);
      upd_store<T, ADT, SAVE, BRES>(c, in, q, res, adt_out, qold_out, bi, k, acc);
    }
  }
  acc = block_reduce<kRedSum, NT>(acc, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// Batched update: as k_update (ADT, SAVE and BRES included), with the squared moves summed per mesh
// (rms[cmap[c]] += ...).  cmap is non-decreasing, so each mesh is a run of
// cells; workgroup b covers cells [b*CH, (b+1)*CH) and, for every mesh run
// its range meets, writes one partial (block tree over its threads, in a
// fixed order) at slot pslot[m] + b - (first workgroup of mesh m).
// k_fold_seg adds a mesh's partials in workgroup order.
template <int NT, typename T, int UPC, bool ADT, bool SAVE, bool BRES>
__global__ __launch_bounds__(NT) void k_update_seg(const T *qold, T *q,
                                                   T *__restrict__ res, const T *adt,
                                                   int ncell, const int *__restrict__ cmap,
                                                   const int *__restrict__ seg_off,
                                                   const int *__restrict__ pslot,
                                                   double *__restrict__ partials,
                                                   unsigned *__restrict__ log_count, int nseg,
                                                   unsigned *__restrict__ log_base,
                                                   const int4 *__restrict__ pcell,
                                                   const T *__restrict__ x, T *adt_out,
                                                   T *qold_out, BresIdx<T> bi,
                                                   AirfoilConsts<T> k) {
  __shared__ double lds[NT / kWave];
  // reserve the log slots of this call's nseg results for k_fold_seg
  if (log_count && blockIdx.x == 0 && threadIdx.x == 0) *log_base = atomicAdd(log_count, (unsigned)nseg);
  constexpr int CH = NT * UPC;
  const int b = blockIdx.x;
  const int c_lo = b * CH;
  const int c_hi = min(c_lo + CH, ncell);
  double s[UPC];
#pragma unroll
  for (int j = 0; j < UPC; j++) {
    const int c = c_lo + threadIdx.x + j * NT;
    s[j] = 0.0;
    if (c >= c_hi) continue;
    UpdIn<T, ADT> in;
    upd_load<T, ADT, BRES>(c, qold, res, adt, pcell, x, bi, in);
    upd_store<T, ADT, SAVE, BRES>(c, in, q, res, adt_out, qold_out, bi, k, s[j]);
  }
  const int m_lo = cmap[c_lo], m_hi = cmap[c_hi - 1];
  for (int m = m_lo; m <= m_hi; m++) {
    const int a = max(c_lo, seg_off[m]), z = min(c_hi, seg_off[m + 1]);
    if (a >= z) continue;  // a mesh without cells (uniform across the workgroup)
    double v = 0.0;
#pragma unroll
    for (int j = 0; j < UPC; j++) {
      const int c = c_lo + threadIdx.x + j * NT;
      if (c >= a && c < z) v += s[j];
    }
    v = block_reduce<kRedSum, NT>(v, lds);
    if (threadIdx.x == 0) partials[pslot[m] + b - seg_off[m] / CH] = v;
    __syncthreads();
  }
}

// One Runge-Kutta stage of a captured graph as one kernel: res_calc,
// bres_calc, update, and optionally the next iteration's save_soln and the
// adt_calc that follows, on owner-gather tiles that own every cell.  A tile
// holds the complete edge sum of each cell it owns, so it finishes the cell
// in the same workgroup and res never reaches memory.  Other tiles still
// read the cell's old q and adt as halo, so the new q and adt go to other
// buffers (q_out, adt_out; the runtime alternates two buffers per dat).
// Per owned cell, in the order the separate loops use:
//   r = res (or +0 when res is known to hold zeros) + the tile's edge fluxes
//       in the plan's fixed order (res_calc)
//   r += each boundary flux, in edge order (bres_calc)
//   del = r / adt in double, q_out = qold - del rounded once, rms += del^2
//       (update); res = 0 unless the next stage takes it as zero
//   SAVE: qold_out = q_out (save_soln of the next iteration)
//   ADT: adt_out = adt of q_out (adt_calc), else adt_out = adt
// so q, adt and res hold the bits of the separate loops.  The rms sum of the
// tile's cells is one partial per tile (folded by the next kernel); in the
// batched form every tile lies in one mesh and the tiles of mesh m are
// folded in order.
// adt_calc comes in two forms.  kAdtMap gathers each cell's nodes through
// adt_calc's pcell.  kAdtSides takes the cell's four sides from the edges
// the tile has just evaluated: every side of an owned cell is an edge of
// the tile (one of the cell's increment records) or one of its boundary
// edges, and a per-cell recipe (3 bits per side, in pcell's side order:
// 0-3 the increment record, 4-5 the boundary record) says which.  Each
// thread keeps the side vectors (x1 - x2) of the edges it evaluates in
// registers and writes them to LDS once the flux sums are read (no more LDS
// than the fluxes), so adt needs neither pcell nor a second gather of the
// nodes; the tile's positions must fit in kSideTrips passes of the block.
constexpr int kAdtNone = 0, kAdtMap = 1, kAdtSides = 2;
constexpr int kSideTrips = 3;
template <typename T>
struct StageArgs {
  const int *blk_off;
  int nblk;
  const int4 *ids;
  const int *owned_off, *owned;
  const ushort4 *ell;
  const T *x, *q_in, *adt_in, *qold;
  T *res, *q_out, *adt_out, *qold_out;
  const int4 *pcell;  // kAdtMap: the cell -> node map of adt_calc
  const T *xc;        // kAdtMap: the node coordinates adt_calc reads
  const unsigned short *sides;  // kAdtSides: per owned cell (owned order), its side recipe
  double *partials;   // one per tile
  unsigned *log_count, *log_base;  // batched form: log slots of this stage's nseg results
  int nseg;
};

// One edge of a stage tile (its 32-bit position ids).
template <typename T>
__device__ __forceinline__ void stage_edge(const StageArgs<T> &s, const AirfoilConsts<T> &k,
                                           T *fs, int np, int p0, int p, T side[2]) {
  tile_edge<false>(s.ids, nullptr, int4{}, s.x, s.q_in, s.adt_in, k, fs, np, p0, p, side);
}

// With one owned cell per thread the register budget is that of 6 waves
// per SIMD (80 VGPRs): measured faster than the compiler's own choice
// (about 100 VGPRs, 4 waves) by 7% at 2.88M cells, and than 8 waves (which
// spills) by 13%.  Two cells per thread keep 4 waves (no spills).
template <int NT, typename T, int OWN, bool OWNID, bool BRES, bool SAVE, int ADT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(OWN == 1 ? 6 : 4, 8))) void
k_stage(
    StageArgs<T> s, BresIdx<T> bi, AirfoilConsts<T> k, CarriedFold fold, bool zero_res,
    bool store_res) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[NT / kWave];
  if (carries_fold(fold)) carried_fold<NT>(fold, red);
  if (s.log_count && blockIdx.x == 0 && threadIdx.x == 0)
    *s.log_base = atomicAdd(s.log_count, (unsigned)s.nseg);
  const int b = xcd_block(blockIdx.x, s.nblk);
  const int p0 = s.blk_off[b], np = s.blk_off[b + 1] - p0;
  const int o0 = s.owned_off[b], no = s.owned_off[b + 1] - o0;
  T *fs = reinterpret_cast<T *>(smem);  // [4][np]
  const int tid = threadIdx.x;
  // the owned cells' increment records (and side recipes), issued before
  // the flux phase so that their latency overlaps it
  bool own[OWN];
  int c[OWN];
  ushort4 e4[OWN];
  unsigned recipe[OWN];
#pragma unroll
  for (int j = 0; j < OWN; j++) {
    const int i = tid + j * NT;
    own[j] = i < no;
    c[j] = own[j] ? (OWNID ? o0 + i : s.owned[o0 + i]) : 0;
    e4[j] = make_ushort4(0xffff, 0xffff, 0xffff, 0xffff);
    recipe[j] = 0;
    if (own[j]) {
      const op2_v4us v = __builtin_nontemporal_load((const op2_v4us *)(s.ell + o0 + i));
      e4[j] = make_ushort4(v.x, v.y, v.z, v.w);
      if (ADT == kAdtSides) recipe[j] = s.sides[o0 + i];
    }
  }
  T side[kSideTrips][2];
  if (ADT == kAdtSides) {
#pragma unroll
    for (int t = 0; t < kSideTrips; t++) {
      const int p = tid + t * NT;
      if (p < np) stage_edge(s, k, fs, np, p0, p, side[t]);
    }
  } else {
    for (int p = tid; p < np; p += NT) {
      T unused[2];
      stage_edge(s, k, fs, np, p0, p, unused);
    }
  }
  // the owned cells' update inputs (after the flux phase: fewer registers
  // held through it) and, for kAdtMap, the nodes of each cell, gathered
  // while the other waves finish their fluxes
  T qo[OWN][4], a[OWN], xs[OWN][4][2];
  int bidx[OWN];
#pragma unroll
  for (int j = 0; j < OWN; j++) {
    a[j] = T(1);
    bidx[j] = -1;
    if (!own[j]) continue;
    Vec<T>::load4(s.qold + 4 * (size_t)c[j], qo[j]);
    a[j] = s.adt_in[c[j]];
    if (BRES) {
      const uint2 w = bi.word[c[j] >> 5];
      const unsigned bit = 1u << (c[j] & 31);
      bidx[j] = (w.x & bit) ? (int)(w.y + __popc(w.x & (bit - 1u))) : -1;
    }
    if (ADT == kAdtMap) cell_nodes(s.pcell, s.xc, c[j], xs[j]);
  }
  __syncthreads();
  double acc = 0.0;
  T qk[OWN][4];  // kAdtSides: the new q, for adt after the side exchange
#pragma unroll
  for (int j = 0; j < OWN; j++) {
    if (!own[j]) continue;
    const int cj = c[j];
    T sum[4] = {T(0), T(0), T(0), T(0)};
    ell_sum(fs, np, e4[j], sum);
    // res (+0 when it is known zero, the first stage of a graph excepted:
    // read here, not held in registers through the flux phase)
    T r[4] = {T(0), T(0), T(0), T(0)};
    if (!zero_res) Vec<T>::load4(s.res + 4 * (size_t)cj, r);
#pragma unroll
    for (int d = 0; d < 4; d++) r[d] = r[d] + sum[d];
    if (BRES && bidx[j] >= 0) {
      T qc[4];
      Vec<T>::load4(s.q_in + 4 * (size_t)cj, qc);
      const int4 rec[2] = {bi.rec[2 * bidx[j]], bi.rec[2 * bidx[j] + 1]};
#pragma unroll
      for (int m = 0; m < 2; m++) {
        if (rec[m].z < 0) break;
        T x1[2], x2[2], f[4];
        Vec<T>::load2(bi.x + 2 * (size_t)rec[m].x, x1);
        Vec<T>::load2(bi.x + 2 * (size_t)rec[m].y, x2);
        const T *qi = bi.qinf_dat ? bi.qinf_dat + 4 * (size_t)rec[m].w : k.qinf;
        boundary_flux(x1, x2, qc, a[j], bi.bound[rec[m].z], qi, k, f);
#pragma unroll
        for (int d = 0; d < 4; d++) r[d] += f[d];
      }
    }
    const double adti = 1.0 / (double)a[j];
    T qn[4];
#pragma unroll
    for (int d = 0; d < 4; d++) {
      const double del = adti * (double)r[d];
      qn[d] = (T)((double)qo[j][d] - del);
      acc += del * del;
    }
    Vec<T>::store4(s.q_out + 4 * (size_t)cj, qn);
    if (SAVE) Vec<T>::store4(s.qold_out + 4 * (size_t)cj, qn);
    if (ADT == kAdtMap) s.adt_out[cj] = adt_of(xs[j][0], xs[j][1], xs[j][2], xs[j][3], qn, k);
    if (ADT == kAdtNone) s.adt_out[cj] = a[j];
#pragma unroll
    for (int d = 0; d < 4; d++) qk[j][d] = qn[d];
    if (store_res) {
      const T z[4] = {T(0), T(0), T(0), T(0)};
      Vec<T>::store4(s.res + 4 * (size_t)cj, z);
    }
  }
  if (ADT == kAdtSides) {
    // every flux sum is read: the flux rows take the side vectors
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kSideTrips; t++) {
      const int p = tid + t * NT;
      if (p < np) {
        fs[p] = side[t][0];
        fs[np + p] = side[t][1];
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < OWN; j++) {
      if (!own[j]) continue;
      // every side read from LDS without branches (a boundary side reads
      // position 0 and is replaced below: only boundary cells branch)
      T dxs[4], dys[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const unsigned src = (recipe[j] >> (3 * i)) & 7u;
        // (selects, not an indexed array: that would live in scratch)
        const unsigned e = (src & 3u) == 0   ? e4[j].x
                           : (src & 3u) == 1 ? e4[j].y
                           : (src & 3u) == 2 ? e4[j].z
                                             : e4[j].w;
        const int pos = src < 4 ? (int)(e >> 2) : 0;
        dxs[i] = fs[pos];
        dys[i] = fs[np + pos];
      }
      if (BRES && bidx[j] >= 0) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const unsigned src = (recipe[j] >> (3 * i)) & 7u;
          if (src < 4) continue;
          // a boundary edge of the cell (nodes from its record)
          const int4 rec = bi.rec[2 * bidx[j] + (int)(src - 4)];
          T x1[2], x2[2];
          Vec<T>::load2(bi.x + 2 * (size_t)rec.x, x1);
          Vec<T>::load2(bi.x + 2 * (size_t)rec.y, x2);
          dxs[i] = x1[0] - x2[0];
          dys[i] = x1[1] - x2[1];
        }
      }
      s.adt_out[c[j]] = adt_of_sides(dxs, dys, qk[j], k);
    }
  }
  acc = block_reduce<kRedSum, NT>(acc, red);
  if (tid == 0) s.partials[b] = acc;
}

// One workgroup per mesh m: out = sum of its partials (thread-strided, then
// the block tree: a fixed order); the dat gets dat[m] += out and, with
// deferred reductions, the log slot reserved by k_update_seg gets out (the
// results of one call in mesh order).
__global__ __launch_bounds__(kFixedBlock) void k_fold_seg(const double *__restrict__ partials,
                                                 const int *__restrict__ pslot,
                                                 double *__restrict__ dat,
                                                 double *__restrict__ log,
                                                 const unsigned *__restrict__ log_base) {
  constexpr int NT = kFixedBlock;
  __shared__ double lds[NT / kWave];
  const int m = blockIdx.x;
  double acc = strided_sum(partials, pslot[m] + threadIdx.x, pslot[m + 1], NT);
  acc = block_reduce<kRedSum, NT>(acc, lds);
  if (threadIdx.x == 0) {
    dat[m] += acc;
    if (log) log[(*log_base + m) & (kLogCap - 1)] = acc;
  }
}

// Folds the per-workgroup partials of NSLOT reductions in workgroup order
// (deterministic for a given grid): slot s sums, or takes the maximum if bit
// s of maxmask is set, partials[s*nparts .. s*nparts+nparts) into out[s].
// One workgroup; the kernel boundary after the producer makes the partials
// visible without device-scope fences.
template <int NSLOT>
__global__ __launch_bounds__(kFixedBlock) void k_fold(const double *__restrict__ partials, int nparts,
                                             unsigned maxmask, double *__restrict__ out,
                                             double *__restrict__ log,
                                             unsigned *__restrict__ log_count) {
  constexpr int NT = kFixedBlock;
  __shared__ double lds[NT / kWave];
#pragma unroll
  for (int s = 0; s < NSLOT; s++) {
    const bool mx = (maxmask >> s) & 1u;
    double acc;
    if (mx) {
      acc = red_identity<kRedMax>();
      for (int i = threadIdx.x; i < nparts; i += NT)
        acc = red_apply<kRedMax>(acc, partials[s * nparts + i]);
    } else {
      acc = strided_sum(partials + (size_t)s * nparts, threadIdx.x, nparts, NT);
    }
    acc = mx ? block_reduce<kRedMax, NT>(acc, lds) : block_reduce<kRedSum, NT>(acc, lds);
    if (threadIdx.x == 0) {
      out[s] = acc;
      if (log) {
        const unsigned i = atomicAdd(log_count, 1u);
        log[i & (kLogCap - 1)] = acc;
      }
    }
    __syncthreads();
  }
}

// Staged Jacobi edge loop (jac1 res): du[pp0] += beta * A * u[pp1]; one map,
// slots (pp, 1) for u and (pp, 0) for du.
template <int NT, typename T, bool COLOURED>
__global__ __launch_bounds__(NT) void k_jac_res(
    const int *__restrict__ blk_off, const int *__restrict__ blkmap, int blk_first, int nblk,
    const int *__restrict__ perm, const int *__restrict__ node_off,
    const int *__restrict__ node_list, const ushort2 *__restrict__ loc,
    const int *__restrict__ col_off, const int *__restrict__ col_base,
    const int *__restrict__ blk_ncol, const T *__restrict__ A, const T *__restrict__ u,
    T *__restrict__ du, T beta) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int b = blkmap[blk_first + xcd_block(blockIdx.x, nblk)];
  const int e0 = blk_off[b], ne = blk_off[b + 1] - e0;
  const int n0 = node_off[b], nn = node_off[b + 1] - n0;
  T *us = reinterpret_cast<T *>(smem);
  T *ds = us + nn;
  // nodes this block increments: the staged list also holds nodes it only
  // reads (pp[2e+1]), which a block of the same colour may be incrementing,
  // so only flagged nodes are written back
  unsigned char *inc = reinterpret_cast<unsigned char *>(ds + nn);
  const int tid = threadIdx.x;
  for (int i = tid; i < nn; i += NT) {
    us[i] = u[node_list[n0 + i]];
    ds[i] = T(0);
    inc[i] = 0;
  }
  __syncthreads();
  const int cb = COLOURED ? col_base[b] : 0;
  const int ncol = COLOURED ? blk_ncol[b] : 1;
  for (int c = 0; c < ncol; c++) {
    const int cs = COLOURED ? col_off[cb + c] - e0 : 0;
    const int ce = COLOURED ? col_off[cb + c + 1] - e0 : ne;
    for (int e = cs + tid; e < ce; e += NT) {
      const ushort2 L = loc[e0 + e];
      const int el = perm ? perm[e0 + e] : e0 + e;
      const T v = beta * A[el] * us[L.x];
      inc[L.y] = 1;
      if (COLOURED)
        ds[L.y] += v;
      else
        atomicAdd(ds + L.y, v);
    }
    __syncthreads();
  }
  for (int i = tid; i < nn; i += NT) {
    if (!inc[i]) continue;
    const int g = node_list[n0 + i];
    if (COLOURED)
      du[g] += ds[i];
    else
      atomicAdd(du + g, ds[i]);
  }
}

// Owner-compute Jacobi edge loop (jac1 res): nodes cut into runs, each
// workgroup sums the contributions to its nodes in a fixed order.
template <int NT, typename T>
__global__ __launch_bounds__(NT) void k_jac_res_owner(
    const int *__restrict__ blk_off, int nblk, const int *__restrict__ perm,
    const int *__restrict__ node_off, const int *__restrict__ node_list,
    const ushort2 *__restrict__ loc, const int *__restrict__ owned_off,
    const int *__restrict__ owned, const int *__restrict__ csr_off,
    const uint16_t *__restrict__ csr, const T *__restrict__ A, const T *__restrict__ u,
    T *__restrict__ du, T beta) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int b = xcd_block(blockIdx.x, nblk);
  const int p0 = blk_off[b], np = blk_off[b + 1] - p0;
  const int n0 = node_off[b], nn = node_off[b + 1] - n0;
  const int o0 = owned_off[b], no = owned_off[b + 1] - o0;
  T *us = reinterpret_cast<T *>(smem);
  T *fs = us + nn;
  const int tid = threadIdx.x;
  for (int i = tid; i < nn; i += NT) us[i] = u[node_list[n0 + i]];
  __syncthreads();
  for (int p = tid; p < np; p += NT) fs[p] = beta * A[perm[p0 + p]] * us[loc[p0 + p].x];
  __syncthreads();
  for (int i = tid; i < no; i += NT) {
    T acc = T(0);
    const int j1 = csr_off[o0 + i + 1];
    for (int j = csr_off[o0 + i]; j < j1; j++) acc += fs[csr[j] >> 2];
    du[owned[o0 + i]] += acc;
  }
}

// jac1 update: u += du + alpha r, du = 0, u_sum += u^2, u_max = max(u_max, u)
// (partials: u_sum of each workgroup, then u_max of each workgroup)
template <int NT, typename T>
__global__ __launch_bounds__(NT) void k_jac_update(const T *__restrict__ r, T *__restrict__ du,
                                                   T *__restrict__ u, int n, T alpha,
                                                   double *__restrict__ partials) {
  __shared__ double lds[NT / kWave];
  double s = 0.0, m = red_identity<kRedMax>();
  for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
    T ui = u[i];
    ui += du[i] + alpha * r[i];
    u[i] = ui;
    du[i] = T(0);
    s += (double)ui * (double)ui;
    m = (m > (double)ui) ? m : (double)ui;
  }
  s = block_reduce<kRedSum, NT>(s, lds);
  __syncthreads();
  m = block_reduce<kRedMax, NT>(m, lds);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = s;
    partials[gridDim.x + blockIdx.x] = m;
  }
}

// ============================================================================
// host side
// ============================================================================

namespace {

bool is_type(const op_arg &a, const char *t) { return a.type && !strcmp(a.type, t); }

bool direct(const op_arg &a, int dim, op_access acc, op_set set) {
  return a.argtype == OP_ARG_DAT && !a.map && a.idx == -1 && a.dim == dim && a.acc == acc &&
         a.dat && a.dat->set == set && a.dat->dim == dim;
}

bool via(const op_arg &a, op_map m, int idx, int dim, op_access acc) {
  return a.argtype == OP_ARG_DAT && a.map == m && m && a.idx == idx && a.dim == dim &&
         a.acc == acc && a.dat && a.dat->dim == dim && a.dat->set == m->to;
}

bool gbl(const op_arg &a, int dim, op_access acc) {
  return a.argtype == OP_ARG_GBL && a.dim == dim && a.acc == acc && a.data;
}

// element type of the dat arguments: 8 (double), 4 (float) or 0 (mixed /
// other)
int real_bytes(int nargs, const op_arg *args) {
  int rb = -1;
  for (int i = 0; i < nargs; i++) {
    if (args[i].argtype != OP_ARG_DAT) continue;
    if (!strcmp(args[i].type, "int")) continue;
    int t = is_type(args[i], "double") ? 8 : is_type(args[i], "float") ? 4 : 0;
    if (rb >= 0 && t != rb) return 0;
    rb = t;
  }
  return rb < 0 ? 0 : rb;
}

// true if an argument reaches its dat through a map
bool through_map(int nargs, const op_arg *args) {
  for (int i = 0; i < nargs; i++)
    if (args[i].argtype == OP_ARG_DAT && args[i].map) return true;
  return false;
}

// Algorithmic bytes of one call: each touched dat once per direction (read
// for READ/RW/INC, write for WRITE/RW/INC), each map array once; an
// indirect dat counts min(target set size, slots x loop set size) elements.
double loop_bytes(op_set set, int nargs, const op_arg *args) {
  double bytes = 0;
  std::vector<op_dat> dats;
  std::vector<op_map> maps;
  for (int i = 0; i < nargs; i++) {
    const op_arg &a = args[i];
    if (a.argtype != OP_ARG_DAT) continue;
    if (a.map && std::find(maps.begin(), maps.end(), a.map) == maps.end()) {
      maps.push_back(a.map);
      bytes += 4.0 * a.map->dim * a.map->from->size;
    }
    if (std::find(dats.begin(), dats.end(), a.dat) != dats.end()) continue;
    dats.push_back(a.dat);
    bool rd = false, wr = false;
    std::vector<std::pair<op_map, int>> slots;
    for (int j = 0; j < nargs; j++) {
      const op_arg &b = args[j];
      if (b.argtype != OP_ARG_DAT || b.dat != a.dat) continue;
      rd |= b.acc == OP_READ || b.acc == OP_RW || b.acc == OP_INC;
      wr |= b.acc == OP_WRITE || b.acc == OP_RW || b.acc == OP_INC;
      std::pair<op_map, int> s{b.map, b.idx};
      if (std::find(slots.begin(), slots.end(), s) == slots.end()) slots.push_back(s);
    }
    double elems = a.map ? std::min<double>(a.dat->set->size, (double)slots.size() * set->size)
                         : set->size;
    bytes += elems * a.dat->size * ((rd ? 1 : 0) + (wr ? 1 : 0));
  }
  return bytes;
}

template <typename T>
AirfoilConsts<T> airfoil_consts() {
  AirfoilConsts<T> k;
  k.gam = (T)get_const("gam").as_double(0);
  k.gm1 = (T)get_const("gm1").as_double(0);
  k.cfl = (T)get_const("cfl").as_double(0);
  k.eps = (T)get_const("eps").as_double(0);
  const Const &qi = get_const("qinf");
  for (int i = 0; i < 4; i++) k.qinf[i] = (T)qi.as_double(i);
  return k;
}

// Threads per workgroup of a loop's kernels: the loop's configured block
// size, else OP_block_size (OP_BLOCK_SIZE at op_init; 256 by default).
int block_of(const std::string &loop) {
  const int b = config_for(loop).block_size;
  return b ? b : OP_block_size;
}

// Calls f(std::integral_constant<int, B>()) with B = bs, the block size the
// kernels are instantiated for (128, 256 or 512).
template <class F>
auto with_block(int bs, F &&f) {
  switch (bs) {
    case 128: return f(std::integral_constant<int, 128>());
    case 512: return f(std::integral_constant<int, 512>());
    case 256: return f(std::integral_constant<int, 256>());
  }
  fail("op2: block size %d is not supported (128, 256 or 512 threads)", bs);
}

int grid_for(long long n, int cap, int bs = kFixedBlock) {
  long long g = (n + bs - 1) / bs;
  if (g < 1) g = 1;
  return (int)std::min<long long>(g, cap);
}

void check_launch(const char *loop) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) fail("op2: kernel launch for loop %s failed: %s", loop, hipGetErrorString(e));
}

// Default strategy when none is configured: owner-compute tiles, gathering
// from global memory where the loop has that kernel (res_calc), staged
// otherwise: one launch, no atomics, deterministic, and the fastest
// strategy at every mesh size measured, in both precisions.
int strategy_of(const std::string &loop) {
  int s = config_for(loop).strategy;
  if (s) return s;
  return OP2_STRATEGY_OWNER_GATHER;
}

int part_of(const std::string &loop, int dflt) {
  int p = config_for(loop).part_size;
  if (!p) p = OP_part_size;
  return p ? p : dflt;
}

// Tiles of the fused stage kernels (op2_set_stage_config).  By default 256
// cells on 256 threads whatever res_calc's own configuration: a stage tile
// runs flux, update and adt_calc phases, and more, smaller workgroups per CU
// overlap those phases better (2.88M cells: 256/256 beat res_calc's tuned
// 512/512 by 5% in FP64 and 4% in the FP32 form; 128-cell tiles exceed the
// side form of adt_calc).
int stage_part_of(const std::string &loop) {
  const Runtime &r = rt();
  return r.stage_part < 0 ? part_of(loop, 256) : r.stage_part ? r.stage_part : 256;
}
int stage_block_of(const std::string &loop) {
  const Runtime &r = rt();
  return r.stage_part < 0 ? block_of(loop) : r.stage_block ? r.stage_block : 256;
}

Plan *plan_for(const char *name, op_set set, int nargs, op_arg *args, int strategy, int part) {
  PlanRequest req{name, set, nargs, args, strategy, part};
  Plan *p = get_plan(req);
  p->upload();
  return p;
}

// reads `n` reduction results back to the host (waits for the stream)
const double *fetch_results(const double *d_out, int n) {
  Runtime &r = rt();
  OP2_HIP_CHECK(hipMemcpyAsync(r.h_result, d_out, n * sizeof(double), hipMemcpyDeviceToHost,
                               r.stream));
  OP2_HIP_CHECK(hipStreamSynchronize(r.stream));
  return r.h_result;
}

void gbl_inc(const op_arg &a, double v) {
  if (is_type(a, "double"))
    *(double *)a.data += v;
  else if (is_type(a, "float"))
    *(float *)a.data = (float)((double)*(float *)a.data + v);
  else
    fail("op2: OP_INC global of type %s is not supported", a.type);
}

void gbl_max(const op_arg &a, double v) {
  if (is_type(a, "double")) {
    double *p = (double *)a.data;
    *p = (*p > v) ? *p : v;
  } else if (is_type(a, "float")) {
    float *p = (float *)a.data;
    float fv = (float)v;
    *p = (*p > fv) ? *p : fv;
  } else {
    fail("op2: OP_MAX global of type %s is not supported", a.type);
  }
}

double gbl_read(const op_arg &a) {
  if (is_type(a, "double")) return *(const double *)a.data;
  if (is_type(a, "float")) return *(const float *)a.data;
  fail("op2: OP_READ global of type %s is not supported", a.type);
}

// ---- per-loop launchers ------------------------------------------------------

template <int NT, typename T>
void run_save_soln(op_set set, op_arg *a) {
  const long long chunks = (long long)set->size * 4 * sizeof(T) / 16;
  hipLaunchKernelGGL((k_copy16<NT>), dim3(grid_for(chunks, 1 << 20, NT)), dim3(NT), 0, rt().stream,
                     (const uint4 *)a[0].dat->data_d, (uint4 *)a[1].dat->data_d, chunks);
}

// save_soln (a) followed by adt_calc (adt_a) on the q it copies
template <int NT, typename T>
void run_save_adt(op_set set, const op_arg *a, const op_arg *adt_a) {
  hipLaunchKernelGGL((k_save_adt<NT, T>), dim3(grid_for(set->size, 1 << 30, NT)), dim3(NT), 0,
                     rt().stream, (const int4 *)adt_a[0].map->map_d,
                     (const T *)adt_a[0].dat->data_d, (const T *)a[0].dat->data_d,
                     (T *)a[1].dat->data_d, (T *)adt_a[5].dat->data_d, set->size,
                     airfoil_consts<T>());
}

template <int NT, typename T>
void run_adt_calc(op_set set, op_arg *a) {
  hipLaunchKernelGGL((k_adt_calc<NT, T>), dim3(grid_for(set->size, 1 << 30, NT)), dim3(NT), 0,
                     rt().stream, (const int4 *)a[0].map->map_d, (const T *)a[0].dat->data_d,
                     (const T *)a[4].dat->data_d, (T *)a[5].dat->data_d, set->size,
                     airfoil_consts<T>());
}

// Dats known to hold zeros (an update launched earlier zeroed its res, and
// no loop has written it since), so that the next owner-compute res_calc
// need not read them: in eager calls, and in a graph capture with fusion on.
// Empty at the start of every capture (a graph may be replayed after any
// eager work) and after anything that writes dats outside op_par_loop.
std::set<op_dat> &zero_dats() {
  static std::set<op_dat> z;
  return z;
}

// dats whose device pointer the program holds (op2_dat_device_ptr): other
// code may write them at any time
std::set<op_dat> &exposed_dats() {
  static std::set<op_dat> z;
  return z;
}

bool known_zero(op_dat d) {
  const Runtime &r = rt();
  if (r.capturing && !r.graph_fusion) return false;
  return zero_dats().count(d) && !exposed_dats().count(d);
}

// A launched loop wrote the dats of its non-READ arguments.
void note_writes(int nargs, const op_arg *a) {
  for (int m = 0; m < nargs; m++)
    if (a[m].argtype == OP_ARG_DAT && a[m].acc != OP_READ) zero_dats().erase(a[m].dat);
}

// The fold of the last update launched in a graph capture with fusion on,
// not yet launched: the next res_calc kernel carries it (take_fold); any
// other launch first launches it on its own (flush_fold), which also keeps
// the next reduction from overwriting the partials before they are summed.
struct PendingFold {
  bool active = false;
  CarriedFold f{};
};

PendingFold &pending_fold() {
  static PendingFold p;
  return p;
}

void flush_fold() {
  PendingFold &p = pending_fold();
  if (!p.active) return;
  p.active = false;
  if (p.f.nseg > 0)
    hipLaunchKernelGGL(k_fold_seg, dim3(p.f.nseg), dim3(kFixedBlock), 0, rt().stream, p.f.partials,
                       p.f.pslot, p.f.dat, p.f.log, p.f.log_base);
  else
    hipLaunchKernelGGL((k_fold<1>), dim3(1), dim3(kFixedBlock), 0, rt().stream, p.f.partials, p.f.n, 0u,
                       p.f.out, p.f.log, p.f.log_count);
  check_launch("update");
}

// The pending fold for a kernel of nblocks workgroups (which must cover
// every workgroup the fold needs), else none (launched on its own first).
CarriedFold take_fold(int nblocks) {
  PendingFold &p = pending_fold();
  if (p.active && nblocks < std::max(p.f.nseg, 1)) flush_fold();
  if (!p.active) return CarriedFold{};
  p.active = false;
  return p.f;
}

template <int NT, typename T, int EPT>
void launch_res_calc(Plan *p, op_arg *a, bool coloured) {
  const PlanGroup &nodes = p->groups[p->slots[0].first];
  const PlanGroup &cells = p->groups[p->slots[2].first];
  // (at least the carried fold's NT / 64 doubles)
  const size_t lds = std::max(
      sizeof(T) * (2 * ((nodes.max_count + 1) & ~1) + 5 * (size_t)cells.max_count),
      sizeof(double) * (NT / kWave));
  auto kern = coloured ? k_res_calc<NT, T, EPT, true> : k_res_calc<NT, T, EPT, false>;
  if (lds > 64 * 1024)
    OP2_HIP_CHECK(hipFuncSetAttribute((const void *)kern,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  AirfoilConsts<T> k = airfoil_consts<T>();
  bool first_launch = true;
  for (int c = 0; c < p->ncolors; c++) {
    const int first = p->colblk_off[c], nb = p->colblk_off[c + 1] - first;
    if (nb == 0) continue;
    const CarriedFold fold = first_launch ? take_fold(nb) : CarriedFold{};
    first_launch = false;
    hipLaunchKernelGGL(kern, dim3(nb), dim3(NT), lds, rt().stream, p->d_blk_off, p->d_blkmap,
                       first, nb, nodes.d_off, nodes.d_list, cells.d_off, cells.d_list,
                       (const ushort4 *)p->d_loc, p->d_col_off, p->d_col_base, p->d_blk_ncol,
                       (const T *)a[0].dat->data_d, (const T *)a[2].dat->data_d,
                       (const T *)a[4].dat->data_d, (T *)a[6].dat->data_d, k, fold);
  }
  flush_fold();  // a plan without blocks carries nothing
}

// LDS bytes of an owner-compute block: staged groups, padded, plus one
// increment record per element position
template <typename T>
size_t owner_lds(const Plan *p, int node_words, int cell_words, int inc_words) {
  constexpr int pad = 16 / sizeof(T);
  size_t w = 0;
  if (node_words) w += (size_t)node_words * ((p->groups[p->slots[0].first].max_count + 1) & ~1);
  if (cell_words) {
    const int mc = p->groups[p->owner_group].max_count;
    w += (size_t)cell_words * mc + ((mc + pad - 1) / pad) * pad;
  }
  w += (size_t)inc_words * p->max_positions;
  return w * sizeof(T);
}

template <typename K>
void allow_lds(K kern, size_t lds) {
  if (lds > 160 * 1024) fail("op2: block needs %zu bytes of LDS; reduce the part size", lds);
  if (lds > 64 * 1024)
    OP2_HIP_CHECK(hipFuncSetAttribute((const void *)kern,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
}

template <int NT, typename T>
void run_res_calc(const char *name, op_set set, int nargs, op_arg *a) {
  const int strategy = strategy_of(name);
  if (strategy == OP2_STRATEGY_OWNER) {
    Plan *p = plan_for(name, set, nargs, a, strategy, part_of(name, 256));
    const PlanGroup &nodes = p->groups[p->slots[0].first];
    const PlanGroup &cells = p->groups[p->owner_group];
    const int own = (p->max_owned + NT - 1) / NT;
    if (own > 4) fail("op2: owner tile of %s exceeds %d targets", name, 4 * NT);
    const size_t lds = owner_lds<T>(p, 2, 4, 4);
    const bool ell = p->d_ell != nullptr;
    auto kern = ell ? (own <= 1 ? k_res_calc_owner<NT, T, 1, true>
                                : own <= 2 ? k_res_calc_owner<NT, T, 2, true>
                                           : k_res_calc_owner<NT, T, 4, true>)
                    : (own <= 1 ? k_res_calc_owner<NT, T, 1, false>
                                : own <= 2 ? k_res_calc_owner<NT, T, 2, false>
                                           : k_res_calc_owner<NT, T, 4, false>);
    allow_lds(kern, std::max(lds, (size_t)64));
    if (p->nblocks == 0) {
      flush_fold();
      return;
    }
    hipLaunchKernelGGL(kern, dim3(p->nblocks), dim3(NT), std::max(lds, (size_t)64),
                       rt().stream, p->d_blk_off, p->nblocks, nodes.d_off, nodes.d_list,
                       cells.d_off, cells.d_list, (const ushort4 *)p->d_loc, p->d_owned_off,
                       p->d_owned, p->d_csr_off, p->d_csr, (const ushort4 *)p->d_ell,
                       (const T *)a[0].dat->data_d, (const T *)a[2].dat->data_d,
                       (const T *)a[4].dat->data_d, (T *)a[6].dat->data_d, airfoil_consts<T>(),
                       take_fold(p->nblocks), known_zero(a[6].dat));
    return;
  }
  if (strategy == OP2_STRATEGY_OWNER_GATHER) {
    Plan *p = plan_for(name, set, nargs, a, strategy, part_of(name, 256));
    p->airfoil_res = true;  // a stage kernel of a capture may run on it
    const size_t lds = sizeof(T) * 4 * (size_t)p->max_positions;
    const int own = (p->max_owned + NT - 1) / NT;
    if (own > 4) fail("op2: owner tile of %s exceeds %d targets", name, 4 * NT);
    const bool ell = p->d_ell != nullptr;
    const bool oid = p->owned_identity;
    decltype(&k_res_calc_owner_direct<NT, T, 1, true, true>) kern;
#define OP2_G(O, E) \
  (oid ? k_res_calc_owner_direct<NT, T, O, E, true> : k_res_calc_owner_direct<NT, T, O, E, false>)
    if (ell)
      kern = own <= 1 ? OP2_G(1, true) : own <= 2 ? OP2_G(2, true) : OP2_G(4, true);
    else
      kern = own <= 1 ? OP2_G(1, false) : own <= 2 ? OP2_G(2, false) : OP2_G(4, false);
#undef OP2_G
    allow_lds(kern, std::max(lds, (size_t)64));
    if (p->nblocks == 0) {
      flush_fold();
      return;
    }
    hipLaunchKernelGGL(kern, dim3(p->nblocks), dim3(NT), std::max(lds, (size_t)64),
                       rt().stream, p->d_blk_off, p->nblocks, (const int4 *)p->d_ids,
                       p->d_ids16, (const int4 *)p->d_id_win, p->d_owned_off, p->d_owned,
                       p->d_csr_off, p->d_csr,
                       (const ushort4 *)p->d_ell, (const T *)a[0].dat->data_d,
                       (const T *)a[2].dat->data_d, (const T *)a[4].dat->data_d,
                       (T *)a[6].dat->data_d, airfoil_consts<T>(), take_fold(p->nblocks),
                       known_zero(a[6].dat));
    return;
  }
  if (strategy == OP2_STRATEGY_DIRECT) {
    const int grid = grid_for(set->size, 1 << 30, NT);
    hipLaunchKernelGGL((k_res_calc_direct<NT, T>), dim3(grid), dim3(NT), 0, rt().stream,
                       (const int2 *)a[0].map->map_d, (const int2 *)a[2].map->map_d,
                       (const T *)a[0].dat->data_d, (const T *)a[2].dat->data_d,
                       (const T *)a[4].dat->data_d, (T *)a[6].dat->data_d, set->size,
                       airfoil_consts<T>(), take_fold(grid));
    return;
  }
  const int part = part_of(name, sizeof(T) == 8 ? 256 : 512);
  if (part > 4 * NT) fail("op2: part size %d of %s exceeds %d", part, name, 4 * NT);
  Plan *p = plan_for(name, set, nargs, a, strategy, part);
  const bool col = strategy == OP2_STRATEGY_COLOURED;
  if (part <= NT)
    launch_res_calc<NT, T, 1>(p, a, col);
  else if (part <= 2 * NT)
    launch_res_calc<NT, T, 2>(p, a, col);
  else
    launch_res_calc<NT, T, 4>(p, a, col);
}

// loops with fewer kernel forms than res_calc map the other strategies on
// their nearest form
int narrow_strategy(int s) {
  if (s == OP2_STRATEGY_OWNER_GATHER) return OP2_STRATEGY_OWNER;
  if (s == OP2_STRATEGY_DIRECT) return OP2_STRATEGY_ATOMIC;
  return s;
}

// Boundary edges per cell of a bres_calc (its cell map pbecell), for the
// update that applies it: per 32 cells a (mask, boundary cells before)
// record, and per boundary cell two (node 1, node 2, edge, mesh) records
// from pbedge and, in the batched form, bedge -> mesh (edge -1: absent).
// Built when bres_calc runs outside a graph capture (the capture cannot
// copy to the device), cached per loop's maps.  `usable` is false when a
// cell has more than two boundary edges: that bres_calc is not fused.
struct BresIndex {
  op_map pbedge = nullptr, pbecell = nullptr, bmesh = nullptr;
  bool usable = false;
  uint2 *d_word = nullptr;
  int4 *d_rec = nullptr;
};

std::vector<BresIndex> &bres_cache() {
  static std::vector<BresIndex> cache;
  return cache;
}

const BresIndex *find_bres_index(int nargs, const op_arg *a) {
  const op_map bm = nargs == 7 ? a[6].map : nullptr;
  for (auto &b : bres_cache())
    if (b.pbedge == a[0].map && b.pbecell == a[2].map && b.bmesh == bm) return &b;
  return nullptr;
}

void build_bres_index(int nargs, const op_arg *a) {
  if (find_bres_index(nargs, a) || rt().capturing) return;
  BresIndex b;
  b.pbedge = a[0].map;
  b.pbecell = a[2].map;
  b.bmesh = nargs == 7 ? a[6].map : nullptr;
  const int ne = b.pbecell->from->size, nc = b.pbecell->to->size;
  const int *cell = map_host(b.pbecell), *nodes = map_host(b.pbedge);
  const int *mesh = b.bmesh ? map_host(b.bmesh) : nullptr;
  std::vector<int> cnt(nc, 0);
  for (int e = 0; e < ne; e++) cnt[cell[e]]++;
  b.usable = std::all_of(cnt.begin(), cnt.end(), [](int n) { return n <= 2; });
  if (b.usable) {
    std::vector<uint2> word(std::max((nc + 31) / 32, 1), make_uint2(0u, 0u));
    std::vector<int> slot(nc, -1);
    int nbc = 0;
    for (int c = 0; c < nc; c++) {
      if (c % 32 == 0) word[c / 32].y = (unsigned)nbc;
      if (!cnt[c]) continue;
      word[c / 32].x |= 1u << (c % 32);
      slot[c] = nbc++;
    }
    std::vector<int4> rec(2 * (size_t)std::max(nbc, 1), make_int4(0, 0, -1, 0));
    std::vector<int> used(nbc, 0);
    for (int e = 0; e < ne; e++) {  // edge order within each cell
      const int i = slot[cell[e]];
      rec[2 * (size_t)i + used[i]++] =
          make_int4(nodes[2 * (size_t)e], nodes[2 * (size_t)e + 1], e, mesh ? mesh[e] : 0);
    }
    b.d_word = (uint2 *)dev_alloc(sizeof(uint2) * word.size());
    h2d(b.d_word, word.data(), sizeof(uint2) * word.size());
    b.d_rec = (int4 *)dev_alloc(sizeof(int4) * rec.size());
    h2d(b.d_rec, rec.data(), sizeof(int4) * rec.size());
  }
  bres_cache().push_back(b);
}

template <typename T>
BresIdx<T> bres_idx(const op_arg *b, int nb, bool keep_res) {
  BresIdx<T> bi{};
  bi.keep_res = keep_res;
  if (!b) return bi;
  const BresIndex *ix = find_bres_index(nb, b);
  if (!ix || !ix->usable) fail("op2: no boundary index for map %s", b[2].map->name);
  bi.word = ix->d_word;
  bi.rec = ix->d_rec;
  bi.x = (const T *)b[0].dat->data_d;
  bi.q = (const T *)b[2].dat->data_d;
  bi.bound = (const int *)b[5].dat->data_d;
  bi.qinf_dat = nb == 7 ? (const T *)b[6].dat->data_d : nullptr;
  return bi;
}

template <int NT, typename T>
void run_bres_calc(const char *name, op_set set, int nargs, op_arg *a) {
  if (rt().graph_fusion) build_bres_index(nargs, a);  // for captures that fuse it
  const int requested = strategy_of(name);
  const int strategy = narrow_strategy(requested);
  // batched form: the free stream of each boundary edge's mesh, through a map
  const T *qd = nargs == 7 ? (const T *)a[6].dat->data_d : nullptr;
  const int *bm = nargs == 7 ? a[6].map->map_d : nullptr;
  if (requested == OP2_STRATEGY_DIRECT) {
    if (set->size == 0) return;
    hipLaunchKernelGGL((k_bres_calc_direct<NT, T>), dim3((set->size + NT - 1) / NT), dim3(NT), 0,
                       rt().stream, (const int2 *)a[0].map->map_d, (const int *)a[2].map->map_d,
                       (const T *)a[0].dat->data_d, (const T *)a[2].dat->data_d,
                       (const T *)a[3].dat->data_d, (T *)a[4].dat->data_d,
                       (const int *)a[5].dat->data_d, qd, bm, set->size, airfoil_consts<T>());
    return;
  }
  if (strategy == OP2_STRATEGY_OWNER) {
    Plan *p = plan_for(name, set, nargs, a, strategy, part_of(name, NT));
    const PlanGroup &nodes = p->groups[p->slots[0].first];
    const PlanGroup &cells = p->groups[p->owner_group];
    const size_t lds = owner_lds<T>(p, 2, 4, 4);
    allow_lds(k_bres_calc_owner<NT, T>, lds);
    if (p->nblocks == 0) return;
    hipLaunchKernelGGL((k_bres_calc_owner<NT, T>), dim3(p->nblocks), dim3(NT), lds, rt().stream,
                       p->d_blk_off, p->nblocks, p->d_perm, nodes.d_off, nodes.d_list,
                       cells.d_off, cells.d_list, (const ushort4 *)p->d_loc, p->d_owned_off,
                       p->d_owned, p->d_csr_off, p->d_csr, (const T *)a[0].dat->data_d,
                       (const T *)a[2].dat->data_d, (const T *)a[3].dat->data_d,
                       (T *)a[4].dat->data_d, (const int *)a[5].dat->data_d, qd, bm,
                       airfoil_consts<T>());
    return;
  }
  Plan *p = plan_for(name, set, nargs, a, strategy, part_of(name, NT));
  const bool col = strategy == OP2_STRATEGY_COLOURED;
  const PlanGroup &nodes = p->groups[p->slots[0].first];
  const PlanGroup &cells = p->groups[p->slots[2].first];
  const size_t lds = sizeof(T) * (2 * ((nodes.max_count + 1) & ~1) + 4 * (size_t)cells.max_count +
                                  ((cells.max_count + 3) & ~3) + 4 * (size_t)cells.max_count);
  auto kern = col ? k_bres_calc<NT, T, true> : k_bres_calc<NT, T, false>;
  if (lds > 64 * 1024)
    OP2_HIP_CHECK(hipFuncSetAttribute((const void *)kern,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  AirfoilConsts<T> k = airfoil_consts<T>();
  for (int c = 0; c < p->ncolors; c++) {
    const int first = p->colblk_off[c], nb = p->colblk_off[c + 1] - first;
    if (nb == 0) continue;
    hipLaunchKernelGGL(kern, dim3(nb), dim3(NT), lds, rt().stream, p->d_blk_off, p->d_blkmap,
                       first, nb, p->d_perm, nodes.d_off, nodes.d_list, cells.d_off,
                       cells.d_list, (const ushort4 *)p->d_loc, p->d_col_off, p->d_col_base,
                       p->d_blk_ncol, (const T *)a[0].dat->data_d, (const T *)a[2].dat->data_d,
                       (const T *)a[3].dat->data_d, (T *)a[4].dat->data_d,
                       (const int *)a[5].dat->data_d, qd, bm, k);
  }
}

constexpr int kReduceBlocks = 2048;

struct ReduceScratch {
  double *partials;
  double *out;
};

// layout: out (256 B), then nslots x nparts partials
ReduceScratch reduce_scratch(int nslots, int nparts = kReduceBlocks) {
  char *base = (char *)reduct_scratch(256 + sizeof(double) * (size_t)nparts * nslots);
  return {(double *)(base + 256), (double *)base};
}

#define OP2_UPD(A, S, X, B) k_update<NT, T, A, S, X, B>
template <int NT, typename T>
auto update_kernel(bool adt, bool save, bool xcd, bool bres) {
  decltype(&OP2_UPD(false, false, false, false)) table[16] = {
      OP2_UPD(false, false, false, false), OP2_UPD(false, false, false, true),
      OP2_UPD(false, false, true, false),  OP2_UPD(false, false, true, true),
      OP2_UPD(false, true, false, false),  OP2_UPD(false, true, false, true),
      OP2_UPD(false, true, true, false),   OP2_UPD(false, true, true, true),
      OP2_UPD(true, false, false, false),  OP2_UPD(true, false, false, true),
      OP2_UPD(true, false, true, false),   OP2_UPD(true, false, true, true),
      OP2_UPD(true, true, false, false),   OP2_UPD(true, true, false, true),
      OP2_UPD(true, true, true, false),    OP2_UPD(true, true, true, true)};
  return table[adt * 8 + save * 4 + xcd * 2 + bres];
}
#undef OP2_UPD

// update; fused with the bres_calc of bres_a (nb arguments) before it and
// the save_soln of save_a and/or the adt_calc of adt_a after it (all on the
// same cells and state) when given.  use: reduction scratch of the caller
// (result slot and kReduceBlocks partials) instead of the runtime's.
template <int NT, typename T>
const double *run_update(op_set set, op_arg *a, const op_arg *adt_a = nullptr,
                         const op_arg *save_a = nullptr, const op_arg *bres_a = nullptr,
                         int nb = 0, bool keep_res = false, const ReduceScratch *use = nullptr) {
  const int grid = grid_for(set->size, kReduceBlocks, NT);
  ReduceScratch s = use ? *use : reduce_scratch(1, grid);
  const bool xcd = sizeof(T) == 4 || (long long)grid * NT >= set->size;
  auto kern = update_kernel<NT, T>(adt_a, save_a, xcd, bres_a);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), 0, rt().stream,
                     (const T *)a[0].dat->data_d, (T *)a[1].dat->data_d, (T *)a[2].dat->data_d,
                     (const T *)a[3].dat->data_d, set->size, s.partials,
                     adt_a ? (const int4 *)adt_a[0].map->map_d : nullptr,
                     adt_a ? (const T *)adt_a[0].dat->data_d : nullptr,
                     adt_a ? (T *)adt_a[5].dat->data_d : nullptr,
                     save_a ? (T *)save_a[1].dat->data_d : nullptr,
                     bres_idx<T>(bres_a, nb, keep_res),
                     adt_a || bres_a ? airfoil_consts<T>() : AirfoilConsts<T>{});
  Runtime &r = rt();
  CarriedFold f{};
  f.partials = s.partials;
  f.n = grid;
  f.out = s.out;
  f.log = r.deferred ? r.d_log : nullptr;
  f.log_count = r.d_log_count;
  if (r.capturing && r.graph_fusion) {
    pending_fold() = PendingFold{true, f};  // the next res_calc carries it
  } else {
    hipLaunchKernelGGL((k_fold<1>), dim3(1), dim3(kFixedBlock), 0, r.stream, f.partials, f.n, 0u, f.out,
                       f.log, f.log_count);
  }
  check_launch("update");
  return s.out;
}

// Segments of a non-decreasing dim-1 map (cell -> mesh) for the batched
// update: run offsets and the partial slots of each run, cached per map.
struct SegPlan {
  op_map map = nullptr;
  int ch = 0;  // cells per workgroup of the update it was made for
  int nseg = 0, nslots = 0;
  int *d_seg_off = nullptr, *d_pslot = nullptr;
  double *d_partials = nullptr;
  unsigned *d_log_base = nullptr;  // log slot of this map's results in the current call
};

constexpr int kSegUpc = 4;  // cells per thread in k_update_seg

std::vector<SegPlan> &seg_cache() {
  static std::vector<SegPlan> cache;
  return cache;
}

SegPlan &seg_plan(op_map m, int ch) {
  std::vector<SegPlan> &cache = seg_cache();
  for (auto &sp : cache)
    if (sp.map == m && sp.ch == ch) return sp;
  const int n = m->from->size, nseg = m->to->size;
  const int *mm = map_host(m);
  std::vector<int> off(nseg + 1, 0);
  for (int c = 0; c < n; c++) {
    if (c > 0 && mm[c] < mm[c - 1])
      fail("op2: batched update needs the cell -> mesh map %s to be non-decreasing", m->name);
    off[mm[c] + 1]++;
  }
  for (int i = 0; i < nseg; i++) off[i + 1] += off[i];
  std::vector<int> pslot(nseg + 1, 0);
  for (int i = 0; i < nseg; i++)
    pslot[i + 1] = pslot[i] + (off[i + 1] > off[i] ? (off[i + 1] - 1) / ch - off[i] / ch + 1 : 0);
  SegPlan sp;
  sp.map = m;
  sp.ch = ch;
  sp.nseg = nseg;
  sp.nslots = pslot[nseg];
  sp.d_seg_off = (int *)dev_alloc(sizeof(int) * off.size());
  h2d(sp.d_seg_off, off.data(), sizeof(int) * off.size());
  sp.d_pslot = (int *)dev_alloc(sizeof(int) * pslot.size());
  h2d(sp.d_pslot, pslot.data(), sizeof(int) * pslot.size());
  sp.d_partials = (double *)dev_alloc(sizeof(double) * std::max(sp.nslots, 1));
  sp.d_log_base = (unsigned *)dev_alloc(256);
  cache.push_back(sp);
  return cache.back();
}

}  // namespace

// Frees the cached segment plans and boundary indexes (their maps are about
// to be freed).
namespace {

// Buffers of the stage kernels (k_stage) on one owner-gather res_calc plan,
// made before a capture because a capture cannot allocate: two sets of
// per-tile rms partials and of batched log slots (a stage writes one set
// while its workgroups fold the previous stage's), the single form's result
// slot, and for each cell -> mesh map of a batched update the first tile of
// each mesh (empty when a tile holds cells of two meshes).  Side recipes
// (kAdtSides) per (adt_calc cell -> node map, bres_calc maps): one 16-bit
// recipe per owned cell, or none when some side of some cell is neither an
// edge of its increment records nor one of its boundary edges.
struct SideRecipe {
  op_map pcell = nullptr, pbedge = nullptr, pbecell = nullptr, bmesh = nullptr;
  unsigned short *d = nullptr;
};
struct StageBufs {
  const Plan *plan = nullptr;
  double *partials[2] = {nullptr, nullptr};
  unsigned *log_base[2] = {nullptr, nullptr};
  double *out = nullptr;
  int parity = 0;
  std::vector<std::pair<op_map, int *>> tile_off;
  std::vector<SideRecipe> sides;
};

std::vector<StageBufs> &stage_bufs() {
  static std::vector<StageBufs> v;
  return v;
}

StageBufs *bufs_of(const Plan *p) {
  for (auto &b : stage_bufs())
    if (b.plan == p) return &b;
  return nullptr;
}

const int *tiles_of_meshes(const StageBufs &b, op_map m) {
  for (auto &t : b.tile_off)
    if (t.first == m) return t.second;
  return nullptr;
}

// The second buffer of every dat a stage writes (q, adt), and, during a
// capture, the buffer that holds each such dat's current values: null for
// its own; else its second buffer or, after a stage that also ran the next
// save_soln, the buffer of that save_soln's target (qold), which then holds
// the same values.
std::map<op_dat, void *> &shadow_bufs() {
  static std::map<op_dat, void *> m;
  return m;
}

std::map<op_dat, char *> &live_buf() {
  static std::map<op_dat, char *> m;
  return m;
}

void free_stage_bufs(StageBufs &b, std::vector<void *> *keep) {
  auto drop = [keep](void *p) {
    if (!p) return;
    if (keep)
      keep->push_back(p);
    else
      dev_free(p);
  };
  for (int i = 0; i < 2; i++) {
    drop(b.partials[i]);
    drop(b.log_base[i]);
  }
  drop(b.out);
  for (auto &t : b.tile_off) drop(t.second);
  for (auto &s : b.sides) drop(s.d);
}

// The side recipe of every owned cell of owner plan p (res_calc: args 0
// and 1 are the edge's nodes through pedge) for adt_calc's pcell and the
// boundary edges of a bres_calc on pbedge / pbecell: side i of cell c runs
// from pcell node i to node i+1 (mod 4); its 3-bit source is the first
// unused increment record of the cell whose edge has those two nodes (0-3),
// else the first unused boundary edge of the cell, in edge order, with those
// nodes (4-5: the cell's boundary record).  Empty when a side has no source.
std::vector<unsigned short> side_recipes(const Plan *p, op_map pcell, op_map pbedge,
                                         op_map pbecell) {
  std::vector<unsigned short> out;
  const op_map pedge = p->maps[0];
  if (p->ell.empty() || !pedge || p->maps[1] != pedge || pcell->dim != 4 || pbedge->dim < 2)
    return out;
  if (p->perm.empty() || pbecell->dim != 1 || pcell->from != pbecell->to) return out;
  const int nc = pcell->from->size;
  const int *en = map_host(pedge), *cn = map_host(pcell), *bn = map_host(pbedge),
            *bc = map_host(pbecell);
  const int i0 = p->idxs[0], i1 = p->idxs[1];
  // boundary edges of each cell in edge order (the order of the boundary
  // records, build_bres_index)
  std::vector<int> boff(nc + 1, 0);
  for (int e = 0; e < pbecell->from->size; e++) boff[bc[e] + 1]++;
  for (int c = 0; c < nc; c++) boff[c + 1] += boff[c];
  std::vector<int> bedges(boff[nc]);
  {
    std::vector<int> fill(boff.begin(), boff.end() - 1);
    for (int e = 0; e < pbecell->from->size; e++) bedges[fill[bc[e]]++] = e;
  }
  auto same = [](int a, int b, int u, int v) { return (a == u && b == v) || (a == v && b == u); };
  out.assign(p->owned.size(), 0);
  for (int t = 0; t < p->nblocks; t++)
    for (int i = p->owned_off[t]; i < p->owned_off[t + 1]; i++) {
      const int c = p->owned[i];
      if (boff[c + 1] - boff[c] > 2) return {};
      unsigned used = 0, recipe = 0;
      for (int k = 0; k < 4; k++) {
        const int u = cn[4 * (size_t)c + k], v = cn[4 * (size_t)c + (k + 1) % 4];
        int src = -1;
        for (int m = 0; m < 4 && src < 0; m++) {
          const unsigned short e = p->ell[4 * (size_t)i + m];
          if (e == 0xffff || (used >> m & 1)) continue;
          const int edge = p->perm[p->blk_off[t] + (e >> 2)];
          if (same(en[pedge->dim * (size_t)edge + i0], en[pedge->dim * (size_t)edge + i1], u, v))
            src = m;
        }
        for (int m = 0; m < boff[c + 1] - boff[c] && src < 0; m++) {
          const int be = bedges[boff[c] + m];
          if (!(used >> (4 + m) & 1) &&
              same(bn[pbedge->dim * (size_t)be], bn[pbedge->dim * (size_t)be + 1], u, v))
            src = 4 + m;
        }
        if (src < 0) return {};
        used |= 1u << src;
        recipe |= (unsigned)src << (3 * k);
      }
      out[i] = (unsigned short)recipe;
    }
  return out;
}

}  // namespace

void release_exec_caches() {
  for (auto &sp : seg_cache()) {
    dev_free(sp.d_seg_off);
    dev_free(sp.d_pslot);
    dev_free(sp.d_partials);
    dev_free(sp.d_log_base);
  }
  seg_cache().clear();
  for (auto &b : bres_cache()) {
    dev_free(b.d_word);
    dev_free(b.d_rec);
  }
  bres_cache().clear();
  for (auto &b : stage_bufs()) free_stage_bufs(b, nullptr);
  stage_bufs().clear();
  for (auto &s : shadow_bufs()) dev_free(s.second);
  shadow_bufs().clear();
  live_buf().clear();
}

void forget_plan_caches(const Plan *p) {
  auto &v = stage_bufs();
  for (size_t i = 0; i < v.size(); i++)
    if (v[i].plan == p) {
      free_stage_bufs(v[i], live_graphs() ? &rt().retired : nullptr);
      v.erase(v.begin() + i);
      return;
    }
}

// Before a capture with stage fusion: for every res_calc plan of the
// owner-gather strategy whose tiles own every cell, the stage buffers, the
// second buffers of its q and adt dats, and the tiles of each mesh of every
// batched update's cell -> mesh map.
void prepare_capture() {
  Runtime &r = rt();
  for (auto &kv : live_buf()) kv.second = nullptr;
  if (r.graph_fusion < 2) return;
  // the stage tiles of every res_calc that ran on gather tiles, built before
  // the capture (no allocation inside it)
  std::vector<Plan *> res_plans;
  for (auto &up : r.plans)
    if (up->airfoil_res && up->strategy == OP2_STRATEGY_OWNER_GATHER)
      res_plans.push_back(up.get());
  for (Plan *p : res_plans) {
    const int part = stage_part_of(p->loop);
    if (part == p->part_size) continue;
    Plan *sp = plan_for(p->loop.c_str(), p->set, (int)p->args.size(), p->args.data(),
                        OP2_STRATEGY_OWNER_GATHER, part);
    sp->airfoil_res = true;
  }
  for (auto &up : r.plans) {
    const Plan *p = up.get();
    if (!p->airfoil_res || p->strategy != OP2_STRATEGY_OWNER_GATHER || !p->d_ell ||
        p->max_owned > 2 * std::max(block_of(p->loop), stage_block_of(p->loop)))
      continue;
    const op_set cells = p->groups[p->owner_group].map->to;
    if ((int)p->owned.size() != cells->size) continue;
    StageBufs *b = bufs_of(p);
    if (!b) {
      StageBufs nb;
      nb.plan = p;
      for (int i = 0; i < 2; i++) {
        nb.partials[i] = (double *)dev_alloc(sizeof(double) * std::max(p->nblocks, 1));
        nb.log_base[i] = (unsigned *)dev_alloc(256);
      }
      nb.out = (double *)dev_alloc(256);
      stage_bufs().push_back(nb);
      b = &stage_bufs().back();
    }
    for (int a : {2, 4}) {  // q and adt
      op_dat d = p->dats[a];
      if (!shadow_bufs().count(d))
        shadow_bufs()[d] = dev_alloc((size_t)d->size * d->set->size);
      live_buf()[d] = nullptr;
    }
    for (auto &sp : seg_cache()) {
      bool known = false;
      for (auto &t : b->tile_off) known |= t.first == sp.map;
      if (sp.map->from != cells || known) continue;
      const int *mm = map_host(sp.map);
      std::vector<int> first(sp.nseg + 1, p->nblocks);
      bool ok = true;
      int last = -1;
      for (int t = 0; t < p->nblocks && ok; t++) {
        const int lo = p->owned_off[t], hi = p->owned_off[t + 1];
        const int m = mm[p->owned[lo]];
        for (int i = lo; i < hi && ok; i++) ok = mm[p->owned[i]] == m;
        ok = ok && m >= last;
        if (m != last)
          for (int k = last + 1; k <= m; k++) first[k] = t;
        last = m;
      }
      int *d = nullptr;
      if (ok) {
        d = (int *)dev_alloc(sizeof(int) * first.size());
        h2d(d, first.data(), sizeof(int) * first.size());
      }
      b->tile_off.emplace_back(sp.map, d);
    }
    // side recipes for each cell -> 4 nodes map and each boundary index on
    // these cells (adt_calc and bres_calc joining this plan's stages)
    for (int mi = 0; mi < OP_map_index; mi++) {
      op_map pc = OP_map_list[mi];
      if (pc->from != cells || pc->dim != 4) continue;
      for (auto &bx : bres_cache()) {
        if (!bx.usable || bx.pbecell->to != cells) continue;
        bool known = false;
        for (auto &s : b->sides)
          known |= s.pcell == pc && s.pbedge == bx.pbedge && s.pbecell == bx.pbecell &&
                   s.bmesh == bx.bmesh;
        if (known) continue;
        SideRecipe sr{pc, bx.pbedge, bx.pbecell, bx.bmesh, nullptr};
        const std::vector<unsigned short> h = side_recipes(p, pc, bx.pbedge, bx.pbecell);
        if (!h.empty()) {
          sr.d = (unsigned short *)dev_alloc(sizeof(unsigned short) * h.size());
          h2d(sr.d, h.data(), sizeof(unsigned short) * h.size());
        }
        b->sides.push_back(sr);
      }
    }
  }
}

namespace {

#define OP2_UPDS(A, S, B) k_update_seg<NT, T, kSegUpc, A, S, B>
template <int NT, typename T>
auto update_seg_kernel(bool adt, bool save, bool bres) {
  decltype(&OP2_UPDS(false, false, false)) table[8] = {
      OP2_UPDS(false, false, false), OP2_UPDS(false, false, true), OP2_UPDS(false, true, false),
      OP2_UPDS(false, true, true),   OP2_UPDS(true, false, false), OP2_UPDS(true, false, true),
      OP2_UPDS(true, true, false),   OP2_UPDS(true, true, true)};
  return table[adt * 4 + save * 2 + bres];
}
#undef OP2_UPDS

template <int NT, typename T>
void run_update_batch(op_set set, op_arg *a, const op_arg *adt_a = nullptr,
                      const op_arg *save_a = nullptr, const op_arg *bres_a = nullptr,
                      int nb = 0, bool keep_res = false) {
  const int ch = NT * kSegUpc;
  SegPlan &sp = seg_plan(a[4].map, ch);
  const int grid = (set->size + ch - 1) / ch;
  Runtime &r = rt();
  auto kern = update_seg_kernel<NT, T>(adt_a, save_a, bres_a);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), 0, r.stream,
                     (const T *)a[0].dat->data_d, (T *)a[1].dat->data_d, (T *)a[2].dat->data_d,
                     (const T *)a[3].dat->data_d, set->size, (const int *)a[4].map->map_d,
                     (const int *)sp.d_seg_off, (const int *)sp.d_pslot, sp.d_partials,
                     r.deferred ? r.d_log_count : nullptr, sp.nseg, sp.d_log_base,
                     adt_a ? (const int4 *)adt_a[0].map->map_d : nullptr,
                     adt_a ? (const T *)adt_a[0].dat->data_d : nullptr,
                     adt_a ? (T *)adt_a[5].dat->data_d : nullptr,
                     save_a ? (T *)save_a[1].dat->data_d : nullptr,
                     bres_idx<T>(bres_a, nb, keep_res),
                     adt_a || bres_a ? airfoil_consts<T>() : AirfoilConsts<T>{});
  if (sp.nseg > 0) {
    CarriedFold f{};
    f.partials = sp.d_partials;
    f.pslot = sp.d_pslot;
    f.nseg = sp.nseg;
    f.dat = (double *)a[4].dat->data_d;
    f.log = r.deferred ? r.d_log : nullptr;
    f.log_base = sp.d_log_base;
    if (r.capturing && r.graph_fusion)
      pending_fold() = PendingFold{true, f};  // the next res_calc carries it
    else
      hipLaunchKernelGGL(k_fold_seg, dim3(f.nseg), dim3(kFixedBlock), 0, r.stream, f.partials, f.pslot,
                         f.dat, f.log, f.log_base);
  }
}

template <int NT, typename T>
void run_jac_res(const char *name, op_set set, int nargs, op_arg *a) {
  const int strategy = narrow_strategy(strategy_of(name));
  if (strategy == OP2_STRATEGY_OWNER) {
    Plan *p = plan_for(name, set, nargs, a, strategy, part_of(name, NT));
    if (p->arg_slot[1] != 0 || p->arg_slot[2] != 1) fail("op2: unexpected slot layout for %s", name);
    const PlanGroup &nodes = p->groups[0];
    const size_t lds = sizeof(T) * ((size_t)nodes.max_count + p->max_positions);
    allow_lds(k_jac_res_owner<NT, T>, lds);
    if (p->nblocks == 0) return;
    hipLaunchKernelGGL((k_jac_res_owner<NT, T>), dim3(p->nblocks), dim3(NT), lds, rt().stream,
                       p->d_blk_off, p->nblocks, p->d_perm, nodes.d_off, nodes.d_list,
                       (const ushort2 *)p->d_loc, p->d_owned_off, p->d_owned, p->d_csr_off,
                       p->d_csr, (const T *)a[0].dat->data_d, (const T *)a[1].dat->data_d,
                       (T *)a[2].dat->data_d, (T)gbl_read(a[3]));
    return;
  }
  Plan *p = plan_for(name, set, nargs, a, strategy, part_of(name, 1024));
  const bool col = strategy == OP2_STRATEGY_COLOURED;
  const PlanGroup &nodes = p->groups[0];
  const size_t lds = (sizeof(T) * 2 + 1) * (size_t)nodes.max_count;  // us, ds, inc flags
  auto kern = col ? k_jac_res<NT, T, true> : k_jac_res<NT, T, false>;
  if (lds > 64 * 1024)
    OP2_HIP_CHECK(hipFuncSetAttribute((const void *)kern,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  // slot order: first use is arg 1 (u via idx 1), then arg 2 (du via idx 0)
  if (p->arg_slot[1] != 0 || p->arg_slot[2] != 1) fail("op2: unexpected slot layout for %s", name);
  const T beta = (T)gbl_read(a[3]);
  for (int c = 0; c < p->ncolors; c++) {
    const int first = p->colblk_off[c], nb = p->colblk_off[c + 1] - first;
    if (nb == 0) continue;
    hipLaunchKernelGGL(kern, dim3(nb), dim3(NT), lds, rt().stream, p->d_blk_off, p->d_blkmap,
                       first, nb, p->d_perm, nodes.d_off, nodes.d_list, (const ushort2 *)p->d_loc,
                       p->d_col_off, p->d_col_base, p->d_blk_ncol, (const T *)a[0].dat->data_d,
                       (const T *)a[1].dat->data_d, (T *)a[2].dat->data_d, beta);
  }
}

// use: as for run_update (two results of kReduceBlocks partials each)
template <int NT, typename T>
const double *run_jac_update(op_set set, op_arg *a, const ReduceScratch *use = nullptr) {
  ReduceScratch s = use ? *use : reduce_scratch(2);
  const int grid = grid_for(set->size, kReduceBlocks, NT);
  const T alpha = (T)get_const("alpha").as_double(0);
  hipLaunchKernelGGL((k_jac_update<NT, T>), dim3(grid), dim3(NT), 0, rt().stream,
                     (const T *)a[0].dat->data_d, (T *)a[1].dat->data_d, (T *)a[2].dat->data_d,
                     set->size, alpha, s.partials);
  hipLaunchKernelGGL((k_fold<2>), dim3(1), dim3(kFixedBlock), 0, rt().stream, (const double *)s.partials,
                     grid, 2u, s.out, rt().deferred ? rt().d_log : nullptr, rt().d_log_count);
  check_launch("update");
  return s.out;
}

enum class Loop {
  kUnknown, kSaveSoln, kAdtCalc, kResCalc, kBresCalc, kUpdate, kUpdateBatch, kJacRes, kJacUpdate
};

Loop identify(const char *name, op_set set, int nargs, const op_arg *a) {
  auto is = [&](const char *n) { return !strcmp(name, n); };
  if (is("save_soln") && nargs == 2 && direct(a[0], 4, OP_READ, set) &&
      direct(a[1], 4, OP_WRITE, set))
    return Loop::kSaveSoln;
  if (is("adt_calc") && nargs == 6 && a[0].map && a[0].map->dim == 4 && a[0].map->from == set) {
    bool ok = true;
    for (int i = 0; i < 4; i++) ok &= via(a[i], a[0].map, i, 2, OP_READ) && a[i].dat == a[0].dat;
    if (ok && direct(a[4], 4, OP_READ, set) && direct(a[5], 1, OP_WRITE, set))
      return Loop::kAdtCalc;
  }
  if (is("res_calc") && nargs == 8 && a[0].map && a[2].map && a[0].map->dim == 2 &&
      a[2].map->dim == 2 && a[0].map->from == set && a[2].map->from == set) {
    op_map pe = a[0].map, pc = a[2].map;
    bool ok = via(a[0], pe, 0, 2, OP_READ) && via(a[1], pe, 1, 2, OP_READ) &&
              a[1].dat == a[0].dat && via(a[2], pc, 0, 4, OP_READ) &&
              via(a[3], pc, 1, 4, OP_READ) && a[3].dat == a[2].dat &&
              via(a[4], pc, 0, 1, OP_READ) && via(a[5], pc, 1, 1, OP_READ) &&
              a[5].dat == a[4].dat && via(a[6], pc, 0, 4, OP_INC) &&
              via(a[7], pc, 1, 4, OP_INC) && a[7].dat == a[6].dat && pe != pc;
    if (ok) return Loop::kResCalc;
  }
  if (is("bres_calc") && nargs == 6 && a[0].map && a[2].map && a[0].map->dim == 2 &&
      a[2].map->dim == 1 && a[0].map->from == set && a[2].map->from == set) {
    op_map pb = a[0].map, pc = a[2].map;
    bool ok = via(a[0], pb, 0, 2, OP_READ) && via(a[1], pb, 1, 2, OP_READ) &&
              a[1].dat == a[0].dat && via(a[2], pc, 0, 4, OP_READ) &&
              via(a[3], pc, 0, 1, OP_READ) && via(a[4], pc, 0, 4, OP_INC) &&
              direct(a[5], 1, OP_READ, set) && is_type(a[5], "int") && pb != pc;
    if (ok) return Loop::kBresCalc;
  }
  // batched airfoil: each boundary edge reads its mesh's free stream
  if (is("bres_calc") && nargs == 7 && a[0].map && a[2].map && a[6].map &&
      a[0].map->dim == 2 && a[2].map->dim == 1 && a[6].map->dim == 1 &&
      a[0].map->from == set && a[2].map->from == set && a[6].map->from == set) {
    op_map pb = a[0].map, pc = a[2].map;
    bool ok = via(a[0], pb, 0, 2, OP_READ) && via(a[1], pb, 1, 2, OP_READ) &&
              a[1].dat == a[0].dat && via(a[2], pc, 0, 4, OP_READ) &&
              via(a[3], pc, 0, 1, OP_READ) && via(a[4], pc, 0, 4, OP_INC) &&
              direct(a[5], 1, OP_READ, set) && is_type(a[5], "int") && pb != pc &&
              via(a[6], a[6].map, 0, 4, OP_READ) && a[6].map != pb && a[6].map != pc;
    if (ok) return Loop::kBresCalc;
  }
  if (is("update") && nargs == 5 && direct(a[0], 4, OP_READ, set) &&
      direct(a[1], 4, OP_WRITE, set) && direct(a[2], 4, OP_RW, set) &&
      direct(a[3], 1, OP_READ, set) && gbl(a[4], 1, OP_INC))
    return Loop::kUpdate;
  // batched airfoil: the rms of each mesh, incremented through cell -> mesh
  if (is("update") && nargs == 5 && direct(a[0], 4, OP_READ, set) &&
      direct(a[1], 4, OP_WRITE, set) && direct(a[2], 4, OP_RW, set) &&
      direct(a[3], 1, OP_READ, set) && a[4].map && a[4].map->from == set &&
      a[4].map->dim == 1 && via(a[4], a[4].map, 0, 1, OP_INC) && is_type(a[4], "double"))
    return Loop::kUpdateBatch;
  if (is("res") && nargs == 4 && direct(a[0], 1, OP_READ, set) && a[1].map &&
      a[1].map->dim == 2 && via(a[1], a[1].map, 1, 1, OP_READ) &&
      via(a[2], a[1].map, 0, 1, OP_INC) && gbl(a[3], 1, OP_READ))
    return Loop::kJacRes;
  if (is("update") && nargs == 5 && direct(a[0], 1, OP_READ, set) &&
      direct(a[1], 1, OP_RW, set) && direct(a[2], 1, OP_INC, set) && gbl(a[3], 1, OP_INC) &&
      gbl(a[4], 1, OP_MAX))
    return Loop::kJacUpdate;
  return Loop::kUnknown;
}

template <typename T>
void dispatch(Loop l, const char *name, op_set set, int nargs, op_arg *a) {
  with_block(block_of(name), [&](auto B) {
    constexpr int BS = decltype(B)::value;
    switch (l) {
      case Loop::kSaveSoln: run_save_soln<BS, T>(set, a); break;
      case Loop::kAdtCalc: run_adt_calc<BS, T>(set, a); break;
      case Loop::kResCalc: run_res_calc<BS, T>(name, set, nargs, a); break;
      case Loop::kBresCalc: run_bres_calc<BS, T>(name, set, nargs, a); break;
      case Loop::kJacRes: run_jac_res<BS, T>(name, set, nargs, a); break;
      default: break;
    }
  });
}

// Candidate configurations of a tunable loop, in trial order.
std::vector<LoopConfig> tune_candidates(Loop l) {
  std::vector<LoopConfig> c;
  if (l == Loop::kResCalc) {
    // owner-gather tiles of 128-512 cells, each with workgroups of 128-512
    // threads holding one owned cell each: the form in which the stage
    // kernels of captured graphs run fastest (two cells per thread keep 4
    // waves per SIMD and, with more positions than three passes of the
    // workgroup, take adt_calc's nodes through pcell: FP32 stage 86 against
    // 78 us at 2.88M cells)
    for (int part : {128, 256, 512})
      for (int block : {128, 256, 512})
        if (part <= block) c.push_back({OP2_STRATEGY_OWNER_GATHER, part, block});
    for (int part : {128, 256}) c.push_back({OP2_STRATEGY_OWNER, part, 0});
  } else {
    for (int part : {128, 256, 512}) c.push_back({OP2_STRATEGY_OWNER, part, 0});
    if (l == Loop::kBresCalc) c.push_back({OP2_STRATEGY_DIRECT, 0, 0});
  }
  c.push_back({OP2_STRATEGY_ATOMIC, 256, 0});
  for (int part : {128, 256}) c.push_back({OP2_STRATEGY_COLOURED, part, 0});
  return c;
}

void dispatch_any(Loop l, int rb, const char *name, op_set set, int nargs, op_arg *a) {
  if (rb == 8)
    dispatch<double>(l, name, set, nargs, a);
  else
    dispatch<float>(l, name, set, nargs, a);
}

// One call of a loop under online tuning: runs the current candidate (the
// first call of each is untimed: it builds the plan), times it with events
// on the runtime stream, and after the last candidate keeps the fastest.
// A candidate that cannot run (a plan limit) is skipped and the call runs
// under the loop's previous configuration instead.
void online_call(const std::string &name, OnlineTune &ot, Loop l, int rb, op_set set, int nargs,
                 op_arg *a) {
  Runtime &r = rt();
  if (l != Loop::kResCalc && l != Loop::kBresCalc && l != Loop::kJacRes)
    fail("op2: loop \"%s\" cannot be autotuned (only indirect loops without global results)",
         name.c_str());
  if (r.capturing)
    fail("op2: online autotuning of loop \"%s\" during graph capture", name.c_str());
  if (!ot.started) {
    ot.started = true;
    ot.cands = tune_candidates(l);
    ot.us.assign(ot.cands.size(), 0.0);
    ot.events.assign(ot.cands.size(), {});
    auto it = r.loop_configs.find(name);
    ot.best = it != r.loop_configs.end() ? it->second : LoopConfig{};
  }
  const int per = ot.reps + 1;
  const int ncand = (int)ot.cands.size();
  const int idx = ot.call / per, sub = ot.call % per;
  const LoopConfig before = ot.best;
  r.loop_configs[name] = ot.cands[idx];
  const int keep_mode = r.error_mode;
  bool ran = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (sub > 0) {
    OP2_HIP_CHECK(hipEventCreate(&e0));
    OP2_HIP_CHECK(hipEventCreate(&e1));
    OP2_HIP_CHECK(hipEventRecord(e0, r.stream));
  }
  const std::string keep_err = r.last_error;
  try {
    r.error_mode = OP2_ERRORS_RETURN;
    dispatch_any(l, rb, name.c_str(), set, nargs, a);
    ran = true;
  } catch (const op2::Error &) {
    (void)hipGetLastError();
    r.last_error = keep_err;  // a candidate that cannot run is not an error of the call
  }
  r.error_mode = keep_mode;
  if (sub > 0) {
    OP2_HIP_CHECK(hipEventRecord(e1, r.stream));
    ot.events[idx].push_back({e0, e1});
  }
  if (!ran) {
    // skip the candidate and run this call under the previous configuration
    ot.us[idx] = 1e300;
    ot.call = (idx + 1) * per;
    r.loop_configs[name] = before;
    if (!before.strategy) r.loop_configs.erase(name);
    dispatch_any(l, rb, name.c_str(), set, nargs, a);
  } else {
    ot.call++;
  }
  if (ot.call < ncand * per) return;
  // all candidates ran: pick the fastest
  OP2_HIP_CHECK(hipStreamSynchronize(r.stream));
  int bi = -1;
  for (int i = 0; i < ncand; i++) {
    double sum = 0;
    for (auto &e : ot.events[i]) {
      float ms = 0;
      if (ot.us[i] < 1e300) OP2_HIP_CHECK(hipEventElapsedTime(&ms, e.first, e.second));
      sum += ms;
      OP2_HIP_CHECK(hipEventDestroy(e.first));
      OP2_HIP_CHECK(hipEventDestroy(e.second));
    }
    ot.events[i].clear();
    if (ot.us[i] < 1e300) ot.us[i] = 1e3 * sum / ot.reps;
    if (OP_diags > 0)
      printf(" online autotune %s: strategy %d part %d block %d: %.1f us\n", name.c_str(),
             ot.cands[i].strategy, ot.cands[i].part_size, ot.cands[i].block_size, ot.us[i]);
    if (ot.us[i] < 1e300 && (bi < 0 || ot.us[i] < ot.us[bi])) bi = i;
  }
  ot.done = true;
  if (bi < 0) {
    r.loop_configs[name] = before;
    if (!before.strategy) r.loop_configs.erase(name);
    fail("op2: no candidate configuration of loop \"%s\" ran", name.c_str());
  }
  ot.best = ot.cands[bi];
  ot.best_us = ot.us[bi];
  r.loop_configs[name] = ot.best;
  for (size_t i = 0; i < r.plans.size();) {
    Plan *p = r.plans[i].get();
    if (p->loop == name && p->set == set &&
        !(p->strategy == ot.best.strategy && p->part_size == ot.best.part_size))
      drop_plan(p);
    else
      i++;
  }
}

}  // namespace

namespace {
int run_loop(const char *name, op_set set, int nargs, op_arg *args);

// Loop fusion inside a graph capture.  bres_calc, save_soln and update are
// held back when called during capture.  An update on the cells whose res a
// held bres_calc increments takes the bres_calc in (BRES: each boundary cell
// adds its boundary fluxes before updating); a save_soln of the q a held
// update writes (the next iteration starting) joins it; then, if the next
// loop is an adt_calc on the same cells that reads that q (save_soln's
// source, update's result), one kernel does them all (k_save_adt, k_update /
// k_update_seg with BRES, SAVE and ADT), so q is not read again; any other
// next call first launches what is held as it is.  An update completed by
// its adt_calc stays held for one more call: when that is a res_calc whose
// owner-compute tiles own every cell of the res update zeroes, update does
// not store the zeros (keep_res) and res_calc takes res as zero without
// reading it (the zeros are never observed).  Each cell's arithmetic,
// and update's cells per workgroup, are those of the separate kernels, so
// the results are the same bit for bit, except that a cell with more than
// one boundary edge adds their increments one by one in edge order (a
// bres_calc strategy that sums them first may round differently).
struct Held {
  bool active = false;
  Loop l = Loop::kUnknown;
  int rb = 0;
  std::string name;
  op_set set = nullptr;
  std::vector<op_arg> args;       // dats, maps and dims only are used
  std::vector<op_arg> save_args;  // a save_soln joined to a held update
  std::vector<op_arg> bres_args;  // a bres_calc taken in by a held update
  std::vector<op_arg> adt_args;   // the adt_calc that completes a held update
};

Held &held() {
  static Held h;
  return h;
}

template <typename T>
void launch_held(Held &h, const op_arg *adt_a, bool keep_res) {
  const op_arg *save_a = h.save_args.empty() ? nullptr : h.save_args.data();
  const op_arg *bres_a = h.bres_args.empty() ? nullptr : h.bres_args.data();
  const int nb = (int)h.bres_args.size();
  // the fused kernels run with the held loop's block size
  with_block(block_of(h.name), [&](auto B) {
    constexpr int BS = decltype(B)::value;
    switch (h.l) {
      case Loop::kBresCalc:
        run_bres_calc<BS, T>(h.name.c_str(), h.set, (int)h.args.size(), h.args.data());
        break;
      case Loop::kSaveSoln:
        if (adt_a)
          run_save_adt<BS, T>(h.set, h.args.data(), adt_a);
        else
          run_save_soln<BS, T>(h.set, h.args.data());
        break;
      case Loop::kUpdate:
        run_update<BS, T>(h.set, h.args.data(), adt_a, save_a, bres_a, nb, keep_res);
        break;
      case Loop::kUpdateBatch:
        run_update_batch<BS, T>(h.set, h.args.data(), adt_a, save_a, bres_a, nb, keep_res);
        break;
      default: fail("op2: loop %s cannot be held for fusion", h.name.c_str());
    }
  });
}

// Launches the held loop, fused with the adt_calc of adt_a if given (or
// the one held with it).  keep_res: the res_calc that follows overwrites
// every cell's res without reading it, so update's zeros are not stored.
void release_held(const op_arg *adt_a, bool keep_res = false) {
  Held &h = held();
  if (!h.active) return;
  h.active = false;
  flush_fold();
  std::vector<op_arg> adt;
  adt.swap(h.adt_args);
  if (!adt_a && !adt.empty()) adt_a = adt.data();
  if (h.rb == 8)
    launch_held<double>(h, adt_a, keep_res);
  else
    launch_held<float>(h, adt_a, keep_res);
  check_launch(h.name.c_str());
  note_writes((int)h.bres_args.size(), h.bres_args.data());
  note_writes((int)h.args.size(), h.args.data());
  note_writes((int)h.save_args.size(), h.save_args.data());
  if (adt_a) note_writes(6, adt_a);
  if (h.l == Loop::kUpdate || h.l == Loop::kUpdateBatch) zero_dats().insert(h.args[2].dat);
  h.save_args.clear();
  h.bres_args.clear();
}

// A res_calc call (l, set, a) that follows a held update + adt_calc
// and overwrites every cell of the res that update zeroes without reading
// it: an owner-compute plan (the kernels that take res as known zero) whose
// tiles own every cell.
bool overwrites_held_res(const char *name, Loop l, op_set set, int nargs, op_arg *a) {
  const Held &h = held();
  if (!h.active || h.adt_args.empty() || l != Loop::kResCalc) return false;
  if (h.l != Loop::kUpdate && h.l != Loop::kUpdateBatch) return false;
  if (a[6].dat != h.args[2].dat || a[6].map->to != h.set) return false;
  auto ot = rt().online.find(name);
  if (ot != rt().online.end() && !ot->second.done) return false;
  const int strategy = strategy_of(name);
  if (strategy != OP2_STRATEGY_OWNER && strategy != OP2_STRATEGY_OWNER_GATHER) return false;
  const Plan *p = plan_for(name, set, nargs, a, strategy, part_of(name, 256));
  return p->owned.size() == (size_t)h.set->size;
}

// An update call (l, rb, set, a) on the cells whose res a held bres_calc
// increments.
bool takes_held_bres(Loop l, int rb, op_set set, const op_arg *a) {
  const Held &h = held();
  // the fused kernel applies the boundary flux with update's adt (a[3]), so
  // the bres_calc must read that same dat
  return h.active && h.l == Loop::kBresCalc && (l == Loop::kUpdate || l == Loop::kUpdateBatch) &&
         h.rb == rb && h.args[2].map->to == set && h.args[4].dat == a[2].dat &&
         h.args[3].dat == a[3].dat;
}

// A save_soln call (l, rb, set, a) that copies the q a held update writes.
bool joins_held(Loop l, int rb, op_set set, const op_arg *a) {
  const Held &h = held();
  return h.active && l == Loop::kSaveSoln && h.save_args.empty() && h.adt_args.empty() &&
         (h.l == Loop::kUpdate || h.l == Loop::kUpdateBatch) && h.set == set && h.rb == rb &&
         a[0].dat == h.args[1].dat;
}

// An adt_calc call (l, rb, set, a) that completes the held loop.
bool fuses_with_held(Loop l, int rb, op_set set, const op_arg *a) {
  const Held &h = held();
  if (!h.active || !h.adt_args.empty() || l != Loop::kAdtCalc || h.set != set || h.rb != rb)
    return false;
  const op_dat q = h.l == Loop::kSaveSoln ? h.args[0].dat : h.args[1].dat;
  return a[4].dat == q;
}

// ---- one kernel per Runge-Kutta stage (graph fusion level 2) ----------------
//
// During a capture a res_calc on an owner-gather plan whose tiles own every
// cell is held as the start of a stage; the bres_calc, the update and the
// next save_soln and adt_calc on the same cells and dats join it, and the
// stage runs as one k_stage kernel when the next call does not join.  The
// stage reads q and adt from the buffer that holds their current values and
// writes them to the other one (two stages per airfoil iteration bring them
// back); any other loop first copies such dats back (sync_shadows).  When
// the call that ends a stage starts the next stage on the same res, the
// stage does not store res's zeros and the next one takes res as zero.

struct Stage {
  bool active = false;
  int rb = 0;
  Plan *plan = nullptr;
  std::string name;
  op_set eset = nullptr;
  std::vector<op_arg> res, bres, upd, save, adt;
  bool batch = false;
  bool zero_in = false;  // res holds zeros at the stage's start
};

Stage &stage() {
  static Stage s;
  return s;
}

char *live_ptr(op_dat d) {
  auto it = live_buf().find(d);
  return it != live_buf().end() && it->second ? it->second : d->data_d;
}

// a buffer of d that does not hold its current values
char *other_ptr(op_dat d) {
  return live_ptr(d) == d->data_d ? (char *)shadow_bufs().at(d) : d->data_d;
}

template <typename T>
__global__ __launch_bounds__(kFixedBlock) void k_copy_elems(const T *__restrict__ src, T *__restrict__ dst,
                                                   long long n) {
  constexpr int NT = kFixedBlock;
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < n;
       i += (long long)gridDim.x * NT)
    dst[i] = src[i];
}

// Copies every dat whose current values are in its second buffer back to its
// own (before a loop other than a stage reads or writes dats, and at the end
// of a capture).
void sync_shadows() {
  for (auto &kv : live_buf()) {
    if (!kv.second) continue;
    op_dat d = kv.first;
    const long long bytes = (long long)d->size * d->set->size;
    if (bytes % 16 == 0) {  // 16-byte elements (device buffers are 256-byte aligned)
      const long long n = bytes / 16;
      hipLaunchKernelGGL((k_copy_elems<uint4>), dim3(grid_for(n, 1 << 16)), dim3(kFixedBlock), 0,
                         rt().stream, (const uint4 *)kv.second, (uint4 *)d->data_d, n);
    } else {
      const long long words = bytes / 4;
      hipLaunchKernelGGL((k_copy_elems<unsigned>), dim3(grid_for(words, 1 << 16)),
                         dim3(kFixedBlock), 0, rt().stream, (const unsigned *)kv.second,
                         (unsigned *)d->data_d, words);
    }
    check_launch("graph stage");
    kv.second = nullptr;
  }
}

// The owner-gather plan of a res_calc call that can start a stage, else null.
Plan *starts_stage(const char *name, Loop l, int rb, op_set set, int nargs, op_arg *a) {
  const Runtime &r = rt();
  if (r.graph_fusion < 2 || l != Loop::kResCalc) return nullptr;
  auto ot = r.online.find(name);
  if (ot != r.online.end() && !ot->second.done) return nullptr;
  if (strategy_of(name) != OP2_STRATEGY_OWNER_GATHER) return nullptr;
  // (built by prepare_capture: nothing is built inside a capture)
  Plan *p = find_plan({name, set, nargs, a, OP2_STRATEGY_OWNER_GATHER, stage_part_of(name)});
  if (!p || !bufs_of(p) || !shadow_bufs().count(a[2].dat) || !shadow_bufs().count(a[4].dat))
    return nullptr;
  if (p->max_owned > 2 * stage_block_of(name)) return nullptr;  // at most two cells per thread
  const op_dat q = a[2].dat, adt = a[4].dat, res = a[6].dat;
  if (q == adt || q == res || adt == res || a[0].dat == q || a[0].dat == adt) return nullptr;
  (void)rb;
  return p;
}

// A call that joins the held stage (returns true).
bool stage_takes(Loop l, int rb, op_set set, int nargs, op_arg *a) {
  Stage &st = stage();
  if (!st.active || rb != st.rb) return false;
  const op_dat q = st.res[2].dat, adt = st.res[4].dat, res = st.res[6].dat;
  const op_set cells = st.res[6].map->to;
  if (l == Loop::kBresCalc && st.bres.empty() && st.upd.empty()) {
    const BresIndex *ix = find_bres_index(nargs, a);
    if (!ix || !ix->usable || a[2].map->to != cells || a[4].dat != res || a[2].dat != q ||
        a[3].dat != adt || a[0].dat == q || a[0].dat == adt)
      return false;
    st.bres.assign(a, a + nargs);
    return true;
  }
  if ((l == Loop::kUpdate || l == Loop::kUpdateBatch) && st.upd.empty()) {
    if (set != cells || a[2].dat != res || a[1].dat != q || a[3].dat != adt)
      return false;
    if (a[0].dat == q || a[0].dat == adt || a[0].dat == res) return false;
    if (l == Loop::kUpdateBatch && !tiles_of_meshes(*bufs_of(st.plan), a[4].map)) return false;
    st.upd.assign(a, a + nargs);
    st.batch = l == Loop::kUpdateBatch;
    return true;
  }
  if (l == Loop::kSaveSoln && !st.upd.empty() && st.save.empty() && st.adt.empty()) {
    if (set != cells || a[0].dat != q || a[1].dat == q || a[1].dat == adt || a[1].dat == res)
      return false;
    st.save.assign(a, a + nargs);
    return true;
  }
  if (l == Loop::kAdtCalc && !st.upd.empty() && st.adt.empty()) {
    if (set != cells || a[4].dat != q || a[5].dat != adt || a[0].dat == q || a[0].dat == adt)
      return false;
    st.adt.assign(a, a + nargs);
    return true;
  }
  return false;
}

#define OP2_STG(W, O, B, S, A) k_stage<NT, T, W, O, B, S, A>
#define OP2_STG3(W, O, B, S) OP2_STG(W, O, B, S, 0), OP2_STG(W, O, B, S, 1), OP2_STG(W, O, B, S, 2)
#define OP2_STG24(W)                                                                      \
  OP2_STG3(W, false, false, false), OP2_STG3(W, false, false, true),                       \
      OP2_STG3(W, false, true, false), OP2_STG3(W, false, true, true),                     \
      OP2_STG3(W, true, false, false), OP2_STG3(W, true, false, true),                     \
      OP2_STG3(W, true, true, false), OP2_STG3(W, true, true, true)
template <int NT, typename T>
auto stage_kernel(int own, bool oid, bool bres, bool save, int adt) {
  decltype(&OP2_STG(1, true, true, true, 0)) table[48] = {OP2_STG24(1), OP2_STG24(2)};
  return table[(own > 1) * 24 + oid * 12 + bres * 6 + save * 3 + adt];
}
#undef OP2_STG24
#undef OP2_STG3
#undef OP2_STG

template <int NT, typename T>
void run_stage(Stage &st, bool store_res) {
  Runtime &r = rt();
  Plan *p = st.plan;
  StageBufs *sb = bufs_of(p);
  const op_dat q = st.res[2].dat, adt = st.res[4].dat, res = st.res[6].dat;
  const bool save = !st.save.empty(), has_adt = !st.adt.empty(), has_bres = !st.bres.empty();
  StageArgs<T> a{};
  a.blk_off = p->d_blk_off;
  a.nblk = p->nblocks;
  a.ids = (const int4 *)p->d_ids;
  a.owned_off = p->d_owned_off;
  a.owned = p->d_owned;
  a.ell = (const ushort4 *)p->d_ell;
  a.x = (const T *)live_ptr(st.res[0].dat);
  a.q_in = (const T *)live_ptr(q);
  a.adt_in = (const T *)live_ptr(adt);
  a.qold = (const T *)live_ptr(st.upd[0].dat);
  a.res = (T *)res->data_d;
  // with the next save_soln joined, the new q is the next iteration's qold:
  // it is written once, over the qold the update reads (each cell's own,
  // read by its own thread before the store), and q then lives in qold's
  // buffer; not when q already lives there (other tiles read it as halo)
  const op_dat qold = st.upd[0].dat;
  const bool save_in_place = save && st.save[1].dat == qold && qold != q && qold != adt &&
                             !shadow_bufs().count(qold) && live_ptr(q) != qold->data_d;
  char *const q_dst = save_in_place ? qold->data_d : other_ptr(q);
  a.q_out = (T *)q_dst;
  a.adt_out = (T *)other_ptr(adt);
  a.qold_out = save && !save_in_place ? (T *)st.save[1].dat->data_d : nullptr;
  a.pcell = has_adt ? (const int4 *)st.adt[0].map->map_d : nullptr;
  a.xc = has_adt ? (const T *)live_ptr(st.adt[0].dat) : nullptr;
  const int par = sb->parity;
  sb->parity ^= 1;
  a.partials = sb->partials[par];
  a.log_count = st.batch && r.deferred ? r.d_log_count : nullptr;
  a.log_base = sb->log_base[par];
  const int *tiles = st.batch ? tiles_of_meshes(*sb, st.upd[4].map) : nullptr;
  a.nseg = st.batch ? st.upd[4].map->to->size : 0;
  BresIdx<T> bi = has_bres ? bres_idx<T>(st.bres.data(), (int)st.bres.size(), false) : BresIdx<T>{};
  const int own = (p->max_owned + NT - 1) / NT;
  if (own > 2) fail("op2: stage kernel with %d cells per thread", own);
  // adt_calc from the tile's edges when a side recipe exists for these maps,
  // every node is read from the one x dat, and the tile's positions fit the
  // kernel's passes; else through pcell
  int adt_mode = has_adt ? kAdtMap : kAdtNone;
  if (has_adt && has_bres && p->max_positions <= kSideTrips * NT &&
      st.adt[0].dat == st.res[0].dat && st.bres[0].dat == st.res[0].dat) {
    bool in_order = true;
    for (int i = 0; i < 4; i++) in_order &= st.adt[i].map == st.adt[0].map && st.adt[i].idx == i;
    const op_map bm = st.bres.size() == 7 ? st.bres[6].map : nullptr;
    for (auto &sr : sb->sides)
      if (in_order && sr.d && sr.pcell == st.adt[0].map && sr.pbedge == st.bres[0].map &&
          sr.pbecell == st.bres[2].map && sr.bmesh == bm) {
        a.sides = sr.d;
        adt_mode = kAdtSides;
      }
  }
  auto kern = stage_kernel<NT, T>(own, p->owned_identity, has_bres, save && !save_in_place,
                                  adt_mode);
  const size_t lds = std::max(sizeof(T) * 4 * (size_t)p->max_positions, (size_t)64);
  allow_lds(kern, lds);
  const CarriedFold fold = take_fold(p->nblocks);
  const int mark = mark_begin(has_bres * 1 + save * 2 + has_adt * 4 + !st.zero_in * 8 +
                              store_res * 16 + (adt_mode == kAdtSides) * 32 +
                              save_in_place * 64 + ((char *)a.qold == (char *)a.q_in) * 128);
  hipLaunchKernelGGL(kern, dim3(p->nblocks), dim3(NT), lds, r.stream, a, bi, airfoil_consts<T>(),
                     fold, st.zero_in, store_res);
  check_launch("graph stage");
  mark_end(mark);
  live_buf()[q] = q_dst == q->data_d ? nullptr : q_dst;
  live_buf()[adt] = (char *)a.adt_out == adt->data_d ? nullptr : (char *)a.adt_out;
  CarriedFold f{};
  f.partials = sb->partials[par];
  f.log = r.deferred ? r.d_log : nullptr;
  if (st.batch) {
    f.pslot = tiles;
    f.nseg = a.nseg;
    f.dat = (double *)st.upd[4].dat->data_d;
    f.log_base = sb->log_base[par];
  } else {
    f.n = p->nblocks;
    f.out = sb->out;
    f.log_count = r.d_log_count;
  }
  if (f.nseg > 0 || f.n > 0) pending_fold() = PendingFold{true, f};
  zero_dats().erase(q);
  zero_dats().erase(adt);
  if (save) zero_dats().erase(st.save[1].dat);
  zero_dats().insert(res);
}

// Launches the held stage: as one k_stage kernel once an update has joined,
// else as the separate res_calc (and bres_calc) kernels.  next_takes_res:
// the call that ends the stage starts a stage on the same res (which will
// take it as zero), so the zeros need not be stored.
void launch_stage(bool next_takes_res) {
  Stage &st = stage();
  if (!st.active) return;
  st.active = false;
  if (st.upd.empty()) {
    sync_shadows();
    const int nr = (int)st.res.size();
    if (st.rb == 8)
      dispatch<double>(Loop::kResCalc, st.name.c_str(), st.eset, nr, st.res.data());
    else
      dispatch<float>(Loop::kResCalc, st.name.c_str(), st.eset, nr, st.res.data());
    check_launch(st.name.c_str());
    note_writes(nr, st.res.data());
    if (!st.bres.empty()) {
      flush_fold();
      const int nb = (int)st.bres.size();
      const op_set bset = st.bres[2].map->from;
      if (st.rb == 8)
        dispatch<double>(Loop::kBresCalc, "bres_calc", bset, nb, st.bres.data());
      else
        dispatch<float>(Loop::kBresCalc, "bres_calc", bset, nb, st.bres.data());
      check_launch("bres_calc");
      note_writes(nb, st.bres.data());
    }
    return;
  }
  with_block(stage_block_of(st.name), [&](auto B) {
    constexpr int BS = decltype(B)::value;
    if (st.rb == 8)
      run_stage<BS, double>(st, !next_takes_res);
    else
      run_stage<BS, float>(st, !next_takes_res);
  });
}
}  // namespace

void fusion_flush() {
  launch_stage(false);
  release_held(nullptr);
  sync_shadows();
  flush_fold();
  zero_dats().clear();
}
void forget_zero_dats() { zero_dats().clear(); }

void expose_dat(op_dat d) {
  exposed_dats().insert(d);
  zero_dats().erase(d);
}

void fusion_drop() {
  exposed_dats().clear();
  stage().active = false;
  for (auto &kv : live_buf()) kv.second = nullptr;
  held().active = false;
  held().save_args.clear();
  held().bres_args.clear();
  held().adt_args.clear();
  pending_fold().active = false;
  zero_dats().clear();
}

int par_loop(const char *name, op_set set, int nargs, op_arg *args) {
  if (!name || !set || nargs <= 0 || !args) fail("op_par_loop error -- invalid arguments");
  if (checkpoint_skip(name, nargs, args)) return 0;
  const int rc = run_loop(name, set, nargs, args);
  checkpoint_after(name, nargs, args);
  return rc;
}

namespace {

int run_loop(const char *name, op_set set, int nargs, op_arg *args) {
  ensure_device();
  if (rt().auto_renumber) auto_renumber(set, nargs, args);
  if (OP_diags > 0) {
    int ninds = 0;
    for (int m = 0; m < nargs; m++) op_arg_check(set, m, args[m], &ninds, name);
  }
  for (int m = 0; m < nargs; m++) {
    if (args[m].argtype != OP_ARG_DAT) continue;
    if (!args[m].dat || !args[m].dat->data_d)
      fail("op_par_loop error -- arg %d of %s has no device data (declare dats with op_decl_dat)",
           m, name);
    if (args[m].map) ensure_map_device(args[m].map);
  }
  // a loop registered by a translation unit of the generic path runs the
  // user's own kernel, even where a built-in kernel has the same name and
  // argument signature (the user's body may differ from the built-in one)
  auto reg = rt().loops.find(name);
  const Loop l = reg != rt().loops.end() ? Loop::kUnknown : identify(name, set, nargs, args);
  if (l == Loop::kUnknown) {
    release_held(nullptr);
    flush_fold();
    auto it = reg;
    if (it == rt().loops.end())
      fail("op_par_loop error -- no device kernel for loop \"%s\" with this argument "
           "signature, and no registered generic loop of that name", name);
    LoopTiming &t = timing_for(name, loop_bytes(set, nargs, args), through_map(nargs, args));
    hipEvent_t start;
    timing_begin(t, &start);
    const int rc = it->second(name, set, nargs, args);
    if (rc != 0) fail("op_par_loop error -- generic loop \"%s\": %s", name, rt().last_error.c_str());
    check_launch(name);
    note_writes(nargs, args);
    timing_end(t, start);
    return 0;
  }
  // the batched update's per-mesh rms is double whatever the state type
  const int rb = real_bytes(l == Loop::kUpdateBatch ? 4 : nargs, args);
  if (rb != 4 && rb != 8)
    fail("op_par_loop error -- loop \"%s\" needs all its real dats to be double or all float",
         name);
  if (OP_diags > 2) printf(" kernel routine on device: %s \n", name);

  Runtime &r = rt();
  const bool fuse = r.capturing && r.graph_fusion;
  if (fuse) {
    if (stage_takes(l, rb, set, nargs, args)) return 0;
    Plan *sp = starts_stage(name, l, rb, set, nargs, args);
    launch_stage(sp && stage().res.size() > 6 && stage().res[6].dat == args[6].dat);
    if (sp) {
      // a stage reads q and adt where they are (the previous stage's
      // outputs); anything held from before is launched first
      release_held(nullptr, overwrites_held_res(name, l, set, nargs, args));
      Stage &st = stage();
      st = Stage{};
      st.active = true;
      st.rb = rb;
      st.plan = sp;
      st.name = name;
      st.eset = set;
      st.res.assign(args, args + nargs);
      st.zero_in = known_zero(args[6].dat);
      return 0;
    }
    sync_shadows();
  }
  if (fuse && fuses_with_held(l, rb, set, args)) {
    Held &h = held();
    if (h.l == Loop::kUpdate || h.l == Loop::kUpdateBatch) {
      h.adt_args.assign(args, args + nargs);  // held until the next call (keep_res)
      return 0;
    }
    release_held(args);
    return 0;
  }
  if (fuse && joins_held(l, rb, set, args)) {
    held().save_args.assign(args, args + nargs);
    return 0;
  }
  if (fuse && takes_held_bres(l, rb, set, args)) {
    Held &h = held();
    h.bres_args.swap(h.args);
    h.l = l;
    h.name = name;
    h.set = set;
    h.args.assign(args, args + nargs);
    return 0;
  }
  release_held(nullptr, fuse && overwrites_held_res(name, l, set, nargs, args));
  if (fuse && (l == Loop::kSaveSoln || l == Loop::kUpdate || l == Loop::kUpdateBatch ||
               (l == Loop::kBresCalc && find_bres_index(nargs, args) &&
                find_bres_index(nargs, args)->usable))) {
    // results go to the reduction log (capture needs deferred reductions)
    Held &h = held();
    h.active = true;
    h.l = l;
    h.rb = rb;
    h.name = name;
    h.set = set;
    h.args.assign(args, args + nargs);
    return 0;
  }

  if (l != Loop::kResCalc) flush_fold();  // res_calc may carry it
  LoopTiming &t = timing_for(name, loop_bytes(set, nargs, args), through_map(nargs, args));
  hipEvent_t start;
  timing_begin(t, &start);
  if (l == Loop::kUpdate) {
    const double *d = with_block(block_of(name), [&](auto B) {
      constexpr int BS = decltype(B)::value;
      return rb == 8 ? run_update<BS, double>(set, args) : run_update<BS, float>(set, args);
    });
    note_writes(nargs, args);
    if (!rt().capturing) zero_dats().insert(args[2].dat);  // update leaves res zero
    timing_end(t, start);
    if (rt().deferred) return 0;
    gbl_inc(args[4], fetch_results(d, 1)[0]);
    return 0;
  }
  if (l == Loop::kUpdateBatch) {
    with_block(block_of(name), [&](auto B) {
      constexpr int BS = decltype(B)::value;
      if (rb == 8)
        run_update_batch<BS, double>(set, args);
      else
        run_update_batch<BS, float>(set, args);
    });
    check_launch(name);
    note_writes(nargs, args);
    if (!rt().capturing) zero_dats().insert(args[2].dat);
    timing_end(t, start);
    return 0;
  }
  if (l == Loop::kJacUpdate) {
    const double *d = with_block(block_of(name), [&](auto B) {
      constexpr int BS = decltype(B)::value;
      return rb == 8 ? run_jac_update<BS, double>(set, args) : run_jac_update<BS, float>(set, args);
    });
    note_writes(nargs, args);
    timing_end(t, start);
    if (rt().deferred) return 0;
    const double *h = fetch_results(d, 2);
    gbl_inc(args[3], h[0]);
    gbl_max(args[4], h[1]);
    return 0;
  }
  auto ot = rt().online.find(name);
  if (ot != rt().online.end() && !ot->second.done) {
    online_call(ot->first, ot->second, l, rb, set, nargs, args);
  } else {
    if (rb == 8)
      dispatch<double>(l, name, set, nargs, args);
    else
      dispatch<float>(l, name, set, nargs, args);
  }
  check_launch(name);
  note_writes(nargs, args);  // after the launch: res_calc checked its res first
  timing_end(t, start);
  return 0;
}

}  // namespace

// Autotuner: times candidate (strategy, part size) pairs of one loop on the
// live mesh and keeps the fastest as the loop's configuration.  Every dat
// the loop writes is saved before and restored after, so the state is
// unchanged; plans of the losing candidates are freed.  Loops with global
// reductions are not tuned (their results would be consumed).
int autotune(const char *name, op_set set, int nargs, op_arg *args, int reps, int *best_strategy,
             int *best_part, double *best_us) {
  ensure_device();
  refuse_with_live_graphs("op2_autotune_loop");
  forget_zero_dats();  // the trials write and restore the loop's dats
  const Loop l = rt().loops.count(name) ? Loop::kUnknown : identify(name, set, nargs, args);
  if (l != Loop::kResCalc && l != Loop::kBresCalc && l != Loop::kJacRes)
    fail("op2: loop \"%s\" cannot be autotuned (only indirect loops without global results)",
         name);
  // the batched update's per-mesh rms is double whatever the state type
  const int rb = real_bytes(l == Loop::kUpdateBatch ? 4 : nargs, args);
  if (rb != 4 && rb != 8) fail("op2: loop \"%s\" mixes real types", name);
  for (int m = 0; m < nargs; m++)
    if (args[m].argtype == OP_ARG_DAT && args[m].map) ensure_map_device(args[m].map);
  Runtime &r = rt();
  if (r.capturing) fail("op2: autotuning during graph capture");
  // save written dats
  std::vector<std::pair<op_dat, void *>> saved;
  for (int m = 0; m < nargs; m++) {
    const op_arg &a = args[m];
    if (a.argtype != OP_ARG_DAT || a.acc == OP_READ) continue;
    bool have = false;
    for (auto &sv : saved) have |= sv.first == a.dat;
    if (have) continue;
    const size_t bytes = (size_t)a.dat->size * a.dat->set->size;
    void *tmp = dev_alloc(bytes);
    OP2_HIP_CHECK(hipMemcpyAsync(tmp, a.dat->data_d, bytes, hipMemcpyDeviceToDevice, r.stream));
    saved.emplace_back(a.dat, tmp);
  }
  const LoopConfig keep = config_for(name);
  const int keep_mode = r.error_mode;
  const bool keep_timing = r.timing_enabled;
  r.error_mode = OP2_ERRORS_RETURN;
  r.timing_enabled = false;
  hipEvent_t e0, e1;
  OP2_HIP_CHECK(hipEventCreate(&e0));
  OP2_HIP_CHECK(hipEventCreate(&e1));
  double best = 1e300;
  int bs = 0, bp = 0, bb = 0;
  const std::string keep_err = r.last_error;
  auto trial = [&](int strategy, int part, int block) {
    r.loop_configs[name] = LoopConfig{strategy, part, block};
    const size_t nplans = r.plans.size();
    double us = 1e300;
    try {
      if (rb == 8)
        dispatch<double>(l, name, set, nargs, args);
      else
        dispatch<float>(l, name, set, nargs, args);
      check_launch(name);
      OP2_HIP_CHECK(hipEventRecord(e0, r.stream));
      for (int i = 0; i < reps; i++) {
        if (rb == 8)
          dispatch<double>(l, name, set, nargs, args);
        else
          dispatch<float>(l, name, set, nargs, args);
      }
      OP2_HIP_CHECK(hipEventRecord(e1, r.stream));
      OP2_HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0;
      OP2_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      us = 1e3 * ms / reps;
    } catch (const op2::Error &) {
      (void)hipGetLastError();
      r.last_error = keep_err;
    }
    if (OP_diags > 0)
      printf(" autotune %s: strategy %d part %d block %d: %.1f us\n", name, strategy, part,
             block_of(name), us);
    if (us < best) {
      best = us;
      bs = strategy;
      bp = part;
      bb = block;
    }
    // free the plans this trial built unless it is the best so far
    while (r.plans.size() > nplans) {
      Plan *p = r.plans.back().get();
      if (p->strategy == bs && p->part_size == bp) break;
      drop_plan(p);
    }
  };
  const int strategies[] = {OP2_STRATEGY_OWNER_GATHER, OP2_STRATEGY_OWNER, OP2_STRATEGY_ATOMIC,
                            OP2_STRATEGY_COLOURED};
  for (int st : strategies) {
    if (l != Loop::kResCalc && st == OP2_STRATEGY_OWNER_GATHER) continue;
    trial(st, 256, 0);
  }
  for (int part : {128, 512}) trial(bs, part, 0);
  // then the workgroup size of the best (strategy, part size)
  for (int block : {128, 256, 512})
    if (block != OP_block_size) trial(bs, bp, block);
  OP2_HIP_CHECK(hipEventDestroy(e0));
  OP2_HIP_CHECK(hipEventDestroy(e1));
  // drop every losing plan of this loop that is still cached
  for (size_t i = 0; i < r.plans.size();) {
    Plan *p = r.plans[i].get();
    if (p->loop == name && p->set == set && !(p->strategy == bs && p->part_size == bp))
      drop_plan(p);
    else
      i++;
  }
  for (auto &sv : saved) {
    const size_t bytes = (size_t)sv.first->size * sv.first->set->size;
    OP2_HIP_CHECK(hipMemcpyAsync(sv.first->data_d, sv.second, bytes, hipMemcpyDeviceToDevice,
                                 r.stream));
  }
  OP2_HIP_CHECK(hipStreamSynchronize(r.stream));
  for (auto &sv : saved) dev_free(sv.second);
  r.error_mode = keep_mode;
  r.timing_enabled = keep_timing;
  if (best >= 1e300) {
    r.loop_configs[name] = keep;
    fail("op2: no candidate configuration of loop \"%s\" ran", name);
  }
  r.loop_configs[name] = LoopConfig{bs, bp, bb};
  if (best_strategy) *best_strategy = bs;
  if (best_part) *best_part = bp;
  if (best_us) *best_us = best;
  return 0;
}

// ---- allocation-free loop entry point --------------------------------------
//
// op2_par_loop_noalloc runs a loop whose only device scratch is the
// reduction partials (airfoil and jac1 update) or none (save_soln, adt_calc)
// without allocating device memory: the partials live in `scratch`, a
// file-scope device array every GPU holds once, whose address is taken with
// hipGetSymbolAddress on the first call on each device (the name itself is
// only a host-side handle).  Maximum: 2 reduction results of kReduceBlocks
// workgroup partials each, plus a 256-byte result slot; the partials are
// fully written by each call (nothing needs clearing).  Calls that overlap
// on one GPU share the array and race.

namespace {

constexpr int kNoAllocSlots = 2;
constexpr size_t kScratchDoubles = 256 / sizeof(double) + (size_t)kNoAllocSlots * kReduceBlocks;
constexpr int kMaxDevices = 64;

__device__ __align__(256) double scratch[kScratchDoubles];

double *scratch_of_device[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up once).
ReduceScratch noalloc_scratch() {
  int dev = 0;
  OP2_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) fail("op2_par_loop_noalloc: device %d out of range", dev);
  if (!scratch_of_device[dev]) {
    void *p = nullptr;
    OP2_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    scratch_of_device[dev] = (double *)p;
  }
  double *base = scratch_of_device[dev];
  return {base + 256 / sizeof(double), base};
}

}  // namespace

int par_loop_noalloc(const char *name, op_set set, int nargs, op_arg *args) {
  Runtime &r = rt();
  if (!name || !set || nargs <= 0 || !args) fail("op2_par_loop_noalloc: invalid arguments");
  if (!r.device_ready)
    fail("op2_par_loop_noalloc: the runtime has no device yet (call op_init / op2_init)");
  if (r.capturing) fail("op2_par_loop_noalloc: not during graph capture");
  if (OP_diags > 0) {
    int ninds = 0;
    for (int m = 0; m < nargs; m++) op_arg_check(set, m, args[m], &ninds, name);
  }
  for (int m = 0; m < nargs; m++) {
    if (args[m].argtype != OP_ARG_DAT) continue;
    if (!args[m].dat || !args[m].dat->data_d)
      fail("op2_par_loop_noalloc: arg %d of %s has no device data", m, name);
    if (args[m].map && !args[m].map->map_d)
      fail("op2_par_loop_noalloc: map %s of loop %s is not on the device yet (run the loop "
           "once through op2_par_loop)", args[m].map->name, name);
  }
  if (r.loops.count(name))
    fail("op2_par_loop_noalloc: loop %s runs on the generic path, which allocates", name);
  const Loop l = identify(name, set, nargs, args);
  if (l != Loop::kUpdate && l != Loop::kJacUpdate && l != Loop::kSaveSoln &&
      l != Loop::kAdtCalc)
    fail("op2_par_loop_noalloc: loop %s needs plans or buffers the runtime allocates (only "
         "save_soln, adt_calc and the airfoil and jac1 update loops run without)", name);
  const int rb = real_bytes(nargs, args);
  if (rb != 4 && rb != 8)
    fail("op2_par_loop_noalloc: loop \"%s\" needs all its real dats to be double or all float",
         name);
  ReduceScratch s = noalloc_scratch();
  if (l == Loop::kSaveSoln || l == Loop::kAdtCalc) {
    if (rb == 8)
      dispatch<double>(l, name, set, nargs, args);
    else
      dispatch<float>(l, name, set, nargs, args);
    check_launch(name);
    note_writes(nargs, args);
    return 0;
  }
  note_writes(nargs, args);
  const double *d = with_block(block_of(name), [&](auto B) {
    constexpr int BS = decltype(B)::value;
    if (l == Loop::kUpdate)
      return rb == 8 ? run_update<BS, double>(set, args, nullptr, nullptr, nullptr, 0, false, &s)
                     : run_update<BS, float>(set, args, nullptr, nullptr, nullptr, 0, false, &s);
    return rb == 8 ? run_jac_update<BS, double>(set, args, &s)
                   : run_jac_update<BS, float>(set, args, &s);
  });
  if (r.deferred) return 0;
  if (l == Loop::kUpdate) {
    gbl_inc(args[4], fetch_results(d, 1)[0]);
  } else {
    const double *h = fetch_results(d, 2);
    gbl_inc(args[3], h[0]);
    gbl_max(args[4], h[1]);
  }
  return 0;
}

}  // namespace op2

using namespace op2;

extern "C" int op2_par_loop_noalloc(char const *name, op_set set, int nargs, op_arg *args) {
  try {
    return par_loop_noalloc(name, set, nargs, args);
  } catch (const op2::Error &) {
    return -1;
  }
}
