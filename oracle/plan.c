/*
 * Oracle restatement of the reference execution-plan builder and checker.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows common/op_lib_core.c:
 *   block size / count ............ :663-673
 *   per-block renumbering .......... :768-835  (sort + dedupe per indirect set)
 *   thread colouring ............... :839-880  (OP_INC args, 32-bit masks, ffs)
 *   block colouring ................ :892-939
 *   ncolblk / blkmap ............... :944-965
 *   nshared ........................ :972-985
 *   transfer model ................. :989-1056 (float accumulation, same order)
 *   plan checks .................... :469-612
 * The reference leaves its colouring check empty (:615-619); orc_plan_check
 * adds an exact one.
 */
#define _DEFAULT_SOURCE
#include "oracle.h"

#include <stdlib.h>
#include <string.h>
#include <strings.h>

#define OP2_ROUND_UP(b) (((b) + 15) & ~15)
#define CACHE_LINE 128 /* OP_cache_line_size default, op_lib_core.c:43 */

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

static inline int tgt(const orc_plan_in *in, int a, int e) {
  return in->map[a][in->idx[a] + e * in->map_dim[a]];
}

static int lead_arg(const orc_plan_in *in, int m) {
  for (int a = 0; a < in->nargs; a++)
    if (in->inds[a] == m) return a;
  return -1;
}

static int is_inc(const orc_plan_in *in, int a) {
  return in->inds[a] >= 0 && in->acc[a] == ORC_INC;
}

void orc_plan_free(orc_plan *p) {
  free(p->nthrcol); free(p->thrcol); free(p->offset); free(p->nelems);
  free(p->ncolblk); free(p->blkmap); free(p->blkcol);
  free(p->ind_offs); free(p->ind_sizes);
  for (int m = 0; m < ORC_MAXARGS; m++) { free(p->ind_maps[m]); free(p->loc_maps[m]); }
  memset(p, 0, sizeof *p);
}

int orc_plan_build(const orc_plan_in *in, orc_plan *p) {
  memset(p, 0, sizeof *p);
  const int n = in->set_size, nargs = in->nargs, ninds = in->ninds;
  if (n <= 0 || nargs <= 0 || nargs > ORC_MAXARGS || ninds > ORC_MAXARGS) return -1;

  int staged_bytes = 0;
  for (int a = 0; a < nargs; a++)
    if (in->inds[a] >= 0) staged_bytes += in->dat_size[a];
  int bsize = in->part_size;
  if (bsize == 0) {
    if (staged_bytes == 0) return -1;
    bsize = (48 * 1024 / (64 * staged_bytes)) * 64;
  }
  if (bsize <= 0) return -1;
  const int nblocks = (n - 1) / bsize + 1;

  p->bsize = bsize; p->nblocks = nblocks; p->ninds = ninds; p->nargs = nargs;
  p->set_size = n;
  p->nthrcol = malloc(sizeof(int) * nblocks);
  p->thrcol = malloc(sizeof(int) * n);
  p->offset = malloc(sizeof(int) * nblocks);
  p->nelems = malloc(sizeof(int) * nblocks);
  p->blkmap = malloc(sizeof(int) * nblocks);
  p->blkcol = malloc(sizeof(int) * nblocks);
  p->ind_offs = malloc(sizeof(int) * (size_t)nblocks * (ninds ? ninds : 1));
  p->ind_sizes = malloc(sizeof(int) * (size_t)nblocks * (ninds ? ninds : 1));

  /* per indirect set: local number of each target element (renumbering) and
   * colour bit masks (colouring) */
  int *local[ORC_MAXARGS];
  unsigned *mask[ORC_MAXARGS];
  for (int m = 0; m < ninds; m++) {
    int a = lead_arg(in, m);
    int cnt = 0;
    for (int a2 = 0; a2 < nargs; a2++) cnt += (in->inds[a2] == m);
    p->ind_maps[m] = malloc(sizeof(int) * (size_t)cnt * n);
    local[m] = malloc(sizeof(int) * in->to_size[a]);
    mask[m] = calloc(in->to_size[a], sizeof(unsigned));
  }
  for (int a = 0; a < nargs; a++)
    p->loc_maps[a] = in->inds[a] >= 0 ? malloc(sizeof(short) * n) : NULL;
  int *buf = malloc(sizeof(int) * (size_t)nargs * bsize);

  for (int b = 0; b < nblocks; b++) {
    const int lo = b * bsize;
    const int cnt = (n - lo < bsize) ? n - lo : bsize;
    p->offset[b] = lo;
    p->nelems[b] = cnt;

    /* local renumbering: sorted distinct targets of each indirect set */
    for (int m = 0; m < ninds; m++) {
      int ne = 0;
      for (int a = 0; a < nargs; a++)
        if (in->inds[a] == m)
          for (int e = lo; e < lo + cnt; e++) buf[ne++] = tgt(in, a, e);
      qsort(buf, ne, sizeof(int), cmp_int);
      int nu = 0;
      for (int k = 0; k < ne; k++)
        if (nu == 0 || buf[k] != buf[nu - 1]) buf[nu++] = buf[k];
      const int base = p->nindirect[m];
      for (int k = 0; k < nu; k++) {
        p->ind_maps[m][base + k] = buf[k];
        local[m][buf[k]] = k;
      }
      p->nindirect[m] = base + nu;
      for (int a = 0; a < nargs; a++)
        if (in->inds[a] == m)
          for (int e = lo; e < lo + cnt; e++)
            p->loc_maps[a][e] = (short)local[m][tgt(in, a, e)];
      p->ind_offs[m + b * ninds] = base;
      p->ind_sizes[m + b * ninds] = nu;
    }

    /* thread colours: greedy, 32 colours per pass */
    for (int e = lo; e < lo + cnt; e++) p->thrcol[e] = -1;
    int pass_base = 0, ncol = 0, pending = 1;
    while (pending) {
      pending = 0;
      for (int a = 0; a < nargs; a++)
        if (in->inds[a] >= 0)
          for (int e = lo; e < lo + cnt; e++) mask[in->inds[a]][tgt(in, a, e)] = 0;
      for (int e = lo; e < lo + cnt; e++) {
        if (p->thrcol[e] != -1) continue;
        unsigned used = 0;
        for (int a = 0; a < nargs; a++)
          if (is_inc(in, a)) used |= mask[in->inds[a]][tgt(in, a, e)];
        int c = ffs((int)~used) - 1;
        if (c < 0) { pending = 1; continue; }
        p->thrcol[e] = pass_base + c;
        if (pass_base + c + 1 > ncol) ncol = pass_base + c + 1;
        for (int a = 0; a < nargs; a++)
          if (is_inc(in, a)) mask[in->inds[a]][tgt(in, a, e)] |= 1u << c;
      }
      pass_base += 32;
    }
    p->nthrcol[b] = ncol;
  }

  /* block colours: greedy over blocks, 32 colours per pass */
  for (int b = 0; b < nblocks; b++) p->blkcol[b] = -1;
  int pass_base = 0, ncolors = 0, pending = 1;
  while (pending) {
    pending = 0;
    for (int m = 0; m < ninds; m++)
      memset(mask[m], 0, sizeof(unsigned) * in->to_size[lead_arg(in, m)]);
    for (int b = 0; b < nblocks; b++) {
      if (p->blkcol[b] != -1) continue;
      const int lo = p->offset[b], cnt = p->nelems[b];
      unsigned used = 0;
      for (int a = 0; a < nargs; a++)
        if (is_inc(in, a))
          for (int e = lo; e < lo + cnt; e++) used |= mask[in->inds[a]][tgt(in, a, e)];
      int c = ffs((int)~used) - 1;
      if (c < 0) { pending = 1; continue; }
      p->blkcol[b] = pass_base + c;
      if (pass_base + c + 1 > ncolors) ncolors = pass_base + c + 1;
      for (int a = 0; a < nargs; a++)
        if (is_inc(in, a))
          for (int e = lo; e < lo + cnt; e++) mask[in->inds[a]][tgt(in, a, e)] |= 1u << c;
    }
    pass_base += 32;
  }
  p->ncolors = ncolors;
  p->ncolblk = calloc(ncolors, sizeof(int));
  for (int b = 0; b < nblocks; b++) p->ncolblk[p->blkcol[b]]++;
  {
    int *start = malloc(sizeof(int) * ncolors);
    int s = 0;
    for (int c = 0; c < ncolors; c++) { start[c] = s; s += p->ncolblk[c]; }
    for (int b = 0; b < nblocks; b++) p->blkmap[start[p->blkcol[b]]++] = b;
    free(start);
  }

  /* staged bytes per block, maximum over blocks */
  p->nshared = 0;
  for (int b = 0; b < nblocks; b++) {
    int bytes = 0;
    for (int m = 0; m < ninds; m++)
      bytes += OP2_ROUND_UP(p->ind_sizes[m + b * ninds] * in->dat_size[lead_arg(in, m)]);
    if (bytes > p->nshared) p->nshared = bytes;
  }

  /* byte model, accumulated in float in the reference's order */
  float tr = 0, tr2 = 0, tr3 = 0;
  for (int b = 0; b < nblocks; b++) {
    for (int a = 0; a < nargs; a++) {
      if (in->inds[a] < 0) {
        float fac = (in->acc[a] == ORC_READ) ? 1.0f : 2.0f;
        if (in->dat_size[a] > 0) {
          float v = fac * p->nelems[b] * in->dat_size[a];
          tr += v; tr2 += v; tr3 += v;
        }
      } else {
        size_t v = p->nelems[b] * sizeof(short);
        tr += v; tr2 += v; tr3 += v;
      }
    }
    for (int m = 0; m < ninds; m++) {
      const int a = lead_arg(in, m);
      const int dsz = in->dat_size[a], ddim = in->dat_dim[a];
      const int isz = p->ind_sizes[m + b * ninds];
      float fac = (in->acc[a] == ORC_READ) ? 1.0f : 2.0f;
      tr += fac * isz * dsz;
      const int e0 = p->ind_offs[m + b * ninds], e1 = e0 + isz;
      int l_old = -1, l_new;
      for (int e = e0; e < e1; e++) {
        const int g = p->ind_maps[m][e];
        l_new = (g * dsz) / CACHE_LINE;
        if (l_new > l_old) tr2 += fac * CACHE_LINE;
        l_old = l_new;
        l_new = ((g + 1) * dsz - 1) / CACHE_LINE;
        tr2 += fac * (l_new - l_old) * CACHE_LINE;
        l_old = l_new;
      }
      l_old = -1;
      for (int e = e0; e < e1; e++) {
        const int g = p->ind_maps[m][e];
        l_new = (g * dsz) / (ddim * CACHE_LINE);
        if (l_new > l_old) tr3 += fac * ddim * CACHE_LINE;
        l_old = l_new;
        l_new = ((g + 1) * dsz - 1) / (ddim * CACHE_LINE);
        tr3 += fac * (l_new - l_old) * ddim * CACHE_LINE;
        l_old = l_new;
      }
      fac = (in->acc[a] == ORC_RW) ? 2.0f : 1.0f;
      tr += fac * isz * sizeof(int);
      tr2 += fac * isz * sizeof(int);
      tr3 += fac * isz * sizeof(int);
    }
  }
  p->transfer = tr; p->transfer2 = tr2; p->transfer3 = tr3;

  for (int m = 0; m < ninds; m++) { free(local[m]); free(mask[m]); }
  free(buf);
  return 0;
}

typedef struct { int set, target, colour, owner; } trip;

static int cmp_trip(const void *a, const void *b) {
  const trip *x = a, *y = b;
  if (x->set != y->set) return x->set < y->set ? -1 : 1;
  if (x->target != y->target) return x->target < y->target ? -1 : 1;
  if (x->colour != y->colour) return x->colour < y->colour ? -1 : 1;
  return (x->owner > y->owner) - (x->owner < y->owner);
}

static int count_clashes(trip *t, size_t k) {
  int err = 0;
  qsort(t, k, sizeof(trip), cmp_trip);
  for (size_t i = 1; i < k; i++)
    err += (t[i].set == t[i - 1].set && t[i].target == t[i - 1].target &&
            t[i].colour == t[i - 1].colour && t[i].owner != t[i - 1].owner);
  return err;
}

int orc_plan_check(const orc_plan_in *in, const orc_plan *p) {
  int err = 0;
  const int nb = p->nblocks, ninds = p->ninds;
  int tot = 0;
  for (int c = 0; c < p->ncolors; c++) tot += p->ncolblk[c];
  if (tot != nb) err++;
  /* sizes and offsets */
  int nel = 0;
  for (int b = 0; b < nb; b++) { err += (p->offset[b] != nel); nel += p->nelems[b]; }
  err += (nel != in->set_size);
  /* blkmap is a permutation, grouped by colour */
  {
    char *seen = calloc(nb, 1);
    int k = 0;
    for (int c = 0; c < p->ncolors; c++)
      for (int i = 0; i < p->ncolblk[c]; i++, k++) {
        int b = p->blkmap[k];
        if (b < 0 || b >= nb || seen[b]) { err++; continue; }
        seen[b] = 1;
        err += (p->blkcol[b] != c);
      }
    free(seen);
  }
  /* ind_offs / ind_sizes consistent; ind_maps strictly increasing, in range */
  for (int m = 0; m < ninds; m++) {
    const int a = lead_arg(in, m);
    int run = 0;
    for (int b = 0; b < nb; b++) {
      err += (p->ind_offs[m + b * ninds] != run);
      int last = -1;
      for (int e = run; e < run + p->ind_sizes[m + b * ninds]; e++) {
        err += (p->ind_maps[m][e] <= last);
        last = p->ind_maps[m][e];
      }
      err += (last >= in->to_size[a]);
      run += p->ind_sizes[m + b * ninds];
    }
    err += (run != p->nindirect[m]);
  }
  /* local -> staged -> global round trip reproduces the user's map */
  for (int a = 0; a < in->nargs; a++) {
    if (in->inds[a] < 0) continue;
    const int m = in->inds[a];
    for (int b = 0; b < nb; b++)
      for (int e = p->offset[b]; e < p->offset[b] + p->nelems[b]; e++) {
        int l = p->loc_maps[a][e];
        if (l < 0 || l >= p->ind_sizes[m + b * ninds]) { err++; continue; }
        err += (p->ind_maps[m][p->ind_offs[m + b * ninds] + l] != tgt(in, a, e));
      }
  }
  /* exact colouring check over OP_INC targets: sort (target, colour, owner)
   * triples and look for two different owners of one colour on one target */
  {
    size_t cap = 0;
    for (int a = 0; a < in->nargs; a++) cap += is_inc(in, a) ? (size_t)in->set_size : 0;
    trip *t = malloc(sizeof(trip) * (cap ? cap : 1));
    /* thread colours, block by block */
    for (int b = 0; b < nb; b++) {
      size_t k = 0;
      for (int e = p->offset[b]; e < p->offset[b] + p->nelems[b]; e++) {
        err += (p->thrcol[e] < 0 || p->thrcol[e] >= p->nthrcol[b]);
        for (int a = 0; a < in->nargs; a++)
          if (is_inc(in, a)) {
            t[k].set = in->inds[a]; t[k].target = tgt(in, a, e);
            t[k].colour = p->thrcol[e]; t[k].owner = e; k++;
          }
      }
      err += count_clashes(t, k);
    }
    /* block colours, over the whole set */
    size_t k = 0;
    for (int b = 0; b < nb; b++)
      for (int e = p->offset[b]; e < p->offset[b] + p->nelems[b]; e++)
        for (int a = 0; a < in->nargs; a++)
          if (is_inc(in, a)) {
            t[k].set = in->inds[a]; t[k].target = tgt(in, a, e);
            t[k].colour = p->blkcol[b]; t[k].owner = b; k++;
          }
    err += count_clashes(t, k);
    free(t);
  }
  return err;
}
