/*
 * Dumps the reference execution plans (common/op_lib_core.c:627-1089) of the
 * three indirect airfoil loops and the Jacobi res loop on a given mesh.
 * Test infrastructure only.
 *
 * Usage: harness_plan mesh.bin jacobi.bin real_bytes out.bin [OP_PART_SIZE=n]
 * Output: per plan, int32 header {nblocks, ncolors, nshared, nargs, ninds,
 * set_size}, float {transfer, transfer2}, then nthrcol, thrcol, offset,
 * nelems, ncolblk, blkmap, ind_offs, ind_sizes, nindirect, every ind_map and
 * every loc_map (int16, only for indirect args).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "op_lib.h"
#include "harness_common.h"

static void dump(FILE *f, op_plan *p) {
  int h[6] = {p->nblocks, p->ncolors, p->nshared, p->nargs, p->ninds, p->set->size};
  fwrite(h, sizeof(int), 6, f);
  float t[2] = {p->transfer, p->transfer2};
  fwrite(t, sizeof(float), 2, f);
  fwrite(p->nthrcol, sizeof(int), p->nblocks, f);
  fwrite(p->thrcol, sizeof(int), p->set->size, f);
  fwrite(p->offset, sizeof(int), p->nblocks, f);
  fwrite(p->nelems, sizeof(int), p->nblocks, f);
  fwrite(p->ncolblk, sizeof(int), p->ncolors, f);
  fwrite(p->blkmap, sizeof(int), p->nblocks, f);
  fwrite(p->ind_offs, sizeof(int), p->nblocks * p->ninds, f);
  fwrite(p->ind_sizes, sizeof(int), p->nblocks * p->ninds, f);
  fwrite(p->nindirect, sizeof(int), p->ninds, f);
  for (int m = 0; m < p->ninds; m++) fwrite(p->ind_maps[m], sizeof(int), p->nindirect[m], f);
  for (int a = 0; a < p->nargs; a++)
    if (p->loc_maps[a]) fwrite(p->loc_maps[a], sizeof(short), p->set->size, f);
}

int main(int argc, char **argv) {
  if (argc < 5) {
    printf("usage: %s mesh.bin jacobi.bin real_bytes out.bin [OP_PART_SIZE=n]\n", argv[0]);
    return 2;
  }
  FILE *fm = h_open(argv[1], "rb");
  int *hdr = (int *)h_read(fm, 8 * sizeof(int));
  const int nnode = hdr[2], ncell = hdr[3], nedge = hdr[4], nbedge = hdr[5];
  free(h_read(fm, sizeof(double) * 2 * nnode));
  int *cell = (int *)h_read(fm, sizeof(int) * 4 * ncell);
  int *edge = (int *)h_read(fm, sizeof(int) * 2 * nedge);
  int *ecell = (int *)h_read(fm, sizeof(int) * 2 * nedge);
  int *bedge = (int *)h_read(fm, sizeof(int) * 2 * nbedge);
  int *becell = (int *)h_read(fm, sizeof(int) * nbedge);
  int *bound = (int *)h_read(fm, sizeof(int) * nbedge);
  fclose(fm);
  FILE *fj = h_open(argv[2], "rb");
  int *jh = (int *)h_read(fj, 4 * sizeof(int));
  const int jn = jh[2], je = jh[3];
  int *pp = (int *)h_read(fj, sizeof(int) * 2 * je);
  fclose(fj);
  const int rb = atoi(argv[3]);
  const char *tn = rb == 8 ? "double" : "float";

  op_init(argc - 4, argv + 4, 1);
  op_set nodes = op_decl_set(nnode, "nodes");
  op_set edges = op_decl_set(nedge, "edges");
  op_set bedges = op_decl_set(nbedge, "bedges");
  op_set cells = op_decl_set(ncell, "cells");
  op_map pedge = op_decl_map(edges, nodes, 2, edge, "pedge");
  op_map pecell = op_decl_map(edges, cells, 2, ecell, "pecell");
  op_map pbedge = op_decl_map(bedges, nodes, 2, bedge, "pbedge");
  op_map pbecell = op_decl_map(bedges, cells, 1, becell, "pbecell");
  op_map pcell = op_decl_map(cells, nodes, 4, cell, "pcell");
  /* plans only read the dats' sizes, so the data pointers are placeholders */
  char *dummy = (char *)calloc(1, 64);
  op_dat p_bound = op_decl_dat_core(bedges, 1, "int", sizeof(int), (char *)bound, "p_bound");
  op_dat p_x = op_decl_dat_core(nodes, 2, tn, rb, dummy, "p_x");
  op_dat p_q = op_decl_dat_core(cells, 4, tn, rb, dummy, "p_q");
  op_dat p_adt = op_decl_dat_core(cells, 1, tn, rb, dummy, "p_adt");
  op_dat p_res = op_decl_dat_core(cells, 4, tn, rb, dummy, "p_res");
  int part = OP_part_size;

  FILE *fo = h_open(argv[4], "wb");
  {
    op_arg a[6] = {op_arg_dat(p_x, 0, pcell, 2, tn, OP_READ),
                   op_arg_dat(p_x, 1, pcell, 2, tn, OP_READ),
                   op_arg_dat(p_x, 2, pcell, 2, tn, OP_READ),
                   op_arg_dat(p_x, 3, pcell, 2, tn, OP_READ),
                   op_arg_dat(p_q, -1, OP_ID, 4, tn, OP_READ),
                   op_arg_dat(p_adt, -1, OP_ID, 1, tn, OP_WRITE)};
    int inds[6] = {0, 0, 0, 0, -1, -1};
    dump(fo, op_plan_get("adt_calc", cells, part, 6, a, 1, inds));
  }
  {
    op_arg a[8] = {op_arg_dat(p_x, 0, pedge, 2, tn, OP_READ),
                   op_arg_dat(p_x, 1, pedge, 2, tn, OP_READ),
                   op_arg_dat(p_q, 0, pecell, 4, tn, OP_READ),
                   op_arg_dat(p_q, 1, pecell, 4, tn, OP_READ),
                   op_arg_dat(p_adt, 0, pecell, 1, tn, OP_READ),
                   op_arg_dat(p_adt, 1, pecell, 1, tn, OP_READ),
                   op_arg_dat(p_res, 0, pecell, 4, tn, OP_INC),
                   op_arg_dat(p_res, 1, pecell, 4, tn, OP_INC)};
    int inds[8] = {0, 0, 1, 1, 2, 2, 3, 3};
    dump(fo, op_plan_get("res_calc", edges, part, 8, a, 4, inds));
  }
  {
    op_arg a[6] = {op_arg_dat(p_x, 0, pbedge, 2, tn, OP_READ),
                   op_arg_dat(p_x, 1, pbedge, 2, tn, OP_READ),
                   op_arg_dat(p_q, 0, pbecell, 4, tn, OP_READ),
                   op_arg_dat(p_adt, 0, pbecell, 1, tn, OP_READ),
                   op_arg_dat(p_res, 0, pbecell, 4, tn, OP_INC),
                   op_arg_dat(p_bound, -1, OP_ID, 1, "int", OP_READ)};
    int inds[6] = {0, 0, 1, 2, 3, -1};
    dump(fo, op_plan_get("bres_calc", bedges, part, 6, a, 4, inds));
  }
  {
    op_set jnodes = op_decl_set(jn, "jnodes");
    op_set jedges = op_decl_set(je, "jedges");
    op_map ppedge = op_decl_map(jedges, jnodes, 2, pp, "ppedge");
    op_dat p_A = op_decl_dat_core(jedges, 1, tn, rb, dummy, "p_A");
    op_dat p_u = op_decl_dat_core(jnodes, 1, tn, rb, dummy, "p_u");
    op_dat p_du = op_decl_dat_core(jnodes, 1, tn, rb, dummy, "p_du");
    double beta = 0;
    op_arg a[4] = {op_arg_dat(p_A, -1, OP_ID, 1, tn, OP_READ),
                   op_arg_dat(p_u, 1, ppedge, 1, tn, OP_READ),
                   op_arg_dat(p_du, 0, ppedge, 1, tn, OP_INC),
                   op_arg_gbl_core((char *)&beta, 1, tn, OP_READ)};
    int inds[4] = {-1, 0, 1, -1};
    dump(fo, op_plan_get("res", jedges, part, 4, a, 2, inds));
  }
  fclose(fo);
  return 0;
}
