"""Execution plans of the runtime, built and verified on the host.

The runtime's own verifier must report zero violations, and an independent
numpy check here must agree: positions are a permutation of each block's
elements, no two same-colour elements of a block share an OP_INC target, and
no two same-colour blocks share one.
"""

import ctypes as C

import numpy as np
import pytest

from oplus2_matlab_amd import mesh as M
from oplus2_matlab_amd import op2


@pytest.fixture
def host_runtime():
    L = op2.lib()
    argv = (C.c_char_p * 1)(b"t")
    L.op_init_core(1, argv, 0)
    yield
    op2.op_exit()


def declare(mesh):
    nodes = op2.op_decl_set(mesh.nnode, "nodes")
    edges = op2.op_decl_set(mesh.nedge, "edges")
    bedges = op2.op_decl_set(mesh.nbedge, "bedges")
    cells = op2.op_decl_set(mesh.ncell, "cells")
    pedge = op2.op_decl_map(edges, nodes, 2, mesh.edge, "pedge")
    pecell = op2.op_decl_map(edges, cells, 2, mesh.ecell, "pecell")
    pbedge = op2.op_decl_map(bedges, nodes, 2, mesh.bedge, "pbedge")
    pbecell = op2.op_decl_map(bedges, cells, 1, mesh.becell, "pbecell")
    pcell = op2.op_decl_map(cells, nodes, 4, mesh.cell, "pcell")
    x = op2.op_decl_dat_host(nodes, 2, "double", mesh.x, "x")
    q = op2.op_decl_dat_host(cells, 4, "double", np.zeros(4 * mesh.ncell), "q")
    adt = op2.op_decl_dat_host(cells, 1, "double", np.zeros(mesh.ncell), "adt")
    res = op2.op_decl_dat_host(cells, 4, "double", np.zeros(4 * mesh.ncell), "res")
    bound = op2.op_decl_dat_host(bedges, 1, "int", mesh.bound, "bound")
    R, I = op2.OP_READ, op2.OP_INC
    res_args = [op2.op_arg_dat(x, 0, pedge, 2, "double", R),
                op2.op_arg_dat(x, 1, pedge, 2, "double", R),
                op2.op_arg_dat(q, 0, pecell, 4, "double", R),
                op2.op_arg_dat(q, 1, pecell, 4, "double", R),
                op2.op_arg_dat(adt, 0, pecell, 1, "double", R),
                op2.op_arg_dat(adt, 1, pecell, 1, "double", R),
                op2.op_arg_dat(res, 0, pecell, 4, "double", I),
                op2.op_arg_dat(res, 1, pecell, 4, "double", I)]
    bres_args = [op2.op_arg_dat(x, 0, pbedge, 2, "double", R),
                 op2.op_arg_dat(x, 1, pbedge, 2, "double", R),
                 op2.op_arg_dat(q, 0, pbecell, 4, "double", R),
                 op2.op_arg_dat(adt, 0, pbecell, 1, "double", R),
                 op2.op_arg_dat(res, 0, pbecell, 4, "double", I),
                 op2.op_arg_dat(bound, -1, None, 1, "int", R)]
    return dict(edges=edges, bedges=bedges, nodes=nodes, cells=cells, pedge=pedge, pecell=pecell,
                pbedge=pbedge, pbecell=pbecell, pcell=pcell, res_args=res_args,
                bres_args=bres_args)


def check_plan(exp, targets, coloured):
    """targets: (n, k) INC targets of each element."""
    perm, thr, off, col = exp["perm"], exp["thrcol"], exp["blk_off"], exp["blkcol"]
    n = perm.size
    assert off[0] == 0 and off[-1] == n
    for b in range(off.size - 1):
        lo, hi = off[b], off[b + 1]
        assert np.array_equal(np.sort(perm[lo:hi]), np.arange(lo, hi))
    if not coloured:
        return
    blk_of = np.repeat(np.arange(off.size - 1), np.diff(off))
    k = targets.shape[1]
    # thread colours: (block, colour, target) must be owned by one element
    e = perm
    key = np.stack([np.repeat(blk_of, k), np.repeat(thr, k), targets[e].reshape(-1),
                    np.repeat(e, k)], axis=1)
    key = np.unique(key, axis=0)  # one element touching a target twice is fine
    _, cnt = np.unique(key[:, :3], axis=0, return_counts=True)
    assert cnt.max() == 1, "two elements of one thread colour share a target"
    # block colours: (colour, target) owned by one block
    bkey = np.unique(np.stack([np.repeat(col[blk_of], k), targets[e].reshape(-1),
                               np.repeat(blk_of, k)], axis=1), axis=0)
    _, cnt = np.unique(bkey[:, :2], axis=0, return_counts=True)
    assert cnt.max() == 1, "two blocks of one colour share a target"


@pytest.mark.parametrize("ordering", ["cartesian", "random", "rcm"])
@pytest.mark.parametrize("strategy", [op2.STRATEGY_COLOURED, op2.STRATEGY_ATOMIC])
@pytest.mark.parametrize("part", [64, 100, 256, 1000])
def test_res_calc_plans(host_runtime, ordering, strategy, part):
    mesh = M.o_mesh(60, 30, ordering, seed=2)
    d = declare(mesh)
    s = op2.plan_summary("res_calc", d["edges"], d["res_args"], strategy, part)
    assert s["verify_errors"] == 0
    assert s["nblocks"] == -(-mesh.nedge // part)
    exp = op2.plan_export("res_calc", d["edges"], d["res_args"], strategy, part)
    check_plan(exp, mesh.ecell, strategy == op2.STRATEGY_COLOURED)
    if strategy == op2.STRATEGY_COLOURED:
        # every cell has four edges: blocks that hold several of them need
        # several thread colours, and neighbouring blocks need block colours
        assert s["mean_thread_colours"] >= (2 if ordering != "random" else 1)
        assert s["ncolors"] >= 2


@pytest.mark.parametrize("strategy", [op2.STRATEGY_COLOURED, op2.STRATEGY_ATOMIC])
def test_bres_calc_plan(host_runtime, strategy):
    mesh = M.o_mesh(60, 1, "random", seed=4)  # one ring: cells carry two bedges
    d = declare(mesh)
    s = op2.plan_summary("bres_calc", d["bedges"], d["bres_args"], strategy, 32)
    assert s["verify_errors"] == 0
    exp = op2.plan_export("bres_calc", d["bedges"], d["bres_args"], strategy, 32)
    check_plan(exp, mesh.becell.reshape(-1, 1), strategy == op2.STRATEGY_COLOURED)


def test_single_element_and_small_sets(host_runtime):
    a = op2.op_decl_set(1, "one")
    b = op2.op_decl_set(1, "tgt")
    m = op2.op_decl_map(a, b, 2, np.array([0, 0], np.int32), "m")
    u = op2.op_decl_dat_host(b, 1, "double", np.zeros(1), "u")
    du = op2.op_decl_dat_host(b, 1, "double", np.zeros(1), "du")
    A = op2.op_decl_dat_host(a, 1, "double", np.zeros(1), "A")
    beta = np.zeros(1)
    args = [op2.op_arg_dat(A, -1, None, 1, "double", op2.OP_READ),
            op2.op_arg_dat(u, 1, m, 1, "double", op2.OP_READ),
            op2.op_arg_dat(du, 0, m, 1, "double", op2.OP_INC),
            op2.op_arg_gbl(beta, 1, "double", op2.OP_READ)]
    s = op2.plan_summary("res", a, args, op2.STRATEGY_COLOURED, 64)
    assert s["nblocks"] == 1 and s["verify_errors"] == 0 and s["ncolors"] == 1


def test_plan_is_cached(host_runtime):
    mesh = M.o_mesh(12, 6, "random", seed=1)
    d = declare(mesh)
    op2.plan_summary("res_calc", d["edges"], d["res_args"], op2.STRATEGY_COLOURED, 64)
    n0 = op2.plan_summary("res_calc", d["edges"], d["res_args"], op2.STRATEGY_COLOURED,
                          64)["cached_plans"]
    s = op2.plan_summary("res_calc", d["edges"], d["res_args"], op2.STRATEGY_COLOURED, 64)
    assert s["cached_plans"] == n0
    s = op2.plan_summary("res_calc", d["edges"], d["res_args"], op2.STRATEGY_ATOMIC, 64)
    assert s["cached_plans"] == n0 + 1


@pytest.mark.parametrize("ordering", ["cartesian", "random", "rcm"])
@pytest.mark.parametrize("tile", [32, 100, 256])
def test_owner_plans(host_runtime, ordering, tile):
    """Owner-compute plans: the runtime's verifier checks that every
    increment is applied once, by the owner of its target."""
    mesh = M.o_mesh(60, 30, ordering, seed=2)
    d = declare(mesh)
    s = op2.plan_summary("res_calc", d["edges"], d["res_args"], op2.STRATEGY_OWNER, tile)
    assert s["verify_errors"] == 0
    assert s["max_incs"] == 4  # quad cells: packed 4-entry records
    # every cell is owned once: about ncell / tile tiles, each within the limit
    assert s["nblocks"] >= mesh.ncell // tile
    assert s["max_block"] <= 3 * tile + 8
    b = op2.plan_summary("bres_calc", d["bedges"], d["bres_args"], op2.STRATEGY_OWNER, tile)
    assert b["verify_errors"] == 0


def test_owner_tiles_are_compact(host_runtime):
    """Breadth-first tiles on a cartesian mesh cut few edges: elements per
    tile stay well below the 3x of a one-row strip."""
    mesh = M.o_mesh(240, 120, "cartesian")
    d = declare(mesh)
    s = op2.plan_summary("res_calc", d["edges"], d["res_args"], op2.STRATEGY_OWNER, 256)
    assert s["verify_errors"] == 0
    assert s["max_block"] < 2.5 * 256


def test_short_position_records_where_tiles_fit_two_windows(host_runtime):
    """Gather tiles store 16-bit (window, offset) records when each map's
    targets in the tile fit two windows of 32768 ids, and keep 32-bit ids
    otherwise: a random numbering of 320k cells scatters every tile's
    neighbours over the whole set (only the few-cell tiles left over by
    breadth-first growth fit), the renumbered mesh keeps them close.  The
    verifier decodes every short record against the map.  (A layout of this
    runtime: the reference's plans keep 32-bit maps, common/op_lib_core.c:
    627-1089.)"""
    mesh = M.o_mesh(800, 400, "random", seed=5)
    d = declare(mesh)
    far = op2.plan_summary("res_calc", d["edges"], d["res_args"], op2.STRATEGY_OWNER_GATHER, 256)
    assert far["verify_errors"] == 0 and far["short_ids"] < 0.25 * far["nblocks"]
    op2.renumber(d["pecell"])
    near = op2.plan_summary("res_calc", d["edges"], d["res_args"],
                            op2.STRATEGY_OWNER_GATHER, 256)
    assert near["verify_errors"] == 0
    assert 0.9 * near["nblocks"] <= near["short_ids"] <= near["nblocks"]
    # other strategies keep no position ids
    col = op2.plan_summary("res_calc", d["edges"], d["res_args"], op2.STRATEGY_COLOURED, 256)
    assert col["short_ids"] == 0


# ---------------------------------------------------------------------------
# renumbering (host side: permutations, internal maps, compact owner tiles)

@pytest.mark.parametrize("ordering", ["random", "cartesian"])
def test_renumber_permutations_and_maps(host_runtime, ordering):
    """Every set reached from the cells is permuted (a bijection), the
    internal maps are the user's maps under the permutations, and owner
    tiles of the renumbered cells are compact whatever the user ordering."""
    mesh = M.o_mesh(64, 32, ordering, seed=3)
    d = declare(mesh)
    before = op2.plan_summary("res_calc", d["edges"], d["res_args"], op2.STRATEGY_OWNER, 128)
    op2.renumber(d["pecell"])
    perms = {}
    for name in ("cells", "edges", "nodes", "bedges"):
        p = op2.renumber_export(d[name])
        assert p is not None and np.array_equal(np.sort(p), np.arange(d[name].size))
        perms[name] = p
    for mname, frm, to in (("pecell", "edges", "cells"), ("pedge", "edges", "nodes"),
                           ("pcell", "cells", "nodes"), ("pbecell", "bedges", "cells")):
        m = d[mname]
        core = C.cast(m._p, C.POINTER(op2._MapCore)).contents
        n = d[frm].size
        internal = np.ctypeslib.as_array(C.cast(core.map_i, C.POINTER(C.c_int)), (n * m.dim,))
        user = getattr(mesh, {"pecell": "ecell", "pedge": "edge", "pcell": "cell",
                              "pbecell": "becell"}[mname]).reshape(n, -1)
        want = np.empty_like(user)
        want[perms[frm]] = perms[to][user]
        assert np.array_equal(internal.reshape(n, -1), want), mname
    s = op2.plan_summary("res_calc", d["edges"], d["res_args"], op2.STRATEGY_OWNER, 128)
    assert s["verify_errors"] == 0 and s["max_incs"] == 4
    # aligned runs of 128 renumbered cells: full tiles with few cut edges
    # (an 11 x 11 square of cells touches 2.2 x 128 edges)
    assert s["nblocks"] == -(-mesh.ncell // 128) <= before["nblocks"]
    assert s["max_block"] < 2.6 * 128
    g = op2.plan_summary("res_calc", d["edges"], d["res_args"], op2.STRATEGY_OWNER_GATHER, 256)
    assert g["verify_errors"] == 0
    c = op2.plan_summary("res_calc", d["edges"], d["res_args"], op2.STRATEGY_COLOURED, 256)
    assert c["verify_errors"] == 0


def test_renumber_refusals(host_runtime):
    mesh = M.o_mesh(16, 8, "random", seed=1)
    d = declare(mesh)
    op2.renumber(d["pecell"])
    with pytest.raises(op2.OP2Error, match="already renumbered"):
        op2.renumber(d["pecell"])


def test_renumber_keeps_components_in_order(host_runtime):
    """Disconnected meshes (a batch) stay contiguous and in their order, so
    a cell -> mesh map stays non-decreasing."""
    b = M.mesh_batch(nmesh=4, seed=2, cells_lo=200, cells_hi=900, ordering="random")
    d = declare(b.union)
    op2.renumber(d["pecell"])
    p = op2.renumber_export(d["cells"])
    new_mesh = np.empty(b.union.ncell, np.int64)
    new_mesh[p] = b.cmesh
    assert np.all(np.diff(new_mesh) >= 0)


def test_generic_plan_colours_every_indirect_write(host_runtime):
    """The generic loop path's plan colours on OP_RW / OP_WRITE / OP_MIN /
    OP_MAX as well as OP_INC, and takes more than four (map, index) pairs;
    the exact verifier checks both colourings."""
    mesh = M.o_mesh(40, 20, "random", seed=7)
    d = declare(mesh)
    ng = 37
    groups = op2.op_decl_set(ng, "groups")
    rng = np.random.Generator(np.random.PCG64(1))
    cg = np.ascontiguousarray(rng.integers(0, ng, (mesh.ncell, 2)).astype(np.int32))
    pcg = op2.op_decl_map(d["cells"], groups, 2, cg, "pcg")
    v = op2.op_decl_dat_host(d["nodes"], 1, "double", np.zeros(mesh.nnode), "v")
    g = op2.op_decl_dat_host(groups, 1, "int", np.zeros(ng, np.int32), "g")
    c = op2.op_decl_dat_host(d["cells"], 1, "float", np.zeros(mesh.ncell, np.float32), "c")
    args = [op2.op_arg_dat(c, -1, None, 1, "float", op2.OP_READ)]
    args += [op2.op_arg_dat(v, k, d["pcell"], 1, "double", op2.OP_RW) for k in range(4)]
    args += [op2.op_arg_dat(g, 0, pcg, 1, "int", op2.OP_MAX),
             op2.op_arg_dat(g, 1, pcg, 1, "int", op2.OP_MIN)]
    s = op2.plan_summary("gen_cell", d["cells"], args, op2.PLAN_GENERIC, 128)
    assert s["verify_errors"] == 0
    assert s["ncolors"] > 1 and s["mean_thread_colours"] > 1
    # the same loop on the device kernels' strategies is refused (six pairs)
    with pytest.raises(op2.OP2Error, match="at most 4"):
        op2.plan_summary("gen_cell", d["cells"], args, op2.STRATEGY_COLOURED, 128)


# ---------------------------------------------------------------------------
# op_plan_core / op_plan_get: the plan in the reference's layout

def _view_args(mesh, prob, rb):
    """(name, set, args, inds) of the reference's four indirect loops, with
    the inds of airfoil/<loop>_kernel.cpp and jac1/res_kernel.cpp, on dats
    of rb-byte reals."""
    t = "double" if rb == 8 else "float"
    real = np.float64 if rb == 8 else np.float32
    d = declare(mesh)
    x = op2.op_decl_dat_host(d["nodes"], 2, t, mesh.x.astype(real), "x")
    q = op2.op_decl_dat_host(d["cells"], 4, t, np.zeros(4 * mesh.ncell, real), "q")
    adt = op2.op_decl_dat_host(d["cells"], 1, t, np.zeros(mesh.ncell, real), "adt")
    res = op2.op_decl_dat_host(d["cells"], 4, t, np.zeros(4 * mesh.ncell, real), "res")
    bound = op2.op_decl_dat_host(d["bedges"], 1, "int", mesh.bound, "bound")
    R, W, I = op2.OP_READ, op2.OP_WRITE, op2.OP_INC
    A = op2.op_arg_dat
    adt_args = [A(x, k, d["pcell"], 2, t, R) for k in range(4)] + [
        A(q, -1, None, 4, t, R), A(adt, -1, None, 1, t, W)]
    res_args = [A(x, 0, d["pedge"], 2, t, R), A(x, 1, d["pedge"], 2, t, R),
                A(q, 0, d["pecell"], 4, t, R), A(q, 1, d["pecell"], 4, t, R),
                A(adt, 0, d["pecell"], 1, t, R), A(adt, 1, d["pecell"], 1, t, R),
                A(res, 0, d["pecell"], 4, t, I), A(res, 1, d["pecell"], 4, t, I)]
    bres_args = [A(x, 0, d["pbedge"], 2, t, R), A(x, 1, d["pbedge"], 2, t, R),
                 A(q, 0, d["pbecell"], 4, t, R), A(adt, 0, d["pbecell"], 1, t, R),
                 A(res, 0, d["pbecell"], 4, t, I), A(bound, -1, None, 1, "int", R)]
    jn = op2.op_decl_set(prob.nnode, "jnodes")
    je = op2.op_decl_set(prob.nedge, "jedges")
    pp = op2.op_decl_map(je, jn, 2, prob.pp, "pp")
    jA = op2.op_decl_dat_host(je, 1, t, prob.A.astype(real), "A")
    u = op2.op_decl_dat_host(jn, 1, t, np.zeros(prob.nnode, real), "u")
    du = op2.op_decl_dat_host(jn, 1, t, np.zeros(prob.nnode, real), "du")
    beta = np.zeros(1, real)
    jac_args = [A(jA, -1, None, 1, t, R), A(u, 1, pp, 1, t, R), A(du, 0, pp, 1, t, I),
                op2.op_arg_gbl(beta, 1, t, R)]
    return [("adt_calc", d["cells"], adt_args, [0, 0, 0, 0, -1, -1]),
            ("res_calc", d["edges"], res_args, [0, 0, 1, 1, 2, 2, 3, 3]),
            ("bres_calc", d["bedges"], bres_args, [0, 0, 1, 2, 3, -1]),
            ("jac_res", je, jac_args, [-1, 0, 1, -1])]


@pytest.mark.parametrize("rb", [4, 8])
def test_op_plan_get_matches_reference_plans(host_runtime, golden_dir, rb):
    """op_plan_core through the C ABI returns, for adt_calc, res_calc,
    bres_calc and jac1 res (part sizes 0 and 64), the plan the reference's
    own op_plan_core built for the same mesh (tests/golden fixtures from the
    reference build): every array and the float byte model bit for bit, and
    the oracle's restated op_plan_check finds no error in it."""
    import os

    from oracle import oracle as O

    fx = np.load(os.path.join(golden_dir, "plans_24x12_random_s1.npz"))
    mesh = M.o_mesh(24, 12, "random", seed=1)
    prob = M.jacobi_grid(16, seed=1)
    for name, set_, args, inds in _view_args(mesh, prob, rb):
        for part in (0, 64):
            pre = f"r{rb}_p{part}_{name}_"
            v = op2.op_plan_get(name, set_, args, inds, part)
            nb, nc, nsh = fx[pre + "header"][:3]
            assert (v["nblocks"], v["ncolors"], v["nshared"]) == (nb, nc, nsh), pre
            assert np.array_equal(np.array([v["transfer"], v["transfer2"]], np.float32),
                                  fx[pre + "transfer"]), pre
            for k in ("nthrcol", "thrcol", "offset", "nelems", "ncolblk", "blkmap", "ind_offs",
                      "ind_sizes", "nindirect"):
                assert np.array_equal(v[k], fx[pre + k]), pre + k
            ninds = max(inds) + 1
            for m in range(ninds):
                assert np.array_equal(v["ind_maps"][m], fx[pre + f"ind_map{m}"]), pre
            for a, lm in enumerate(v["loc_maps"]):
                assert (lm is None) == (inds[a] < 0)
                if lm is not None:
                    assert np.array_equal(lm, fx[pre + f"loc_map{a}"]), pre + str(a)
            assert list(v["idxs"]) == [a.idx for a in args]
            assert list(v["accs"]) == [a.acc for a in args]
            if name == "jac_res":
                _, oargs, _ = O.jacobi_plan_args(prob, rb)
            else:
                _, oargs, _ = O.airfoil_plan_args(mesh, name, rb)
            assert O.check_plan(set_.size, oargs, ninds, v) == 0, pre


def test_op_plan_get_cache_and_globals(host_runtime):
    """A second call with the same inputs returns the same plan (count 2);
    OP_plans / OP_plan_index list the views; invalid inds are an error."""
    mesh = M.o_mesh(12, 6, "random", seed=1)
    views = _view_args(mesh, M.jacobi_grid(8, seed=1), 8)
    name, set_, args, inds = views[1]
    L = op2.lib()
    p1 = op2.op_plan_get(name, set_, args, inds, 64)
    p2 = op2.op_plan_get(name, set_, args, inds, 64)
    assert p2["count"] == 2 and p1["count"] == 1
    assert C.c_int.in_dll(L, "OP_plan_index").value == 1
    op2.op_plan_get(name, set_, args, inds, 32)
    assert C.c_int.in_dll(L, "OP_plan_index").value == 2
    plans = C.POINTER(op2.OpPlan).in_dll(L, "OP_plans")
    assert plans[0].nblocks == p1["nblocks"] and plans[0].count == 2
    assert plans[1].part_size == 32
    bad = list(inds)
    bad[2] = -2
    with pytest.raises(op2.OP2Error, match="out of range"):
        op2.op_plan_get(name, set_, args, bad, 64)
