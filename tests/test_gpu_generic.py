"""GPU tests of the generic loop path (include/op_hip_loop.h): loops the
runtime has no built-in kernel for, compiled from device user functions in
tests/generic_loops.hip and registered by name, checked against the
sequential semantics of op_par_loop (common/op_seq.h) restated in numpy.
Exact where the result does not depend on the order of the elements
(writes, RW maxima, MIN/MAX reductions), within float rounding for sums."""

import ctypes as C
import os

import numpy as np
import pytest

from oplus2_matlab_amd import mesh as M
from oplus2_matlab_amd import op2
from oplus2_matlab_amd.op2 import (OP_ID, OP_INC, OP_MAX, OP_MIN, OP_READ, OP_RW, OP_WRITE,
                                   op_arg_dat, op_arg_gbl)

pytestmark = pytest.mark.gpu

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
LOOPS = os.path.join(ROOT, "oplus2_matlab_amd", "lib", "libop2_testloops.so")


def note(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def loops():
    if not os.path.exists(LOOPS):
        pytest.fail(f"{LOOPS} not built (make -C oplus2_matlab_amd/csrc testloops)")
    op2.lib()
    return C.CDLL(LOOPS, mode=C.RTLD_GLOBAL)  # registers gen_direct, gen_edge, gen_cell


@pytest.fixture
def runtime(loops):
    op2.op_init([], 0)
    yield
    op2.op_exit()


def test_direct_loop_every_global_access(runtime):
    note("generic direct loop: mixed types, dims 3/2, OP_GBL READ/INC/MAX/MIN")
    n = 100_003
    rng = np.random.Generator(np.random.PCG64(4))
    a = np.ascontiguousarray(rng.random((n, 3)))
    b = np.zeros((n, 2), np.float32)
    cells = op2.op_decl_set(n, "cells")
    pa = op2.op_decl_dat(cells, 3, "double", a, "pa")
    pb = op2.op_decl_dat(cells, 2, "float", b, "pb")
    op2.op_decl_const2("gen_scale", 1, "double", np.array([2.0]))
    s = np.array([1.5, -0.5])
    tot = np.array([0.25, 1.0], np.float32)
    imax = np.array([-5], np.int32)
    dmin = np.array([10.0])
    op2.op_par_loop(None, "gen_direct", cells,
                    op_arg_dat(pa, -1, OP_ID, 3, "double", OP_READ),
                    op_arg_dat(pb, -1, OP_ID, 2, "float", OP_WRITE),
                    op_arg_gbl(s, 2, "double", OP_READ),
                    op_arg_gbl(tot, 2, "float", OP_INC),
                    op_arg_gbl(imax, 1, "int", OP_MAX),
                    op_arg_gbl(dmin, 1, "double", OP_MIN))
    b0 = (s[0] * a[:, 0] + a[:, 1]).astype(np.float32)
    b1 = (s[1] * a[:, 2] * 2.0).astype(np.float32)
    got = op2.op_fetch_data(pb).reshape(n, 2)
    assert np.array_equal(got[:, 0], b0) and np.array_equal(got[:, 1], b1)
    want = np.array([0.25 + b0.astype(np.float64).sum(),
                     1.0 + (b1.astype(np.float64) ** 2).sum()])
    np.testing.assert_allclose(tot, want, rtol=2e-6)
    assert imax[0] == max(-5, int((a[:, 0] * 100.0).astype(np.int32).max()))
    assert dmin[0] == min(10.0, float((a[:, 1] - a[:, 2]).min()))


@pytest.mark.parametrize("renumber", [False, True])
def test_indirect_increments(runtime, renumber):
    note(f"generic edge loop: float dim-3 increments through both map indices, renumber={renumber}")
    mesh = M.o_mesh(120, 60, "random", seed=3)
    rng = np.random.Generator(np.random.PCG64(5))
    nodes = op2.op_decl_set(mesh.nnode, "nodes")
    edges = op2.op_decl_set(mesh.nedge, "edges")
    pedge = op2.op_decl_map(edges, nodes, 2, mesh.edge, "pedge")
    x = np.ascontiguousarray(mesh.x)
    nd = np.ascontiguousarray(rng.random((mesh.nnode, 3)).astype(np.float32))
    nd[:, 2] = 0.0  # counts the edges at each node: small integers, exact
    nd0 = nd.copy()
    w = np.ascontiguousarray(rng.random(mesh.nedge).astype(np.float32))
    ln = np.zeros(mesh.nedge)
    px = op2.op_decl_dat(nodes, 2, "double", x, "px")
    pn = op2.op_decl_dat(nodes, 3, "float", nd, "pn")
    pw = op2.op_decl_dat(edges, 1, "float", w, "pw")
    pl = op2.op_decl_dat(edges, 1, "double", ln, "pl")
    if renumber:
        op2.renumber(pedge)
    args = (op_arg_dat(px, 0, pedge, 2, "double", OP_READ),
            op_arg_dat(px, 1, pedge, 2, "double", OP_READ),
            op_arg_dat(pn, 0, pedge, 3, "float", OP_INC),
            op_arg_dat(pn, 1, pedge, 3, "float", OP_INC),
            op_arg_dat(pw, -1, OP_ID, 1, "float", OP_READ),
            op_arg_dat(pl, -1, OP_ID, 1, "double", OP_WRITE))
    for _ in range(3):
        op2.op_par_loop(None, "gen_edge", edges, *args)
    e1, e2 = mesh.edge[:, 0], mesh.edge[:, 1]
    dx = x[e2, 0] - x[e1, 0]
    dy = x[e2, 1] - x[e1, 1]
    want = nd0.astype(np.float64)
    for _ in range(3):
        np.add.at(want[:, 0], e1, w.astype(np.float64))
        np.add.at(want[:, 1], e1, dx.astype(np.float32).astype(np.float64))
        np.add.at(want[:, 2], e1, 1.0)
        np.add.at(want[:, 0], e2, -w.astype(np.float64))
        np.add.at(want[:, 1], e2, -dy.astype(np.float32).astype(np.float64))
        np.add.at(want[:, 2], e2, 1.0)
    got = op2.op_fetch_data(pn).reshape(-1, 3)
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-5)
    assert np.array_equal(got[:, 2], want[:, 2].astype(np.float32))
    np.testing.assert_allclose(op2.op_fetch_data(pl), np.sqrt(dx * dx + dy * dy), rtol=1e-15)


@pytest.mark.parametrize("renumber", [False, True])
def test_indirect_rw_min_max_six_map_slots(runtime, renumber):
    note(f"generic cell loop: OP_RW x4, OP_MAX/OP_MIN through a second map, renumber={renumber}")
    mesh = M.o_mesh(96, 48, "random", seed=8)
    rng = np.random.Generator(np.random.PCG64(6))
    ng = 101
    nodes = op2.op_decl_set(mesh.nnode, "nodes")
    cells = op2.op_decl_set(mesh.ncell, "cells")
    groups = op2.op_decl_set(ng, "groups")
    pcell = op2.op_decl_map(cells, nodes, 4, mesh.cell, "pcell")
    cg = np.ascontiguousarray(rng.integers(0, ng, (mesh.ncell, 2)).astype(np.int32))
    pcg = op2.op_decl_map(cells, groups, 2, cg, "pcg")
    c = np.ascontiguousarray((rng.random(mesh.ncell) * 50).astype(np.float32))
    v = np.ascontiguousarray(rng.random(mesh.nnode) * 40)
    v0 = v.copy()
    lo_init = np.full(ng, 300, np.int32)
    hi = np.zeros(ng, np.int32)
    lo = lo_init.copy()
    wc = np.zeros(mesh.ncell, np.float32)
    pc = op2.op_decl_dat(cells, 1, "float", c, "pc")
    pv = op2.op_decl_dat(nodes, 1, "double", v, "pv")
    phi = op2.op_decl_dat(groups, 1, "int", hi, "phi")
    plo = op2.op_decl_dat(groups, 1, "int", lo, "plo")
    pwc = op2.op_decl_dat(cells, 1, "float", wc, "pwc")
    if renumber:
        op2.renumber(pcell)
    op2.op_par_loop(None, "gen_cell", cells,
                    op_arg_dat(pc, -1, OP_ID, 1, "float", OP_READ),
                    *[op_arg_dat(pv, k, pcell, 1, "double", OP_RW) for k in range(4)],
                    op_arg_dat(phi, 0, pcg, 1, "int", OP_MAX),
                    op_arg_dat(plo, 1, pcg, 1, "int", OP_MIN),
                    op_arg_dat(pwc, -1, OP_ID, 1, "float", OP_WRITE))
    t = c.astype(np.float64)
    want_v = v0.copy()
    for k in range(4):
        np.maximum.at(want_v, mesh.cell[:, k], t + k)
    want_hi = np.zeros(ng, np.int32)
    want_lo = lo_init.copy()
    np.maximum.at(want_hi, cg[:, 0], (t * 10.0).astype(np.int32))
    np.minimum.at(want_lo, cg[:, 1], (t * 10.0).astype(np.int32))
    assert np.array_equal(op2.op_fetch_data(pv), want_v)
    assert np.array_equal(op2.op_fetch_data(phi), want_hi)
    assert np.array_equal(op2.op_fetch_data(plo), want_lo)
    assert np.array_equal(op2.op_fetch_data(pwc), c)


def test_generic_errors(runtime):
    note("generic loop errors: wrong dat type, unknown loop")
    cells = op2.op_decl_set(10, "cells")
    pa = op2.op_decl_dat(cells, 3, "float", np.zeros(30, np.float32), "pa")
    pb = op2.op_decl_dat(cells, 2, "float", np.zeros(20, np.float32), "pb")
    s = np.zeros(2)
    tot = np.zeros(2, np.float32)
    imax = np.zeros(1, np.int32)
    dmin = np.zeros(1)
    with pytest.raises(op2.OP2Error, match="type does not match"):
        op2.op_par_loop(None, "gen_direct", cells,
                        op_arg_dat(pa, -1, OP_ID, 3, "float", OP_READ),
                        op_arg_dat(pb, -1, OP_ID, 2, "float", OP_WRITE),
                        op_arg_gbl(s, 2, "double", OP_READ), op_arg_gbl(tot, 2, "float", OP_INC),
                        op_arg_gbl(imax, 1, "int", OP_MAX), op_arg_gbl(dmin, 1, "double", OP_MIN))
    with pytest.raises(op2.OP2Error, match="no registered generic loop"):
        op2.op_par_loop(None, "never_registered", cells,
                        op_arg_dat(pb, -1, OP_ID, 2, "float", OP_WRITE))


def test_reductions_in_an_indirect_loop(runtime):
    """OP_GBL INC / MAX inside a coloured (indirect-increment) loop."""
    note("generic edge loop with OP_GBL INC and MAX and int increments")
    mesh = M.o_mesh(64, 32, "random", seed=9)
    nodes = op2.op_decl_set(mesh.nnode, "nodes")
    edges = op2.op_decl_set(mesh.nedge, "edges")
    pedge = op2.op_decl_map(edges, nodes, 2, mesh.edge, "pedge")
    px = op2.op_decl_dat(nodes, 2, "double", np.ascontiguousarray(mesh.x), "px")
    cnt = np.zeros(mesh.nnode, np.int32)
    pc = op2.op_decl_dat(nodes, 1, "int", cnt, "pc")
    total = np.array([1.0])
    longest = np.array([0.0])
    op2.op_par_loop(None, "gen_edge_sum", edges,
                    op_arg_dat(px, 0, pedge, 2, "double", OP_READ),
                    op_arg_dat(px, 1, pedge, 2, "double", OP_READ),
                    op_arg_dat(pc, 0, pedge, 1, "int", OP_INC),
                    op_arg_dat(pc, 1, pedge, 1, "int", OP_INC),
                    op_arg_gbl(total, 1, "double", OP_INC),
                    op_arg_gbl(longest, 1, "double", OP_MAX))
    e1, e2 = mesh.edge[:, 0], mesh.edge[:, 1]
    ln = np.hypot(mesh.x[e2, 0] - mesh.x[e1, 0], mesh.x[e2, 1] - mesh.x[e1, 1])
    want_cnt = np.bincount(e1, minlength=mesh.nnode) + np.bincount(e2, minlength=mesh.nnode)
    assert np.array_equal(op2.op_fetch_data(pc), want_cnt.astype(np.int32))
    np.testing.assert_allclose(total[0], 1.0 + ln.sum(), rtol=1e-13)
    np.testing.assert_allclose(longest[0], ln.max(), rtol=1e-15)


@pytest.mark.parametrize("renumber", [False, True])
def test_generic_res_calc_equals_the_device_kernel(runtime, renumber):
    """airfoil/res_calc.h's arithmetic through the generic path gives the
    hand-written res_calc kernel's increments (to summation order)."""
    note(f"generic res_calc vs the device kernel, 120x60 random, renumber={renumber}")
    from oplus2_matlab_amd import airfoil as AF
    mesh = M.o_mesh(120, 60, "random", seed=4)
    c = M.flow_constants().as_array()
    rng = np.random.Generator(np.random.PCG64(12))
    q0 = np.tile(c[6:10], (mesh.ncell, 1)) * (1 + 0.02 * rng.standard_normal((mesh.ncell, 4)))
    a = AF.Airfoil(mesh, c, "double", q0=q0, renumber=renumber)
    a.adt_calc()
    a.res_calc()
    native = op2.op_fetch_data(a.p_res).reshape(-1, 4).copy()
    res2 = np.zeros(4 * mesh.ncell)
    pres2 = op2.op_decl_dat(a.cells, 4, "double", res2, "pres2")
    args = list(a.res_calc_args())
    args[6] = op_arg_dat(pres2, 0, a.pecell, 4, "double", OP_INC)
    args[7] = op_arg_dat(pres2, 1, a.pecell, 4, "double", OP_INC)
    op2.op_par_loop(None, "gen_res_calc", a.edges, *args)
    got = op2.op_fetch_data(pres2).reshape(-1, 4)
    assert np.max(np.abs(got - native)) <= 1e-13 * np.max(np.abs(native))


def test_generic_res_calc_on_owner_tiles_is_the_sequential_sum(runtime):
    """gen_res_calc has the owner-tile form (its written arguments are the
    two OP_INC res through pecell): each cell adds its edges' increments in
    edge order, so on the mesh's own numbering res is bit for bit the
    oracle's sequential res_calc (common/op_seq.h order)."""
    note("generic res_calc on owner tiles vs the sequential oracle, 60x30 random")
    from oplus2_matlab_amd import airfoil as AF
    from oracle import oracle as O
    mesh = M.o_mesh(60, 30, "random", seed=6)
    c = M.flow_constants().as_array()
    rng = np.random.Generator(np.random.PCG64(13))
    q0 = np.tile(c[6:10], (mesh.ncell, 1)) * (1 + 0.02 * rng.standard_normal((mesh.ncell, 4)))
    a = AF.Airfoil(mesh, c, "double", q0=q0)
    a.adt_calc()
    adt = op2.op_fetch_data(a.p_adt).copy()
    res0 = 1e-3 * rng.standard_normal(4 * mesh.ncell)
    pres = op2.op_decl_dat(a.cells, 4, "double", res0.copy(), "pres")
    args = list(a.res_calc_args())
    args[6] = op_arg_dat(pres, 0, a.pecell, 4, "double", OP_INC)
    args[7] = op_arg_dat(pres, 1, a.pecell, 4, "double", OP_INC)
    arr = (op2.OpArg * 8)(*args)
    info = (C.c_char * 128)()  # op2_generic_owner_plan
    arg_info = (C.c_int * 16)()
    assert op2.lib().op2_generic_owner_plan_get(b"gen_res_calc", a.edges._p, 8, arr, 0, info,
                                                arg_info) == 0
    # OP_INC pairs of res1, res2; target-row places of (pedge 0, 1), (pecell 0, 1)
    assert list(arg_info) == [-1, -1, -1, -1, -1, -1, 0, 1] + [0, 1, 2, 3, 2, 3, 2, 3]
    op2.op_par_loop(None, "gen_res_calc", a.edges, *args)
    want = res0.reshape(-1, 4).copy()
    O.res_calc("f64", mesh, np.ascontiguousarray(mesh.x), np.ascontiguousarray(q0), adt, want, c)
    got = op2.op_fetch_data(pres).reshape(-1, 4)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("renumber,part", [(False, 64), (False, 256), (True, 128)])
def test_owner_tile_increments_in_element_order(runtime, renumber, part):
    """gen_edge_owner on owner tiles: float dim-3 and int increments through
    both indices of the edge -> node map (two dats, two arguments each) and
    an OP_GBL OP_MAX; each node's float sum is the sequential one (element
    order, bit for bit on the mesh's own numbering), the counts and the
    maximum exact.  A loop that also writes a direct dat (gen_edge) or
    increments through two maps keeps the coloured plan."""
    note(f"generic owner tiles: gen_edge_owner, renumber={renumber}, part={part}")
    mesh = M.o_mesh(90, 45, "random", seed=21)
    rng = np.random.Generator(np.random.PCG64(8))
    nodes = op2.op_decl_set(mesh.nnode, "nodes")
    edges = op2.op_decl_set(mesh.nedge, "edges")
    pedge = op2.op_decl_map(edges, nodes, 2, mesh.edge, "pedge")
    x = np.ascontiguousarray(mesh.x)
    nd = np.ascontiguousarray(rng.random((mesh.nnode, 3)).astype(np.float32))
    cnt = np.zeros(mesh.nnode, np.int32)
    w = np.ascontiguousarray(rng.random(mesh.nedge).astype(np.float32))
    px = op2.op_decl_dat(nodes, 2, "double", x, "px")
    pn = op2.op_decl_dat(nodes, 3, "float", nd.copy(), "pn")
    pc = op2.op_decl_dat(nodes, 1, "int", cnt.copy(), "pc")
    pw = op2.op_decl_dat(edges, 1, "float", w, "pw")
    if renumber:
        op2.renumber(pedge)
    op2.set_loop_config("gen_edge_owner", 0, part, 0)
    longest = np.zeros(1)
    args = (op_arg_dat(px, 0, pedge, 2, "double", OP_READ),
            op_arg_dat(px, 1, pedge, 2, "double", OP_READ),
            op_arg_dat(pn, 0, pedge, 3, "float", OP_INC),
            op_arg_dat(pn, 1, pedge, 3, "float", OP_INC),
            op_arg_dat(pc, 0, pedge, 1, "int", OP_INC),
            op_arg_dat(pc, 1, pedge, 1, "int", OP_INC),
            op_arg_dat(pw, -1, OP_ID, 1, "float", OP_READ),
            op_arg_gbl(longest, 1, "double", OP_MAX))
    for _ in range(2):
        op2.op_par_loop(None, "gen_edge_owner", edges, *args)
    e1, e2 = mesh.edge[:, 0], mesh.edge[:, 1]
    dx = x[e2, 0] - x[e1, 0]
    dy = x[e2, 1] - x[e1, 1]
    # the sequential order: edge by edge, node 1's increment then node 2's,
    # each rounded to float
    idx = np.stack([e1, e2], axis=1).reshape(-1)
    want = nd.copy()
    for _ in range(2):
        for comp, v1, v2 in ((0, w, -w), (1, dx.astype(np.float32), -dy.astype(np.float32)),
                             (2, np.ones_like(w), np.ones_like(w))):
            vals = np.stack([v1, v2], axis=1).reshape(-1).astype(np.float32)
            col = want[:, comp].copy()
            np.add.at(col, idx, vals)
            want[:, comp] = col
    got = op2.op_fetch_data(pn).reshape(-1, 3)
    if renumber:
        np.testing.assert_allclose(got, want, rtol=0, atol=2e-5)
    else:
        assert np.array_equal(got, want)
    want_cnt = 2 * (np.bincount(e1, minlength=mesh.nnode) + 2 * np.bincount(e2, minlength=mesh.nnode))
    assert np.array_equal(op2.op_fetch_data(pc), want_cnt.astype(np.int32))
    np.testing.assert_allclose(longest[0], np.sqrt(dx * dx + dy * dy).max(), rtol=1e-15)
    op2.set_loop_config("gen_edge_owner", 0, 0, 0)


def test_reductions_in_a_captured_graph(runtime):
    """With deferred reductions a generic loop's OP_GBL results go to the
    device log (the loop's sum for OP_INC, the result including the caller's
    value for OP_MAX), so the loop can be captured into a graph and replayed."""
    note("generic edge loop with OP_GBL INC and MAX captured in a graph, replayed")
    mesh = M.o_mesh(48, 24, "random", seed=11)
    nodes = op2.op_decl_set(mesh.nnode, "nodes")
    edges = op2.op_decl_set(mesh.nedge, "edges")
    pedge = op2.op_decl_map(edges, nodes, 2, mesh.edge, "pedge")
    px = op2.op_decl_dat(nodes, 2, "double", np.ascontiguousarray(mesh.x), "px")
    pc = op2.op_decl_dat(nodes, 1, "int", np.zeros(mesh.nnode, np.int32), "pc")
    total = np.array([1.0])
    longest = np.array([0.0])

    def loop():
        op2.op_par_loop(None, "gen_edge_sum", edges,
                        op_arg_dat(px, 0, pedge, 2, "double", OP_READ),
                        op_arg_dat(px, 1, pedge, 2, "double", OP_READ),
                        op_arg_dat(pc, 0, pedge, 1, "int", OP_INC),
                        op_arg_dat(pc, 1, pedge, 1, "int", OP_INC),
                        op_arg_gbl(total, 1, "double", OP_INC),
                        op_arg_gbl(longest, 1, "double", OP_MAX))

    loop()  # eager: plan and maps on the device
    e1, e2 = mesh.edge[:, 0], mesh.edge[:, 1]
    ln = np.hypot(mesh.x[e2, 0] - mesh.x[e1, 0], mesh.x[e2, 1] - mesh.x[e1, 1])
    np.testing.assert_allclose(total[0], 1.0 + ln.sum(), rtol=1e-13)
    op2.set_deferred_reductions(True)
    try:
        op2.reduction_log()
        op2.graph_begin()
        loop()
        g = op2.graph_end()
        op2.graph_launch(g, 2)
        loop()  # eager, deferred, between replays
        op2.graph_launch(g, 1)
        log = op2.reduction_log()
    finally:
        op2.set_deferred_reductions(False)
    assert log.size == 8
    np.testing.assert_allclose(log[0::2], ln.sum(), rtol=1e-13)
    assert np.all(log[1::2] == ln.max())
    assert total[0] == 1.0 + ln.sum() or np.isclose(total[0], 1.0 + ln.sum(), rtol=1e-13)
    want_cnt = 5 * (np.bincount(e1, minlength=mesh.nnode) + np.bincount(e2, minlength=mesh.nnode))
    assert np.array_equal(op2.op_fetch_data(pc), want_cnt.astype(np.int32))


def test_registered_loop_takes_precedence_over_built_in(runtime, loops):
    """A loop registered under a built-in's name and argument signature runs
    the registered kernel (here airfoil update's signature with half the
    move); removing the registration brings the built-in kernel back."""
    note("registered 'update' with another body runs instead of the built-in update")
    from oplus2_matlab_amd import airfoil as AF

    mesh = M.o_mesh(24, 12, "random", seed=2)
    c = M.flow_constants().as_array()
    rng = np.random.Generator(np.random.PCG64(8))
    q0 = np.tile(c[6:10], (mesh.ncell, 1)) * (1 + 0.01 * rng.standard_normal((mesh.ncell, 4)))
    res0 = 1e-3 * rng.standard_normal((mesh.ncell, 4))
    a = AF.Airfoil(mesh, c, "double", q0=q0, res0=res0)
    a.save_soln()
    a.adt_calc()
    adt = op2.op_fetch_data(a.p_adt).copy()
    loops.testloops_register_half_update.argtypes = [C.c_int]
    assert loops.testloops_register_half_update(1) == 0
    try:
        rms = a.update()
    finally:
        assert loops.testloops_register_half_update(0) == 0
    q = op2.op_fetch_data(a.p_q).reshape(-1, 4)
    half = 0.5 / adt[:, None] * res0
    np.testing.assert_allclose(q, q0 - half, rtol=1e-15, atol=1e-15)
    assert abs(rms - float((half ** 2).sum())) < 1e-12 * rms
    assert not op2.op_fetch_data(a.p_res).any()
    # the built-in update again: the full move (res is zero now: q = qold)
    a.update()
    assert np.array_equal(op2.op_fetch_data(a.p_q).reshape(-1, 4), q0)
