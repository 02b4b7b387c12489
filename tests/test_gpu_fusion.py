"""Loop fusion inside graph capture: save_soln + adt_calc, and bres_calc +
update + the next iteration's save_soln + adt_calc (or any run of these)
as one kernel each when captured back to back.  The fused
graph must give the unfused graph's results bit for bit (same per-cell
arithmetic, same reduction order), and a held loop that finds no partner
must still run, in call order."""

import dataclasses

import numpy as np
import pytest

from oplus2_matlab_amd import airfoil as AF
from oplus2_matlab_amd import mesh as M
from oplus2_matlab_amd import op2
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture
def runtime():
    """Owner-compute strategies throughout: their sums have a fixed order,
    so two runs can be compared bit for bit (atomic increments cannot)."""
    op2.op_init([], 0)
    for loop in ("res_calc", "bres_calc"):
        op2.set_loop_config(loop, op2.STRATEGY_OWNER_GATHER, 128, 0)
    yield
    op2.set_graph_fusion(True)
    op2.set_stage_config(0)
    op2.set_deferred_reductions(False)
    op2.op_exit()
    op2.set_loop_config(None, 0, 0, 0)


def graph_run(mesh, c, dtype, niter, fusion, renumber=False, graph_iters=1):
    op2.set_graph_fusion(fusion)
    hist, q, adt, res = AF.run_airfoil(mesh, c, niter, dtype, graph=True, renumber=renumber,
                                       graph_iters=graph_iters)
    return hist, q, adt, res


def same_rms(h1, h0, level):
    """Loop fusion keeps the rms bits; stage kernels group the rms sum per
    tile, so it agrees to double rounding."""
    if level == op2.FUSION_STAGES:
        return bool(np.all(np.abs(h1 - h0) <= 1e-14 * np.abs(h0)))
    return np.array_equal(h1, h0)


@pytest.mark.parametrize("level", [op2.FUSION_LOOPS, op2.FUSION_STAGES])
@pytest.mark.parametrize("dtype", ["double", "float"])
@pytest.mark.parametrize("renumber,graph_iters", [(False, 1), (True, 1), (False, 3), (True, 4)])
def test_fused_graph_equals_unfused(runtime, level, dtype, renumber, graph_iters):
    """graph_iters > 1 adds the update + next save_soln + adt_calc fusion;
    29 replayed iterations leave a remainder for the one-iteration graph.
    Stage kernels run on tiles owning consecutive cells (renumbered) or a
    list of cells (as given)."""
    note(f"fused (level {level}) vs unfused graph, 40x20 random, {dtype}, renumber={renumber}, "
         f"{graph_iters} iterations per graph, 30 iterations")
    mesh = M.o_mesh(40, 20, "random", seed=3)
    c = M.flow_constants().as_array()
    h1, q1, a1, r1 = graph_run(mesh, c, dtype, 30, level, renumber, graph_iters)
    op2.op_exit()
    op2.op_init([], 0)
    h0, q0, a0, r0 = graph_run(mesh, c, dtype, 30, False, renumber, graph_iters)
    assert same_rms(h1, h0, level)
    assert np.array_equal(q1, q0) and np.array_equal(a1, a0) and np.array_equal(r1, r0)
    if dtype == "double":
        st, oh = O.airfoil_seq(mesh, c, 30, "f64")
        assert np.max(np.abs(h1 - oh) / oh) < 1e-12


@pytest.mark.parametrize("strategy", [op2.STRATEGY_OWNER, op2.STRATEGY_COLOURED])
def test_fused_equals_unfused_other_strategies(runtime, strategy):
    """The carried fold rides in whichever res_calc kernel follows: the
    staged owner kernel, or the first launch of the coloured one."""
    note(f"fused vs unfused graph, res_calc strategy {strategy}, 2 iterations per graph")
    for loop in ("res_calc", "bres_calc"):
        op2.set_loop_config(loop, strategy, 128, 0)
    mesh = M.o_mesh(40, 20, "random", seed=13)
    c = M.flow_constants().as_array()
    h1, q1, a1, _ = graph_run(mesh, c, "double", 21, True, False, 2)
    op2.op_exit()
    op2.op_init([], 0)
    h0, q0, a0, _ = graph_run(mesh, c, "double", 21, False, False, 2)
    assert np.array_equal(h1, h0) and np.array_equal(q1, q0) and np.array_equal(a1, a0)


@pytest.mark.parametrize("level", [op2.FUSION_LOOPS, op2.FUSION_STAGES])
@pytest.mark.parametrize("renumber", [False, True])
@pytest.mark.parametrize("graph_iters", [1, 4])
def test_fused_batch_graph_equals_unfused(runtime, level, renumber, graph_iters):
    """Renumbered, every tile lies in one mesh and the stage kernels fold
    each mesh's tiles; as given, tiles may straddle meshes and the stages
    fall back to the separate kernels."""
    note(f"fused (level {level}) vs unfused batch graph, 6 meshes, renumber={renumber}, "
         f"12 iterations, {graph_iters} per graph")
    batch = M.mesh_batch(nmesh=6, seed=4, cells_lo=300, cells_hi=3000, ordering="random")
    op2.set_graph_fusion(level)
    h1, q1 = AF.run_airfoil_batch(batch, 12, "double", graph=True, graph_iters=graph_iters,
                                  renumber=renumber)
    op2.op_exit()
    op2.op_init([], 0)
    op2.set_graph_fusion(False)
    h0, q0 = AF.run_airfoil_batch(batch, 12, "double", graph=True, graph_iters=graph_iters,
                                  renumber=renumber)
    assert same_rms(h1, h0, level) and np.array_equal(q1, q0)
    for i in range(batch.nmesh):
        _, h = O.airfoil_seq(batch.meshes[i], batch.consts[i], 12, "f64")
        assert np.max(np.abs(h1[:, i] - h) / h) < 1e-12


@pytest.mark.parametrize("dtype", ["double", "float"])
def test_fused_boundary_cells_with_two_boundary_edges(runtime, dtype):
    """A one-ring O-mesh: every cell has a wall and a far-field edge, so the
    update that takes bres_calc in adds two boundary increments per cell."""
    note(f"fused graph on a 32x1 O-mesh (two boundary edges per cell), {dtype}")
    mesh = M.o_mesh(32, 1, "random", seed=8)
    assert np.all(np.bincount(mesh.becell, minlength=mesh.ncell) == 2)
    c = M.flow_constants().as_array()
    h, q, _, _ = graph_run(mesh, c, dtype, 40, True, False, 4)
    if dtype == "double":
        st, oh = O.airfoil_seq(mesh, c, 40, "f64")
        assert np.max(np.abs(h - oh) / oh) < 1e-12
        assert np.max(np.abs(q - st.q.reshape(-1, 4))) / np.max(np.abs(st.q)) < 1e-12
    else:
        _, oh = O.airfoil_seq(mesh, c, 40, "mix")
        e_mix = np.max(np.abs(h - oh) / oh)
        note(f"two boundary edges, fp32 form: rms rel err vs oracle fp32 form {e_mix:.2e}")
        # a 32-cell ring amplifies float rounding differences (measured 3.2e-5
        # after 40 iterations)
        assert e_mix < 3e-4


@pytest.mark.parametrize("renumber", [False, True])
def test_cell_without_edges_keeps_zeroed_res(runtime, renumber):
    """A cell that no edge touches is not owned by any res_calc tile, so the
    update before res_calc must still store its zero res (res_calc will not
    overwrite it): fused equals unfused, and both match the oracle."""
    note(f"fused graph, 24x12 O-mesh plus a cell without edges, renumber={renumber}")
    m = M.o_mesh(24, 12, "random", seed=10)
    mesh = dataclasses.replace(m, cell=np.ascontiguousarray(np.concatenate([m.cell, m.cell[:1]])))
    c = M.flow_constants().as_array()
    h1, q1, a1, r1 = graph_run(mesh, c, "double", 13, True, renumber, 4)
    op2.op_exit()
    op2.op_init([], 0)
    h0, q0, a0, r0 = graph_run(mesh, c, "double", 13, False, renumber, 4)
    assert np.array_equal(h1, h0) and np.array_equal(q1, q0) and np.array_equal(r1, r0)
    st, oh = O.airfoil_seq(mesh, c, 13, "f64")
    assert np.max(np.abs(h1 - oh) / oh) < 1e-12


def test_cell_with_three_boundary_edges_is_not_fused(runtime):
    """A boundary edge listed twice gives its cell three: that bres_calc
    keeps its own kernel in the graph, and the run still matches the oracle
    (the duplicate's flux is added twice there too)."""
    note("fused graph, 24x1 O-mesh with one boundary edge duplicated")
    m = M.o_mesh(24, 1, "random", seed=9)
    dup = lambda a: np.ascontiguousarray(np.concatenate([a, a[:1]]))
    mesh = dataclasses.replace(m, bedge=dup(m.bedge), becell=dup(m.becell), bound=dup(m.bound))
    assert np.bincount(mesh.becell).max() == 3
    c = M.flow_constants().as_array()
    h, q, _, _ = graph_run(mesh, c, "double", 20, True, False, 4)
    st, oh = O.airfoil_seq(mesh, c, 20, "f64")
    assert np.max(np.abs(h - oh) / oh) < 1e-12
    assert np.max(np.abs(q - st.q.reshape(-1, 4))) / np.max(np.abs(st.q)) < 1e-12


def test_held_loop_without_partner_runs_in_order(runtime):
    """save_soln held, then an adt_calc that reads qold (not the q that
    save_soln copies): no fusion, and the copy must land before adt_calc
    reads it.  A graph that ends on a held loop still runs it."""
    note("held save_soln flushed before a non-matching adt_calc; graph ending on a held loop")
    mesh = M.o_mesh(24, 12, "random", seed=6)
    c = M.flow_constants().as_array()
    rng = np.random.Generator(np.random.PCG64(5))
    q0 = np.tile(c[6:10], (mesh.ncell, 1)) * (1 + 0.01 * rng.standard_normal((mesh.ncell, 4)))
    a = AF.Airfoil(mesh, c, "double", q0=q0)
    zeros = np.zeros(4 * mesh.ncell)
    p_zero = op2.op_decl_dat(a.cells, 4, "double", zeros, "p_zero")
    R, W, ID = op2.OP_READ, op2.OP_WRITE, op2.OP_ID

    def save(src):
        op2.op_par_loop(op2.save_soln, "save_soln", a.cells,
                        op2.op_arg_dat(src, -1, ID, 4, "double", R),
                        op2.op_arg_dat(a.p_qold, -1, ID, 4, "double", W))

    def adt_on_qold():
        op2.op_par_loop(op2.adt_calc, "adt_calc", a.cells,
                        *[op2.op_arg_dat(a.p_x, i, a.pcell, 2, "double", R) for i in range(4)],
                        op2.op_arg_dat(a.p_qold, -1, ID, 4, "double", R),
                        op2.op_arg_dat(a.p_adt, -1, ID, 1, "double", W))

    save(a.p_q)
    adt_on_qold()
    op2.set_deferred_reductions(True)
    op2.graph_begin()
    save(a.p_q)
    adt_on_qold()
    g1 = op2.graph_end()
    op2.graph_begin()
    save(a.p_q)
    g2 = op2.graph_end()
    oadt = np.zeros(mesh.ncell)
    O.adt_calc("f64", mesh, mesh.x.copy(), q0.copy(), oadt, c)
    for g in (g1, g2):
        save(p_zero)  # qold = 0, so a copy that is skipped or late shows
        op2.graph_launch(g, 1)
        op2.sync()
        _, qold, adt, _ = a.fetch()
        assert np.array_equal(qold, q0), g
        if g == g1:
            assert np.max(np.abs(adt - oadt) / oadt) < 1e-15


def test_back_to_back_updates_fold_in_order(runtime):
    """Two updates captured with nothing between them: the first one's
    reduction (carried over to the next kernel when fusion is on) must be
    folded before the second overwrites the partials.  Same log and state
    as the capture without fusion, bit for bit with loop fusion, and to
    double rounding of the last rms with a stage kernel."""
    note("graph of update, update, res_calc: carried folds vs separate folds")
    mesh = M.o_mesh(40, 20, "random", seed=12)
    c = M.flow_constants().as_array()
    out = {}
    for fusion in (op2.FUSION_STAGES, op2.FUSION_LOOPS, op2.FUSION_OFF):
        op2.op_init([], 0)
        for loop in ("res_calc", "bres_calc"):
            op2.set_loop_config(loop, op2.STRATEGY_OWNER_GATHER, 128, 0)
        op2.set_graph_fusion(fusion)
        a = AF.Airfoil(mesh, c, "double")
        a.iterate()
        a.save_soln()
        a.adt_calc()
        a.res_calc()
        op2.set_deferred_reductions(True)
        op2.graph_begin()
        a.update()
        a.update()
        a.res_calc()
        a.update()
        g = op2.graph_end()
        op2.reduction_log()
        op2.graph_launch(g, 3)
        log = op2.reduction_log()
        q, _, adt, res = a.fetch()
        out[fusion] = (log, q.copy(), res.copy())
        op2.set_deferred_reductions(False)
        op2.op_exit()
    op2.op_init([], 0)
    off = out[op2.FUSION_OFF]
    assert off[0].size == 9
    for x, y in zip(out[op2.FUSION_LOOPS], off):
        assert np.array_equal(x, y)
    # the last update joins res_calc as a stage (one per graph: its q and adt
    # come back from the second buffers at the end of the graph)
    st = out[op2.FUSION_STAGES]
    assert same_rms(st[0], off[0], op2.FUSION_STAGES)
    assert np.array_equal(st[1], off[1]) and np.array_equal(st[2], off[2])


def test_fusion_switch_refused_during_capture(runtime):
    mesh = M.o_mesh(12, 6, "random", seed=2)
    a = AF.Airfoil(mesh, M.flow_constants().as_array(), "double")
    a.iterate()
    op2.set_deferred_reductions(True)
    op2.graph_begin()
    with pytest.raises(op2.OP2Error, match="graph capture"):
        op2.set_graph_fusion(False)
    op2.graph_end()


def test_bres_calc_on_another_adt_is_not_fused(runtime):
    """bres_calc reading an adt dat other than update's is not taken into the
    update (the fused kernel applies the boundary flux with update's adt):
    fused and unfused graphs give the same bits."""
    note("fused graph with bres_calc on a second adt dat")
    mesh = M.o_mesh(24, 12, "random", seed=4)
    c = M.flow_constants().as_array()
    out = []
    for fuse in (True, False):
        op2.op_exit()
        op2.op_init([], 0)
        op2.set_graph_fusion(fuse)
        a = AF.Airfoil(mesh, c, "double")
        adt2 = np.full(mesh.ncell, 3.0)
        a.p_adt2 = op2.op_decl_dat(a.cells, 1, "double", adt2, "p_adt2")
        base = a.bres_calc_args
        a.bres_calc_args = lambda: base()[:3] + (
            op2.op_arg_dat(a.p_adt2, 0, a.pbecell, 1, "double", op2.OP_READ),) + base()[4:]
        a.iterate()
        gid = a.capture_iteration(3)
        a.run_graph(gid, 6)
        op2.set_deferred_reductions(False)
        q, _, adt, res = a.fetch()
        out.append((q.copy(), adt.copy(), res.copy()))
    op2.set_graph_fusion(True)
    for x, y in zip(*out):
        assert np.array_equal(x, y)


def test_autotuning_refused_while_graphs_exist(runtime):
    """Autotuning frees the plans of the losing configurations, which a
    captured graph may use: both tuners refuse while a graph exists (the
    graph is not replayed) and run again once it is destroyed."""
    note("autotune after graph capture is refused")
    mesh = M.o_mesh(24, 12, "random", seed=4)
    c = M.flow_constants().as_array()
    a = AF.Airfoil(mesh, c, "double")
    a.iterate()
    gid = a.capture_iteration(1)
    with pytest.raises(op2.OP2Error, match="captured graphs exist"):
        op2.autotune("res_calc", a.edges, a.res_calc_args(), 1)
    with pytest.raises(op2.OP2Error, match="captured graphs exist"):
        op2.autotune_online("res_calc", 1)
    op2.graph_destroy(gid)
    op2.set_deferred_reductions(False)
    got = op2.autotune("res_calc", a.edges, a.res_calc_args(), 1)
    assert got["strategy"] in range(1, 6)


@pytest.mark.parametrize("block", [128, 512])
@pytest.mark.parametrize("dtype", ["double", "float"])
def test_fused_graphs_at_other_block_sizes(runtime, block, dtype):
    """Stage kernels and fused update kernels with 128 and 512 threads per
    workgroup (256-cell tiles: two cells per thread, or half the threads
    idle in the update phase): fused equals unfused, and FP64 matches the
    oracle."""
    note(f"fused graphs with {block}-thread workgroups, 40x20 random renumbered, {dtype}")
    mesh = M.o_mesh(40, 20, "random", seed=3)
    c = M.flow_constants().as_array()

    def run(level):
        op2.set_loop_config(None, 0, 0, block)
        op2.set_stage_config(-1)  # stage tiles as res_calc's
        for loop in ("res_calc", "bres_calc"):
            op2.set_loop_config(loop, op2.STRATEGY_OWNER_GATHER, 256, 0)
        return graph_run(mesh, c, dtype, 17, level, True, 4)

    h1, q1, a1, r1 = run(op2.FUSION_STAGES)
    op2.op_exit()
    op2.op_init([], 0)
    h0, q0, a0, r0 = run(op2.FUSION_OFF)
    assert same_rms(h1, h0, op2.FUSION_STAGES)
    assert np.array_equal(q1, q0) and np.array_equal(a1, a0) and np.array_equal(r1, r0)
    if dtype == "double":
        st, oh = O.airfoil_seq(mesh, c, 17, "f64")
        assert np.max(np.abs(h1 - oh) / oh) < 1e-12


@pytest.mark.parametrize("block", [128, 512])
def test_fused_batch_at_other_block_sizes(runtime, block):
    """The batched update's per-mesh partial slots depend on its workgroup
    size (cells per workgroup): 6 meshes against their oracle runs."""
    note(f"batch graph with {block}-thread workgroups, 6 meshes, 9 iterations")
    batch = M.mesh_batch(nmesh=6, seed=4, cells_lo=300, cells_hi=3000, ordering="random")
    op2.set_loop_config(None, 0, 0, block)
    op2.set_stage_config(-1)  # stage tiles as res_calc's
    for loop in ("res_calc", "bres_calc"):
        op2.set_loop_config(loop, op2.STRATEGY_OWNER_GATHER, 256, 0)
    for level in (op2.FUSION_STAGES, op2.FUSION_LOOPS):
        op2.set_graph_fusion(level)
        h, _ = AF.run_airfoil_batch(batch, 9, "double", graph=True, graph_iters=4, renumber=True)
        for i in range(batch.nmesh):
            _, oh = O.airfoil_seq(batch.meshes[i], batch.consts[i], 9, "f64")
            assert np.max(np.abs(h[:, i] - oh) / oh) < 1e-12
        op2.op_exit()
        op2.op_init([], 0)
        op2.set_loop_config(None, 0, 0, block)
        for loop in ("res_calc", "bres_calc"):
            op2.set_loop_config(loop, op2.STRATEGY_OWNER_GATHER, 256, 0)


@pytest.mark.parametrize("dtype", ["double", "float"])
@pytest.mark.parametrize("renumber", [False, True])
def test_stage_kernels_with_two_cells_per_thread(runtime, dtype, renumber):
    """Owner-gather tiles of 512 cells: the stage kernel's threads own two
    cells each; fused equals unfused (q, adt and res bit for bit)."""
    note(f"stage kernels on 512-cell tiles, 60x30 random, {dtype}, renumber={renumber}")
    op2.set_stage_config(512, 256)
    mesh = M.o_mesh(60, 30, "random", seed=14)
    c = M.flow_constants().as_array()
    h1, q1, a1, r1 = graph_run(mesh, c, dtype, 11, op2.FUSION_STAGES, renumber, 4)
    op2.op_exit()
    op2.op_init([], 0)
    h0, q0, a0, r0 = graph_run(mesh, c, dtype, 11, op2.FUSION_OFF, renumber, 4)
    assert same_rms(h1, h0, op2.FUSION_STAGES)
    assert np.array_equal(q1, q0) and np.array_equal(a1, a0) and np.array_equal(r1, r0)


@pytest.mark.parametrize("dtype", ["double", "float"])
@pytest.mark.parametrize("tiles", [(0, 0), (192, 256), (128, 128), (512, 512)])
def test_stage_tiles_have_their_own_configuration(runtime, dtype, tiles):
    """The stage kernels run on tiles of their own configuration (default
    256 cells on 256 threads) while res_calc's eager launches keep 128-cell
    tiles: the stage plan is built at graph_begin, the stage kernels run
    (their device times are recorded) and fused equals unfused bit for bit."""
    part, block = tiles
    note(f"stage tiles {part}/{block}, res_calc 128-cell tiles, 60x30 random renumbered, "
         f"{dtype}")
    mesh = M.o_mesh(60, 30, "random", seed=6)
    c = M.flow_constants().as_array()
    op2.set_stage_config(part, block)
    op2.set_graph_kernel_timing(1)
    try:
        h1, q1, a1, r1 = graph_run(mesh, c, dtype, 9, op2.FUSION_STAGES, True, 4)
    finally:
        op2.set_graph_kernel_timing(0)
    op2.op_exit()
    op2.op_init([], 0)
    h0, q0, a0, r0 = graph_run(mesh, c, dtype, 9, op2.FUSION_OFF, True, 4)
    assert same_rms(h1, h0, op2.FUSION_STAGES)
    assert np.array_equal(q1, q0) and np.array_equal(a1, a0) and np.array_equal(r1, r0)
    op2.op_exit()
    op2.op_init([], 0)
    op2.set_graph_fusion(op2.FUSION_STAGES)
    af = AF.Airfoil(mesh, c, dtype, renumber=True)
    af.iterate()
    n_before = op2.plan_summary("res_calc", af.edges, af.res_calc_args(),
                                op2.STRATEGY_OWNER_GATHER, 128)["cached_plans"]
    op2.set_graph_kernel_timing(1)
    g = af.capture_iteration(2)
    op2.set_graph_kernel_timing(0)
    op2.graph_launch(g, 1)
    times = op2.graph_kernel_times(g)
    assert len(times) == 4 and all(0 < ms < 1000 for _, ms in times)
    # the stage plan exists (summary builds nothing more) unless it is res_calc's
    s = op2.plan_summary("res_calc", af.edges, af.res_calc_args(), op2.STRATEGY_OWNER_GATHER,
                         part or 256)
    assert s["verify_errors"] == 0
    assert s["cached_plans"] == n_before + (part != 128)
    op2.reduction_log()
    op2.set_deferred_reductions(False)


@pytest.mark.parametrize("dtype", ["double", "float"])
def test_stage_kernel_times_inside_graphs(runtime, dtype):
    """op2_set_graph_kernel_timing(n) brackets the n-th, 2n-th, ... stage
    kernel of each capture with event-record nodes.  A 3-iteration graph on a
    renumbered mesh holds six stage kernels, whose kinds follow the fusion
    (the first reads res, the second stage of an iteration joins the next
    save_soln and adt_calc, the last stores res); each bracketed one has a
    positive device time after a replay; a graph captured with the timing off
    has none; and the nodes change no result bit."""
    note(f"graph kernel timing, 60x30 random renumbered, {dtype}")
    mesh = M.o_mesh(60, 30, "random", seed=5)
    c = M.flow_constants().as_array()
    op2.set_graph_kernel_timing(1)
    try:
        h1, q1, a1, r1 = graph_run(mesh, c, dtype, 10, op2.FUSION_STAGES, True, 3)
    finally:
        op2.set_graph_kernel_timing(0)
    op2.op_exit()
    op2.op_init([], 0)
    h0, q0, a0, r0 = graph_run(mesh, c, dtype, 10, op2.FUSION_STAGES, True, 3)
    assert np.array_equal(h1, h0)
    assert np.array_equal(q1, q0) and np.array_equal(a1, a0) and np.array_equal(r1, r0)
    op2.op_exit()
    op2.op_init([], 0)
    af = AF.Airfoil(mesh, c, dtype, renumber=True)
    af.iterate()
    op2.set_graph_kernel_timing(1)
    every = af.capture_iteration(3)
    op2.set_graph_kernel_timing(2)
    second = af.capture_iteration(3)
    op2.set_graph_kernel_timing(0)
    plain = af.capture_iteration(1)
    for g in (every, second, plain):
        op2.graph_launch(g, 2)
    times = op2.graph_kernel_times(every)
    note(f"stage kernels {times}")
    # adt_calc from the tiles' edges (128-cell tiles fit the kernel's passes);
    # the new q written over qold when save_soln joins, so that the next
    # stage reads its qold and its q from one buffer
    B, S, A, RR, SR, Q = (op2.KIND_BRES, op2.KIND_SAVE | op2.KIND_SAVE_IN_PLACE,
                          op2.KIND_ADT | op2.KIND_ADT_SIDES, op2.KIND_READS_RES,
                          op2.KIND_STORES_RES, op2.KIND_QOLD_IS_Q)
    kinds = [B | A | RR, B | S | A, B | A | Q, B | S | A, B | A | Q, B | SR]
    assert [k for k, _ in times] == kinds
    assert all(0 < ms < 1000 for _, ms in times)
    times2 = op2.graph_kernel_times(second)
    assert [k for k, _ in times2] == kinds[1::2]
    assert all(0 < ms < 1000 for _, ms in times2)
    assert op2.graph_kernel_times(plain) == []
    op2.reduction_log()
    op2.set_deferred_reductions(False)


@pytest.mark.parametrize("dtype", ["double", "float"])
@pytest.mark.parametrize("tail", ["iteration", "save_soln"])
def test_save_soln_joined_in_place_keeps_every_dat(runtime, dtype, tail):
    """With the next save_soln joined to a stage the new q is stored once,
    over qold, and q lives in qold's buffer until the next stage: after the
    graph q, qold, adt and res hold the unfused graph's bits, whether the
    capture ends with a whole iteration (qold = the last iteration's start)
    or right after a save_soln (qold = q)."""
    note(f"save_soln in place: graphs of 3 iterations ending with {tail}, {dtype}")
    mesh = M.o_mesh(40, 20, "random", seed=17)
    c = M.flow_constants().as_array()
    out = []
    for i, fusion in enumerate((op2.FUSION_STAGES, op2.FUSION_OFF)):
        if i:
            op2.op_init([], 0)
        for loop in ("res_calc", "bres_calc"):
            op2.set_loop_config(loop, op2.STRATEGY_OWNER_GATHER, 128, 0)
        op2.set_graph_fusion(fusion)
        af = AF.Airfoil(mesh, c, dtype, renumber=True)
        af.iterate()
        op2.set_deferred_reductions(True)
        op2.graph_begin()
        for _ in range(3):
            af.save_soln()
            for _ in range(2):
                af.adt_calc()
                af.res_calc()
                af.bres_calc()
                af.update()
        if tail == "save_soln":
            af.save_soln()
        gid = op2.graph_end()
        op2.graph_launch(gid, 3)
        op2.reduction_log()
        op2.set_deferred_reductions(False)
        out.append([v.copy() for v in af.fetch()])
        op2.op_exit()
    op2.op_init([], 0)
    for got, want in zip(out[0], out[1]):
        assert np.array_equal(got, want)
    if tail == "save_soln":
        assert np.array_equal(out[0][0], out[0][1])
