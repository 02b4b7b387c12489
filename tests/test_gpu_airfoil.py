"""GPU parity: the HIP runtime against the oracle and the reference fixtures.

Every loop runs through the C ABI (op2_par_loop) on cuda:0.  FP64 results
must match to 1e-12 relative (the runtime sums increments in another order
than the reference, nothing else differs); the FP32 form is compared with
the oracle's FP32 form and its error against FP64 is reported.
"""

import os

import numpy as np
import pytest

from oplus2_matlab_amd import airfoil as AF
from oplus2_matlab_amd import mesh as M
from oplus2_matlab_amd import op2
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
STRATS = [op2.STRATEGY_ATOMIC, op2.STRATEGY_COLOURED, op2.STRATEGY_DIRECT, op2.STRATEGY_OWNER,
          op2.STRATEGY_OWNER_GATHER]


def note(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture
def runtime():
    op2.op_init([], 0)
    yield
    op2.op_exit()
    op2.set_loop_config(None, 0, 0, 0)


def configure(strategy, part=0):
    for loop in ("res_calc", "bres_calc", "res"):
        op2.set_loop_config(loop, strategy, part, 0)


def relmax(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


# ---------------------------------------------------------------------------
# single loops

@pytest.mark.parametrize("strategy", STRATS)
def test_single_loops_fp64(runtime, strategy):
    note(f"single loops fp64 strategy={strategy} mesh 48x24 random")
    configure(strategy, 128)
    mesh = M.o_mesh(48, 24, "random", seed=9)
    c = M.flow_constants().as_array()
    rng = np.random.Generator(np.random.PCG64(3))
    q0 = np.tile(c[6:10], (mesh.ncell, 1)) * (1 + 0.02 * rng.standard_normal((mesh.ncell, 4)))
    res0 = 1e-3 * rng.standard_normal((mesh.ncell, 4))
    b = AF.Airfoil(mesh, c, "double", q0=q0, res0=res0)

    ox, oq = mesh.x.copy(), q0.copy()
    oqold = np.zeros_like(oq)
    O.save_soln("f64", oq, oqold)
    b.save_soln()
    assert np.array_equal(op2.op_fetch_data(b.p_qold).reshape(-1, 4), oqold)

    oadt = np.zeros(mesh.ncell)
    O.adt_calc("f64", mesh, ox, oq, oadt, c)
    b.adt_calc()
    adt = op2.op_fetch_data(b.p_adt)
    assert relmax(adt, oadt) < 1e-15
    note(f"adt_calc bitwise equal to the oracle: {np.array_equal(adt, oadt)}")

    ores = res0.copy()
    O.res_calc("f64", mesh, ox, oq, oadt, ores, c)
    b.res_calc()
    assert relmax(op2.op_fetch_data(b.p_res).reshape(-1, 4), ores) < 1e-13

    O.bres_calc("f64", mesh, ox, oq, oadt, ores, c)
    b.bres_calc()
    res = op2.op_fetch_data(b.p_res).reshape(-1, 4)
    assert relmax(res, ores) < 1e-13

    rms_o = O.update("f64", oqold, oq, ores, oadt)
    rms_g = b.update()
    assert relmax(op2.op_fetch_data(b.p_q).reshape(-1, 4), oq) < 1e-12
    assert abs(rms_g - rms_o) / rms_o < 1e-12
    assert not op2.op_fetch_data(b.p_res).any()


EPS32 = float(np.finfo(np.float32).eps)


def abs_increments(mesh, x, q, adt, c, res0):
    """Per cell and component: |res0| plus the sum of |f| over the float
    res_calc and bres_calc increments the cell receives, each increment
    taken from the f32 oracle on a one-edge slice of the maps.  A float sum
    of these terms in any order differs from another order by at most
    (terms - 1) x eps32 x this sum."""
    tot = np.abs(res0).astype(np.float64)
    tmp = np.zeros_like(res0)
    for e in range(mesh.nedge):
        one = M.AirfoilMesh(x=mesh.x, cell=mesh.cell, edge=mesh.edge[e:e + 1],
                            ecell=mesh.ecell[e:e + 1], bedge=mesh.bedge[:0],
                            becell=mesh.becell[:0], bound=mesh.bound[:0])
        c1, c2 = mesh.ecell[e]
        tmp[c1] = tmp[c2] = 0
        O.res_calc("f32", one, x, q, adt, tmp, c)
        tot[c1] += np.abs(tmp[c1])
        tot[c2] += np.abs(tmp[c2])
    for e in range(mesh.nbedge):
        one = M.AirfoilMesh(x=mesh.x, cell=mesh.cell, edge=mesh.edge[:0], ecell=mesh.ecell[:0],
                            bedge=mesh.bedge[e:e + 1], becell=mesh.becell[e:e + 1],
                            bound=mesh.bound[e:e + 1])
        c1 = mesh.becell[e]
        tmp[c1] = 0
        O.bres_calc("f32", one, x, q, adt, tmp, c)
        tot[c1] += np.abs(tmp[c1])
    return tot


@pytest.mark.parametrize("strategy", STRATS)
def test_single_loops_fp32_form(runtime, strategy):
    """Each float kernel once against the precision-form oracle
    (orc_*_mix): save_soln and adt_calc bit for bit; res_calc and bres_calc
    (the same float flux per edge, summed in another order) within
    4 eps32 of each cell's sum of |increments|; update bit for bit on the
    same res (move and new q in double, rounded once) and its rms to double
    rounding."""
    note(f"single loops fp32 form strategy={strategy} mesh 48x24 random")
    configure(strategy, 128)
    mesh = M.o_mesh(48, 24, "random", seed=9)
    c = M.flow_constants().as_array()
    rng = np.random.Generator(np.random.PCG64(3))
    q0 = (np.tile(c[6:10], (mesh.ncell, 1)) *
          (1 + 0.02 * rng.standard_normal((mesh.ncell, 4)))).astype(np.float32)
    res0 = (1e-3 * rng.standard_normal((mesh.ncell, 4))).astype(np.float32)
    b = AF.Airfoil(mesh, c, "float", q0=q0, res0=res0)

    ox, oq = mesh.x.astype(np.float32), q0.copy()
    oqold = np.zeros_like(oq)
    O.save_soln("mix", oq, oqold)
    b.save_soln()
    assert np.array_equal(op2.op_fetch_data(b.p_qold).reshape(-1, 4), oqold)

    oadt = np.zeros(mesh.ncell, np.float32)
    O.adt_calc("mix", mesh, ox, oq, oadt, c)
    b.adt_calc()
    adt = op2.op_fetch_data(b.p_adt)
    assert adt.dtype == np.float32 and np.array_equal(adt, oadt)

    bound = 4 * EPS32 * abs_increments(mesh, ox, oq, oadt, c, res0)
    ores = res0.copy()
    O.res_calc("mix", mesh, ox, oq, oadt, ores, c)
    b.res_calc()
    res = op2.op_fetch_data(b.p_res).reshape(-1, 4)
    err = np.abs(res.astype(np.float64) - ores)
    note(f"res_calc: max err / bound {np.max(err / np.maximum(bound, 1e-300)):.3f}, "
         f"bitwise equal cells {np.mean(np.all(res == ores, axis=1)):.3f}")
    assert np.all(err <= bound)

    O.bres_calc("mix", mesh, ox, oq, oadt, ores, c)
    b.bres_calc()
    res = op2.op_fetch_data(b.p_res).reshape(-1, 4)
    assert np.all(np.abs(res.astype(np.float64) - ores) <= bound)

    # update on the runtime's own res, so that only update's arithmetic is
    # compared
    ures = res.copy()
    rms_o = O.update("mix", oqold, oq, ures, oadt)
    rms_g = b.update()
    assert np.array_equal(op2.op_fetch_data(b.p_q).reshape(-1, 4), oq)
    assert abs(rms_g - rms_o) / rms_o < 1e-13
    assert not op2.op_fetch_data(b.p_res).any()


def test_fp32_res_calc_short_records_keep_the_bits(runtime):
    """The FP32 eager res_calc on gather tiles reads 16-bit position records
    where a tile's targets fit two windows and 32-bit ids elsewhere (a
    renumbered 180k-cell mesh has both).  Each cell sums its increments in
    its edge order whatever the tiling, so plans of 128, 256 and 512 cells,
    whose short and full tiles differ, give the same bits."""
    note("fp32 res_calc, short and full position records, 600x300 random renumbered")
    mesh = M.o_mesh(600, 300, "random", seed=1)
    c = M.flow_constants().as_array()
    rng = np.random.Generator(np.random.PCG64(5))
    q0 = (np.tile(c[6:10], (mesh.ncell, 1)) *
          (1 + 0.02 * rng.standard_normal((mesh.ncell, 4)))).astype(np.float32)
    res0 = (1e-3 * rng.standard_normal((mesh.ncell, 4))).astype(np.float32)
    outs = []
    for part in (128, 256, 512):
        configure(op2.STRATEGY_OWNER_GATHER, part)
        b = AF.Airfoil(mesh, c, "float", q0=q0, res0=res0, renumber=True)
        b.adt_calc()
        b.res_calc()
        s = op2.plan_summary("res_calc", b.edges, b.res_calc_args(),
                             op2.STRATEGY_OWNER_GATHER, part)
        note(f"part {part}: {s['short_ids']} of {s['nblocks']} tiles with 16-bit records")
        assert s["verify_errors"] == 0 and 0 < s["short_ids"] < s["nblocks"]
        outs.append(op2.op_fetch_data(b.p_res).reshape(-1, 4))
        op2.op_exit()
        op2.op_init([], 0)
    assert all(np.array_equal(o, outs[0]) for o in outs[1:])
    assert np.any(outs[0] != res0)


def test_eager_res_calc_takes_updated_res_as_zero(runtime):
    """An eager update leaves res zero, so the next owner-gather res_calc
    starts its sums from +0 without reading res: the same bits as a run in
    which res is read (its device pointer taken, which stops the runtime from
    assuming anything about it).  With the pointer taken, values written
    through it between update and res_calc are added to."""
    import ctypes as C
    note("eager res_calc after update, 48x24 random, owner-gather, fp64")
    mesh = M.o_mesh(48, 24, "random", seed=9)
    c = M.flow_constants().as_array()

    def run(expose, fill=None):
        configure(op2.STRATEGY_OWNER_GATHER, 128)
        b = AF.Airfoil(mesh, c, "double")
        if expose:
            ptr = b.p_res.device_ptr()
        b.save_soln()
        for _ in range(2):
            b.adt_calc()
            b.res_calc()
            b.bres_calc()
            b.update()
        if fill is not None:
            # the HIP runtime the op2 library loaded
            path = next((ln.split()[-1] for ln in open("/proc/self/maps")
                         if "libamdhip64" in ln), "libamdhip64.so")
            hip = C.CDLL(path)
            src = np.ascontiguousarray(fill)
            op2.sync()
            assert hip.hipMemcpy(C.c_void_p(ptr), src.ctypes.data_as(C.c_void_p),
                                 C.c_size_t(src.nbytes), 1) == 0
        b.adt_calc()
        b.res_calc()
        res = op2.op_fetch_data(b.p_res).reshape(-1, 4).copy()
        op2.op_exit()
        op2.op_init([], 0)
        return res

    skipped = run(False)
    read = run(True)
    assert np.array_equal(skipped, read) and np.any(skipped != 0)
    ones = np.ones((mesh.ncell, 4))
    added = run(True, ones)
    assert np.max(np.abs((added - 1.0) - skipped)) <= 1e-15


# ---------------------------------------------------------------------------
# full runs against the reference fixtures

@pytest.mark.parametrize("strategy", STRATS)
@pytest.mark.parametrize("part", [64, 256])
def test_airfoil_24x12_matches_reference(runtime, golden_dir, strategy, part):
    note(f"airfoil fp64 24x12 random, 20 iterations, strategy={strategy} part={part}")
    configure(strategy, part)
    fx = np.load(os.path.join(golden_dir, "airfoil_24x12_random_s1.npz"))
    mesh = M.o_mesh(24, 12, "random", seed=1)
    hist, q, adt, res = AF.run_airfoil(mesh, fx["f64_consts"], int(fx["niter"]), "double")
    assert relmax(hist, fx["f64_seq_rms"]) < TOL64
    assert np.max(np.abs(hist - fx["f64_seq_rms"]) / fx["f64_seq_rms"]) < TOL64
    assert relmax(q, fx["f64_seq_q"]) < TOL64
    assert relmax(adt, fx["f64_seq_adt"]) < TOL64


@pytest.mark.parametrize("block", [128, 512])
@pytest.mark.parametrize("strategy", STRATS)
def test_airfoil_block_sizes_match_reference(runtime, golden_dir, strategy, block):
    """Every loop's kernels with 128 and 512 threads per workgroup (the
    indirect loops on 256-element parts: two elements per thread, or half
    the threads idle in the owner phase) against the reference fixture."""
    note(f"airfoil fp64 24x12 random, 20 iterations, strategy={strategy} block={block}")
    op2.set_loop_config(None, 0, 0, block)
    configure(strategy, 256)
    assert op2.get_loop_config("res_calc")["block_size"] == block
    fx = np.load(os.path.join(golden_dir, "airfoil_24x12_random_s1.npz"))
    mesh = M.o_mesh(24, 12, "random", seed=1)
    hist, q, adt, res = AF.run_airfoil(mesh, fx["f64_consts"], int(fx["niter"]), "double")
    assert np.max(np.abs(hist - fx["f64_seq_rms"]) / fx["f64_seq_rms"]) < TOL64
    assert relmax(q, fx["f64_seq_q"]) < TOL64
    assert relmax(adt, fx["f64_seq_adt"]) < TOL64


@pytest.mark.parametrize("strategy", STRATS)
def test_airfoil_180k_1000_iterations_matches_reference(runtime, golden_dir, strategy):
    note(f"airfoil fp64 600x300 random, 1000 iterations, strategy={strategy}")
    configure(strategy, 0)
    fx = np.load(os.path.join(golden_dir, "airfoil_600x300_random_s1_f64.npz"))
    mesh = M.o_mesh(600, 300, "random", seed=1)
    hist, q, adt, _ = AF.run_airfoil(mesh, fx["consts"], int(fx["niter"]), "double")
    rel = np.abs(hist - fx["rms"]) / np.abs(fx["rms"])
    note(f"rms history max rel err {rel.max():.3e}")
    assert rel.max() < TOL64
    s = fx["sample"]
    assert relmax(q[s], fx["q_sample"]) < TOL64
    assert relmax(adt[s], fx["adt_sample"]) < TOL64
    assert np.allclose(q.sum(axis=0), fx["q_sum"], rtol=TOL64, atol=0)


def test_airfoil_fp32_form(runtime):
    """FP32 state and flux, FP64 update and rms: matches the oracle's FP32
    form to float rounding and stays close to the FP64 oracle."""
    note("airfoil fp32 form 48x24 random, 200 iterations")
    mesh = M.o_mesh(48, 24, "random", seed=3)
    c = M.flow_constants().as_array()
    hist, q, _, _ = AF.run_airfoil(mesh, c, 200, "float")
    _, hmix = O.airfoil_seq(mesh, c, 200, "mix")
    st64, h64 = O.airfoil_seq(mesh, c, 200, "f64")
    e_mix = np.max(np.abs(hist - hmix) / np.abs(hmix))
    e64 = np.max(np.abs(hist - h64) / np.abs(h64))
    dq = np.max(np.abs(q - st64.q))
    note(f"fp32 form: rms rel err vs oracle fp32 form {e_mix:.2e}, vs fp64 {e64:.2e}, "
         f"max|q-q64| {dq:.2e}")
    # owner-gather tiles on the mesh as given add each cell's float
    # increments in the sequential order, so only the double rms sums differ
    # (measured 6e-15); an update done in float would differ by ~1e-7
    assert e_mix < 1e-12
    # the form's error against FP64 (measured 8.6e-7)
    assert e64 < 1e-5
    assert dq < 5e-6


@pytest.mark.parametrize("dtype", ["double", "float"])
def test_default_strategy_is_owner_gather(runtime, dtype):
    """With no loop configuration res_calc and bres_calc run owner-compute
    tiles: two runs give the same bits (no atomics), the same bits as
    owner-gather with 256-cell tiles configured explicitly, and FP64 matches
    the oracle."""
    note(f"airfoil {dtype} 60x30 random, 30 iterations: default strategy twice, owner-gather")
    mesh = M.o_mesh(60, 30, "random", seed=5)
    c = M.flow_constants().as_array()
    runs = []
    for explicit in (False, False, True):
        op2.op_exit()
        op2.op_init([], 0)
        if explicit:
            configure(op2.STRATEGY_OWNER_GATHER, 256)
        runs.append(AF.run_airfoil(mesh, c, 30, dtype))
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            assert np.array_equal(a, b)
    if dtype == "double":
        st, h64 = O.airfoil_seq(mesh, c, 30, "f64")
        assert relmax(runs[0][0], h64) < TOL64
        assert relmax(runs[0][1], st.q) < TOL64


# ---------------------------------------------------------------------------
# Jacobi

@pytest.mark.parametrize("strategy", STRATS)
@pytest.mark.parametrize("part", [128, 1024])
def test_jacobi_matches_reference(runtime, golden_dir, strategy, part):
    """Small blocks share many read-only nodes with blocks of the same
    colour: only the nodes a block increments may be written back."""
    note(f"jacobi fp64 64x64 random, 200 iterations, strategy={strategy} part={part}")
    configure(strategy, part if strategy != op2.STRATEGY_OWNER else 256)
    fx = np.load(os.path.join(golden_dir, "jacobi_64_random_s1_f64.npz"))
    prob = M.jacobi_grid(64, seed=1)
    hs, hm, u, du = AF.run_jacobi(prob, int(fx["niter"]), 0.25, -0.25, "double")
    assert relmax(hs, fx["seq_u_sum"]) < TOL64
    assert np.array_equal(hm, fx["seq_u_max"]) or relmax(hm, fx["seq_u_max"]) < TOL64
    assert relmax(u, fx["seq_u"]) < TOL64
    assert not du.any()


@pytest.mark.parametrize("block", [128, 512])
@pytest.mark.parametrize("strategy", STRATS)
def test_jacobi_block_sizes_match_reference(runtime, golden_dir, strategy, block):
    note(f"jacobi fp64 64x64 random, 200 iterations, strategy={strategy} block={block}")
    op2.set_loop_config(None, 0, 0, block)
    configure(strategy, 256)
    fx = np.load(os.path.join(golden_dir, "jacobi_64_random_s1_f64.npz"))
    prob = M.jacobi_grid(64, seed=1)
    hs, hm, u, du = AF.run_jacobi(prob, int(fx["niter"]), 0.25, -0.25, "double")
    assert relmax(hs, fx["seq_u_sum"]) < TOL64
    assert relmax(hm, fx["seq_u_max"]) < TOL64
    assert relmax(u, fx["seq_u"]) < TOL64
    assert not du.any()


def test_jac1_known_answer(runtime):
    """jac1 as shipped (NN=6, alpha = beta = 1): exact dyadic values."""
    note("jac1 NN=6 known answer")
    prob = M.jacobi_reference_problem(6)
    hs, hm, u, _ = AF.run_jacobi(prob, 2, 1.0, 1.0, "double")
    assert list(hm) == [0.5, 0.625]
    assert abs(np.sqrt(hs[0] / 25) - 0.264575) < 1e-6 and abs(np.sqrt(hs[1] / 25) - 0.386491) < 1e-6
    grid = u.reshape(5, 5)
    assert grid[0, 0] == 0.625 and grid[2, 2] == 0.0 and grid[1, 1] == 0.125


# ---------------------------------------------------------------------------
# full size: properties and one-call parity at 2.88M cells

@pytest.mark.parametrize("ordering", ["cartesian", "random"])
@pytest.mark.parametrize("strategy", [op2.STRATEGY_ATOMIC, op2.STRATEGY_OWNER])
def test_full_size_res_calc(runtime, ordering, strategy):
    note(f"res_calc fp64 2400x1200 {ordering} strategy={strategy}: one call vs oracle")
    configure(strategy, 256)
    mesh = M.o_mesh(2400, 1200, ordering, seed=1)
    c = M.flow_constants().as_array()
    rng = np.random.Generator(np.random.PCG64(5))
    q0 = np.tile(c[6:10], (mesh.ncell, 1)) * (1 + 0.01 * rng.standard_normal((mesh.ncell, 4)))
    b = AF.Airfoil(mesh, c, "double", q0=q0)
    b.adt_calc()
    b.res_calc()
    res = op2.op_fetch_data(b.p_res).reshape(-1, 4)
    adt = op2.op_fetch_data(b.p_adt)
    oadt = np.zeros(mesh.ncell)
    O.adt_calc("f64", mesh, mesh.x, q0, oadt, c)
    ores = np.zeros_like(q0)
    O.res_calc("f64", mesh, mesh.x, q0, oadt, ores, c)
    assert relmax(adt, oadt) < 1e-15
    assert relmax(res, ores) < 1e-12
    assert np.abs(res.sum(axis=0)).max() < 1e-11 * np.abs(res).sum()


# ---------------------------------------------------------------------------
# errors

def test_unknown_loop_is_an_error(runtime):
    s = op2.op_decl_set(4, "s")
    d = op2.op_decl_dat(s, 1, "double", np.zeros(4), "d")
    with pytest.raises(op2.OP2Error, match="no device kernel"):
        op2.op_par_loop(None, "my_loop", s, op2.op_arg_dat(d, -1, None, 1, "double", op2.OP_RW))


# ---------------------------------------------------------------------------
# deferred reductions and hipGraph replay

@pytest.mark.parametrize("strategy", STRATS)
def test_airfoil_graph_replays_match_reference(runtime, golden_dir, strategy):
    """One iteration captured as a hipGraph and replayed, with eager
    iterations and host synchronisation between the replays; every
    iteration's rms must match the reference history."""
    note(f"airfoil fp64 24x12 graph replay, strategy={strategy}")
    configure(strategy, 64)
    fx = np.load(os.path.join(golden_dir, "airfoil_24x12_random_s1.npz"))
    mesh = M.o_mesh(24, 12, "random", seed=1)
    a = AF.Airfoil(mesh, fx["f64_consts"], "double")
    ref = fx["f64_seq_rms"]
    hist = [a.iterate()]                       # eager: builds plans
    gid = a.capture_iteration()
    hist += list(a.run_graph(gid, 5))          # replays 1-5
    for _ in range(2):                         # eager, deferred, between replays
        a.iterate()
    log = op2.reduction_log()
    hist += list(np.sqrt(log[1::2] / mesh.ncell))
    op2.sync()
    hist += list(a.run_graph(gid, int(fx["niter"]) - len(hist)))
    op2.set_deferred_reductions(False)
    hist = np.array(hist)
    assert hist.size == ref.size
    assert np.max(np.abs(hist - ref) / ref) < TOL64
    q, _, _, _ = a.fetch()
    assert relmax(q, fx["f64_seq_q"]) < TOL64


@pytest.mark.parametrize("strategy", [op2.STRATEGY_COLOURED, op2.STRATEGY_OWNER])
def test_jacobi_graph_replays_match_reference(runtime, golden_dir, strategy):
    note(f"jacobi fp64 graph replay strategy={strategy}")
    configure(strategy, 1024 if strategy == op2.STRATEGY_COLOURED else 256)
    fx = np.load(os.path.join(golden_dir, "jacobi_64_random_s1_f64.npz"))
    prob = M.jacobi_grid(64, seed=1)
    j = AF.Jacobi(prob, 0.25, "double")
    s0, m0 = j.iterate(-0.25)
    op2.set_deferred_reductions(True)
    op2.graph_begin()
    j.iterate(-0.25)
    gid = op2.graph_end()
    n = int(fx["niter"]) - 1
    op2.graph_launch(gid, 100)
    log1 = op2.reduction_log()
    op2.graph_launch(gid, n - 100)
    log = np.concatenate([log1, op2.reduction_log()])
    op2.set_deferred_reductions(False)
    sums = np.concatenate([[s0], log[0::2]])
    maxs = np.concatenate([[m0], log[1::2]])
    assert relmax(sums, fx["seq_u_sum"]) < TOL64
    assert relmax(maxs, fx["seq_u_max"]) < TOL64
    op2.op_fetch_data(j.p_u)
    assert relmax(j.u, fx["seq_u"]) < TOL64


def test_graph_capture_needs_prepared_loops(runtime):
    """A loop whose plan does not exist yet cannot be captured (its upload
    would become part of the graph): the runtime refuses with an error."""
    mesh = M.o_mesh(24, 12, "random", seed=1)
    a = AF.Airfoil(mesh, M.flow_constants().as_array(), "double")
    configure(op2.STRATEGY_COLOURED, 64)
    op2.set_deferred_reductions(True)
    op2.graph_begin()
    with pytest.raises(op2.OP2Error, match="graph capture"):
        a.res_calc()
    op2.graph_end()
    op2.set_deferred_reductions(False)


@pytest.mark.parametrize("strategy", [op2.STRATEGY_OWNER, op2.STRATEGY_OWNER_GATHER])
@pytest.mark.parametrize("part", [32, 512, 1024])
def test_owner_tile_sizes(runtime, strategy, part):
    """Owner-compute tiles of every kernel shape (positions and owned cells
    per thread) against the oracle."""
    note(f"owner tiles fp64 48x24 random, strategy={strategy} part={part}")
    configure(strategy, part)
    mesh = M.o_mesh(48, 24, "random", seed=4)
    c = M.flow_constants().as_array()
    hist, q, _, _ = AF.run_airfoil(mesh, c, 10, "double")
    st, oh = O.airfoil_seq(mesh, c, 10, "f64")
    assert np.max(np.abs(hist - oh) / oh) < TOL64
    assert relmax(q, st.q) < TOL64


def test_autotuner_keeps_state_and_parity(runtime):
    """The autotuner runs res_calc and bres_calc under every candidate on
    the live mesh; the state must be unchanged afterwards and the tuned run
    must still match the oracle."""
    note("autotune res_calc/bres_calc on 48x24 random")
    mesh = M.o_mesh(48, 24, "random", seed=6)
    c = M.flow_constants().as_array()
    a = AF.Airfoil(mesh, c, "double")
    a.iterate()
    before = [op2.op_fetch_data(d).copy() for d in (a.p_q, a.p_res, a.p_adt)]
    tuned = a.autotune(reps=2)
    note(f"tuned: {tuned}")
    after = [op2.op_fetch_data(d).copy() for d in (a.p_q, a.p_res, a.p_adt)]
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    assert tuned["res_calc"]["strategy"] in (1, 2, 4, 5)
    hist = [a.iterate() for _ in range(9)]
    st, oh = O.airfoil_seq(mesh, c, 10, "f64")
    assert np.max(np.abs(np.array(hist) - oh[1:]) / oh[1:]) < TOL64


@pytest.mark.parametrize("strategy", [op2.STRATEGY_OWNER, op2.STRATEGY_OWNER_GATHER])
def test_owner_cells_with_more_than_four_increments(runtime, strategy):
    """Edges listed twice give cells 5-8 increments, beyond the packed
    4-entry records of the owner plans: those plans keep per-cell lists."""
    import dataclasses
    note(f"res_calc with repeated edges (cells of up to 8 increments), strategy={strategy}")
    configure(strategy, 128)
    base = M.o_mesh(48, 24, "random", seed=5)
    mesh = dataclasses.replace(base, edge=np.ascontiguousarray(np.concatenate([base.edge, base.edge])),
                               ecell=np.ascontiguousarray(np.concatenate([base.ecell, base.ecell])))
    c = M.flow_constants().as_array()
    rng = np.random.Generator(np.random.PCG64(8))
    q0 = np.tile(c[6:10], (mesh.ncell, 1)) * (1 + 0.02 * rng.standard_normal((mesh.ncell, 4)))
    b = AF.Airfoil(mesh, c, "double", q0=q0)
    s = op2.plan_summary("res_calc", b.edges, b.res_calc_args(), strategy, 128)
    assert s["verify_errors"] == 0 and s["max_incs"] == 8
    oadt = np.zeros(mesh.ncell)
    O.adt_calc("f64", mesh, mesh.x.copy(), q0.copy(), oadt, c)
    ores = np.zeros((mesh.ncell, 4))
    O.res_calc("f64", mesh, mesh.x.copy(), q0.copy(), oadt, ores, c)
    b.adt_calc()
    b.res_calc()
    assert relmax(op2.op_fetch_data(b.p_res).reshape(-1, 4), ores) < 1e-13


def test_online_autotuner_runs_real_iterations(runtime):
    """Online tuning times each candidate inside real iterations, so the
    whole run, tuning included, must match the oracle; graph capture is
    refused until the tuning has finished."""
    note("online autotune res_calc/bres_calc on 48x24 random, 30 iterations")
    mesh = M.o_mesh(48, 24, "random", seed=6)
    c = M.flow_constants().as_array()
    a = AF.Airfoil(mesh, c, "double")
    for n in ("res_calc", "bres_calc"):
        op2.autotune_online(n, 2)
    hist = [a.iterate()]
    assert op2.autotune_online_result("res_calc") is None
    op2.set_deferred_reductions(True)
    with pytest.raises(op2.OP2Error, match="unfinished"):
        op2.graph_begin()
    op2.set_deferred_reductions(False)
    hist += [a.iterate() for _ in range(29)]
    r, b = op2.autotune_online_result("res_calc"), op2.autotune_online_result("bres_calc")
    note(f"online choices: res_calc {r}, bres_calc {b}")
    assert r["strategy"] in (1, 2, 4, 5) and b["strategy"] in (1, 2, 3, 4)
    _, oh = O.airfoil_seq(mesh, c, 30, "f64")
    assert np.max(np.abs(np.array(hist) - oh) / oh) < TOL64


def test_online_autotuner_jacobi_and_refusals(runtime, golden_dir):
    note("online autotune of jac1 res; direct loop refused")
    fx = np.load(os.path.join(golden_dir, "jacobi_64_random_s1_f64.npz"))
    op2.autotune_online("res", 1)
    prob = M.jacobi_grid(64, seed=1)
    hs, _, u, _ = AF.run_jacobi(prob, int(fx["niter"]), 0.25, -0.25, "double")
    assert op2.autotune_online_result("res") is not None
    assert relmax(hs, fx["seq_u_sum"]) < TOL64
    assert relmax(u, fx["seq_u"]) < TOL64
    mesh = M.o_mesh(24, 12, "random", seed=1)
    a = AF.Airfoil(mesh, M.flow_constants().as_array(), "double")
    op2.autotune_online("save_soln", 1)
    with pytest.raises(op2.OP2Error, match="cannot be autotuned"):
        a.save_soln()
