"""GPU parity with the runtime's renumbering (op2_renumber): the loops run on
renumbered sets, and everything fetched must still match the oracle in the
mesh's own numbering."""

import os

import numpy as np
import pytest

from oplus2_matlab_amd import airfoil as AF
from oplus2_matlab_amd import mesh as M
from oplus2_matlab_amd import op2
from oplus2_matlab_amd.op2 import OP_ID, OP_READ, OP_WRITE, op_arg_dat
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
    op2.set_deferred_reductions(False)
    op2.op_exit()
    op2.set_loop_config(None, 0, 0, 0)


def configure(strategy, part=0):
    for loop in ("res_calc", "bres_calc", "res"):
        op2.set_loop_config(loop, strategy, part, 0)


def relmax(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("strategy", STRATS)
def test_renumbered_airfoil_matches_oracle(runtime, strategy):
    note(f"renumbered airfoil fp64 48x24 random, 10 iterations, strategy={strategy}")
    configure(strategy, 128)
    mesh = M.o_mesh(48, 24, "random", seed=9)
    c = M.flow_constants().as_array()
    a = AF.Airfoil(mesh, c, "double", renumber=True)
    assert op2.renumber_export(a.cells) is not None
    hist = np.array([a.iterate() for _ in range(10)])
    q, _, adt, res = a.fetch()
    st, oh = O.airfoil_seq(mesh, c, 10, "f64")
    assert np.max(np.abs(hist - oh) / oh) < TOL64
    assert relmax(q, st.q) < TOL64
    assert relmax(adt, st.adt) < TOL64
    assert not res.any()
    # a dat that no loop writes comes back exactly as declared
    assert np.array_equal(op2.op_fetch_data(a.p_x).reshape(-1, 2), mesh.x)


@pytest.mark.parametrize("strategy", [op2.STRATEGY_ATOMIC, op2.STRATEGY_OWNER_GATHER])
def test_renumbered_180k_1000_iterations(runtime, golden_dir, strategy):
    note(f"renumbered airfoil fp64 600x300 random, 1000 iterations, strategy={strategy}")
    configure(strategy, 0)
    fx = np.load(os.path.join(golden_dir, "airfoil_600x300_random_s1_f64.npz"))
    mesh = M.o_mesh(600, 300, "random", seed=1)
    hist, q, adt, _ = AF.run_airfoil(mesh, fx["consts"], int(fx["niter"]), "double", graph=True,
                                     renumber=True)
    rel = np.abs(hist - fx["rms"]) / np.abs(fx["rms"])
    note(f"rms history max rel err {rel.max():.3e}")
    assert rel.max() < TOL64
    s = fx["sample"]
    assert relmax(q[s], fx["q_sample"]) < TOL64
    assert relmax(adt[s], fx["adt_sample"]) < TOL64


def test_renumbered_batch(runtime):
    note("renumbered batch of 6 random meshes, graph replay 12 iterations")
    batch = M.mesh_batch(nmesh=6, seed=11, cells_lo=300, cells_hi=5000, ordering="random")
    hist, q = AF.run_airfoil_batch(batch, 12, "double", graph=True, renumber=True)
    for i in range(batch.nmesh):
        st, h = O.airfoil_seq(batch.meshes[i], batch.consts[i], 12, "f64")
        qi = q[batch.cell_off[i]:batch.cell_off[i + 1]]
        assert relmax(hist[:, i], h) < TOL64 and relmax(qi, st.q.reshape(-1, 4)) < TOL64, i


@pytest.mark.parametrize("strategy", [op2.STRATEGY_COLOURED, op2.STRATEGY_OWNER])
def test_renumbered_jacobi(runtime, golden_dir, strategy):
    note(f"renumbered jacobi 64x64 random, strategy={strategy}")
    configure(strategy, 256)
    fx = np.load(os.path.join(golden_dir, "jacobi_64_random_s1_f64.npz"))
    prob = M.jacobi_grid(64, seed=1)
    hs, hm, u, du = AF.run_jacobi(prob, int(fx["niter"]), 0.25, -0.25, "double", renumber=True)
    assert relmax(hs, fx["seq_u_sum"]) < TOL64
    assert relmax(hm, fx["seq_u_max"]) < TOL64
    assert relmax(u, fx["seq_u"]) < TOL64
    assert not du.any()


def test_declarations_after_renumbering(runtime):
    """A dat and a map declared after op2_renumber are translated: a copy
    of q declared later and a second cell -> node map drive save_soln and
    adt_calc to the oracle's results."""
    note("dat and map declared after renumbering")
    mesh = M.o_mesh(32, 16, "random", seed=4)
    c = M.flow_constants().as_array()
    rng = np.random.Generator(np.random.PCG64(2))
    q0 = np.tile(c[6:10], (mesh.ncell, 1)) * (1 + 0.01 * rng.standard_normal((mesh.ncell, 4)))
    a = AF.Airfoil(mesh, c, "double", q0=q0, renumber=True)
    q2 = np.ascontiguousarray(q0.reshape(-1).copy())
    p_q2 = op2.op_decl_dat(a.cells, 4, "double", q2, "p_q2")
    pcell2 = op2.op_decl_map(a.cells, a.nodes, 4, np.ascontiguousarray(mesh.cell), "pcell2")
    out = np.zeros_like(q2)
    p_out = op2.op_decl_dat(a.cells, 4, "double", out, "p_out")
    op2.op_par_loop(op2.save_soln, "save_soln", a.cells,
                    op_arg_dat(p_q2, -1, OP_ID, 4, "double", OP_READ),
                    op_arg_dat(p_out, -1, OP_ID, 4, "double", OP_WRITE))
    assert np.array_equal(op2.op_fetch_data(p_out).reshape(-1, 4), q0)
    op2.op_par_loop(op2.adt_calc, "adt_calc", a.cells,
                    *[op_arg_dat(a.p_x, i, pcell2, 2, "double", OP_READ) for i in range(4)],
                    op_arg_dat(p_q2, -1, OP_ID, 4, "double", OP_READ),
                    op_arg_dat(a.p_adt, -1, OP_ID, 1, "double", OP_WRITE))
    oadt = np.zeros(mesh.ncell)
    O.adt_calc("f64", mesh, mesh.x.copy(), q0.copy(), oadt, c)
    assert relmax(op2.op_fetch_data(a.p_adt), oadt) < 1e-15


def test_renumbered_fp32_form(runtime):
    """The FP32 form on renumbered sets against the oracle's FP32 form."""
    note("renumbered airfoil fp32 form 48x24 random, 100 iterations (graph)")
    configure(op2.STRATEGY_OWNER_GATHER, 256)
    mesh = M.o_mesh(48, 24, "random", seed=5)
    c = M.flow_constants().as_array()
    hist, q, _, _ = AF.run_airfoil(mesh, c, 100, "float", graph=True, renumber=True)
    _, hmix = O.airfoil_seq(mesh, c, 100, "mix")
    st64, h64 = O.airfoil_seq(mesh, c, 100, "f64")
    e_mix = np.max(np.abs(hist - hmix) / np.abs(hmix))
    e64 = np.max(np.abs(hist - h64) / np.abs(h64))
    dq = np.max(np.abs(q - st64.q))
    note(f"renumbered fp32 form: rms rel err vs oracle fp32 form {e_mix:.2e}, vs fp64 {e64:.2e}, "
         f"max|q-q64| {dq:.2e}")
    # renumbering changes the order of each cell's float increments (measured
    # 2.8e-7 against the oracle's FP32 form, 2.0e-7 against FP64)
    assert e_mix < 5e-6
    assert e64 < 5e-6
    assert dq < 5e-6
