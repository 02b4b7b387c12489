"""GPU parity of the batched airfoil: many meshes with their own free stream
advanced as one union mesh, every mesh checked against the oracle run of
that mesh alone (FP64 to 1e-12 relative, FP32 form against the oracle's
FP32 form)."""

import numpy as np
import pytest

from oplus2_matlab_amd import airfoil as AF
from oplus2_matlab_amd import mesh as M
from oplus2_matlab_amd import op2
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


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
    for loop in ("res_calc", "bres_calc"):
        op2.set_loop_config(loop, strategy, part, 0)


def relmax(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def check_meshes(batch, hist, q, niter, form, tol, which=None):
    """Every mesh's rms history and q against the oracle run of that mesh
    alone; returns the largest relative errors (rms, q)."""
    worst = (0.0, 0.0)
    for i in which if which is not None else range(batch.nmesh):
        st, h = O.airfoil_seq(batch.meshes[i], batch.consts[i], niter, form)
        qi = q[batch.cell_off[i]:batch.cell_off[i + 1]]
        eh, eq = relmax(hist[:, i], h), relmax(qi, st.q.reshape(-1, 4))
        assert eh < tol and eq < tol, (i, eh, eq)
        worst = (max(worst[0], eh), max(worst[1], eq))
    return worst


STRATS = [op2.STRATEGY_ATOMIC, op2.STRATEGY_COLOURED, op2.STRATEGY_DIRECT, op2.STRATEGY_OWNER,
          op2.STRATEGY_OWNER_GATHER]


@pytest.mark.parametrize("strategy", STRATS)
def test_batch_fp64_matches_each_mesh_alone(runtime, strategy):
    note(f"batch fp64 5 meshes random ordering, 6 iterations, strategy={strategy}")
    configure(strategy, 128)
    batch = M.mesh_batch(nmesh=5, seed=3, cells_lo=150, cells_hi=2500, ordering="random")
    assert len({tuple(c[6:10]) for c in batch.consts}) == 5  # free streams differ
    hist, q = AF.run_airfoil_batch(batch, 6, "double")
    assert hist.shape == (6, 5)
    check_meshes(batch, hist, q, 6, "f64", TOL64)


def test_batch_graph_replay_64_meshes(runtime):
    """64 meshes whose runs straddle the update's workgroups; the iteration
    is captured once and replayed."""
    note("batch fp64 64 meshes 1k-20k cells, graph replay 25 iterations")
    batch = M.mesh_batch(nmesh=64, seed=5, cells_lo=1000, cells_hi=20000)
    hist, q = AF.run_airfoil_batch(batch, 25, "double", graph=True)
    assert hist.shape == (25, 64)
    check_meshes(batch, hist, q, 25, "f64", TOL64)


def test_batch_fp32_form(runtime):
    note("batch fp32 form 4 meshes, 50 iterations")
    batch = M.mesh_batch(nmesh=4, seed=7, cells_lo=500, cells_hi=4000)
    hist, q = AF.run_airfoil_batch(batch, 50, "float", graph=True)
    # increments in the sequential order (meshes as given): only the double
    # rms sums differ (measured 4.7e-14, q bit for bit)
    eh, eq = check_meshes(batch, hist, q, 50, "mix", 1e-12)
    note(f"batch fp32 form vs oracle fp32 form: rms rel {eh:.2e}, q rel {eq:.2e}")


def test_batch_one_mesh_equals_single_program(runtime):
    """A batch of one mesh is the airfoil program with qinf as a dat."""
    note("batch of one mesh vs the single-mesh loops")
    mesh = M.o_mesh(40, 20, "random", seed=2)
    c = M.flow_constants(mach=0.5, alpha_deg=1.5).as_array()
    hist, q = AF.run_airfoil_batch(M.concat_meshes([mesh], [c]), 8, "double")
    op2.op_exit()
    op2.op_init([], 0)
    h1, q1, _, _ = AF.run_airfoil(mesh, c, 8, "double")
    assert relmax(hist[:, 0], h1) < TOL64
    assert relmax(q, q1) < TOL64


def test_batch_update_needs_sorted_mesh_map(runtime):
    note("batch with a decreasing cell -> mesh map is refused")
    batch = M.mesh_batch(nmesh=3, seed=3, cells_lo=150, cells_hi=400)
    rev = M.concat_meshes(batch.meshes[::-1], batch.consts[::-1])
    rev.cmesh = np.ascontiguousarray((rev.nmesh - 1) - rev.cmesh)  # decreasing mesh ids
    a = AF.AirfoilBatch(rev, "double")
    with pytest.raises(op2.OP2Error, match="non-decreasing"):
        a.iterate()
