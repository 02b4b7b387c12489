"""GPU tests of checkpoint and fast-forward restart (doc/dev.tex:1080-1087):
a restarted run skips every loop before the checkpoint, hands back the
recorded reduction values, restores the dats at the checkpoint, and then
continues exactly as the recorded run did."""

import numpy as np
import pytest

from oplus2_matlab_amd import airfoil as AF
from oplus2_matlab_amd import mesh as M
from oplus2_matlab_amd import op2

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture
def runtime():
    op2.op_init([], 0)
    yield
    op2.op_exit()
    op2.set_loop_config(None, 0, 0, 0)


def airfoil_run(mesh, c, niter, at):
    """The program: niter iterations, op_checkpoint after each iteration in
    `at`; returns the rms history, final q and which checkpoints restored."""
    op2.set_loop_config("res_calc", op2.STRATEGY_OWNER_GATHER, 128, 0)
    op2.set_loop_config("bres_calc", op2.STRATEGY_OWNER, 128, 0)
    a = AF.Airfoil(mesh, c, "double")
    hist, restored = [], []
    for it in range(1, niter + 1):
        hist.append(a.iterate())
        if it in at:
            restored.append(op2.op_checkpoint())
    q, _, _, _ = a.fetch()
    return np.array(hist), q.copy(), restored


def test_airfoil_restart_is_exact(tmp_path):
    note("airfoil 48x24 random: record 12 iterations with checkpoints at 4 and 8, restart")
    path = str(tmp_path / "airfoil.ck")
    mesh = M.o_mesh(48, 24, "random", seed=2)
    c = M.flow_constants().as_array()
    op2.op_init([], 0)
    op2.checkpoint_record(path)
    h1, q1, r1 = airfoil_run(mesh, c, 12, (4, 8))
    op2.op_exit()
    assert r1 == [False, False]
    op2.op_init([], 0)
    try:
        op2.checkpoint_restart(path)
        h2, q2, r2 = airfoil_run(mesh, c, 12, (4, 8))
        # res_calc ran only after the checkpoint: 4 iterations x 2 stages
        assert op2.loop_stats("res_calc")[0] == 8
    finally:
        op2.op_exit()
    assert r2 == [False, True]
    assert np.array_equal(h2, h1)
    assert np.array_equal(q2, q1)


def test_jacobi_restart_with_max_reduction(tmp_path, golden_dir):
    note("jacobi 64x64: record 30 iterations, checkpoint at 20, restart")
    path = str(tmp_path / "jac.ck")
    prob = M.jacobi_grid(64, seed=1)

    def program():
        op2.set_loop_config("res", op2.STRATEGY_OWNER, 256, 0)
        j = AF.Jacobi(prob, 0.25, "double")
        sums, maxs = [], []
        for it in range(1, 31):
            s, m = j.iterate(-0.25)
            sums.append(s)
            maxs.append(m)
            if it == 20:
                op2.op_checkpoint()
        op2.op_fetch_data(j.p_u)
        return np.array(sums), np.array(maxs), j.u.copy()

    op2.op_init([], 0)
    op2.checkpoint_record(path)
    s1, m1, u1 = program()
    op2.op_exit()
    op2.op_init([], 0)
    try:
        op2.checkpoint_restart(path)
        s2, m2, u2 = program()
    finally:
        op2.op_exit()
        op2.set_loop_config(None, 0, 0, 0)
    assert np.array_equal(s1, s2) and np.array_equal(m1, m2) and np.array_equal(u1, u2)


def test_capture_before_checkpoint_is_refused(tmp_path):
    """Loops skipped by a restart launch nothing, so capturing them into a
    graph would record nothing: the runtime refuses."""
    note("restart: graph capture before the checkpoint is refused")
    path = str(tmp_path / "cap.ck")
    mesh = M.o_mesh(24, 12, "random", seed=3)
    c = M.flow_constants().as_array()
    op2.op_init([], 0)
    op2.checkpoint_record(path)
    airfoil_run(mesh, c, 4, (3,))
    op2.op_exit()
    op2.op_init([], 0)
    try:
        op2.checkpoint_restart(path)
        a = AF.Airfoil(mesh, c, "double")
        a.iterate()  # skipped: before the checkpoint
        op2.set_deferred_reductions(True)
        with pytest.raises(op2.OP2Error, match="before the checkpoint"):
            a.capture_iteration()
    finally:
        op2.set_deferred_reductions(False)
        op2.op_exit()
        op2.set_loop_config(None, 0, 0, 0)
