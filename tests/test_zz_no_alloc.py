"""The allocation-free entry point op2_par_loop_noalloc against
op2_par_loop: the same loops on the same state give the same bits.  Two
calls in a row on one device, the second on a smaller set, so the second
finds the first one's partials in the shared device array."""

import numpy as np
import pytest

from oplus2_matlab_amd import airfoil as AF
from oplus2_matlab_amd import mesh as M
from oplus2_matlab_amd import op2
from oplus2_matlab_amd.op2 import (OP_ID, OP_INC, OP_MAX, OP_READ, OP_RW, OP_WRITE, op_arg_dat,
                                   op_arg_gbl)

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[gpu] {msg}", flush=True)


def perturbed(mesh, c, dtype, seed):
    real = np.float64 if dtype == "double" else np.float32
    rng = np.random.Generator(np.random.PCG64(seed))
    q0 = np.tile(c[6:10], (mesh.ncell, 1)) * (1 + 0.02 * rng.standard_normal((mesh.ncell, 4)))
    res0 = 1e-3 * rng.standard_normal((mesh.ncell, 4))
    return q0.astype(real), res0.astype(real)


def airfoil_step(a, noalloc):
    """save_soln, adt_calc and update through the chosen entry point;
    returns update's rms."""
    t = a.t
    save = (op_arg_dat(a.p_q, -1, OP_ID, 4, t, OP_READ),
            op_arg_dat(a.p_qold, -1, OP_ID, 4, t, OP_WRITE))
    adt = tuple(op_arg_dat(a.p_x, i, a.pcell, 2, t, OP_READ) for i in range(4)) + (
        op_arg_dat(a.p_q, -1, OP_ID, 4, t, OP_READ),
        op_arg_dat(a.p_adt, -1, OP_ID, 1, t, OP_WRITE))
    rms = np.zeros(1)
    upd = (op_arg_dat(a.p_qold, -1, OP_ID, 4, t, OP_READ),
           op_arg_dat(a.p_q, -1, OP_ID, 4, t, OP_WRITE),
           op_arg_dat(a.p_res, -1, OP_ID, 4, t, OP_RW),
           op_arg_dat(a.p_adt, -1, OP_ID, 1, t, OP_READ),
           op_arg_gbl(rms, 1, "double", OP_INC))
    run = op2.par_loop_noalloc if noalloc else (lambda n, s, *g: op2.op_par_loop(n, n, s, *g))
    run("save_soln", a.cells, *save)
    run("adt_calc", a.cells, *adt)
    run("update", a.cells, *upd)
    return float(rms[0])


@pytest.mark.parametrize("dtype", ["double", "float"])
def test_noalloc_airfoil_loops_equal_par_loop(dtype):
    meshes = [M.o_mesh(240, 120, "random", seed=1), M.o_mesh(60, 30, "random", seed=2)]
    c = M.flow_constants().as_array()
    out = {}
    for noalloc in (True, False):
        op2.op_init([], 0)
        try:
            res = []
            for i, mesh in enumerate(meshes):
                note(f"airfoil {dtype} {mesh.ncell} cells, noalloc={noalloc}")
                q0, res0 = perturbed(mesh, c, dtype, 10 + i)
                a = AF.Airfoil(mesh, c, dtype, q0=q0, res0=res0)
                a.adt_calc()  # through op2_par_loop: puts pcell on the device
                rms = airfoil_step(a, noalloc)
                q, qold, adt, r = a.fetch()
                res.append((q.copy(), qold.copy(), adt.copy(), r.copy(), rms))
            out[noalloc] = res
        finally:
            op2.op_exit()
    for got, want in zip(out[True], out[False]):
        for x, y in zip(got[:4], want[:4]):
            assert np.array_equal(x, y)
        assert got[4] == want[4] and got[4] > 0
        assert not got[3].any()


def test_noalloc_jacobi_update_equals_par_loop():
    probs = [M.jacobi_grid(256, seed=3), M.jacobi_grid(32, seed=4)]
    out = {}
    for noalloc in (True, False):
        op2.op_init([], 0)
        try:
            res = []
            for prob in probs:
                note(f"jacobi update {prob.nnode} nodes, noalloc={noalloc}")
                j = AF.Jacobi(prob, 0.25, "double")
                rng = np.random.Generator(np.random.PCG64(prob.nnode))
                du = rng.standard_normal(prob.nnode)
                p_du = op2.op_decl_dat(j.nodes, 1, "double", du, "du_rand")
                s = np.zeros(1)
                m = np.full(1, -1e300)
                args = (op_arg_dat(j.p_r, -1, OP_ID, 1, "double", OP_READ),
                        op_arg_dat(p_du, -1, OP_ID, 1, "double", OP_RW),
                        op_arg_dat(j.p_u, -1, OP_ID, 1, "double", OP_INC),
                        op_arg_gbl(s, 1, "double", OP_INC),
                        op_arg_gbl(m, 1, "double", OP_MAX))
                if noalloc:
                    op2.par_loop_noalloc("update", j.nodes, *args)
                else:
                    op2.op_par_loop(op2.update, "update", j.nodes, *args)
                u = op2.op_fetch_data(j.p_u).copy()
                res.append((u, op2.op_fetch_data(p_du).copy(), s[0], m[0]))
            out[noalloc] = res
        finally:
            op2.op_exit()
    for got, want in zip(out[True], out[False]):
        assert np.array_equal(got[0], want[0]) and not got[1].any()
        assert got[2] == want[2] and got[3] == want[3]


def test_noalloc_refuses_loops_that_allocate():
    op2.op_init([], 0)
    try:
        mesh = M.o_mesh(24, 12, "random", seed=1)
        a = AF.Airfoil(mesh, M.flow_constants().as_array(), "double")
        with pytest.raises(op2.OP2Error, match="not on the device yet"):
            op2.par_loop_noalloc(
                "adt_calc", a.cells,
                *[op_arg_dat(a.p_x, i, a.pcell, 2, "double", OP_READ) for i in range(4)],
                op_arg_dat(a.p_q, -1, OP_ID, 4, "double", OP_READ),
                op_arg_dat(a.p_adt, -1, OP_ID, 1, "double", OP_WRITE))
        a.adt_calc()
        a.res_calc()
        with pytest.raises(op2.OP2Error, match="needs plans or buffers"):
            op2.par_loop_noalloc("res_calc", a.edges, *a.res_calc_args())
    finally:
        op2.op_exit()
